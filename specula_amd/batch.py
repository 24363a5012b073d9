"""Uneven multi-sensor batches: one persistent HIP launch for the SH forward of every sensor
frame, the per-sensor CoG launches after it, all capturable in one HIP graph.

A sensor entry owns its plan, static pupil, phase frames, CCD frames and (optionally) a
ShSlopec plan with slope outputs.  Buffers are allocated once; ``forward()`` only launches.
Sensors with field resampling take their phase on the input grid (the batch resamples it in the
same launch sequence); sensors whose image is normalised by its own sum (FoV cut or clipping
mask) are rescaled after the SH launches.
"""
import ctypes
from dataclasses import dataclass, field

import numpy as np
import torch

from specula_amd import _native
from specula_amd.lib.subaps import detect_subaps
from specula_amd.processing_objects.sh import SHPlan
from specula_amd.processing_objects.sh_slopec import SlopecPlan, compute_xy_weights


@dataclass
class SensorEntry:
    name: str
    plan: SHPlan
    pupil: torch.Tensor            # (ef, ef) float32, static
    phase: torch.Tensor            # (F, ef, ef) float32, filled by the caller
    ccd: torch.Tensor              # (F, ccd, ccd) float32
    S0: float = 1.0
    slopec: SlopecPlan = None
    slopes: torch.Tensor = None    # (F, 2 n_sub)
    flux: torch.Tensor = None
    totals: torch.Tensor = None
    meta: dict = field(default_factory=dict)

    @property
    def n_frames(self):
        return self.phase.shape[0]


def make_entry(name, plan, pupil, n_frames, S0=1.0, with_slopes=True, energy_th=0.5, device=0):
    dev = torch.device(f'cuda:{device}')
    g = plan.geom
    A = torch.as_tensor(np.asarray(pupil, np.float32), device=dev).contiguous()
    if g.needs_interpolation and not getattr(plan, 'resampled', False):
        plan.set_resampling(np.asarray(pupil))      # edge-extrapolation maps from the pupil (sh.py:452-458)
    ph = torch.zeros((n_frames, g.ef_size, g.ef_size), dtype=torch.float32, device=dev)
    ccd = torch.empty((n_frames, g.ccd_side, g.ccd_side), dtype=torch.float32, device=dev)
    e = SensorEntry(name, plan, A, ph, ccd, S0)
    if with_slopes:
        img = torch.empty((1, g.ccd_side, g.ccd_side), dtype=torch.float32, device=dev)
        plan.forward(A, ph[:1], img, 1.0, n_frames=1)
        idxs, _ = detect_subaps(img[0].cpu().numpy(), g.n_lenses, energy_th)
        e.slopec = SlopecPlan(idxs, (g.ccd_side, g.ccd_side), compute_xy_weights(g.subap_npx),
                              max_frames=n_frames, device=device)
        n_sub = idxs.shape[0]
        e.slopes = torch.empty((n_frames, 2 * n_sub), dtype=torch.float32, device=dev)
        e.flux = torch.empty((n_frames, n_sub), dtype=torch.float64, device=dev)
        e.totals = torch.empty((n_frames, 2), dtype=torch.float64, device=dev)
    return e


class MultiSensorBatch:
    """shwfs_batch over a list of SensorEntry (see include/shwfs.h)."""

    def __init__(self, entries, grid_blocks=0, order=0, device=0, fuse_slopes=True):
        self.entries = list(entries)
        self.device = device
        n = len(self.entries)
        P = ctypes.c_void_p * n
        plans = P(*[e.plan._h.value for e in self.entries])
        frames = (ctypes.c_int32 * n)(*[e.n_frames for e in self.entries])
        A = P(*[e.pupil.data_ptr() for e in self.entries])
        ph = P(*[e.phase.data_ptr() for e in self.entries])
        ccd = P(*[e.ccd.data_ptr() for e in self.entries])
        S0 = (ctypes.c_double * n)(*[float(e.S0) for e in self.entries])
        h = ctypes.c_void_p()
        _native.check(_native.lib().shwfs_batch_create(n, plans, frames, A, ph, ccd, S0, int(grid_blocks),
                                                       int(order), device, ctypes.byref(h)),
                      'shwfs_batch_create')
        self._h = h
        ni, gr = ctypes.c_int32(), ctypes.c_int32()
        _native.check(_native.lib().shwfs_batch_info(h, ctypes.byref(ni), ctypes.byref(gr)), 'shwfs_batch_info')
        self.n_items, self.grid = ni.value, gr.value
        self.graph = None
        # sensors whose ShSlopec can run in the SH launch's epilogue (Parseval normalisation; NGS MFMA
        # np 8 with 8 or 16 px per subaperture, or the LGS np 16 body) get it there (the same slopes
        # as SHPlan.forward_slopes); the others keep a CoG launch of their own after the SH launches
        self.fused = [False] * n
        if fuse_slopes:
            for k, e in enumerate(self.entries):
                if e.slopec is None or not e.plan.bind_slopec(e.slopec):
                    continue
                rc = _native.lib().shwfs_batch_set_slopes(h, k, 0.0, 0, _native.ptr(e.slopes), e.slopes.shape[1],
                                                          _native.ptr(e.flux), _native.ptr(e.totals))
                if rc == _native.SHWFS_ERR_UNSUPPORTED:
                    continue
                _native.check(rc, 'shwfs_batch_set_slopes')
                self.fused[k] = True

    def forward(self, stream=None, slopes=True):
        """The SH launches (with the fused sensors' CoG and their slopes), then, with ``slopes``,
        the CoG launches of the other sensors."""
        st = _native.stream_handle(stream)
        _native.check(_native.lib().shwfs_batch_forward(self._h, st), 'shwfs_batch_forward')
        if slopes:
            for e, fused in zip(self.entries, self.fused):
                if e.slopec is not None and not fused:
                    e.slopec.forward(e.ccd, e.slopes, e.flux, e.totals, n_frames=e.n_frames, stream=stream)

    def latencies_us(self, stream=None):
        """Per-sensor completion time of the SH forward of the last launch [us]."""
        out = (ctypes.c_double * len(self.entries))()
        _native.check(_native.lib().shwfs_batch_latencies(self._h, out, _native.stream_handle(stream)),
                      'shwfs_batch_latencies')
        return np.array(out[:])

    def capture(self, slopes=True):
        """Capture forward() (persistent SH launch + CoG launches) in a HIP graph of the native
        library (shwfs_graph_*)."""
        s = torch.cuda.Stream(device=self.device)
        with torch.cuda.device(self.device):
            with torch.cuda.stream(s):
                self.forward(stream=s, slopes=slopes)
            s.synchronize()
            self.graph = _native.HipGraph().capture(lambda: self.forward(stream=s, slopes=slopes), s)
        return self.graph

    def replay(self, stream=None):
        self.graph.replay(stream)

    def __del__(self):
        h = getattr(self, '_h', None)
        if h is not None and h.value:
            try:
                _native.lib().shwfs_batch_destroy(h)
            except Exception:
                pass
            self._h = None
