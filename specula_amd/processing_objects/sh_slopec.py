"""Drop-in ShSlopec (thresholded centre of gravity) running on gfx950 HIP kernels.

Constructor keywords, ports and outputs of specula/processing_objects/sh_slopec.py:25-70;
``trigger_code`` replaces calc_slopes_nofor (sh_slopec.py:308-438) by one ``shslopec_forward``
call (gather, FP64 flux/CoG moments, mean-based dim-subaperture cut-off, slopes).
"""
import ctypes

import numpy as np
import torch

from specula_amd import _native
from specula_amd.base import BaseValue
from specula_amd.data_objects.slopes import Slopes
from specula_amd.data_objects.subap_data import SubapData
from specula_amd.lib.grids import make_mask
from specula_amd.lib.subaps import block_origins
from specula_amd.processing_objects.slopec import Slopec


def compute_xy_weights(np_sub, exp_weight=1.0, weightedPixRad=0.0, quadcell_mode=False, windowing=False):
    """ShSlopec.computeXYweights (sh_slopec.py:86-132) in FP64."""
    x, y = np.meshgrid(np.linspace(-1, 1, np_sub), np.linspace(-1, 1, np_sub))
    if quadcell_mode:
        x = np.where(x > 0, 1.0, -1.0)
        y = np.where(y > 0, 1.0, -1.0)
        xc, yc = x, y
    else:
        xc, yc = x, y
        x = np.where(x > 0, np.power(np.abs(x), exp_weight), -np.power(np.abs(x), exp_weight))
        y = np.where(y > 0, np.power(np.abs(y), exp_weight), -np.power(np.abs(y), exp_weight))
    xc = np.where(x > 0, xc, -np.abs(xc))
    yc = np.where(y > 0, yc, -np.abs(yc))
    if weightedPixRad != 0:
        if not windowing:
            # the reference's weighted-CoG branch calls psf_gaussian with a wrong arity
            # (sh_slopec.py:122 vs 440) and raises; so does this one.
            raise TypeError('weighted CoG (weightedPixRad without windowing) is not callable in the reference')
        mask_weighted = make_mask(np_sub, diaratio=(2.0 * weightedPixRad / np_sub)).astype(np.float64)
        mask_weighted[mask_weighted < 1e-6] = 0.0
        x = x * mask_weighted
        y = y * mask_weighted
    else:
        mask_weighted = np.ones((np_sub, np_sub))
    return {'x': x, 'y': y, 'xc': xc, 'yc': yc, 'mask_weighted': mask_weighted}


class SlopecPlan:
    """Owns a ``shslopec_plan``: subaperture addressing (blocks or index list) and weights."""

    def __init__(self, idxs, frame_shape, weights, max_frames=1, device=0):
        idxs = np.asarray(idxs, dtype=np.int64)
        self.idxs = idxs                      # host copy of the subaperture pixel indices
        self.n_sub, npp = idxs.shape
        self.np_sub = int(round(np.sqrt(npp)))
        self.frame_shape = tuple(frame_shape)
        self.max_frames = max_frames
        H, W = self.frame_shape
        if idxs.min() < 0 or idxs.max() >= H * W:
            raise ValueError('subaperture pixel index outside the pixel frame')
        origins = block_origins(idxs, W)
        self.block_mode = origins is not None
        xw = np.ascontiguousarray(weights['x'].ravel(), dtype=np.float64)
        yw = np.ascontiguousarray(weights['y'].ravel(), dtype=np.float64)
        mw = np.ascontiguousarray(weights['mask_weighted'].ravel(), dtype=np.float64)
        vp = lambda a: a.ctypes.data_as(ctypes.c_void_p)   # noqa: E731
        if self.block_mode:
            org, idx = np.ascontiguousarray(origins), None
        else:
            org, idx = None, np.ascontiguousarray(idxs.astype(np.int32))
        h = ctypes.c_void_p()
        _native.check(_native.lib().shslopec_plan_create(
            self.n_sub, self.np_sub, H, W, vp(org) if org is not None else None,
            vp(idx) if idx is not None else None, vp(xw), vp(yw), vp(mw), max_frames, device,
            ctypes.byref(h)), 'shslopec_plan_create')
        self._h = h

    def forward(self, pixels, slopes, flux, totals, thr=0.0, pedestal=False, n_frames=1, stream=None):
        for t, dt in ((pixels, torch.float32), (slopes, torch.float32), (flux, torch.float64),
                      (totals, torch.float64)):
            if t.dtype != dt or not t.is_contiguous() or not t.is_cuda:
                raise ValueError('ShSlopec buffers: pixels/slopes float32, flux/totals float64, contiguous, on device')
        H, W = self.frame_shape
        if (pixels.numel() != n_frames * H * W or slopes.numel() < n_frames * 2 * self.n_sub
                or flux.numel() < n_frames * self.n_sub or totals.numel() < 2 * n_frames):
            raise ValueError('ShSlopec buffer sizes do not match the plan')
        _native.check(_native.lib().shslopec_forward(
            self._h, n_frames, _native.ptr(pixels), H * W, float(thr), 1 if pedestal else 0,
            _native.ptr(slopes), 2 * self.n_sub, _native.ptr(flux), _native.ptr(totals),
            _native.stream_handle(stream)), 'shslopec_forward')

    def __del__(self):
        h = getattr(self, '_h', None)
        if h is not None and h.value:
            try:
                _native.lib().shslopec_plan_destroy(h)
            except Exception:
                pass
            self._h = None


class ShSlopec(Slopec):
    """specula/processing_objects/sh_slopec.py:25-445 (calc_slopes_nofor path), HIP CoG."""

    def __init__(self, subapdata: SubapData, sn: Slopes = None, thr_value: float = -1,
                 exp_weight: float = 1.0, filtmat=None, corr_template=None,
                 target_device_idx: int = None, precision: int = None):
        super().__init__(sn=sn, filtmat=filtmat, target_device_idx=target_device_idx, precision=precision)
        if self.target_device_idx < 0:
            raise ValueError('ShSlopec runs on the GPU only (target_device_idx >= 0)')
        if self.precision != 1:
            raise ValueError('the HIP ShSlopec path outputs FP32 slopes: use precision=1')
        self.thr_value = thr_value
        self.thr_mask_cube = BaseValue(target_device_idx=self.target_device_idx)
        self.total_counts = BaseValue(target_device_idx=self.target_device_idx)
        self.subap_counts = BaseValue(target_device_idx=self.target_device_idx)
        self.corr_template = corr_template
        self.winMatWindowed = None
        self.vecWeiPixRadT = None
        self.weightedPixRad = 0.0
        self.windowing = False
        self.correlation = False
        self.corrWindowSidePix = 0
        self.thr_ratio_value = 0.0
        self.thr_pedestal = False
        self.mult_factor = 0.0
        self.quadcell_mode = False
        self.two_steps_cog = False
        self.cog_2ndstep_size = 0
        self.dotemplate = False
        self.store_thr_mask_cube = False
        self.exp_weight = exp_weight
        self.subapdata = subapdata
        self.slopes.resize(subapdata.n_subaps * 2)
        self.accumulated_slopes = Slopes(subapdata.n_subaps * 2, target_device_idx=self.target_device_idx,
                                         precision=self.precision)
        self._plan = None
        self._plan_key = None
        self._single_mask = None
        self._flux = torch.zeros(subapdata.n_subaps, dtype=torch.float64, device=self.device)
        self._totals = torch.zeros(2, dtype=torch.float64, device=self.device)
        self.set_xy_weights()
        self.outputs['out_subapdata'] = self.subapdata

    @property
    def subap_idx(self):
        return self.subapdata.idxs

    def set_xy_weights(self):
        if self.subapdata:
            out = compute_xy_weights(self.subapdata.np_sub, self.exp_weight, self.weightedPixRad,
                                     self.quadcell_mode, self.windowing)
            self._weights = out
            self.mask_weighted = out['mask_weighted']
            self.xweights = out['x']
            self.yweights = out['y']
            self.xcweights = out['xc']
            self.ycweights = out['yc']
            self._plan = None

    def _check_supported(self):
        if self.dotemplate or self.correlation or self.two_steps_cog:
            raise NotImplementedError('template / correlation / two-step CoG (calc_slopes_for) are not on the HIP path')
        if self.weight_from_accumulated:
            raise NotImplementedError('weight_from_accumulated is not implemented')   # sh_slopec.py:333-334
        if self.winMatWindowed is not None or self.store_thr_mask_cube or self.vecWeiPixRadT is not None:
            raise NotImplementedError('windowing matrices / threshold-mask cube are not on the HIP path')
        if self.thr_ratio_value > 0:
            raise NotImplementedError('thr_ratio_value is not on the HIP path')

    def _get_plan(self, pixels):
        key = tuple(pixels.shape)
        if self._plan is None or self._plan_key != key:
            self._plan = SlopecPlan(self.subapdata.idxs, key, self._weights, max_frames=1,
                                    device=self.target_device_idx)
            self._plan_key = key
        return self._plan

    def trigger_code(self, accumulated=False):
        self._check_supported()
        pixels = self.local_inputs['in_pixels'].pixels
        if pixels.dtype != torch.float32 or not pixels.is_contiguous() or pixels.device != self.device:
            pixels = pixels.to(self.device, torch.float32).contiguous()
        if self.thr_pedestal or self.thr_value > 0:              # sh_slopec.py:377-380
            thr = float(self.thr_value)
        else:
            thr = 0.0
        plan = self._get_plan(pixels)
        plan.forward(pixels, self.slopes.slopes, self._flux, self._totals, thr=thr,
                     pedestal=self.thr_pedestal, stream=None)
        if self.mult_factor != 0:
            self.slopes.slopes *= self.mult_factor
        if self._single_mask is None or self._single_mask[0] is not self.subapdata:   # built once per SubapData
            self._single_mask = (self.subapdata, self.subapdata.single_mask())
        self.slopes.single_mask = self._single_mask[1]
        self.slopes.display_map = self.subapdata.display_map
        self.slopes.generation_time = self.current_time
        self.flux_per_subaperture_vector.value = self._flux
        self.flux_per_subaperture_vector.generation_time = self.current_time
        self.total_counts.value = self._totals[0]
        self.total_counts.generation_time = self.current_time
        self.subap_counts.value = self._totals[1]
        self.subap_counts.generation_time = self.current_time
