"""Drop-in CCD detector, specula/processing_objects/ccd.py:21-305, on the HIP readout kernels
(``shccd_integrate`` / ``shccd_readout``).

Same keywords, ports (``in_i`` -> ``out_pixels``, ``integrated_i``) and timing: intensities are
integrated every step with factor ``time_step * bandw`` and read out every ``dt``.  When a readout
step has integrated nothing before (dt equal to the loop step, the usual case), the integration
and the readout are one kernel pass.

``cte_noise`` applies the charge-transfer matrices ``cte_mat[:, :, 0] . frame . cte_mat[:, :, 1]``
between the background/dark offset and the CIC / photon noise (``shccd_readout_cte``: two FP32
matrix-core products between two passes of the readout kernel).  'auto' read-out noise and dark
current levels come from the detector model ``name`` at the frame rate 1/dt
(``specula_amd.lib.detector_noise``, calc_detector_noise.py) under the reference's condition: all
three levels given (ccd.py:67-74); an 'auto' background level, which the reference leaves
unresolved, raises.  Not built: the non-uniform QE matrix (no public setter in the reference).
Noise streams are counter-based (see include/shwfs.h): statistically equivalent to the
reference's numpy generators, not the same numbers.
"""
import ctypes

import numpy as np
import torch

from specula_amd import _native
from specula_amd.base import BaseProcessingObj, InputValue
from specula_amd.data_objects.intensity import Intensity
from specula_amd.data_objects.pixels import Pixels
from specula_amd.lib.detector_noise import detector_noise


def resolve_noise_levels(readout_level, darkcurrent_level, background_level, name, dt, binning,
                         readout_used, offset_used):
    """Numeric (readout, dark current, background) levels as ccd.py:67-74 resolves them: when all
    three are given and one is 'auto', the read-out noise and dark current of detector ``name`` at
    frame rate 1/dt replace the 'auto' ones (the background is never resolved).  A level still a
    string where apply_noise would use it raises ValueError; unused or empty levels become 0."""
    if readout_level and darkcurrent_level and background_level and \
            'auto' in (readout_level, darkcurrent_level, background_level):
        ron, dark, _ = detector_noise(1.0 / dt, name, binning)
        if readout_level == 'auto':
            readout_level = ron
        if darkcurrent_level == 'auto':
            darkcurrent_level = dark
    for key, level, used in (('readout_level', readout_level, readout_used),
                             ('darkcurrent_level', darkcurrent_level, offset_used),
                             ('background_level', background_level, offset_used)):
        if isinstance(level, str) and level and used:
            raise ValueError(f"{key}={level!r}: 'auto' levels are resolved from the detector model only when "
                             "readout, dark-current and background levels are all given, and the background "
                             "level never (ccd.py:67-74)")
    return tuple(0.0 if isinstance(v, str) or v is None else float(v)
                 for v in (readout_level, darkcurrent_level, background_level))


class CCD(BaseProcessingObj):
    def __init__(self, simul_params, size, dt: float, bandw: float, name: str = '', binning: int = 1,
                 photon_noise: bool = False, readout_noise: bool = False, excess_noise: bool = False,
                 darkcurrent_noise: bool = False, background_noise: bool = False, cic_noise: bool = False,
                 cte_noise: bool = False, readout_level=0.0, darkcurrent_level=0.0, background_level=0.0,
                 cic_level: float = 0, cte_mat=None, quantum_eff: float = 1.0, pixelGains=None,
                 photon_seed: int = 1, readout_seed: int = 2, excess_seed: int = 3, excess_delta: float = 1.0,
                 start_time: int = 0, ADU_gain: float = 8, ADU_bias: int = 400, emccd_gain: int = 1,
                 target_device_idx: int = None, precision: int = None):
        super().__init__(target_device_idx=target_device_idx, precision=precision)
        if self.target_device_idx < 0:
            raise ValueError('CCD runs on the GPU only (target_device_idx >= 0)')
        if dt <= 0:
            raise ValueError(f'dt (integration time) is {dt} and must be greater than zero')
        self.loop_dt = self.seconds_to_t(simul_params.time_step)
        self._dt = self.seconds_to_t(dt)
        if self._dt % self.loop_dt != 0:
            raise ValueError(f'integration time dt={dt} must be a multiple of the basic simulation '
                             f'time_step={simul_params.time_step}')
        readout_level, darkcurrent_level, background_level = resolve_noise_levels(
            readout_level, darkcurrent_level, background_level, name, dt, binning, readout_noise,
            background_noise or darkcurrent_noise)
        if excess_noise:                       # ccd.py:104-106
            emccd_gain = 400
            ADU_gain = 1 / 20
        self._start_time = self.seconds_to_t(start_time)
        self._photon_noise, self._readout_noise, self._excess_noise = photon_noise, readout_noise, excess_noise
        self._darkcurrent_noise, self._background_noise, self._cic_noise = darkcurrent_noise, background_noise, cic_noise
        self._binning = int(binning)
        self._readout_level = float(readout_level or 0.0)
        self._darkcurrent_level = float(darkcurrent_level or 0.0)
        self._background_level = float(background_level or 0.0)
        self._cic_level = float(cic_level)
        self._qe = float(quantum_eff)
        self._cte_noise = bool(cte_noise)
        self._pixels = Pixels(size[0] // binning, size[1] // binning, target_device_idx=target_device_idx)
        self._cte_rows = self._cte_cols = self._cte_work = None
        if self._cte_noise:
            r, c = self._pixels.size
            # the reference's default cte_mat is all zeros (ccd.py:130)
            m = np.zeros((size[0], size[1], 2), np.float32) if cte_mat is None else np.asarray(cte_mat, np.float32)
            if m.ndim != 3 or m.shape[2] != 2 or m.shape[0] != r or m.shape[1] != r or r != c:
                raise ValueError(f'cte_mat of shape {m.shape} does not multiply the {r} x {c} pixel frame from '
                                 f'both sides: it must be ({r}, {r}, 2) for a square frame')
            self._cte_rows = torch.tensor(np.ascontiguousarray(m[:, :, 0]), device=self.device)
            self._cte_cols = torch.tensor(np.ascontiguousarray(m[:, :, 1]), device=self.device)
            self._cte_work = torch.empty(2 * r * c, dtype=torch.float32, device=self.device)
        s = (self._pixels.size[0] * binning, self._pixels.size[1] * binning)
        self._integrated_i = Intensity(s[0], s[1], target_device_idx=target_device_idx, precision=precision)
        self._photon_seed, self._readout_seed, self._excess_seed = photon_seed, readout_seed, excess_seed
        self._excess_delta = float(excess_delta)
        self._keep_ADU_bias = False
        self._bg_remove_average = False
        self._do_not_remove_dark = False
        self._doNotChangeI = False
        self._ADU_gain, self._ADU_bias, self._emccd_gain = float(ADU_gain), float(ADU_bias), float(emccd_gain)
        self._bandw = bandw
        self._pixelGains = None if pixelGains is None else torch.as_tensor(
            np.asarray(pixelGains), dtype=torch.float32, device=self.device).contiguous()
        self._readouts = 0                     # counter of the noise streams
        self._pending = False                  # integration buffer holds frames not yet read out
        self.inputs['in_i'] = InputValue(type=Intensity)
        self.outputs['out_pixels'] = self._pixels
        self.outputs['integrated_i'] = self._integrated_i

    @property
    def dt(self):
        return self._dt

    @property
    def size(self):
        return self._pixels.size

    @property
    def bandw(self):
        return self._bandw

    @property
    def binning(self):
        return self._binning

    def params(self):
        p = _native.ShccdParams()
        p.in_rows, p.in_cols = self._integrated_i.i.shape
        p.out_rows, p.out_cols = self._pixels.pixels.shape
        p.binning = self._binning
        f = 0
        for flag, on in ((_native.SHCCD_PHOTON, self._photon_noise), (_native.SHCCD_READOUT, self._readout_noise),
                         (_native.SHCCD_EXCESS, self._excess_noise), (_native.SHCCD_DARK, self._darkcurrent_noise),
                         (_native.SHCCD_BACKGROUND, self._background_noise), (_native.SHCCD_CIC, self._cic_noise),
                         (_native.SHCCD_KEEP_BIAS, self._keep_ADU_bias),
                         (_native.SHCCD_BG_REMOVE_AVG, self._bg_remove_average),
                         (_native.SHCCD_KEEP_DARK, self._do_not_remove_dark)):
            if on:
                f |= flag
        p.flags = f
        p.qe = self._qe
        p.background_level, p.darkcurrent_level = self._background_level, self._darkcurrent_level
        p.readout_level, p.cic_level = self._readout_level, self._cic_level
        p.adu_gain, p.adu_bias, p.emccd_gain, p.excess_delta = (self._ADU_gain, self._ADU_bias, self._emccd_gain,
                                                                self._excess_delta)
        p.photon_seed, p.readout_seed, p.excess_seed = self._photon_seed, self._readout_seed, self._excess_seed
        p.cic_seed = self._readout_seed + 0x5851F42D
        return p

    def setup(self):
        super().setup()
        in_i = self.inputs['in_i'].get(self.target_device_idx)
        if in_i is None:
            raise ValueError('Input intensity object has not been set')
        if tuple(in_i.i.shape) != tuple(self._integrated_i.i.shape):
            raise ValueError(f'input intensity {tuple(in_i.i.shape)} does not match the detector '
                             f'{tuple(self._integrated_i.i.shape)} (size x binning)')
        if self._doNotChangeI:
            raise NotImplementedError('_doNotChangeI is not built')
        self._p = self.params()
        self.build_stream(allow_parallel=False, capture=False)

    def _factor(self):
        return self.t_to_seconds(self.loop_dt) * self._bandw

    def trigger_code(self):
        """ccd.py:198-218."""
        if not (self._start_time <= 0 or self.current_time >= self._start_time):
            return
        in_i = self.local_inputs['in_i']
        fresh = in_i.generation_time == self.current_time
        readout = (self.current_time + self.loop_dt - self._dt - self._start_time) % self._dt == 0
        lib = _native.lib()
        st = _native.stream_handle(None)
        integ = self._integrated_i.i
        src = in_i.i if in_i.i.dtype == torch.float32 and in_i.i.is_contiguous() else in_i.i.float().contiguous()
        if readout:
            # the integration buffer is all zeros unless frames are pending (it is zeroed by every
            # readout), so a readout with no pending frame and no fresh input reads a dark frame
            use_integ = self._pending or not fresh
            gains = _native.ptr(self._pixelGains) if self._pixelGains is not None else None
            if self._cte_noise:
                _native.check(lib.shccd_readout_cte(
                    ctypes.byref(self._p), 1, _native.ptr(integ) if use_integ else None, integ.numel(),
                    _native.ptr(src) if fresh else None, src.numel(), self._factor(), _native.ptr(self._cte_rows),
                    _native.ptr(self._cte_cols), _native.ptr(self._cte_work), gains, ctypes.c_uint64(self._readouts),
                    _native.ptr(self._pixels.pixels), self._pixels.pixels.numel(), st), 'shccd_readout_cte')
            else:
                _native.check(lib.shccd_readout(
                    ctypes.byref(self._p), 1, _native.ptr(integ) if use_integ else None, integ.numel(),
                    _native.ptr(src) if fresh else None, src.numel(), self._factor(), gains,
                    ctypes.c_uint64(self._readouts), _native.ptr(self._pixels.pixels), self._pixels.pixels.numel(),
                    st), 'shccd_readout')
            self._pending = False
            self._readouts += 1
            self._pixels.generation_time = self.current_time
        elif fresh:
            _native.check(lib.shccd_integrate(1, _native.ptr(src), src.numel(), self._factor(), _native.ptr(integ),
                                              integ.numel(), integ.numel(), st), 'shccd_integrate')
            self._pending = True


def readout_frames(params, frames, out, factor=1.0, counter=0, pixel_gains=None, stream=None, cte=None):
    """Batched readout (shccd_readout) of F intensity frames (F, in_rows, in_cols) float32 into
    ``out`` (F, out_rows, out_cols): each frame integrated once with ``factor`` and read out with
    the noise of ``params`` (an ``_native.ShccdParams``, e.g. ``CCD.params()``).  ``cte``: a pair
    (rows, cols) of device matrices applies the charge-transfer matrices (shccd_readout_cte)."""
    for t in (frames, out):
        if t.dtype != torch.float32 or not t.is_contiguous() or not t.is_cuda:
            raise ValueError('CCD frames must be contiguous float32 device tensors')
    F = frames.shape[0]
    n_in = params.in_rows * params.in_cols
    n_out = params.out_rows * params.out_cols
    if frames.numel() != F * n_in or out.numel() != F * n_out:
        raise ValueError('frame sizes do not match the detector parameters')
    gains = _native.ptr(pixel_gains) if pixel_gains is not None else None
    if cte is not None:
        rows, cols = cte
        for t, n in ((rows, params.out_rows), (cols, params.out_cols)):
            if t.dtype != torch.float32 or not t.is_contiguous() or not t.is_cuda or tuple(t.shape) != (n, n):
                raise ValueError(f'CTE matrices must be contiguous float32 ({n}, {n}) device tensors')
        work = torch.empty(2 * F * n_out, dtype=torch.float32, device=out.device)
        _native.check(_native.lib().shccd_readout_cte(
            ctypes.byref(params), F, None, 0, _native.ptr(frames), n_in, float(factor), _native.ptr(rows),
            _native.ptr(cols), _native.ptr(work), gains, ctypes.c_uint64(counter), _native.ptr(out), n_out,
            _native.stream_handle(stream)), 'shccd_readout_cte')
        return out
    _native.check(_native.lib().shccd_readout(ctypes.byref(params), F, None, 0, _native.ptr(frames), n_in,
                                              float(factor), gains, ctypes.c_uint64(counter), _native.ptr(out),
                                              n_out, _native.stream_handle(stream)), 'shccd_readout')
    return out
