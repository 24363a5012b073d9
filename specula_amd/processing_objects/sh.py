"""Drop-in SH (Shack-Hartmann WFS) processing object running on gfx950 HIP kernels.

Same constructor keywords (= YAML keys), ports and lifecycle as
specula/processing_objects/sh.py:30-110; the per-frame body (sh.py:348-428: field, FFT, |.|^2,
kernel correlation, mask, cut, CCD rebin, flux normalisation) is one ``shwfs_forward`` call.
"""
import ctypes

import numpy as np
import torch

from specula_amd import RAD2ASEC
from specula_amd import _native
from specula_amd.base import BaseProcessingObj, BaseValue, InputValue
from specula_amd.data_objects.convolution_kernel import ConvolutionKernel, GaussianConvolutionKernel
from specula_amd.data_objects.electric_field import ElectricField
from specula_amd.data_objects.intensity import Intensity
from specula_amd.data_objects.lenslet import Lenslet
from specula_amd.data_objects.laser_launch_telescope import LaserLaunchTelescope
from specula_amd.lib.binning import build_operator_terms
from specula_amd.lib.resample import edge_extrapolation_maps, resampling_map
from specula_amd.lib.geometry import sh_geometry


class SHPlan:
    """Owns a ``shwfs_plan`` (device copies of the binning operators and FP64 scratch)."""

    def __init__(self, geom, lgs=False, max_frames=1, device=0, variant=0, kernels_normalised=True):
        """variant: 0 picks the kernel automatically, 1 forces register-GEMV binning, 2 forces
        FP32 MFMA binning (spec_size 16, np_sub 5..8, subap_npx <= 16), 3 MFMA binning from split
        bf16 pieces (x = hi + lo, all four cross terms; relative error ~2^-17 per operand), 4 the
        any-size path (full M-point LDS FFTs, any np_sub up to 512, any subap_npx; NGS and LGS),
        which variant 0 also takes when no specialised kernel fits the geometry.
        lgs: per-subaperture kernel correlation; spectra come from ``set_kernels``.  With
        ``kernels_normalised`` (unit-sum kernels) the Parseval normalisation applies."""
        rows, cols, M, norm = build_operator_terms(geom, lgs=lgs)
        self.mask_terms = rows.shape[0]                 # > 1: round FoV mask clipping the cut window
        if lgs:
            N = geom.fft_size
            rows = rows * (N * N / float(M * M))        # D = (N^2/M^2) FFT_M(S) K
            if not kernels_normalised:
                norm = None
        row, col = rows[0], cols[0]
        self.geom = geom
        self.lgs = lgs
        self.M = M
        self.L = 2 * geom.ovs_np_sub - 1
        self.parseval = norm is not None
        self.max_frames = max_frames
        self.device = device
        self.row_op = row
        self.col_op = col
        g = _native.ShwfsGeom()
        g.n_lenses = geom.n_lenses
        g.ef_size = geom.ef_ovs if geom.needs_interpolation else geom.ef_size   # field the kernels see
        g.np_sub = geom.ovs_np_sub
        g.fft_size = geom.fft_size
        g.spec_size = M
        g.subap_npx = geom.subap_npx
        g.lgs = 1 if lgs else 0
        g.parseval = 1 if self.parseval else 0
        g.max_frames = max_frames
        g.variant = variant
        g.k_rad_per_nm = 2 * np.pi / geom.wavelengthInNm
        g.tilt_rad = np.pi / geom.fft_size
        g.tilt_a0 = float((-geom.ovs_np_sub) // 2)
        g.norm_const = norm if norm is not None else 0.0
        g.pixel_area = geom.pixel_pitch ** 2
        self._geom_c = g
        self.resampled = False
        self.fpb = 4                                    # frames per block (library default)
        self._slopec = None
        if lgs:     # complex operators as interleaved (re, im)
            rw = np.ascontiguousarray(np.stack([rows.real, rows.imag], axis=-1), dtype=np.float64)
            cw = np.ascontiguousarray(np.stack([cols.real, cols.imag], axis=-1), dtype=np.float64)
        else:
            rw = np.ascontiguousarray(rows, dtype=np.float64)
            cw = np.ascontiguousarray(cols, dtype=np.float64)
        h = ctypes.c_void_p()
        lib = _native.lib()
        _native.check(lib.shwfs_plan_create_masked(ctypes.byref(g), self.mask_terms,
                                                   rw.ctypes.data_as(ctypes.c_void_p),
                                                   cw.ctypes.data_as(ctypes.c_void_p), device, ctypes.byref(h)),
                      'shwfs_plan_create_masked')
        self._h = h

    def set_resampling(self, pupil):
        """Field resampling for fov_ovs != 1, rotation or shift (shwfs_plan_set_resampling): edge-
        extrapolation sources from the pupil mask (sh.py:452-458) and Interp2D's sampling map
        (sh.py:450).  Forward calls then take fields on the input grid."""
        g = self.geom
        mask = np.asarray(pupil.cpu().numpy() if hasattr(pupil, 'cpu') else pupil)
        if mask.shape != (g.ef_size, g.ef_size):
            raise ValueError(f'pupil must be {g.ef_size} x {g.ef_size}')
        src1, src2 = edge_extrapolation_maps(mask, two_pixel=True)
        aff = np.ascontiguousarray(resampling_map((g.ef_size, g.ef_size), (g.ef_ovs, g.ef_ovs), g.rot_deg,
                                                  g.xy_shift[0], g.xy_shift[1]), dtype=np.float64)
        s1 = np.ascontiguousarray(src1, dtype=np.int32)
        s2 = np.ascontiguousarray(src2, dtype=np.int32)
        vp = lambda a: a.ctypes.data_as(ctypes.c_void_p)   # noqa: E731
        _native.check(_native.lib().shwfs_plan_set_resampling(self._h, g.ef_size, vp(s1), vp(s2), vp(aff)),
                      'shwfs_plan_set_resampling')
        self._static_key = None
        self.resampled = True

    def set_frames_per_block(self, fpb):
        """Streaming depth of the NGS MFMA kernel (1 to 64 frames per block)."""
        _native.check(_native.lib().shwfs_plan_set_frames_per_block(self._h, int(fpb)), 'set_frames_per_block')
        self.fpb = int(fpb)

    def set_kernels(self, kspec, stream=None):
        """kspec: (n_sub, L, L) complex64 device tensor of kernel spectra on the lag grid."""
        n_sub = self.geom.n_lenses ** 2
        if (kspec.dtype != torch.complex64 or not kspec.is_cuda or not kspec.is_contiguous()
                or tuple(kspec.shape) != (n_sub, self.L, self.L)):
            raise ValueError(f'kernel spectra must be a contiguous complex64 ({n_sub}, {self.L}, {self.L}) tensor')
        _native.check(_native.lib().shwfs_set_kernels(self._h, _native.ptr(kspec), _native.stream_handle(stream)),
                      'shwfs_set_kernels')

    def set_static_pupil(self, A, S0=1.0, n_frames=1, a_frame_stride=None, stream=None):
        """Register an amplitude that stays constant (shwfs_plan_set_pupil): its statistics are
        computed once and reused by forward calls with the same tensor, stride and S0.  Register
        again after modifying ``A``; ``None`` forgets it."""
        if A is None:
            _native.check(_native.lib().shwfs_plan_set_pupil(self._h, None, 0, 1, 0.0, None), 'set_pupil')
            self._static_key = None
            return
        ef = self.geom.ef_size
        if A.dtype != torch.float32 or not A.is_contiguous() or not A.is_cuda:
            raise ValueError('pupil must be a contiguous float32 device tensor')
        if a_frame_stride is None:
            a_frame_stride = 0 if A.numel() == ef * ef else ef * ef
        _native.check(_native.lib().shwfs_plan_set_pupil(self._h, _native.ptr(A), a_frame_stride, n_frames,
                                                         float(S0), _native.stream_handle(stream)), 'set_pupil')
        self._static_key = (A.data_ptr(), A._version, a_frame_stride, float(S0))

    def static_pupil_is(self, A, S0, a_frame_stride):
        """True when ``A`` (same storage, not modified since) is the registered static pupil; the
        library's own registration (shwfs_plan_pupil_registered) is consulted as well."""
        key = getattr(self, '_static_key', None)
        if key is None or key != (A.data_ptr(), A._version, a_frame_stride, float(S0)):
            return False
        rc = _native.lib().shwfs_plan_pupil_registered(self._h, _native.ptr(A), a_frame_stride, float(S0))
        if rc < 0:
            _native.check(rc, 'shwfs_plan_pupil_registered')
        return rc == 1

    def bind_slopec(self, slopec):
        """Fuse the centre of gravity of a SlopecPlan into this plan's forward
        (shwfs_plan_bind_slopec).  Returns False when the pair cannot be fused; ``None`` unbinds."""
        h = slopec._h if slopec is not None else None
        rc = _native.lib().shwfs_plan_bind_slopec(self._h, h)
        if rc == _native.SHWFS_ERR_UNSUPPORTED:
            self._slopec = None
            return False
        _native.check(rc, 'shwfs_plan_bind_slopec')
        self._slopec = slopec      # keeps the bound plan alive
        return slopec is not None

    def forward_slopes(self, A, phase, out, S0, slopes, flux, totals, thr=0.0, pedestal=False, n_frames=1,
                       a_frame_stride=None, stream=None):
        """SH forward plus the bound ShSlopec in one pass (equal to ``forward`` followed by
        ``SlopecPlan.forward(out, ...)``)."""
        sc = getattr(self, '_slopec', None)
        if sc is None:
            raise RuntimeError('no SlopecPlan bound: call bind_slopec first')
        a_frame_stride = self._check(A, phase, out, n_frames, a_frame_stride)
        for t, dt in ((slopes, torch.float32), (flux, torch.float64), (totals, torch.float64)):
            if t.dtype != dt or not t.is_contiguous() or not t.is_cuda:
                raise ValueError('slopes float32, flux/totals float64, contiguous, on device')
        if (slopes.numel() < n_frames * 2 * sc.n_sub or flux.numel() < n_frames * sc.n_sub
                or totals.numel() < 2 * n_frames):
            raise ValueError('ShSlopec buffer sizes do not match the plan')
        ef, ccd = self.geom.ef_size, self.geom.ccd_side
        _native.check(_native.lib().shwfs_forward_slopes(
            self._h, n_frames, _native.ptr(A), a_frame_stride, _native.ptr(phase), ef * ef, float(S0),
            _native.ptr(out), ccd * ccd, float(thr), 1 if pedestal else 0, _native.ptr(slopes), 2 * sc.n_sub,
            _native.ptr(flux), _native.ptr(totals), _native.stream_handle(stream)), 'shwfs_forward_slopes')

    def _check(self, A, phase, out, n_frames, a_frame_stride):
        for t in (A, phase, out):
            if t.dtype != torch.float32 or not t.is_contiguous() or not t.is_cuda:
                raise ValueError('SH buffers must be contiguous float32 device tensors')
        ef = self.geom.ef_size
        ccd = self.geom.ccd_side
        if phase.numel() != n_frames * ef * ef or out.numel() != n_frames * ccd * ccd:
            raise ValueError('SH buffer sizes do not match the plan geometry')
        if a_frame_stride is None:
            a_frame_stride = 0 if A.numel() == ef * ef else ef * ef
        if A.numel() < (ef * ef if a_frame_stride == 0 else n_frames * ef * ef):
            raise ValueError('amplitude buffer too small')
        return a_frame_stride

    def forward(self, A, phase, out, S0, n_frames=1, a_frame_stride=None, stream=None):
        """A, phase: (n_frames,) ef x ef float32 device tensors (A may be 2-D = shared);
        out: (n_frames,) ccd x ccd float32."""
        a_frame_stride = self._check(A, phase, out, n_frames, a_frame_stride)
        ef, ccd = self.geom.ef_size, self.geom.ccd_side
        _native.check(_native.lib().shwfs_forward(
            self._h, n_frames, _native.ptr(A), a_frame_stride, _native.ptr(phase), ef * ef,
            float(S0), _native.ptr(out), ccd * ccd, _native.stream_handle(stream)), 'shwfs_forward')

    def vjp(self, A, phase, grad_ccd, grad_phase, S0, n_frames=1, a_frame_stride=None, stream=None):
        """dL/dphase (n_frames, ef, ef) from dL/d(CCD intensity) (n_frames, ccd, ccd) at the
        forward's inputs (shwfs_vjp; every plan kind, resampled plans on the input grid)."""
        a_frame_stride = self._check(A, phase, grad_ccd, n_frames, a_frame_stride)
        ef, ccd = self.geom.ef_size, self.geom.ccd_side
        if (grad_phase.dtype != torch.float32 or not grad_phase.is_contiguous() or not grad_phase.is_cuda
                or grad_phase.numel() != n_frames * ef * ef):
            raise ValueError('grad_phase must be a contiguous float32 device tensor of n_frames x ef x ef')
        _native.check(_native.lib().shwfs_vjp(
            self._h, n_frames, _native.ptr(A), a_frame_stride, _native.ptr(phase), ef * ef, float(S0),
            _native.ptr(grad_ccd), ccd * ccd, _native.ptr(grad_phase), ef * ef, _native.stream_handle(stream)),
            'shwfs_vjp')

    def __del__(self):
        h = getattr(self, '_h', None)
        if h is not None and h.value:
            try:
                _native.lib().shwfs_plan_destroy(h)
            except Exception:
                pass
            self._h = None


class SH(BaseProcessingObj):
    """specula/processing_objects/sh.py:30-477, HIP forward."""

    def __init__(self,
                 wavelengthInNm: float,
                 subap_wanted_fov: float,
                 sensor_pxscale: float,
                 subap_on_diameter: int,
                 subap_npx: int,
                 FoVres30mas: bool = False,
                 squaremask: bool = True,
                 fov_ovs_coeff: float = 0,
                 xShiftPhInPixel: float = 0,
                 yShiftPhInPixel: float = 0,
                 aXShiftPhInPixel: float = 0,
                 aYShiftPhInPixel: float = 0,
                 rotAnglePhInDeg: float = 0,
                 aRotAnglePhInDeg: float = 0,
                 do_not_double_fov_ovs: bool = False,
                 set_fov_res_to_turbpxsc: bool = False,
                 laser_launch_tel: LaserLaunchTelescope = None,
                 target_device_idx: int = None,
                 precision: int = None):
        super().__init__(target_device_idx=target_device_idx, precision=precision)
        if self.target_device_idx < 0:
            raise ValueError('SH runs on the GPU only (target_device_idx >= 0)')
        if self.precision != 1:
            raise ValueError('the HIP SH path computes in FP32 (with FP64 reductions): use precision=1')
        self._wavelengthInNm = wavelengthInNm
        self._lenslet = Lenslet(subap_on_diameter)
        self._subap_wanted_fov_arcsec = subap_wanted_fov
        self._subap_wanted_fov = subap_wanted_fov / RAD2ASEC
        self._sensor_pxscale_arcsec = sensor_pxscale
        self._sensor_pxscale = sensor_pxscale / RAD2ASEC
        self._subap_npx = subap_npx
        self._fov_ovs_coeff = fov_ovs_coeff
        self._squaremask = squaremask
        self._FoVres30mas = FoVres30mas
        self._fov_resolution_arcsec = 0.03 if FoVres30mas else 0
        self._rotAnglePhInDeg = rotAnglePhInDeg
        self._aRotAnglePhInDeg = aRotAnglePhInDeg
        self._xShiftPhInPixel = xShiftPhInPixel
        self._yShiftPhInPixel = yShiftPhInPixel
        self._aXShiftPhInPixel = aXShiftPhInPixel
        self._aYShiftPhInPixel = aYShiftPhInPixel
        self._set_fov_res_to_turbpxsc = set_fov_res_to_turbpxsc
        self._do_not_double_fov_ovs = do_not_double_fov_ovs
        self._laser_launch_tel = laser_launch_tel
        self._np_sub = 0
        self._fft_size = 0
        self._fov_ovs = 1
        self._floatShifts = False
        self._plan = None
        self._graph_key = None
        self.use_graph = False

        self._ccd_side = self._subap_npx * self._lenslet.n_lenses
        self._out_i = Intensity(self._ccd_side, self._ccd_side, precision=self.precision,
                                target_device_idx=self.target_device_idx)
        if self._laser_launch_tel is not None:
            self.inputs['sodium_altitude'] = InputValue(type=BaseValue, optional=True)
            self.inputs['sodium_intensity'] = InputValue(type=BaseValue, optional=True)
        self.inputs['in_ef'] = InputValue(type=ElectricField)
        self.outputs['out_i'] = self._out_i

    # ------------------------------------------------------------------ geometry
    def set_in_ef(self, in_ef):
        """sh.py:112-237 (+ calc_trigger_geometry's cut/mask, sh.py:262-282)."""
        if in_ef.size[0] != in_ef.size[1]:
            raise ValueError('SH needs a square ElectricField')
        g = sh_geometry(in_ef.size[0], in_ef.pixel_pitch, self._wavelengthInNm,
                        self._subap_wanted_fov_arcsec, self._sensor_pxscale_arcsec,
                        self._lenslet.n_lenses, self._subap_npx, FoVres30mas=self._FoVres30mas,
                        squaremask=self._squaremask, fov_ovs_coeff=self._fov_ovs_coeff,
                        do_not_double_fov_ovs=self._do_not_double_fov_ovs,
                        set_fov_res_to_turbpxsc=self._set_fov_res_to_turbpxsc,
                        fov_resolution_arcsec=self._fov_resolution_arcsec,
                        xShiftPhInPixel=self._xShiftPhInPixel, yShiftPhInPixel=self._yShiftPhInPixel,
                        aXShiftPhInPixel=self._aXShiftPhInPixel, aYShiftPhInPixel=self._aYShiftPhInPixel,
                        rotAnglePhInDeg=self._rotAnglePhInDeg, aRotAnglePhInDeg=self._aRotAnglePhInDeg)
        self.geometry = g
        self._fov_resolution_arcsec = g.fov_resolution_arcsec
        self._np_sub = g.np_sub
        self._ovs_np_sub = g.ovs_np_sub
        self._fov_ovs = g.fov_ovs
        self._fft_size = g.fft_size
        self._cutpixels = g.cutpixels
        self._cutsize = g.cutsize
        self._fp_mask = g.fp_mask
        self._sensor_pxscale = g.sensor_pxscale

    def _make_plan(self, **kw):
        plan = SHPlan(self.geometry, max_frames=1, device=self.target_device_idx, **kw)
        if self.geometry.needs_interpolation:
            # extrapolation sources from the pupil at set-up time, as the reference (sh.py:452-453)
            plan.set_resampling(self.in_ef.A.float())
        return plan

    def setup(self):
        super().setup()
        in_ef = self.inputs['in_ef'].get(target_device_idx=self.target_device_idx)
        self.set_in_ef(in_ef)
        self.in_ef = in_ef
        self._kernelobj = None
        self._kernel_version = None
        if self._laser_launch_tel is None:
            self._plan = self._make_plan(lgs=False)
        else:
            self._make_kernel_object(in_ef)
        self.build_stream(allow_parallel=False, capture=False)

    def _make_kernel_object(self, in_ef):
        """Kernel object as SH.calc_trigger_geometry builds it (sh.py:290-323)."""
        g = self.geometry
        llt = self._laser_launch_tel
        common = dict(dimx=self._lenslet.dimx, dimy=self._lenslet.dimy,
                      pxscale=g.fp4_pixel_pitch * RAD2ASEC, pupil_size_m=in_ef.pixel_pitch * in_ef.size[0],
                      dimension=g.fft_size, oversampling=1, return_fft=True, positive_shift_tt=True,
                      target_device_idx=-1, precision=self.precision)
        if len(llt.tel_pos) == 0:
            self._kernelobj = GaussianConvolutionKernel(spot_size=llt.spot_size, **common)
        else:
            theta = llt.beacon_tt if len(llt.beacon_tt) != 0 else []
            self._kernelobj = ConvolutionKernel(launcher_pos=llt.tel_pos, seeing=0.0, launcher_size=llt.spot_size,
                                                zfocus=llt.beacon_focus, theta=theta if len(theta) else [0.0, 0.0],
                                                **common)

    def prepare_trigger(self, t):
        super().prepare_trigger(t)
        if self._kernelobj is not None:
            self.prepare_kernels()

    def prepare_kernels(self):
        """sh.py:330-346, then upload the lag-grid spectra when the kernels changed."""
        if len(self._laser_launch_tel.tel_pos) != 0:
            alt = self.local_inputs.get('sodium_altitude')
            inten = self.local_inputs.get('sodium_intensity')
            if alt is None or inten is None:
                raise ValueError('sodium_altitude and sodium_intensity must be provided')
            alt, inten = alt.value, inten.value
            alt = alt.tolist() if hasattr(alt, 'tolist') else alt
            inten = inten.tolist() if hasattr(inten, 'tolist') else inten
        else:
            alt = inten = None
        self._kernelobj.prepare_for_sh(sodium_altitude=alt, sodium_intensity=inten, current_time=self.current_time)
        if self._kernelobj.version != self._kernel_version:
            normalised = self._kernelobj.kernels_normalised()
            if self._plan is None or self._plan.parseval != normalised:
                self._plan = self._make_plan(lgs=True, kernels_normalised=normalised)
                self._graph_key = None
            self._kspec = torch.tensor(self._kernelobj.lag_spectra(self.geometry.ovs_np_sub), device=self.device)
            self._plan.set_kernels(self._kspec, stream=None)
            self._kernel_version = self._kernelobj.version

    # ------------------------------------------------------------------ per frame
    def _inputs_fp32(self):
        ef = self.local_inputs.get('in_ef') or self.in_ef
        A, ph = ef.A, ef.phaseInNm
        if A.dtype != torch.float32 or not A.is_contiguous() or A.device != self.device:
            A = A.to(self.device, torch.float32).contiguous()
        if ph.dtype != torch.float32 or not ph.is_contiguous() or ph.device != self.device:
            ph = ph.to(self.device, torch.float32).contiguous()
        return ef, A, ph

    def trigger_code(self):
        ef, A, ph = self._inputs_fp32()
        if not self.use_graph and not self._plan.static_pupil_is(A, ef.S0, 0):
            # pupil statistics once per distinct amplitude (storage and in-place version)
            self._plan.set_static_pupil(A, ef.S0, stream=None)
        self._plan.forward(A, ph, self._out_i.i, ef.S0, n_frames=1, stream=None)

    def trigger(self):
        if self.ready and self.use_graph:
            ef, A, ph = self._inputs_fp32()
            if not self._plan.static_pupil_is(A, ef.S0, 0):
                # the captured forward reads the pupil statistics and the subaperture order from
                # plan buffers: refresh them here, outside the graph, when the amplitude changes
                self._plan.set_static_pupil(A, ef.S0, stream=None)
            key = (A.data_ptr(), ph.data_ptr(), float(ef.S0), self._out_i.i.data_ptr())
            if key != self._graph_key:
                self._keep = (A, ph)
                self.capture_stream()
                self._graph_key = key
        super().trigger()

    def post_trigger(self):
        super().post_trigger()
        self._out_i.generation_time = self.current_time

    def get_tlt_f(self, p, c):
        """Half-pixel tilt (sh.py:469-477), for inspection; the kernels evaluate it inline."""
        a = np.arange(-p // 2, p // 2)
        xx, yy = np.meshgrid(a, a)
        return np.exp(-2 * np.pi * 1j * (xx + yy) / (2 * (p + c)))
