"""SHSensor: a batched SH + ShSlopec pipeline on preallocated device buffers, with its adjoint.

This is the array-level API under the drop-in processing objects: many frames per launch
(push-pull fields, sequential screens), vector-Jacobian products of the slopes with respect to
the phase, and interaction matrices built from them.

Slopes follow SubapData order (column-major subaperture blocks, sh_subap_calibrator.py:56-78),
[sx..., sy...] (slopes.py:27-32).  Phases are in nm (ElectricField.phaseInNm).
"""

import numpy as np
import torch

from specula_amd import _native
from specula_amd.lib.subaps import detect_subaps
from specula_amd.processing_objects.sh import SHPlan
from specula_amd.processing_objects.sh_slopec import SlopecPlan, compute_xy_weights


class SHSensor:
    def __init__(self, geom, pupil, idxs=None, energy_th=0.5, max_frames=1, device=0, thr=0.0, variant=0,
                 extrapolation_mask=None, kernel=None):
        """extrapolation_mask: with field resampling, the amplitude the edge-extrapolation maps are
        built from (SH.setup reads in_ef.A, sh.py:453; inside a simulation that is the not yet
        propagated all-ones field, so no pixel is extrapolated).  None = ``pupil``.
        kernel: an LGS sensor's convolution kernel (a ConvolutionKernel prepared for this SH,
        sh.py:362-391): its lag-grid spectra are uploaded once; None = NGS."""
        self.geom = geom
        self.device = torch.device(f'cuda:{device}')
        self.max_frames = max_frames
        lgs = kernel is not None
        self.plan = SHPlan(geom, lgs=lgs, max_frames=max_frames, device=device, variant=variant,
                           kernels_normalised=kernel.kernels_normalised() if lgs else True)
        if lgs:
            self.plan.set_kernels(torch.as_tensor(kernel.lag_spectra(geom.ovs_np_sub), device=self.device)
                                  .to(torch.complex64).contiguous())
        ef, ccd = geom.ef_size, geom.ccd_side
        self.A = torch.as_tensor(np.asarray(pupil, dtype=np.float32), device=self.device).contiguous()
        if geom.needs_interpolation:
            self.plan.set_resampling(np.asarray(pupil if extrapolation_mask is None else extrapolation_mask))
        self.thr = float(thr)
        if idxs is None:
            flat = torch.zeros((1, ef, ef), dtype=torch.float32, device=self.device)
            img = torch.empty((1, ccd, ccd), dtype=torch.float32, device=self.device)
            self.plan.forward(self.A, flat, img, 1.0, n_frames=1)
            idxs, dmap = detect_subaps(img[0].cpu().numpy(), geom.n_lenses, energy_th)
            self.display_map = dmap
        self.idxs = np.asarray(idxs, dtype=np.int64)
        self.n_sub = self.idxs.shape[0]
        self.weights = compute_xy_weights(geom.subap_npx)
        self.slopec = SlopecPlan(self.idxs, (ccd, ccd), self.weights, max_frames=max_frames, device=device)
        self.ccd = torch.empty((max_frames, ccd, ccd), dtype=torch.float32, device=self.device)
        self.slopes_buf = torch.empty((max_frames, 2 * self.n_sub), dtype=torch.float32, device=self.device)
        self.flux = torch.empty((max_frames, self.n_sub), dtype=torch.float64, device=self.device)
        self.totals = torch.empty((max_frames, 2), dtype=torch.float64, device=self.device)
        self._gccd = None
        self._slot = None
        self._im_work = None
        # the sensor's pupil is fixed: its statistics once (S0 = 1, the forward default)
        self.plan.set_static_pupil(self.A, 1.0)
        # CoG fused into the SH kernel epilogue when the geometry allows it (same results)
        self.fused = self.plan.bind_slopec(self.slopec)

    # ------------------------------------------------------------------ forward
    def forward(self, phase, S0=1.0, stream=None, two_pass=False):
        """phase: (F, ef, ef) float32 device tensor -> slopes view (F, 2 n_sub) (buffers reused).
        The CoG runs in the SH kernel epilogue when the geometry allows it; ``two_pass`` forces
        the separate ShSlopec pass (FP64 sums per pixel instead of per 32-pixel partial)."""
        F = phase.shape[0]
        if F > self.max_frames:
            raise ValueError(f'{F} frames > max_frames {self.max_frames}')
        if self.fused and not two_pass:
            self.plan.forward_slopes(self.A, phase, self.ccd[:F], S0, self.slopes_buf[:F], self.flux[:F],
                                     self.totals[:F], thr=self.thr, n_frames=F, stream=stream)
        else:
            self.plan.forward(self.A, phase, self.ccd[:F], S0, n_frames=F, stream=stream)
            self.slopec.forward(self.ccd[:F], self.slopes_buf[:F], self.flux[:F], self.totals[:F], thr=self.thr,
                                n_frames=F, stream=stream)
        return self.slopes_buf[:F]

    # ------------------------------------------------------------------ adjoint
    def vjp(self, phase, grad_slopes, stream=None, sensor_grid=False):
        """dL/dphase for cotangents grad_slopes (C, 2 n_sub) at ``phase`` (ef, ef) or (C, ef, ef).
        With field resampling the gradient is on the input grid, or with ``sensor_grid`` with
        respect to the resampled phase (ef_ovs, ef_ovs) that the subapertures sample."""
        ef, ccd = self.geom.ef_size, self.geom.ccd_side
        efs = self.geom.ef_ovs if (sensor_grid and self.geom.needs_interpolation) else ef
        C = grad_slopes.shape[0]
        if C > self.max_frames:
            raise ValueError(f'{C} cotangents > max_frames {self.max_frames}')
        if grad_slopes.dim() != 2 or grad_slopes.shape[1] != 2 * self.n_sub:
            raise ValueError(f'grad_slopes must be (C, {2 * self.n_sub}), got {tuple(grad_slopes.shape)}')
        shared = phase.dim() == 2 or phase.shape[0] == 1
        if tuple(phase.shape[-2:]) != (ef, ef):
            raise ValueError(f'phase must be ({ef}, {ef}) or (C, {ef}, {ef}), got {tuple(phase.shape)}')
        if not shared and phase.shape[0] != C:
            # the VJP kernel reads one phase frame per cotangent
            raise ValueError(f'{phase.shape[0]} phase frames for {C} cotangents (give one phase or one per cotangent)')
        ph = phase.reshape(-1, ef, ef).contiguous()
        ph_fwd = ph[:1].expand(C, ef, ef).contiguous() if shared else ph
        self.forward(ph_fwd, stream=stream)                 # pixels of the current state
        if self._gccd is None or self._gccd.shape[0] < C:
            self._gccd = torch.empty((self.max_frames, ccd, ccd), dtype=torch.float32, device=self.device)
        gs = grad_slopes.to(torch.float32).contiguous()
        gph = torch.empty((C, efs, efs), dtype=torch.float32, device=self.device)
        lib = _native.lib()
        st = _native.stream_handle(stream)
        _native.check(lib.shslopec_vjp(self.slopec._h, C, _native.ptr(self.ccd), ccd * ccd, self.thr, 0,
                                       _native.ptr(gs), 2 * self.n_sub, _native.ptr(self._gccd), st),
                      'shslopec_vjp')
        fn = lib.shwfs_vjp_sensor_grid if sensor_grid else lib.shwfs_vjp
        _native.check(fn(self.plan._h, C, _native.ptr(self.A), 0, _native.ptr(ph), 0 if shared else ef * ef, 1.0,
                         _native.ptr(self._gccd), ccd * ccd, _native.ptr(gph), efs * efs, st), 'shwfs_vjp')
        return gph

    def jacobian_maps(self, phase, stream=None):
        """(Jx, Jy): sensor-grid maps of d slope_x / d phase and d slope_y / d phase of each pixel's
        own subaperture (block-diagonal Jacobian, two VJPs).  The sensor grid is the input grid, or
        for a resampled sensor the resampled one (ef_ovs x ef_ovs)."""
        g = torch.zeros((2, 2 * self.n_sub), dtype=torch.float32, device=self.device)
        g[0, :self.n_sub] = 1.0
        g[1, self.n_sub:] = 1.0
        gph = self.vjp(phase, g, stream=stream, sensor_grid=True)
        return gph[0], gph[1]

    def resample(self, frames, stream=None):
        """Phase-like frames (K, ef, ef) on the input grid -> (K, ef_ovs, ef_ovs) through the
        sensor's field resampling (edge extrapolation + bilinear, shwfs_resample); the frames
        themselves when the sensor does not resample."""
        if not self.geom.needs_interpolation:
            return frames
        ef, eo = self.geom.ef_size, self.geom.ef_ovs
        frames = frames.reshape(-1, ef, ef).to(torch.float32).contiguous()
        out = torch.empty((frames.shape[0], eo, eo), dtype=torch.float32, device=self.device)
        for k0 in range(0, frames.shape[0], 65535):
            k1 = min(frames.shape[0], k0 + 65535)
            _native.check(_native.lib().shwfs_resample(self.plan._h, k1 - k0, _native.ptr(frames[k0:k1]), ef * ef,
                                                       _native.ptr(out[k0:k1]), eo * eo,
                                                       _native.stream_handle(stream)), 'shwfs_resample')
        return out

    def slot_map(self):
        if self._slot is None:
            n, npx, W = self.geom.n_lenses, self.geom.subap_npx, self.geom.ccd_side
            slot = -np.ones(n * n, dtype=np.int32)
            org = self.idxs[:, 0]
            i, j = (org // W) // npx, (org % W) // npx
            slot[i * n + j] = np.arange(self.n_sub, dtype=np.int32)
            self._slot = torch.tensor(slot, device=self.device)
        return self._slot

    def interaction_matrix(self, phase, modes, scale=2.0, jac=None, stream=None):
        """IM (K, 2 n_sub) = scale * J . modes, push-pull convention for scale = 2.  ``modes``
        (K, ef, ef) on the input grid; a resampled sensor contracts its sensor-grid Jacobian with
        the resampled modes (the resampling is linear, so J R m = (J) (R m))."""
        jx, jy = jac if jac is not None else self.jacobian_maps(phase, stream=stream)
        modes = self.resample(modes, stream=stream)
        K = modes.shape[0]
        ef = modes.shape[-1]
        im = torch.empty((K, 2 * self.n_sub), dtype=torch.float32, device=self.device)
        modes = modes.contiguous()
        # workspace of the grid-order pass: K * 2 * n^2 floats + n_sub int32 (kept between calls)
        need = K * 2 * self.geom.n_lenses ** 2 + self.n_sub
        if self._im_work is None or self._im_work.numel() < need:
            self._im_work = torch.empty(need, dtype=torch.float32, device=self.device)
        _native.check(_native.lib().shwfs_im_from_jacobian(
            self.geom.n_lenses, self.geom.ovs_np_sub, ef, _native.ptr(jx), _native.ptr(jy), K,
            _native.ptr(modes), ef * ef, _native.ptr(self.slot_map()), self.n_sub, float(scale),
            _native.ptr(im), _native.ptr(self._im_work), _native.stream_handle(stream)), 'shwfs_im_from_jacobian')
        return im

    def pushpull_interaction_matrix(self, phase, modes, amp, chunk=None, stream=None):
        """Reference convention (im_rec_calibrator.py:58-63): IM[m] = s(phi + a m) / a + s(phi - a m) / (-a),
        evaluated in batches of 2*chunk sensor frames (fields and rows by shwfs_pushpull_*)."""
        K = modes.shape[0]
        chunk = chunk or max(1, self.max_frames // 2)
        ef = self.geom.ef_size
        n = 2 * self.n_sub
        im = torch.empty((K, n), dtype=torch.float32, device=self.device)
        buf = torch.empty((2 * chunk, ef, ef), dtype=torch.float32, device=self.device)
        ph = phase.reshape(ef, ef).to(torch.float32).contiguous()
        modes = modes.reshape(K, ef, ef).to(torch.float32).contiguous()
        lib, st = _native.lib(), _native.stream_handle(stream)
        for m0 in range(0, K, chunk):
            m1 = min(K, m0 + chunk)
            c = m1 - m0
            _native.check(lib.shwfs_pushpull_fields(c, _native.ptr(ph), ef * ef, _native.ptr(modes[m0:m1]), ef * ef,
                                                    float(amp), _native.ptr(buf), ef * ef, st), 'shwfs_pushpull_fields')
            s = self.forward(buf[:2 * c], stream=stream)
            _native.check(lib.shwfs_pushpull_rows(c, _native.ptr(s), n, n, float(amp), _native.ptr(im[m0:m1]), n, st),
                          'shwfs_pushpull_rows')
        return im


class SHSlopes(torch.autograd.Function):
    """slopes = SHSensor.forward(phase) with the hand-written HIP adjoint as backward.

    phase: (ef, ef) or (F, ef, ef) float32 device tensor [nm]; returns (F, 2 n_sub)."""

    @staticmethod
    def forward(ctx, phase, sensor):
        ph = phase.reshape(-1, sensor.geom.ef_size, sensor.geom.ef_size).to(torch.float32).contiguous()
        ctx.sensor = sensor
        ctx.save_for_backward(ph)
        ctx.in_shape = phase.shape
        step = sensor.max_frames
        return torch.cat([sensor.forward(ph[f0:f0 + step]).clone() for f0 in range(0, ph.shape[0], step)])

    @staticmethod
    def backward(ctx, grad_slopes):
        (ph,) = ctx.saved_tensors
        sensor = ctx.sensor
        F = ph.shape[0]
        gs = grad_slopes.reshape(F, -1)
        # frame f's cotangent at frame f's phase: one batched VJP per max_frames frames
        step = sensor.max_frames
        g = torch.cat([sensor.vjp(ph[f0:f0 + step], gs[f0:f0 + step]) for f0 in range(0, F, step)])
        return g.reshape(ctx.in_shape), None


def sh_slopes(phase, sensor):
    return SHSlopes.apply(phase, sensor)


def fit_modes(sensor, modes, target_slopes, phase0=None, coeffs0=None, max_iter=50):
    """Modal least-squares fit through the sensor model: the coefficients c minimising
    ||slopes(phase0 + sum_m c_m modes[m]) - target||^2, by L-BFGS with a strong-Wolfe line search
    on gradients from the hand-written adjoint (SHSlopes.backward).  This is the fit behind
    non-common-path aberration estimates and phase retrieval from measured slopes: the
    non-linear sensor response is modelled, not linearised as an interaction matrix would.

    modes (K, ef, ef) float32 device tensor [nm per unit coefficient]; target_slopes (2 n_sub,)
    in the sensor's slope order; phase0 (ef, ef) known phase [nm] (None: flat); coeffs0 (K,)
    starting point (None: zeros).  Returns (coefficients (K,) float32, final loss)."""
    ef = sensor.geom.ef_size
    K = modes.shape[0]
    if tuple(modes.shape[1:]) != (ef, ef):
        raise ValueError(f'modes must be (K, {ef}, {ef}), got {tuple(modes.shape)}')
    m = modes.reshape(K, ef * ef).to(torch.float32)
    base = (torch.zeros(ef * ef, dtype=torch.float32, device=sensor.device) if phase0 is None
            else phase0.reshape(ef * ef).to(torch.float32))
    tgt = target_slopes.reshape(1, -1).to(torch.float32)
    if tgt.shape[1] != 2 * sensor.n_sub:
        raise ValueError(f'target has {tgt.shape[1]} slopes, the sensor {2 * sensor.n_sub}')
    c = (torch.zeros(K, dtype=torch.float32, device=sensor.device) if coeffs0 is None
         else coeffs0.detach().to(torch.float32).clone())
    c.requires_grad_(True)
    opt = torch.optim.LBFGS([c], lr=1.0, max_iter=max_iter, line_search_fn='strong_wolfe')

    def loss_of(coeffs):
        r = sh_slopes((base + coeffs @ m).reshape(ef, ef), sensor) - tgt
        return (r * r).sum()

    def closure():
        opt.zero_grad()
        loss = loss_of(c)
        loss.backward()
        return loss

    opt.step(closure)
    with torch.no_grad():
        final = float(loss_of(c.detach()))
    return c.detach(), final
