"""Adjoint of SH sensors with field resampling (fov_ovs != 1, rotation, shift: the reference's
default set-ups) and with a FoV cut (image normalised by its own sum, not the Parseval scale).

* The VJP against central differences of the FP64 oracle chain (sh.py:348-428 + sh_slopec.py),
  at the params_scao_sh_test.yml geometry (fov_ovs 2, N 32, cut 24), the 20x20 default
  (fov_ovs 2, np 16, N 144, cut 88), a rotated / shifted field (fov_ovs 1, N 72, cut 16) and the
  same without resampling (cut only).
* The sensor-grid VJP and the input-grid VJP are related by the transpose of the resampling:
  <J^T g, v> on the input grid = <(J_s)^T g, R v> with R the resampling (shwfs_resample), which
  pins the scatter kernel against the forward gather, edge extrapolation included.
* The interaction matrix of the calibration geometry (40 Zernike modes of the 0.1-obstructed
  64 px pupil, the fixture scao_rec_n8_th0.5_ref.fits set-up) from the adjoint against push-pull
  on the GPU, with the Richardson extrapolation (4 IM(a/2) - IM(a))/3 that cancels push-pull's
  O(a^2) term."""
import numpy as np
import pytest
import torch

from oracle import sh_oracle as O
from specula_amd.lib.geometry import sh_geometry
from specula_amd.lib.grids import make_mask
from specula_amd.lib.phasescreen import annular_pupil, von_karman_opd
from specula_amd.sensor import SHSensor

pytestmark = pytest.mark.gpu
DEV = 'cuda:0'

SCAO_TEST = dict(wavelengthInNm=600, subap_wanted_fov=4.0, sensor_pxscale=0.5, subap_on_diameter=8, subap_npx=8)
K20 = dict(wavelengthInNm=750, subap_wanted_fov=2.4, sensor_pxscale=0.3, subap_on_diameter=20, subap_npx=8)
CASES = {
    'scao_test_fov_ovs2_cut24': (64, 0.015625, 0.0, SCAO_TEST),
    '20x20_fov_ovs2_cut88': (160, 0.05, 0.14, K20),
    '20x20_rot_shift_cut16': (160, 0.05, 0.14, dict(K20, fov_ovs_coeff=1.0, rotAnglePhInDeg=10.0,
                                                      xShiftPhInPixel=1.0, yShiftPhInPixel=-2.0)),
    '20x20_cut16_no_resampling': (160, 0.05, 0.14, dict(K20, fov_ovs_coeff=1.0)),
}


def _setup(name, max_frames=4):
    ef, pitch, obs, kw = CASES[name]
    g = sh_geometry(ef, pitch, **kw)
    go = O.sh_geometry(ef_size=ef, pixel_pitch=pitch, **kw)
    assert g.cutpixels > 0
    A = annular_pupil(ef, obs) if obs else O.make_mask(ef, 0.0, 1.0).astype(np.float64)
    sen = SHSensor(g, A.astype(np.float32), max_frames=max_frames, device=0)
    assert not sen.plan.parseval
    ph = (von_karman_opd(1 << int(np.ceil(np.log2(ef + 8))), pitch, seed=3)[:ef, :ef] * 0.5).astype(np.float32)
    return g, go, A, sen, ph


@pytest.mark.parametrize('name', list(CASES))
def test_vjp_matches_fp64_finite_differences(name):
    g, go, A, sen, ph = _setup(name)
    print(f'{name}: resampled {g.needs_interpolation}, np {g.ovs_np_sub}, N {g.fft_size}, cut {g.cutpixels}, '
          f'{sen.n_sub} subapertures', flush=True)
    rng = np.random.default_rng(21)
    gs = rng.standard_normal(2 * sen.n_sub)
    grad = sen.vjp(torch.tensor(ph, device=DEV), torch.tensor(gs[None], dtype=torch.float32, device=DEV))[0]
    grad = grad.double().cpu().numpy()
    assert grad.shape == ph.shape
    h = 1e-3 / (2 * np.pi / go['wavelengthInNm'])            # 1e-3 rad in nm
    errs = []
    for d in range(3):
        v = rng.standard_normal(ph.shape)
        sp = O.shslopec(O.sh_forward(A, ph.astype(np.float64) + h * v, go), sen.idxs)['slopes']
        sm = O.shslopec(O.sh_forward(A, ph.astype(np.float64) - h * v, go), sen.idxs)['slopes']
        fd = gs @ (sp - sm) / (2 * h)
        adj = float((grad * v).sum())
        errs.append(abs(adj - fd) / abs(fd))
        print(f'  direction {d}: fd {fd:.6e} adjoint {adj:.6e} rel {errs[-1]:.2e}', flush=True)
        assert abs(adj - fd) <= 2e-3 * abs(fd) + 1e-6 * np.abs(gs).sum()


@pytest.mark.parametrize('name', [n for n in CASES if 'no_resampling' not in n])
def test_input_grid_vjp_is_the_resampling_transpose_of_the_sensor_grid_vjp(name):
    g, go, A, sen, ph = _setup(name)
    rng = np.random.default_rng(5)
    gsl = torch.tensor(rng.standard_normal((2, 2 * sen.n_sub)), dtype=torch.float32, device=DEV)
    pt = torch.tensor(ph, device=DEV)
    g_in = sen.vjp(pt, gsl).double()
    g_s = sen.vjp(pt, gsl, sensor_grid=True).double()
    assert g_s.shape[-1] == g.ef_ovs and g_in.shape[-1] == g.ef_size
    V = torch.tensor(rng.standard_normal((4, g.ef_size, g.ef_size)), dtype=torch.float32, device=DEV)
    RV = sen.resample(V).double()
    for c in range(2):
        lhs = (g_in[c][None] * V.double()).sum(dim=(1, 2))
        rhs = (g_s[c][None] * RV).sum(dim=(1, 2))
        scale = (g_s[c].abs()[None] * RV.abs()).sum(dim=(1, 2))
        rel = ((lhs - rhs).abs() / scale).max().item()
        print(f'{name}: cotangent {c}: max |<J^T g, v> - <J_s^T g, R v>| / sum|terms| {rel:.2e}', flush=True)
        assert rel < 1e-5


def test_resampled_calibration_im_matches_richardson_pushpull():
    from specula_amd.processing_objects.dm import DM
    from specula_amd.lib import fits
    import os
    golden = os.path.join(os.path.dirname(__file__), 'golden')
    g = sh_geometry(64, 0.015625, **SCAO_TEST)
    dm = DM(pixel_pitch=0.015625, height=0, type_str='zernike', nmodes=40, npixels=64, obsratio=0.1, sign=1,
            target_device_idx=0)
    A = make_mask(64, None, 1.0).astype(np.float32) * dm.layer.A.cpu().numpy()
    idxs = fits.read(os.path.join(golden, 'scao_subaps_n8_th0.5_ref.fits'))[1][0].T
    sen = SHSensor(g, A, idxs=idxs, max_frames=80, device=0, extrapolation_mask=np.ones((64, 64)))
    modes = dm.project(torch.eye(40, device=DEV))                  # (40, 64, 64) phase per unit command
    flat = torch.zeros((64, 64), dtype=torch.float32, device=DEV)
    im_adj = sen.interaction_matrix(flat, modes, scale=2.0).double()
    pp = {a: sen.pushpull_interaction_matrix(flat, modes, amp=a).double() for a in (50.0, 10.0, 5.0)}
    rich = (4 * pp[5.0] - pp[10.0]) / 3

    def rel(a, b):
        return float((a - b).norm() / b.norm())
    r50, r10, rr = rel(im_adj, pp[50.0]), rel(im_adj, pp[10.0]), rel(im_adj, rich)
    print(f'scao calibration IM (40 Zernike, fov_ovs 2, cut 24): adjoint vs push-pull 50 nm {r50:.2e}, '
          f'10 nm {r10:.2e}, Richardson (10, 5 nm) {rr:.2e}', flush=True)
    assert rr < 1e-3
