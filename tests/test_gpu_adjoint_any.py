"""Adjoint of the SH sensors served by the any-size VJP (sh_general_vjp_kernel): LGS sensors (the
kernel convolution of sh.py:362-391, with the specialised np 16 forward and with the any-size
forward), the large subapertures of the reference's own set-ups at reduced lenslet counts
(params_eris.yml:102-113 LGS np 20 / M 64 with fov_ovs 5, params_eris.yml:148-157 NGS np 80 /
M 256, params_morfeo_full.yml:259-268 NGS np 240 / M 512), round FoV masks (K separable terms,
sh.py:282) and an NGS np 7 / M 16 plan whose forward runs on the matrix cores.

* The VJP against central differences of the FP64 oracle chain (sh.py:348-428 + sh_slopec.py)
  along three random directions.
* The interaction matrix of the any-np contraction (im_from_jacobian_any_kernel) against the
  direct FP64 contraction of the same Jacobian maps, and the adjoint IM of an LGS sensor against
  push-pull with the Richardson extrapolation (4 IM(a/2) - IM(a)) / 3."""
import numpy as np
import pytest
import torch

from oracle import sh_oracle as O
from specula_amd import RAD2ASEC, workloads
from specula_amd.data_objects.convolution_kernel import ConvolutionKernel
from specula_amd.lib.geometry import sh_geometry
from specula_amd.lib.phasescreen import annular_pupil, von_karman_opd
from specula_amd.sensor import SHSensor

pytestmark = pytest.mark.gpu
DEV = 'cuda:0'

ERIS_LGS = dict(subap_wanted_fov=5, sensor_pxscale=0.85, subap_npx=6, wavelengthInNm=589)
ERIS_LO = dict(subap_wanted_fov=4.8, sensor_pxscale=0.4, subap_npx=12, wavelengthInNm=700)
MORFEO_NGS1 = dict(subap_wanted_fov=1.8, sensor_pxscale=0.015, subap_npx=120, wavelengthInNm=1650)
_P16 = 39.0 / 640
_LD16 = 589e-9 / (16 * _P16) * RAD2ASEC
LGS16 = dict(wavelengthInNm=589, subap_wanted_fov=16 * _LD16, sensor_pxscale=_LD16, subap_npx=16,
             fov_ovs_coeff=1.0)
NP7 = dict(wavelengthInNm=750, subap_wanted_fov=2.4, sensor_pxscale=0.3, subap_npx=8, fov_ovs_coeff=1.0)

# (name, ef, pitch, obstruction, LGS, SH keywords, expected (np, N, cut, M))
CASES = [
    ('eris_lgs_np20_fov_ovs5', 16, 0.05, 0.0, True, dict(ERIS_LGS, subap_on_diameter=4), (20, 120, 70, 64)),
    ('lgs_np16_specialised_forward', 64, _P16, 0.0, True, dict(LGS16, subap_on_diameter=4), (16, 32, 0, 32)),
    ('eris_lgs_round_mask', 16, 0.05, 0.0, True, dict(ERIS_LGS, subap_on_diameter=4, squaremask=False),
     (20, 120, 70, 64)),
    ('eris_lo_np80', 80, 0.05, 0.0, False, dict(ERIS_LO, subap_on_diameter=2), (80, 160, 28, 256)),
    ('morfeo_ngs1_np240', 480, 0.0802, 0.0, False, dict(MORFEO_NGS1, subap_on_diameter=2), (240, 480, 276, 512)),
    ('20x20_round_mask', 160, 0.05, 0.14, False,
     dict(wavelengthInNm=750, subap_wanted_fov=2.4, sensor_pxscale=0.3, subap_on_diameter=20, subap_npx=8,
          fov_ovs_coeff=1.0, squaremask=False), (8, 72, 16, 16)),
    ('np7_mfma_forward', 28, 0.05, 0.0, False, dict(NP7, subap_on_diameter=4), (7, 21, 5, 16)),
]


def lgs_kernel(g, launcher=(6.0, -3.0, 0.0)):
    """Sodium-profile kernels (11 layers, FWHM 10 km at 90 km) seen from an off-axis launcher."""
    k = ConvolutionKernel(g.n_lenses, g.n_lenses, g.fp4_pixel_pitch * RAD2ASEC, g.pixel_pitch * g.ef_size,
                          g.fft_size, launcher_pos=list(launcher), launcher_size=1.0, target_device_idx=-1,
                          precision=1)
    k.prepare_for_sh(sodium_altitude=workloads.LGS_LAYERS_M.tolist(),
                     sodium_intensity=workloads.sodium_profile().tolist())
    return k


def _setup(case, max_frames=4):
    name, ef, pitch, obs, lgs, kw, shape = case
    g = sh_geometry(ef, pitch, **kw)
    go = O.sh_geometry(ef_size=ef, pixel_pitch=pitch, **kw)
    M = 1 << int(np.ceil(np.log2(2 * g.ovs_np_sub - 1)))
    assert (g.ovs_np_sub, g.fft_size, g.cutpixels, M) == shape
    A = annular_pupil(ef, obs) if obs else O.make_mask(ef, 0.0, 1.0).astype(np.float64)
    kern = lgs_kernel(g) if lgs else None
    sen = SHSensor(g, A.astype(np.float32), max_frames=max_frames, device=0, kernel=kern)
    ph = (von_karman_opd(1 << int(np.ceil(np.log2(ef + 8))), pitch, seed=3)[:ef, :ef] * 0.5).astype(np.float32)
    print(f'{name}: np {g.ovs_np_sub} N {g.fft_size} M {M} cut {g.cutpixels} npx {g.subap_npx} resampled '
          f'{g.needs_interpolation} mask terms {sen.plan.mask_terms} parseval {sen.plan.parseval}, '
          f'{sen.n_sub} valid subapertures', flush=True)
    return g, go, A, kern, sen, ph


@pytest.mark.parametrize('case', CASES, ids=[c[0] for c in CASES])
def test_vjp_matches_fp64_finite_differences(case):
    g, go, A, kern, sen, ph = _setup(case)
    kk = None if kern is None else kern.kernels
    rng = np.random.default_rng(21)
    gs = rng.standard_normal(2 * sen.n_sub)
    grad = sen.vjp(torch.tensor(ph, device=DEV), torch.tensor(gs[None], dtype=torch.float32, device=DEV))[0]
    grad = grad.double().cpu().numpy()
    assert grad.shape == ph.shape and np.isfinite(grad).all()
    h = 1e-3 / (2 * np.pi / go['wavelengthInNm'])            # 1e-3 rad in nm
    for d in range(3):
        v = rng.standard_normal(ph.shape)
        sp = O.shslopec(O.sh_forward(A, ph.astype(np.float64) + h * v, go, kernels=kk), sen.idxs)['slopes']
        sm = O.shslopec(O.sh_forward(A, ph.astype(np.float64) - h * v, go, kernels=kk), sen.idxs)['slopes']
        fd = gs @ (sp - sm) / (2 * h)
        adj = float((grad * v).sum())
        print(f'  direction {d}: fd {fd:.6e} adjoint {adj:.6e} rel {abs(adj - fd) / abs(fd):.2e}', flush=True)
        assert abs(adj - fd) <= 2e-3 * abs(fd) + 1e-6 * np.abs(gs).sum()


def test_cotangent_batch_and_shared_phase():
    """Several cotangents at one phase (phase stride 0) equal one call per cotangent."""
    g, go, A, kern, sen, ph = _setup(CASES[0])
    rng = np.random.default_rng(4)
    gsl = torch.tensor(rng.standard_normal((3, 2 * sen.n_sub)), dtype=torch.float32, device=DEV)
    pt = torch.tensor(ph, device=DEV)
    both = sen.vjp(pt, gsl)
    for c in range(3):
        one = sen.vjp(pt, gsl[c:c + 1])[0]
        err = ((both[c] - one).abs().max() / one.abs().max()).item()
        print(f'cotangent {c}: batched vs single {err:.2e}', flush=True)
        assert err < 1e-5


@pytest.mark.parametrize('case', [CASES[0], CASES[3], CASES[6]], ids=[CASES[i][0] for i in (0, 3, 6)])
def test_any_np_interaction_matrix_is_the_contraction_of_the_jacobian(case):
    g, go, A, kern, sen, ph = _setup(case, max_frames=2)
    rng = np.random.default_rng(9)
    K = 37
    modes = torch.tensor(rng.standard_normal((K, g.ef_size, g.ef_size)), dtype=torch.float32, device=DEV)
    flat = torch.zeros((g.ef_size, g.ef_size), dtype=torch.float32, device=DEV)
    jx, jy = sen.jacobian_maps(flat)
    im = sen.interaction_matrix(flat, modes, scale=2.0, jac=(jx, jy)).double().cpu().numpy()
    rm = sen.resample(modes).double().cpu().numpy()
    npp, n, ef = g.ovs_np_sub, g.n_lenses, rm.shape[-1]
    W = g.ccd_side
    org = sen.idxs[:, 0]
    ii, jj = (org // W) // g.subap_npx, (org % W) // g.subap_npx
    J = [jx.double().cpu().numpy(), jy.double().cpu().numpy()]
    ref = np.zeros((K, 2 * sen.n_sub))
    for s in range(sen.n_sub):
        y0, x0 = ii[s] * npp, jj[s] * npp
        for c in range(2):
            ref[:, c * sen.n_sub + s] = 2.0 * np.einsum('kyx,yx->k', rm[:, y0:y0 + npp, x0:x0 + npp],
                                                        J[c][y0:y0 + npp, x0:x0 + npp])
    err = np.abs(im - ref).max() / np.abs(ref).max()
    print(f'{case[0]}: IM ({K} modes, ef {ef}, np {npp}) vs FP64 contraction {err:.2e}', flush=True)
    assert err < 1e-5


def test_lgs_adjoint_im_matches_richardson_pushpull():
    from specula_amd.processing_objects.dm import DM
    case = CASES[1]
    g, go, A, kern, sen0, ph = _setup(case)
    ef = g.ef_size
    sen = SHSensor(g, A.astype(np.float32), idxs=sen0.idxs, max_frames=60, device=0, kernel=kern)
    dm = DM(pixel_pitch=g.pixel_pitch, height=0, type_str='zernike', nmodes=20, npixels=ef, obsratio=0.0, sign=1,
            target_device_idx=0)
    modes = dm.project(torch.eye(20, device=DEV))
    flat = torch.zeros((ef, ef), dtype=torch.float32, device=DEV)
    im_adj = sen.interaction_matrix(flat, modes, scale=2.0).double()
    pp = {a: sen.pushpull_interaction_matrix(flat, modes, amp=a).double() for a in (10.0, 5.0)}
    rich = (4 * pp[5.0] - pp[10.0]) / 3

    def rel(a, b):
        return float((a - b).norm() / b.norm())
    r10, rr = rel(im_adj, pp[10.0]), rel(im_adj, rich)
    print(f'LGS np 16 calibration IM (20 Zernike): adjoint vs push-pull 10 nm {r10:.2e}, Richardson {rr:.2e}',
          flush=True)
    assert rr < 1e-3


@pytest.mark.parametrize('case', [CASES[0], CASES[1], CASES[3]], ids=[CASES[i][0] for i in (0, 1, 3)])
def test_per_frame_pupils_equal_single_frame_calls(case):
    """Two frames with their own pupil (one with a dark quadrant) and phase in one plan-level VJP
    equal two one-frame calls (per-frame flux scale and cotangent offset)."""
    g, go, A, kern, sen, ph = _setup(case, max_frames=4)
    ef, ccd = g.ef_size, g.ccd_side
    rng = np.random.default_rng(13)
    A2 = np.stack([A, A]).astype(np.float32)
    A2[1, :ef // 2, :ef // 2] = 0.0
    ph2 = np.stack([ph, np.roll(ph, 3, axis=1) * 0.7]).astype(np.float32)
    At = torch.tensor(A2, device=DEV)
    pt = torch.tensor(ph2, device=DEV)
    gc = torch.tensor(rng.standard_normal((2, ccd, ccd)), dtype=torch.float32, device=DEV)
    both = torch.empty((2, ef, ef), dtype=torch.float32, device=DEV)
    sen.plan.vjp(At, pt, gc, both, 1.0, n_frames=2, a_frame_stride=ef * ef)
    for f in range(2):
        one = torch.empty((1, ef, ef), dtype=torch.float32, device=DEV)
        sen.plan.vjp(At[f].contiguous(), pt[f:f + 1].contiguous(), gc[f:f + 1].contiguous(), one, 1.0, n_frames=1)
        err = ((both[f] - one[0]).abs().max() / one.abs().max()).item()
        print(f'{case[0]}: frame {f}: two-frame call vs one-frame call {err:.2e}', flush=True)
        assert err < 1e-5
    assert float(both[1][:ef // 2, :ef // 2].abs().max()) == 0.0 or g.needs_interpolation
