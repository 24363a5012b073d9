"""SH with field resampling (fov_ovs != 1, rotation, shift): edge extrapolation + bilinear
resampling on the GPU (shwfs_plan_set_resampling) against the FP64 oracle and the reference's
fov_ovs = 2 fixtures."""
import os

import numpy as np
import pytest
import torch

from helpers import check_intensity, check_slopes
from oracle import sh_oracle as O
from specula_amd.lib import fits
from specula_amd.lib.geometry import sh_geometry
from specula_amd.lib.phasescreen import annular_pupil, von_karman_opd
from specula_amd.lib.subaps import detect_subaps
from specula_amd.processing_objects.sh import SHPlan
from specula_amd.processing_objects.sh_slopec import SlopecPlan, compute_xy_weights

pytestmark = pytest.mark.gpu
DEV = 'cuda:0'
GOLDEN = os.path.join(os.path.dirname(__file__), 'golden')

SCAO_TEST = dict(wavelengthInNm=600, subap_wanted_fov=4.0, sensor_pxscale=0.5, subap_on_diameter=8, subap_npx=8)
CASES = [
    # (ef, pitch, obstruction, keys): params_scao_sh_test.yml SH (fov_ovs 2, N 32, cut 24, 1:1 CCD)
    (64, 0.015625, 0.0, SCAO_TEST),
    # default fov_ovs_coeff on the 20x20 config: fov_ovs 2, np 16 per subaperture
    (160, 0.05, 0.14, dict(wavelengthInNm=750, subap_wanted_fov=2.4, sensor_pxscale=0.3,
                            subap_on_diameter=20, subap_npx=8)),
    # fov_ovs 1 with a rotation and a shift
    (160, 0.05, 0.14, dict(wavelengthInNm=750, subap_wanted_fov=2.4, sensor_pxscale=0.3,
                            subap_on_diameter=20, subap_npx=8, fov_ovs_coeff=1.0,
                            rotAnglePhInDeg=10.0, xShiftPhInPixel=1.0, yShiftPhInPixel=-2.0)),
]


def run(ef, pitch, obs, kw, n_frames=2, S0=3.0, seed=11):
    g = sh_geometry(ef, pitch, **kw)
    go = O.sh_geometry(ef_size=ef, pixel_pitch=pitch, **kw)
    assert g.needs_interpolation and go['interpolate']
    assert (g.ovs_np_sub, g.fft_size, g.cutpixels, g.ef_ovs) == (go['ovs_np_sub'], go['fft_size'], go['cutpixels'], go['ef_ovs'])
    A = annular_pupil(ef, obs) if obs else O.make_mask(ef, 0.0, 1.0).astype(np.float64)
    scr = von_karman_opd(1 << int(np.ceil(np.log2(ef + 8))), pitch, seed=seed)
    frames = np.stack([scr[:ef, k:k + ef] for k in range(n_frames)]).astype(np.float32)
    plan = SHPlan(g, max_frames=n_frames, device=0)
    plan.set_resampling(A)
    At = torch.tensor(A.astype(np.float32), device=DEV)
    ph = torch.tensor(frames, device=DEV)
    out = torch.empty((n_frames, g.ccd_side, g.ccd_side), dtype=torch.float32, device=DEV)
    plan.forward(At, ph, out, S0, n_frames=n_frames)
    torch.cuda.synchronize()
    return g, go, A, frames, out.cpu().numpy(), plan


@pytest.mark.parametrize('case', range(len(CASES)))
def test_resampled_sh_matches_oracle(case):
    ef, pitch, obs, kw = CASES[case]
    print(f'resample case {case}: ef={ef} kw={kw}', flush=True)
    g, go, A, frames, out, plan = run(ef, pitch, obs, kw)
    for k in range(len(frames)):
        ref = O.sh_forward(A, frames[k].astype(np.float64), go, S0=3.0)
        st = check_intensity(out[k], ref, g.n_lenses, g.subap_npx)
        print(f'  frame {k}: fov_ovs {g.fov_ovs} np {g.ovs_np_sub} N {g.fft_size}: total flux rel err '
              f'{st["total"]:.2e}, normalised image err {st["img"]:.2e}', flush=True)


def test_resampled_grey_pupil_matches_oracle():
    """Grey amplitude (apodised ring): the extrapolation maps weigh neighbours by A, so pixels
    inside the pupil with A < 1 are extrapolated as in the reference."""
    ef, pitch, obs, kw = CASES[1]
    g = sh_geometry(ef, pitch, **kw)
    go = O.sh_geometry(ef_size=ef, pixel_pitch=pitch, **kw)
    A = annular_pupil(ef, obs)
    r = np.hypot(*np.mgrid[-1:1:ef * 1j, -1:1:ef * 1j])
    A = A * np.where(r > 0.8, 0.5, 1.0)
    ph = von_karman_opd(256, pitch, seed=5)[:ef, :ef].astype(np.float32)
    plan = SHPlan(g, max_frames=1, device=0)
    plan.set_resampling(A)
    out = torch.empty((1, g.ccd_side, g.ccd_side), dtype=torch.float32, device=DEV)
    plan.forward(torch.tensor(A.astype(np.float32), device=DEV), torch.tensor(ph[None], device=DEV), out, 1.5)
    torch.cuda.synchronize()
    ref = O.sh_forward(A, ph.astype(np.float64), go, S0=1.5)
    st = check_intensity(out[0].cpu().numpy(), ref, g.n_lenses, g.subap_npx)
    print(f'grey pupil: total flux rel err {st["total"]:.2e}, normalised image err {st["img"]:.2e}', flush=True)


def test_resampled_flat_wavefront_reproduces_scao_fixtures():
    """GPU image of the flat wavefront: the subapertures of scao_subaps_n8_th0.5_ref.fits and the
    near-zero slopes of scao_sn_n8_th0.5_ref.fits."""
    g, go, A, frames, out, plan = run(64, 0.015625, 0.0, SCAO_TEST, n_frames=1)
    At = torch.tensor(A.astype(np.float32), device=DEV)
    flat = torch.zeros((1, 64, 64), dtype=torch.float32, device=DEV)
    img = torch.empty((1, 64, 64), dtype=torch.float32, device=DEV)
    plan.forward(At, flat, img, 1.0)
    torch.cuda.synchronize()
    img = img[0].cpu().numpy()
    idxs_ref = fits.read(os.path.join(GOLDEN, 'scao_subaps_n8_th0.5_ref.fits'))[1][0].T
    idxs, _ = detect_subaps(img, 8, 0.5)
    np.testing.assert_array_equal(idxs, idxs_ref)
    sc = SlopecPlan(idxs, (64, 64), compute_xy_weights(8), max_frames=1, device=0)
    s = torch.empty((1, 2 * idxs.shape[0]), dtype=torch.float32, device=DEV)
    fl = torch.empty((1, idxs.shape[0]), dtype=torch.float64, device=DEV)
    tt = torch.empty((1, 2), dtype=torch.float64, device=DEV)
    sc.forward(torch.tensor(img[None], device=DEV), s, fl, tt)
    torch.cuda.synchronize()
    sn = fits.read(os.path.join(GOLDEN, 'scao_sn_n8_th0.5_ref.fits'))[1][0]
    check_slopes(s[0].cpu().numpy(), sn, 8)
    print(f'flat-wavefront slopes: max |s| {float(s.abs().max()):.2e} (fixture {np.abs(sn).max():.2e})', flush=True)


def test_sh_object_default_fov_ovs_matches_oracle():
    """Drop-in SH with the default fov_ovs_coeff (0 -> fov_ovs 2): set-up builds the extrapolation
    sources from in_ef.A (sh.py:453) and keeps them; two triggers, the second after an in-place
    change of the pupil (new statistics, set-up extrapolation maps)."""
    from specula_amd.data_objects.electric_field import ElectricField
    from specula_amd.processing_objects.sh import SH
    ef_size, pitch, obs, kw = CASES[1]
    sh = SH(target_device_idx=0, **kw)
    ef = ElectricField(ef_size, ef_size, pitch, S0=2.0, target_device_idx=0)
    A = annular_pupil(ef_size, obs)
    A_setup = A.copy()
    ef.A[:] = torch.tensor(A.astype(np.float32), device=DEV)
    scr = von_karman_opd(256, pitch, seed=4)
    sh.inputs['in_ef'].set(ef)
    sh.setup()
    assert sh.geometry.fov_ovs == 2
    go = O.sh_geometry(ef_size=ef_size, pixel_pitch=pitch, **kw)
    for t in range(2):
        ph = scr[:ef_size, 3 * t:3 * t + ef_size].astype(np.float32)
        ef.phaseInNm[:] = torch.tensor(ph, device=DEV)
        if t == 1:
            ef.A.mul_(0.5)
            A = 0.5 * A
        ef.generation_time = t + 1
        sh.check_ready(t + 1)
        sh.trigger()
        sh.post_trigger()
        torch.cuda.synchronize()
        print(f'SH object frame {t}', flush=True)
        ref = O.sh_forward(A, ph.astype(np.float64), go, S0=2.0, ext_mask=A_setup)
        check_intensity(sh.outputs['out_i'].i.cpu().numpy(), ref, 20, 8)


def test_resampled_plan_vjp_through_the_abi():
    """shwfs_vjp on a resampled, non-Parseval plan (the adjoint itself is checked against finite
    differences in test_gpu_adjoint_general.py): a zero cotangent gives an exactly zero gradient
    over the whole input grid (the scatter target is cleared, including pixels no tap reaches),
    and shwfs_resample refuses a plan without resampling."""
    from specula_amd import _native
    g, go, A, frames, out, plan = run(*CASES[0], n_frames=1)
    At = torch.tensor(A.astype(np.float32), device=DEV)
    ph = torch.tensor(frames[:1], device=DEV)
    gccd = torch.zeros((1, 64, 64), dtype=torch.float32, device=DEV)
    gph = torch.full_like(ph, float('nan'))
    rc = _native.lib().shwfs_vjp(plan._h, 1, _native.ptr(At), 0, _native.ptr(ph), 64 * 64, 1.0, _native.ptr(gccd),
                                 64 * 64, _native.ptr(gph), 64 * 64, None)
    torch.cuda.synchronize()
    assert rc == 0 and bool((gph == 0).all())
    plain = SHPlan(sh_geometry(160, 0.05, **dict(CASES[2][3], rotAnglePhInDeg=0.0, xShiftPhInPixel=0.0,
                                                   yShiftPhInPixel=0.0)), max_frames=1, device=0)
    x = torch.zeros((1, 160, 160), dtype=torch.float32, device=DEV)
    rc = _native.lib().shwfs_resample(plain._h, 1, _native.ptr(x), 160 * 160, _native.ptr(x), 160 * 160, None)
    assert rc == _native.SHWFS_ERR_ARG and b'no field resampling' in _native.lib().shwfs_last_error()


def _sh_object(**kw):
    from specula_amd.processing_objects.sh import SH
    return SH(target_device_idx=0, **kw)


def test_reference_sh_flux_at_default_fov_ovs():
    """test_sh.py:18-42 verbatim (default fov_ovs_coeff -> fov_ovs 2, N 36, cut 10): sum of the
    intensity equals S0 * masked area (FP32 image, FP64 normalisation: rtol 1e-6 instead of the
    FP64 test's decimal 7)."""
    from specula_amd.data_objects.electric_field import ElectricField
    sh = _sh_object(wavelengthInNm=500, subap_wanted_fov=3, sensor_pxscale=0.5, subap_on_diameter=20, subap_npx=6)
    ef = ElectricField(120, 120, 0.05, S0=100, target_device_idx=0)
    ef.generation_time = 1
    sh.inputs['in_ef'].set(ef)
    sh.setup()
    g = sh.geometry
    assert (g.fov_ovs, g.fft_size, g.cutpixels) == (2, 36, 10)
    sh.check_ready(1); sh.trigger(); sh.post_trigger()
    tot = float(sh.outputs['out_i'].i.double().sum())
    print(f'test_sh flux: {tot:.9g} vs {100 * ef.masked_area():.9g}', flush=True)
    np.testing.assert_allclose(tot, 100 * ef.masked_area(), rtol=1e-6)
    go = O.sh_geometry(ef_size=120, pixel_pitch=0.05, wavelengthInNm=500, subap_wanted_fov=3, sensor_pxscale=0.5,
                       subap_on_diameter=20, subap_npx=6)
    ref = O.sh_forward(np.ones((120, 120)), np.zeros((120, 120)), go, S0=100, ext_mask=np.ones((120, 120)))
    check_intensity(sh.outputs['out_i'].i.cpu().numpy(), ref, 20, 6)


def test_reference_pixelscale_at_default_fov_ovs():
    """test_sh.py:44-95 verbatim: a tilt of one CCD pixel shifts the spots by one pixel
    (decimal 3), with the default fov_ovs 2 resampling."""
    from specula_amd.data_objects.electric_field import ElectricField
    pxs, ef_px, pitch, npx = 0.5, 120, 0.05, 6
    sh = _sh_object(wavelengthInNm=500, subap_wanted_fov=npx * pxs, sensor_pxscale=pxs, subap_on_diameter=20,
                    subap_npx=npx)
    ef = ElectricField(ef_px, ef_px, pitch, S0=1, target_device_idx=0)
    ef.generation_time = 1
    sh.inputs['in_ef'].set(ef)
    sh.setup()
    sh.check_ready(1); sh.trigger(); sh.post_trigger()
    flat = sh.outputs['out_i'].i.double().cpu().numpy().copy()
    tilt_value = np.radians(ef_px * pitch * 1 / (60 * 60) * pxs)
    tilt = np.linspace(-tilt_value / 2, tilt_value / 2, ef_px)
    ef.phaseInNm[:] = torch.tensor(np.broadcast_to(tilt, (ef_px, ef_px)) * 1e9, dtype=torch.float32, device=DEV)
    ef.generation_time = 2
    sh.check_ready(2); sh.trigger(); sh.post_trigger()
    tilted = sh.outputs['out_i'].i.double().cpu().numpy().copy()
    shifted = np.roll(flat, (0, 1))
    shifted[:, ::npx] = 0
    tilted[:, ::npx] = 0
    print(f'test_sh pixelscale: max |tilted - shifted| {np.abs(tilted - shifted).max():.3e}', flush=True)
    np.testing.assert_array_almost_equal(tilted, shifted, decimal=3)


def test_resampled_sh_graph_replays_match_eager():
    """Captured SH (fov_ovs 2 resampling inside the graph) replayed over several frames, with eager
    work and synchronisations in between, equals the eager object on every frame."""
    from specula_amd.data_objects.electric_field import ElectricField
    ef_size, pitch, obs, kw = CASES[1]
    ef = ElectricField(ef_size, ef_size, pitch, S0=1.0, target_device_idx=0)
    eager, graph = _sh_object(**kw), _sh_object(**kw)
    for o in (eager, graph):
        o.inputs['in_ef'].set(ef)
        o.setup()
    graph.use_graph = True
    graph.graph_min_nodes = 1                       # keep the graph whatever its size
    ef.A[:] = torch.tensor(annular_pupil(ef_size, obs).astype(np.float32), device=DEV)
    scr = von_karman_opd(256, pitch, seed=6)
    for t in range(1, 5):
        ef.phaseInNm[:] = torch.tensor(scr[:ef_size, 2 * t:2 * t + ef_size].astype(np.float32), device=DEV)
        ef.generation_time = t
        for o in (graph, eager):
            o.check_ready(t); o.trigger(); o.post_trigger()
        torch.cuda.synchronize()
        _ = (ef.phaseInNm * 2).sum().item()
        assert torch.equal(graph.outputs['out_i'].i, eager.outputs['out_i'].i), t
    assert graph.cuda_graph is not None
