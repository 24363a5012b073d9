"""SH forward with the ShSlopec centre of gravity fused into the kernel epilogue
(shwfs_forward_slopes) against the two-pass path (shwfs_forward + shslopec_forward) and the
FP64 oracle."""
import numpy as np
import pytest
import torch

from helpers import check_slopes
from oracle import sh_oracle as O
from specula_amd import workloads
from specula_amd.lib.subaps import detect_subaps
from specula_amd.processing_objects.sh import SHPlan
from specula_amd.processing_objects.sh_slopec import SlopecPlan, compute_xy_weights
from specula_amd.sensor import SHSensor

pytestmark = pytest.mark.gpu
DEV = 'cuda:0'


def _setup(wl, F, variant=0):
    g = wl.geometry()
    A = torch.tensor(wl.pupil().astype(np.float32), device=DEV)
    plan = SHPlan(g, max_frames=F, device=0, variant=variant)
    flat = torch.zeros((1, g.ef_size, g.ef_size), dtype=torch.float32, device=DEV)
    img = torch.empty((1, g.ccd_side, g.ccd_side), dtype=torch.float32, device=DEV)
    plan.forward(A, flat, img, 1.0)
    idxs, _ = detect_subaps(img[0].cpu().numpy(), g.n_lenses, 0.5)
    sc = SlopecPlan(idxs, (g.ccd_side, g.ccd_side), compute_xy_weights(g.subap_npx), max_frames=F, device=0)
    return g, A, plan, sc, idxs


def _buffers(g, n_sub, F):
    return (torch.empty((F, g.ccd_side, g.ccd_side), dtype=torch.float32, device=DEV),
            torch.empty((F, 2 * n_sub), dtype=torch.float32, device=DEV),
            torch.empty((F, n_sub), dtype=torch.float64, device=DEV),
            torch.empty((F, 2), dtype=torch.float64, device=DEV))


@pytest.mark.parametrize('variant', [0, 3])
@pytest.mark.parametrize('name', ['elt_ngs_80x80', 'scao_ngs_20x20'])
@pytest.mark.parametrize('thr_rel,pedestal', [(0.0, False), (0.3, False), (0.3, True)])
def test_fused_equals_two_pass(name, thr_rel, pedestal, variant):
    wl = getattr(workloads, name)()
    F = 5                              # with 4 frames per block: one full and one partial group
    g, A, plan, sc, idxs = _setup(wl, F, variant)
    n_sub = idxs.shape[0]
    print(f'fused case {name} thr_rel={thr_rel} pedestal={pedestal} variant={variant} n_sub={n_sub}', flush=True)
    ph = torch.tensor(wl.frames(F).astype(np.float32), device=DEV)
    ccd2, s2, f2, t2 = _buffers(g, n_sub, F)
    plan.forward(A, ph, ccd2, 1.0, n_frames=F)
    thr = float(thr_rel * ccd2.mean()) if thr_rel > 0 else 0.0
    sc.forward(ccd2, s2, f2, t2, thr=thr, pedestal=pedestal, n_frames=F)
    assert plan.bind_slopec(sc)
    ccd1, s1, f1, t1 = _buffers(g, n_sub, F)
    plan.forward_slopes(A, ph, ccd1, 1.0, s1, f1, t1, thr=thr, pedestal=pedestal, n_frames=F)
    torch.cuda.synchronize()
    assert torch.equal(ccd1, ccd2)                       # same kernel arithmetic for the image
    # moments: FP32 partials of 4 x 8 pixels per lane (fused) vs FP64 per pixel (two-pass)
    torch.testing.assert_close(s1, s2, rtol=0, atol=2e-6)
    # flux: the fused path takes the exact Parseval value (fs norm sum(A^2) of the subaperture), the
    # two-pass path sums the image pixels: they differ by the image's own flux error (FP32 MFMA
    # ~1e-7, bf16-split ~1e-5 relative); both are checked against the oracle below
    ftol = 2e-6 if variant != 3 else 2e-5
    torch.testing.assert_close(f1, f2, rtol=ftol, atol=0)
    torch.testing.assert_close(t1, t2, rtol=ftol, atol=0)
    out = ccd1[0].double().cpu().numpy()
    ref = O.shslopec(out, idxs, thr_value=thr if thr > 0 else -1, thr_pedestal=pedestal)
    check_slopes(s1[0].cpu().numpy(), ref['slopes'], g.subap_npx)
    np.testing.assert_allclose(f1[0].cpu().numpy(), ref['flux'], rtol=1e-5, atol=1e-7 * ref['flux'].max())


def test_fused_frames_per_block_and_repeat():
    """Streaming depth does not change the fused result; a second call overwrites the first."""
    wl = workloads.scao_ngs_20x20()
    F = 7
    g, A, plan, sc, idxs = _setup(wl, F)
    n_sub = idxs.shape[0]
    assert plan.bind_slopec(sc)
    ph = torch.tensor(wl.frames(F).astype(np.float32), device=DEV)
    ref = _buffers(g, n_sub, F)
    plan.set_frames_per_block(1)
    plan.forward_slopes(A, ph, ref[0], 1.0, *ref[1:], n_frames=F)
    for fpb in (2, 3, 8, 16):
        out = _buffers(g, n_sub, F)
        plan.set_frames_per_block(fpb)
        plan.forward_slopes(A, ph, out[0], 1.0, *out[1:], n_frames=F)
        plan.forward_slopes(A, ph, out[0], 1.0, *out[1:], n_frames=F)
        torch.cuda.synchronize()
        for a, b in zip(out, ref):
            assert torch.equal(a, b), fpb


def test_fused_bind_rejects_unsupported_pairs():
    wl = workloads.scao_ngs_20x20()
    g, A, plan, sc, idxs = _setup(wl, 1)
    # index-list addressing (permuted pixel order) cannot be fused
    perm = np.random.default_rng(0).permutation(idxs.shape[1])
    w = {k: v.ravel()[perm].reshape(v.shape) for k, v in compute_xy_weights(g.subap_npx).items()}
    sc_idx = SlopecPlan(idxs[:, perm], (g.ccd_side, g.ccd_side), w, max_frames=1, device=0)
    assert not sc_idx.block_mode
    assert plan.bind_slopec(sc_idx) is False
    # register-GEMV plan
    plan_v1 = SHPlan(g, max_frames=1, device=0, variant=1)
    assert plan_v1.bind_slopec(sc) is False
    assert plan.bind_slopec(sc) is True
    assert plan.bind_slopec(None) is False
    with pytest.raises(RuntimeError):
        s = torch.empty((1, 2 * idxs.shape[0]), dtype=torch.float32, device=DEV)
        plan.forward_slopes(A, A[None].contiguous(), torch.empty((1, g.ccd_side, g.ccd_side), device=DEV), 1.0,
                            s, torch.empty((1, idxs.shape[0]), dtype=torch.float64, device=DEV),
                            torch.empty((1, 2), dtype=torch.float64, device=DEV))


def test_sensor_uses_fused_path():
    wl = workloads.elt_ngs_80x80()
    sen = SHSensor(wl.geometry(), wl.pupil(), max_frames=3, device=0)
    assert sen.fused
    ph = torch.tensor(wl.frames(3).astype(np.float32), device=DEV)
    s = sen.forward(ph).clone()
    ccd, s2, f2, t2 = _buffers(sen.geom, sen.n_sub, 3)
    sen.plan.forward(sen.A, ph, ccd, 1.0, n_frames=3)
    sen.slopec.forward(ccd, s2, f2, t2, n_frames=3)
    torch.cuda.synchronize()
    torch.testing.assert_close(s, s2, rtol=0, atol=2e-6)


@pytest.mark.parametrize('variant', [0, 3])
def test_static_pupil_order_gives_identical_results(variant):
    """A registered static pupil regroups the subapertures lit-first (waves of dark subapertures
    only store zeros): the image is bit-identical and the slopes, fluxes and totals equal the
    unregistered run's (ELT pupil: 1600 of the 6400 subapertures are dark)."""
    wl = workloads.elt_ngs_80x80()
    F = 5
    g, A, plan, sc, idxs = _setup(wl, F, variant)
    n_sub = idxs.shape[0]
    ph = torch.tensor(wl.frames(F).astype(np.float32), device=DEV)
    assert plan.bind_slopec(sc)
    ref = _buffers(g, n_sub, F)
    plan.forward_slopes(A, ph, ref[0], 1.0, *ref[1:], n_frames=F)
    plan.set_static_pupil(A, 1.0)
    out = _buffers(g, n_sub, F)
    plan.forward_slopes(A, ph, out[0], 1.0, *out[1:], n_frames=F)
    img = torch.empty_like(ref[0])
    plan.forward(A, ph, img, 1.0, n_frames=F)
    torch.cuda.synchronize()
    assert torch.equal(out[0], ref[0]) and torch.equal(img, ref[0])
    assert torch.equal(out[1], ref[1]) and torch.equal(out[2], ref[2])
    torch.testing.assert_close(out[3], ref[3], rtol=1e-14, atol=0)   # per-wave partials regrouped
    lit = (A.reshape(80, 8, 80, 8) != 0).any(3).any(1).cpu().numpy()
    dark_img = ref[0].reshape(F, 80, 16, 80, 16).abs().amax(dim=(2, 4)).cpu().numpy()
    assert lit.sum() == 4800 and np.all(dark_img[:, ~lit] == 0)


# --------------------------------------------------------------------------- LGS plans
def _lgs_lag_spectra(kernels, np_sub):
    N = kernels.shape[-1]
    lags = np.arange(-(np_sub - 1), np_sub) % N
    return torch.tensor(np.ascontiguousarray(kernels[:, lags][:, :, lags].astype(np.complex64)), device=DEV)


def _lgs_small(ef, npx, fov, F):
    """8 x 8 lenslets of ef / 8 pupil samples, npx px per subaperture over a FoV of `fov`
    diffraction widths that the FFT grid covers exactly (no cut: Parseval normalisation), random
    normalised elongation kernels; returns the geometry, the plan (kernels set), pupil and F phase
    frames."""
    from specula_amd.lib.geometry import sh_geometry
    from specula_amd.lib.phasescreen import annular_pupil, von_karman_opd
    np_sub, pitch = ef // 8, 0.05
    turb = 589e-9 / (np_sub * pitch) * O.RAD2ASEC
    kw = dict(wavelengthInNm=589, subap_wanted_fov=fov * turb, sensor_pxscale=fov * turb / npx,
              subap_on_diameter=8, subap_npx=npx, fov_ovs_coeff=1.0)
    g = sh_geometry(ef, pitch, fov_resolution_arcsec=turb / 4, **kw)
    assert (g.ovs_np_sub, g.subap_npx, g.cutpixels) == (np_sub, npx, 0)
    rng = np.random.default_rng(ef + npx)
    K = O.process_kernels(rng.random((64, g.fft_size, g.fft_size)) ** 6, 8, 8)
    plan = SHPlan(g, lgs=True, max_frames=F, device=0, kernels_normalised=True)
    plan.set_kernels(_lgs_lag_spectra(K, np_sub))
    A = torch.tensor(annular_pupil(ef, 0.2).astype(np.float32), device=DEV)
    scr = von_karman_opd(2 * ef, pitch, seed=9)
    ph = torch.tensor(np.stack([scr[:ef, k:k + ef] for k in range(F)]).astype(np.float32), device=DEV)
    return g, plan, A, ph


def _lgs_slopec(g, plan, A, F):
    flat = torch.zeros((1, g.ef_size, g.ef_size), dtype=torch.float32, device=DEV)
    img = torch.empty((1, g.ccd_side, g.ccd_side), dtype=torch.float32, device=DEV)
    plan.forward(A, flat, img, 1.0)
    idxs, _ = detect_subaps(img[0].cpu().numpy(), g.n_lenses, 0.5)
    return SlopecPlan(idxs, (g.ccd_side, g.ccd_side), compute_xy_weights(g.subap_npx), max_frames=F, device=0), idxs


def _fused_vs_two_pass(g, plan, sc, idxs, A, ph, F, thr_rel, pedestal):
    n_sub = idxs.shape[0]
    ccd2, s2, f2, t2 = _buffers(g, n_sub, F)
    plan.forward(A, ph, ccd2, 1.0, n_frames=F)
    thr = float(thr_rel * ccd2.mean()) if thr_rel > 0 else 0.0
    sc.forward(ccd2, s2, f2, t2, thr=thr, pedestal=pedestal, n_frames=F)
    assert plan.bind_slopec(sc)
    ccd1, s1, f1, t1 = _buffers(g, n_sub, F)
    plan.forward_slopes(A, ph, ccd1, 1.0, s1, f1, t1, thr=thr, pedestal=pedestal, n_frames=F)
    torch.cuda.synchronize()
    assert torch.equal(ccd1, ccd2)                       # the image does not depend on the epilogue
    # FP32 partials of 32 pixels and FP64 across lanes (fused) vs FP64 per pixel (two-pass); flux:
    # the Parseval value vs the sum of the FP32 pixels
    torch.testing.assert_close(s1, s2, rtol=0, atol=2e-6)
    torch.testing.assert_close(f1, f2, rtol=2e-6, atol=1e-12 * float(f2.abs().max()))
    torch.testing.assert_close(t1, t2, rtol=2e-6, atol=0)
    for f in range(F):
        out = ccd1[f].double().cpu().numpy()
        ref = O.shslopec(out, idxs, thr_value=thr if thr > 0 else -1, thr_pedestal=pedestal)
        check_slopes(s1[f].cpu().numpy(), ref['slopes'], g.subap_npx)
        np.testing.assert_allclose(f1[f].cpu().numpy(), ref['flux'], rtol=1e-5, atol=1e-7 * ref['flux'].max())
    return s1, f1


@pytest.mark.parametrize('thr_rel,pedestal', [(0.0, False), (0.3, False), (0.3, True)])
@pytest.mark.parametrize('ef,npx,fov', [(128, 16, 16), (128, 8, 16), (96, 12, 12), (64, 8, 8), (48, 6, 6)])
def test_lgs_fused_equals_two_pass(ef, npx, fov, thr_rel, pedestal):
    """Every LGS kernel shape (np 16 / 12 with M 32 and 8 subapertures per block, np 8 / 6 with
    M 16 and 16 per block: two or four subapertures per wave in the binning phase), npx 16 and
    below (pixels outside the subaperture in the lanes' 16 x 16 block), the three threshold modes;
    3 frames, the first two lenslet rows dark in the third (whole dark tiles: their valid
    subapertures get zero moments, flux and slopes, as the two-pass CoG of their zero image)."""
    F = 3
    g, plan, A, ph = _lgs_small(ef, npx, fov, F)
    sc, idxs = _lgs_slopec(g, plan, A, F)
    print(f'lgs fused case ef={ef} np={g.ovs_np_sub} npx={npx} thr_rel={thr_rel} pedestal={pedestal} '
          f'n_sub={idxs.shape[0]}', flush=True)
    A3 = A.expand(F, -1, -1).clone()
    A3[2, :2 * g.ovs_np_sub, :] = 0.0
    s1, f1 = _fused_vs_two_pass(g, plan, sc, idxs, A3, ph, F, thr_rel, pedestal)
    rows01 = (idxs[:, 0] // g.ccd_side) < 2 * g.subap_npx
    assert rows01.any() and not rows01.all()
    assert torch.all(f1[2, torch.tensor(rows01, device=DEV)] == 0)
    sel = torch.tensor(np.concatenate([rows01, rows01]), device=DEV)
    assert torch.all(s1[2, sel] == 0) and torch.any(s1[1, sel] != 0)


def test_lgs_40x40_fused_static_pupil():
    """The lgs_40x40 workload (np 16, N 128, 16 px, launcher 20 m off axis) with its pupil
    registered (lit-first tiles, 44 of 200 dark): fused equals two-pass, and the slopes of one
    frame match the FP64 oracle chain's (image and CoG) to 1e-3 CCD px."""
    from helpers import oracle_geometry
    wl = workloads.lgs_40x40()
    g = wl.geometry()
    F = 4
    kern = workloads.lgs_kernel(g)
    plan = SHPlan(g, lgs=True, max_frames=F, device=0, kernels_normalised=kern.kernels_normalised())
    plan.set_kernels(torch.as_tensor(kern.lag_spectra(g.ovs_np_sub), device=DEV).to(torch.complex64).contiguous())
    A = torch.tensor(wl.pupil().astype(np.float32), device=DEV)
    sc, idxs = _lgs_slopec(g, plan, A, F)
    plan.set_static_pupil(A, 1.0)
    frames = wl.frames(F).astype(np.float32)
    ph = torch.tensor(frames, device=DEV)
    s1, _ = _fused_vs_two_pass(g, plan, sc, idxs, A, ph, F, 0.0, False)
    ref = O.sh_forward(wl.pupil(), frames[0].astype(np.float64), oracle_geometry(wl), kernels=kern.kernels)
    err = check_slopes(s1[0].cpu().numpy(), O.shslopec(ref, idxs)['slopes'], g.subap_npx)
    print(f'lgs_40x40 fused slopes vs FP64 chain: {err / (2 / (g.subap_npx - 1)):.2e} px', flush=True)


def test_lgs_sensor_uses_fused_path():
    wl = workloads.lgs_40x40()
    g = wl.geometry()
    sen = SHSensor(g, wl.pupil(), max_frames=2, device=0, kernel=workloads.lgs_kernel(g))
    assert sen.fused
    ph = torch.tensor(wl.frames(2).astype(np.float32), device=DEV)
    s = sen.forward(ph).clone()
    s2 = sen.forward(ph, two_pass=True).clone()
    torch.cuda.synchronize()
    torch.testing.assert_close(s, s2, rtol=0, atol=2e-6)
