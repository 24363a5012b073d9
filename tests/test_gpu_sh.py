"""GPU parity of the HIP SH forward and ShSlopec CoG against the FP64 oracle (calls the C ABI)."""
import numpy as np
import pytest
import torch

from helpers import check_intensity, check_slopes, oracle_geometry
from oracle import sh_oracle as O
from specula_amd import workloads
from specula_amd.lib.geometry import sh_geometry
from specula_amd.lib.phasescreen import annular_pupil, von_karman_opd
from specula_amd.lib.subaps import detect_subaps
from specula_amd.processing_objects.sh import SHPlan
from specula_amd.processing_objects.sh_slopec import SlopecPlan, compute_xy_weights

pytestmark = pytest.mark.gpu
DEV = 'cuda:0'


def gpu_sh(geom, A, frames, S0=1.0, a_shared=True, variant=0):
    plan = SHPlan(geom, max_frames=len(frames), device=0, variant=variant)
    ccd = geom.ccd_side
    At = torch.tensor(np.asarray(A, np.float32), device=DEV)
    if not a_shared:
        At = At.expand(len(frames), *At.shape).contiguous()
    ph = torch.tensor(np.asarray(frames, np.float32), device=DEV)
    out = torch.empty((len(frames), ccd, ccd), dtype=torch.float32, device=DEV)
    plan.forward(At, ph, out, S0, n_frames=len(frames))
    torch.cuda.synchronize()
    return out.cpu().numpy(), plan


def run_case(kw, ef, pitch, obs, seed, fov_res=0.0, n_frames=1, S0=1.0, variant=0):
    print(f'case ef={ef} kw={kw} variant={variant}', flush=True)
    g = sh_geometry(ef, pitch, fov_resolution_arcsec=fov_res, **kw)
    go = O.sh_geometry(ef_size=ef, pixel_pitch=pitch, fov_resolution_arcsec=fov_res or None, **kw)
    A = annular_pupil(ef, obs)
    scr = von_karman_opd(1 << int(np.ceil(np.log2(ef + 8))), pitch, seed=seed)
    frames = np.stack([scr[:ef, k:k + ef] for k in range(n_frames)]).astype(np.float32)
    out, plan = gpu_sh(g, A, frames, S0=S0, variant=variant)
    for k in range(n_frames):
        ref = O.sh_forward(A, frames[k].astype(np.float64), go, S0=S0)
        st = check_intensity(out[k], ref, g.n_lenses, g.subap_npx)
        print(f'  frame {k}: total flux rel err {st["total"]:.2e}, normalised image err {st["img"]:.2e}', flush=True)
    return st, plan


@pytest.mark.parametrize('variant', [1, 2, 3])
def test_scao_ngs_20x20_parity(variant):
    wl = workloads.scao_ngs_20x20()
    st, plan = run_case(wl.sh_kwargs, wl.ef_size, wl.pixel_pitch, wl.obstruction, wl.seed, n_frames=3, S0=7.5,
                        variant=variant)
    assert plan.parseval and plan.M == 16


@pytest.mark.parametrize('kw', [
    dict(wavelengthInNm=500, subap_wanted_fov=1.5, sensor_pxscale=0.25, subap_on_diameter=20, subap_npx=6,
         fov_ovs_coeff=1.0),    # N=42, cut 12, clipping FoV mask, 5:1 rebin
    dict(wavelengthInNm=500, subap_wanted_fov=1.5, sensor_pxscale=0.25, subap_on_diameter=20, subap_npx=4,
         fov_ovs_coeff=1.0),    # N=30, cut 16, 14 -> 4 LCM rebin
    dict(wavelengthInNm=500, subap_wanted_fov=1.2, sensor_pxscale=0.2, subap_on_diameter=20, subap_npx=6,
         fov_ovs_coeff=1.0),    # N=24, cut 10, 14 -> 6 LCM rebin
])
@pytest.mark.parametrize('variant', [1, 2, 3])
def test_cut_mask_lcm_parity(kw, variant):
    st, plan = run_case(kw, 120, 0.05, 0.1, seed=5, variant=variant)
    assert not plan.parseval


@pytest.mark.parametrize('variant', [2, 3])
@pytest.mark.parametrize('npx,px', [(8, 0.25), (16, 0.125)])
def test_np8_cut_fast_store_parity(npx, px, variant):
    """np 8 with a FoV cut (N=40, cut 8): the MFMA kernel's non-Parseval partial sums with the
    row-segment CCD stores of 8 and 16 px subapertures, 3 frames streamed through one block."""
    kw = dict(wavelengthInNm=600, subap_wanted_fov=1.6 if npx == 8 else 2.0, sensor_pxscale=px,
              subap_on_diameter=20, subap_npx=npx, fov_ovs_coeff=1.0)
    st, plan = run_case(kw, 160, 0.05, 0.14, seed=7, n_frames=3, variant=variant)
    assert not plan.parseval and plan.geom.ovs_np_sub == 8 and plan.geom.cutpixels == 8


@pytest.mark.parametrize('variant', [1, 2, 3])
def test_elt_ngs_80x80_parity(variant):
    wl = workloads.elt_ngs_80x80()
    g = wl.geometry()
    assert (g.fft_size, g.cutpixels, g.subap_npx, g.ovs_np_sub) == (64, 0, 16, 8)
    frames = wl.frames(2).astype(np.float32)
    A = wl.pupil()
    out, plan = gpu_sh(g, A, frames, variant=variant)
    go = oracle_geometry(wl)
    for k in range(2):
        ref = O.sh_forward(A, frames[k].astype(np.float64), go)
        st = check_intensity(out[k], ref, g.n_lenses, g.subap_npx)
        print(f'  ELT frame {k} variant {variant}: total flux rel err {st["total"]:.2e}, '
              f'normalised image err {st["img"]:.2e}', flush=True)


def test_batch_equals_single_and_stride():
    wl = workloads.scao_ngs_20x20()
    g = wl.geometry()
    A = wl.pupil()
    frames = wl.frames(4).astype(np.float32)
    batch, _ = gpu_sh(g, A, frames)
    per, _ = gpu_sh(g, A, frames, a_shared=False)
    assert np.array_equal(batch, per)
    v1, _ = gpu_sh(g, A, frames, variant=1)
    np.testing.assert_allclose(v1, batch, rtol=1e-4, atol=1e-5 * batch.max())
    for k in range(4):
        single, _ = gpu_sh(g, A, frames[k:k + 1])
        assert np.array_equal(single[0], batch[k])


@pytest.mark.parametrize('kw', [
    dict(wavelengthInNm=500, subap_wanted_fov=1.5, sensor_pxscale=0.25, subap_on_diameter=20, subap_npx=6,
         fov_ovs_coeff=1.0),                                             # summed normalisation (cut 12)
    dict(wavelengthInNm=500, subap_wanted_fov=2.5, sensor_pxscale=2.0625 / 8, subap_on_diameter=20,
         subap_npx=8, fov_ovs_coeff=1.0),                                # Parseval normalisation
])
def test_flux_known_answer(kw):
    """test_sh.py:18-42: sum of the intensity == S0 * masked area, here at fov_ovs == 1."""
    g = sh_geometry(120, 0.05, **kw)
    A = np.ones((120, 120))
    A[:7, :30] = 0
    out, plan = gpu_sh(g, A, np.zeros((1, 120, 120), np.float32), S0=100)
    np.testing.assert_allclose(out[0].astype(np.float64).sum(), 100 * A.sum() * 0.05 ** 2, rtol=1e-6)


def test_one_pixel_tilt_shifts_one_pixel():
    """test_sh.py:44-95 at fov_ovs == 1: a tilt of one CCD pixel moves the spots one pixel in x."""
    ef, pitch, pxs, npx = 120, 0.05, 0.25, 6
    kw = dict(wavelengthInNm=500, subap_wanted_fov=npx * pxs, sensor_pxscale=pxs, subap_on_diameter=20,
              subap_npx=npx, fov_ovs_coeff=1.0)
    g = sh_geometry(ef, pitch, **kw)
    tilt_value = np.radians(ef * pitch * 1 / (60 * 60) * pxs)
    tilt = np.linspace(-tilt_value / 2, tilt_value / 2, ef)
    ph = np.broadcast_to(tilt, (ef, ef)) * 1e9
    out, _ = gpu_sh(g, np.ones((ef, ef)), np.stack([np.zeros((ef, ef)), ph]).astype(np.float32))
    flat, tilted = out[0].astype(np.float64), out[1].astype(np.float64)
    shifted = np.roll(flat, (0, 1))
    shifted[:, ::npx] = 0
    tilted[:, ::npx] = 0
    np.testing.assert_array_almost_equal(tilted, shifted, decimal=3)


def _slopec_case(wl, pedestal=False, thr=0.0, idx_mode=False, exp_weight=1.0):
    g = wl.geometry()
    A = wl.pupil()
    frames = wl.frames(2).astype(np.float32)
    flat, _ = gpu_sh(g, A, np.zeros((1, wl.ef_size, wl.ef_size), np.float32))
    idxs, dmap = detect_subaps(flat[0], g.n_lenses, 0.5)
    out, _ = gpu_sh(g, A, frames)
    H, W = out.shape[1:]
    if idx_mode:   # a permuted pixel order forces the general index-list addressing
        perm = np.random.default_rng(0).permutation(idxs.shape[1])
        idxs_used = idxs[:, perm]
        weights = compute_xy_weights(g.subap_npx, exp_weight)
        weights = {k: v.ravel()[perm].reshape(v.shape) for k, v in weights.items()}
    else:
        idxs_used = idxs
        weights = compute_xy_weights(g.subap_npx, exp_weight)
    plan = SlopecPlan(idxs_used, (H, W), weights, max_frames=2, device=0)
    assert plan.block_mode == (not idx_mode)
    n_sub = idxs.shape[0]
    pix = torch.tensor(out, device=DEV)
    slopes = torch.empty((2, 2 * n_sub), dtype=torch.float32, device=DEV)
    flux = torch.empty((2, n_sub), dtype=torch.float64, device=DEV)
    totals = torch.empty((2, 2), dtype=torch.float64, device=DEV)
    plan.forward(pix, slopes, flux, totals, thr=thr, pedestal=pedestal, n_frames=2)
    torch.cuda.synchronize()
    for k in range(2):
        ref = O.shslopec(out[k].astype(np.float64), idxs, thr_value=thr if thr > 0 else -1,
                         exp_weight=exp_weight, thr_pedestal=pedestal)
        check_slopes(slopes[k].cpu().numpy(), ref['slopes'], g.subap_npx)
        np.testing.assert_allclose(flux[k].cpu().numpy(), ref['flux'], rtol=1e-12, atol=1e-12 * ref['flux'].max())
        np.testing.assert_allclose(totals[k].cpu().numpy(), [ref['total'], ref['mean']], rtol=1e-12)
    return n_sub


def test_slopec_parity_elt():
    n_sub = _slopec_case(workloads.elt_ngs_80x80())
    assert 4000 < n_sub < 5200


@pytest.mark.parametrize('opts', [dict(idx_mode=True), dict(thr=2e-7), dict(thr=2e-7, pedestal=True),
                                  dict(exp_weight=1.5)])
def test_slopec_modes(opts):
    _slopec_case(workloads.scao_ngs_20x20(), **opts)


@pytest.mark.parametrize('fpb', [1, 2, 4, 8])
def test_streaming_frames_per_block_is_exact(fpb):
    wl = workloads.scao_ngs_20x20()
    g = wl.geometry()
    A = torch.tensor(wl.pupil().astype(np.float32), device=DEV)
    ph = torch.tensor(wl.frames(7).astype(np.float32), device=DEV)
    ref = torch.empty((7, g.ccd_side, g.ccd_side), dtype=torch.float32, device=DEV)
    plan = SHPlan(g, max_frames=7, device=0, variant=2)
    plan.set_frames_per_block(1)
    plan.forward(A, ph, ref, 1.0, n_frames=7)
    out = torch.empty_like(ref)
    plan.set_frames_per_block(fpb)
    plan.forward(A, ph, out, 1.0, n_frames=7)
    torch.cuda.synchronize()
    assert torch.equal(out, ref)



def test_per_frame_pupils_with_dark_subapertures():
    """One pupil per frame (a_frame_stride = ef^2): the darkness of each wave is decided per frame
    from that frame's pupil (whole waves and single subapertures go dark in frame 1, all of them in
    frame 2), and each frame has its own flux normalisation."""
    wl = workloads.scao_ngs_20x20()
    g = wl.geometry()
    go = oracle_geometry(wl)
    A = wl.pupil().astype(np.float32)
    A1 = A.copy()
    A1[: g.ef_size // 2] = 0.0                                    # top half dark
    A1[:, 5 * 8: 6 * 8 + 3] = 0.0                                 # a column of subapertures (and a part)
    A2 = np.zeros_like(A)
    pupils = np.stack([A, A1, A2])
    frames = wl.frames(3).astype(np.float32)
    plan = SHPlan(g, max_frames=3, device=0)
    out = torch.empty((3, g.ccd_side, g.ccd_side), dtype=torch.float32, device=DEV)
    plan.forward(torch.tensor(pupils, device=DEV), torch.tensor(frames, device=DEV), out, 1.0, n_frames=3)
    torch.cuda.synchronize()
    img = out.cpu().numpy()
    for k in range(2):
        ref = O.sh_forward(pupils[k].astype(np.float64), frames[k].astype(np.float64), go)
        check_intensity(img[k], ref, g.n_lenses, g.subap_npx)
    lit = (pupils[1].reshape(20, 8, 20, 8) != 0).any(3).any(1)
    tiles = img[1].reshape(20, g.subap_npx, 20, g.subap_npx)
    assert np.all(tiles.transpose(0, 2, 1, 3)[~lit] == 0)
    # no light at all: the reference's normalisation is 0/0, NaN everywhere (the oracle says so too)
    ref2 = O.sh_forward(pupils[2].astype(np.float64), frames[2].astype(np.float64), go)
    assert np.all(np.isnan(ref2)) and np.all(np.isnan(img[2]))

def test_static_pupil_registration():
    """shwfs_plan_set_pupil: registered statistics give the same image as per-call statistics;
    a different amplitude (other storage or re-registered contents) is not served from it."""
    wl = workloads.scao_ngs_20x20()
    g = wl.geometry()
    A = torch.tensor(wl.pupil().astype(np.float32), device=DEV)
    ph = torch.tensor(wl.frames(2).astype(np.float32), device=DEV)
    plan = SHPlan(g, max_frames=2, device=0)
    ref = torch.empty((2, g.ccd_side, g.ccd_side), dtype=torch.float32, device=DEV)
    plan.forward(A, ph, ref, 2.0, n_frames=2)
    plan.set_static_pupil(A, 2.0)
    out = torch.empty_like(ref)
    plan.forward(A, ph, out, 2.0, n_frames=2)
    torch.cuda.synchronize()
    assert torch.equal(out, ref)
    A2 = A.clone()
    A2[: g.ef_size // 2] *= 0.5                      # other storage: recomputed per call
    ref2 = torch.empty_like(ref)
    plan.forward(A2, ph, ref2, 2.0, n_frames=2)
    assert plan.static_pupil_is(A, 2.0, 0)          # another amplitude leaves the registration
    plan.forward(A, ph, out, 2.0, n_frames=2)      # back to A: served from the registration
    torch.cuda.synchronize()
    assert torch.equal(out, ref) and not torch.equal(ref2, ref)
    assert abs(float(ref2.sum()) / float(ref.sum()) - float(A2.sum()) / float(A.sum())) < 1e-5
    plan.set_static_pupil(A2, 2.0)
    plan.forward(A2, ph, out, 2.0, n_frames=2)
    torch.cuda.synchronize()
    assert torch.equal(out, ref2)



def test_registration_survives_other_calls_and_graph_replays():
    """A forward with another S0 or amplitude computes its statistics in buffers of its own: the
    registered pupil stays registered (library query), and a graph captured on the registered
    pupil replays the same image after such calls."""
    wl = workloads.elt_ngs_80x80()
    g = wl.geometry()
    A = torch.tensor(wl.pupil().astype(np.float32), device=DEV)
    ph = torch.tensor(wl.frames(2).astype(np.float32), device=DEV)
    plan = SHPlan(g, max_frames=2, device=0)
    plan.set_static_pupil(A, 1.0)
    ref = torch.empty((2, g.ccd_side, g.ccd_side), dtype=torch.float32, device=DEV)
    plan.forward(A, ph, ref, 1.0, n_frames=2)
    other = torch.empty_like(ref)
    plan.forward(A, ph, other, 3.0, n_frames=2)                      # same A, other S0
    A_half = A * 0.5
    plan.forward(A_half, ph, other, 1.0, n_frames=2)                  # other amplitude
    assert plan.static_pupil_is(A, 1.0, 0) and not plan.static_pupil_is(A, 3.0, 0)
    out = torch.empty_like(ref)
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        plan.forward(A, ph, out, 1.0, n_frames=2, stream=s)
    s.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph, stream=s):
        plan.forward(A, ph, out, 1.0, n_frames=2, stream=s)
    for k in range(3):
        out.zero_()
        plan.forward(A_half, ph, other, 5.0 + k, n_frames=2)          # eager calls between replays
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        assert torch.equal(out, ref), k
    torch.testing.assert_close(other.sum(), ref.sum() * 7.0 * 0.5, rtol=1e-5, atol=0)


@pytest.mark.parametrize('np_sub', [5, 6, 7])
def test_static_pupil_order_other_wave_widths(np_sub):
    """MFMA path with 12, 10 and 9 subapertures per wave (np 5, 6, 7; Parseval geometry, generic
    stores): the lit-first order of a registered pupil gives the unregistered image bit for bit,
    and both match the oracle."""
    ef, pitch = 20 * np_sub, 0.05
    turb = 500e-9 / (np_sub * pitch) * O.RAD2ASEC
    kw = dict(wavelengthInNm=500, subap_wanted_fov=np_sub * turb, sensor_pxscale=turb, subap_on_diameter=20,
              subap_npx=np_sub, fov_ovs_coeff=1.0)
    g = sh_geometry(ef, pitch, fov_resolution_arcsec=turb / 4, **kw)
    go = O.sh_geometry(ef_size=ef, pixel_pitch=pitch, fov_resolution_arcsec=turb / 4, **kw)
    A = annular_pupil(ef, 0.3)
    scr = von_karman_opd(256, pitch, seed=9)
    frames = np.stack([scr[:ef, k:k + ef] for k in range(3)]).astype(np.float32)
    plan = SHPlan(g, max_frames=3, device=0)
    assert plan.parseval
    At = torch.tensor(A.astype(np.float32), device=DEV)
    ph = torch.tensor(frames, device=DEV)
    ref = torch.empty((3, g.ccd_side, g.ccd_side), dtype=torch.float32, device=DEV)
    plan.forward(At, ph, ref, 1.0, n_frames=3)
    plan.set_static_pupil(At, 1.0)
    out = torch.empty_like(ref)
    plan.forward(At, ph, out, 1.0, n_frames=3)
    torch.cuda.synchronize()
    assert torch.equal(out, ref)
    for k in range(3):
        check_intensity(out[k].cpu().numpy(), O.sh_forward(A, frames[k].astype(np.float64), go), g.n_lenses,
                        g.subap_npx)

def test_sh_object_follows_in_place_pupil_change():
    """The drop-in SH keys its pupil statistics on the tensor's storage and in-place version."""
    from specula_amd.data_objects.electric_field import ElectricField
    from specula_amd.processing_objects.sh import SH
    wl = workloads.scao_ngs_20x20()
    sh = SH(target_device_idx=0, **wl.sh_kwargs)
    ef = ElectricField(wl.ef_size, wl.ef_size, wl.pixel_pitch, S0=1.0, target_device_idx=0)
    ef.A[:] = torch.tensor(wl.pupil().astype(np.float32), device=DEV)
    ef.phaseInNm[:] = torch.tensor(wl.frames(1)[0].astype(np.float32), device=DEV)
    sh.inputs['in_ef'].set(ef)
    sh.setup()
    totals = []
    for t, scale in enumerate((1.0, 0.5, 0.5)):
        if scale != 1.0 and t == 1:
            ef.A.mul_(scale)                         # in place: same storage, new version
        ef.generation_time = t + 1
        sh.check_ready(t + 1)
        sh.trigger()
        sh.post_trigger()
        torch.cuda.synchronize()
        totals.append(float(sh.outputs['out_i'].i.sum()))
    masked = float(torch.tensor(wl.pupil()).sum()) * wl.pixel_pitch ** 2
    np.testing.assert_allclose(totals, [masked, 0.5 * masked, 0.5 * masked], rtol=1e-5)


def test_frame_count_outside_plan_is_refused_before_any_launch():
    """n_frames 0 and above max_frames: the C ABI returns SHWFS_ERR_ARG with its message (no
    launch), and the output buffer keeps its contents; the plan still works afterwards."""
    import ctypes
    from specula_amd import _native
    wl = workloads.scao_ngs_20x20()
    g = wl.geometry()
    plan = SHPlan(g, max_frames=2, device=0)
    ef, ccd = g.ef_size, g.ccd_side
    A = torch.tensor(wl.pupil().astype(np.float32), device=DEV)
    ph = torch.tensor(wl.frames(3).astype(np.float32), device=DEV)
    out = torch.full((3, ccd, ccd), -7.0, device=DEV)
    lib = _native.lib()
    for n in (0, 3, -1):
        rc = lib.shwfs_forward(plan._h, n, _native.ptr(A), 0, _native.ptr(ph), ef * ef, 1.0, _native.ptr(out),
                               ccd * ccd, _native.stream_handle(None))
        assert rc == _native.SHWFS_ERR_ARG and b'outside [1, max_frames]' in lib.shwfs_last_error()
    torch.cuda.synchronize()
    assert bool((out == -7.0).all())
    plan.forward(A, ph[:2], out[:2], 1.0, n_frames=2)
    torch.cuda.synchronize()
    assert bool(torch.isfinite(out[:2]).all()) and bool((out[2] == -7.0).all())
