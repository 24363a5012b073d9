"""The any-size SH path (sh_general_kernel, plan variant 4, taken automatically when no
specialised kernel fits): parity with the FP64 oracle at the subaperture shapes of the
reference's own SH set-ups, each with fewer lenslets than the full sensor (same pupil pixels per
lenslet, pitch, wavelength, FoV and pixel scale, so the same np, FFT size, cut, CCD pixels and
field oversampling per subaperture):

  params_eris.yml:102-113         sh_lgs   LGS  np 20,  N 120, M 64,   cut 70,  6 px,  fov_ovs 5
  params_eris.yml:148-157         sh_lo    NGS  np 80,  N 160, M 256,  cut 28,  12 px, fov_ovs 2
  params_morfeo_ltao.yml:127-138  sh1      LGS  np 119, N 238, M 256,  cut 90,  14 px, fov_ovs 17
  params_morfeo_full.yml:259-268  sh_ngs1  NGS  np 240, N 480, M 512,  cut 276, 120 px
  params_morfeo_full.yml:306-315  sh_ref1  NGS  np 300, N 600, M 1024, cut 314, 30 px, fov_ovs 5

(the YAML values are restated below; the full MORFEO LGS sensor has 68 lenslets and np 121), plus
the reference's known answer of test/test_sh_slopec.py:21-114 (np 120, N 360, 120 px, Gaussian
LLT) on the HIP path, and equality with the specialised kernels where both apply.
"""
import numpy as np
import pytest
import torch

from helpers import check_intensity, check_slopes
from oracle import sh_oracle as O
from specula_amd.data_objects.convolution_kernel import ConvolutionKernel, GaussianConvolutionKernel
from specula_amd.lib.geometry import sh_geometry
from specula_amd.lib.phasescreen import annular_pupil, von_karman_opd
from specula_amd.lib.subaps import detect_subaps
from specula_amd.processing_objects.sh import SHPlan
from specula_amd.processing_objects.sh_slopec import SlopecPlan, compute_xy_weights
from specula_amd import workloads, RAD2ASEC

pytestmark = pytest.mark.gpu
DEV = 'cuda:0'

# (name, pupil px per lenslet, pixel pitch [m], lenslets in the test, LGS, SH keywords)
CASES = [
    ('eris_sh_lgs', 4, 0.05, 4, True,
     dict(subap_wanted_fov=5, sensor_pxscale=0.85, subap_npx=6, wavelengthInNm=589)),
    ('eris_sh_lo', 40, 0.05, 2, False,
     dict(subap_wanted_fov=4.8, sensor_pxscale=0.4, subap_npx=12, wavelengthInNm=700)),
    ('morfeo_ltao_sh1', 7, 0.0802, 3, True,
     dict(subap_wanted_fov=16.1, sensor_pxscale=1.15, subap_npx=14, wavelengthInNm=589, fov_ovs_coeff=1.6)),
    ('morfeo_sh_ngs1', 240, 0.0802, 2, False,
     dict(subap_wanted_fov=1.8, sensor_pxscale=0.015, subap_npx=120, wavelengthInNm=1650)),
    ('morfeo_sh_ref1', 60, 0.0802, 2, False,
     dict(subap_wanted_fov=4.6, sensor_pxscale=0.153, subap_npx=30, wavelengthInNm=750)),
]
SHAPES = {'eris_sh_lgs': (20, 120, 70, 6), 'eris_sh_lo': (80, 160, 28, 12), 'morfeo_ltao_sh1': (119, 238, 90, 14),
          'morfeo_sh_ngs1': (240, 480, 276, 120), 'morfeo_sh_ref1': (300, 600, 314, 30)}


def lgs_kernels(g, launcher=(6.0, -3.0, 0.0)):
    """Sodium-profile kernels (11 layers, FWHM 10 km at 90 km) seen from an off-axis launcher."""
    k = ConvolutionKernel(g.n_lenses, g.n_lenses, g.fp4_pixel_pitch * RAD2ASEC, g.pixel_pitch * g.ef_size,
                          g.fft_size, launcher_pos=list(launcher), launcher_size=1.0, target_device_idx=-1,
                          precision=1)
    k.prepare_for_sh(sodium_altitude=workloads.LGS_LAYERS_M.tolist(),
                     sodium_intensity=workloads.sodium_profile().tolist())
    return k


def run_plan(g, A, frames, lgs=False, kern=None, variant=0):
    plan = SHPlan(g, lgs=lgs, max_frames=len(frames), device=0, variant=variant,
                  kernels_normalised=True if kern is None else kern.kernels_normalised())
    if g.needs_interpolation:
        plan.set_resampling(A)
    if lgs:
        plan.set_kernels(torch.tensor(kern.lag_spectra(g.ovs_np_sub), device=DEV))
    ef_in = A.shape[0]
    out = torch.empty((len(frames), g.ccd_side, g.ccd_side), dtype=torch.float32, device=DEV)
    plan.forward(torch.tensor(A.astype(np.float32), device=DEV),
                 torch.tensor(np.asarray(frames, np.float32).reshape(len(frames), ef_in, ef_in), device=DEV),
                 out, 1.0, n_frames=len(frames))
    torch.cuda.synchronize()
    return out.cpu().numpy(), plan


@pytest.mark.parametrize('case', CASES, ids=[c[0] for c in CASES])
def test_general_path_matches_oracle_at_reference_geometries(case):
    name, ppl, pitch, n, lgs, kw = case
    ef = ppl * n
    kw = dict(kw, subap_on_diameter=n)
    g = sh_geometry(ef, pitch, **kw)
    go = O.sh_geometry(ef_size=ef, pixel_pitch=pitch, **kw)
    assert (g.ovs_np_sub, g.fft_size, g.cutpixels, g.subap_npx) == SHAPES[name]
    print(f'{name}: np {g.ovs_np_sub} N {g.fft_size} cut {g.cutpixels} npx {g.subap_npx} n {n} '
          f'fov_ovs {g.fov_ovs} lgs {lgs}', flush=True)
    A = annular_pupil(ef, 0.15 if n > 2 else 0.0)
    scr = von_karman_opd(max(64, 1 << int(np.ceil(np.log2(ef + 4)))), pitch, seed=21)
    frames = np.stack([scr[:ef, k:k + ef] * 0.3 for k in range(2)]).astype(np.float32)
    kern = lgs_kernels(g) if lgs else None
    out, plan = run_plan(g, A, frames, lgs, kern)
    ref0 = None
    for k in range(2):
        ref = O.sh_forward(A, frames[k].astype(np.float64), go, kernels=None if kern is None else kern.kernels)
        st = check_intensity(out[k], ref, n, g.subap_npx)
        print(f'  frame {k}: flux {st["total"]:.1e}, normalised image {st["img"]:.1e}', flush=True)
        ref0 = ref if ref0 is None else ref0
    # slopes of the GPU images against the oracle chain on the FP64 image
    idxs, _ = detect_subaps(O.sh_forward(A, np.zeros_like(A), go, kernels=None if kern is None else kern.kernels),
                            n, 0.5)
    sp = SlopecPlan(idxs, (g.ccd_side, g.ccd_side), compute_xy_weights(g.subap_npx), max_frames=1, device=0)
    s = torch.empty(2 * idxs.shape[0], dtype=torch.float32, device=DEV)
    fl = torch.empty(idxs.shape[0], dtype=torch.float64, device=DEV)
    tt = torch.empty(2, dtype=torch.float64, device=DEV)
    sp.forward(torch.tensor(out[0], device=DEV), s, fl, tt)
    err = check_slopes(s.cpu().numpy(), O.shslopec(ref0, idxs)['slopes'], g.subap_npx)
    print(f'  slopes: {err / (2 / (g.subap_npx - 1)):.1e} CCD px', flush=True)


def test_reference_known_answer_on_the_general_path():
    """test_sh_slopec.py:21-114 on the HIP path: 20 px pupil of 1 m, 2 x 2 lenslets, 500 nm,
    0.1 arcsec pixels, 120 px per subaperture (np 120 after the 12x field oversampling, N 360,
    M 256), Gaussian LLT of 0.1 arcsec; a one-pixel tilt gives x slopes of 1/(npx/2) (the
    reference's rtol/atol 1e-2) and the images match the oracle."""
    pp, pitch, lam, ps, n, npx = 20, 0.05, 500, 0.1, 2, 120
    kw = dict(wavelengthInNm=lam, subap_wanted_fov=npx * ps, sensor_pxscale=ps, subap_on_diameter=n, subap_npx=npx)
    g = sh_geometry(pp, pitch, **kw)
    go = O.sh_geometry(ef_size=pp, pixel_pitch=pitch, **kw)
    assert (g.ovs_np_sub, g.fft_size, g.needs_interpolation) == (120, 360, True)
    gk = GaussianConvolutionKernel(n, n, g.fp4_pixel_pitch * RAD2ASEC, g.fft_size, ps, pupil_size_m=pp * pitch,
                                   target_device_idx=-1, precision=1)
    gk.prepare_for_sh()
    A = np.ones((pp, pp))
    tv = np.radians(pp * pitch / 3600 * ps)
    tilt = np.broadcast_to(np.linspace(-tv / 2, tv / 2, pp), (pp, pp)) * 1e9
    out, plan = run_plan(g, A, np.stack([np.zeros((pp, pp)), tilt]), lgs=True, kern=gk)
    W = n * npx
    idxs = np.array([np.ravel_multi_index(np.mgrid[i * npx:(i + 1) * npx, j * npx:(j + 1) * npx].reshape(2, -1),
                                          (W, W)) for i in range(n) for j in range(n)])
    sp = SlopecPlan(idxs, (W, W), compute_xy_weights(npx), max_frames=2, device=0)
    s = torch.empty((2, 2 * n * n), dtype=torch.float32, device=DEV)
    sp.forward(torch.tensor(out, device=DEV), s, torch.empty((2, n * n), dtype=torch.float64, device=DEV),
               torch.empty((2, 2), dtype=torch.float64, device=DEV), n_frames=2)
    s = s.cpu().numpy()
    print(f'known answer: x slopes {s[1, :4]}, expected {1 / (npx / 2):.5f}', flush=True)
    np.testing.assert_allclose(s[1, :4], 1.0 / (npx / 2), rtol=1e-2, atol=1e-2)
    for k, ph in enumerate((np.zeros((pp, pp)), tilt)):
        ref = O.sh_forward(A, ph, go, kernels=gk.kernels)
        check_intensity(out[k], ref, n, npx)
        check_slopes(s[k], O.shslopec(ref, idxs)['slopes'], npx)


def test_general_path_equals_specialised_kernels_where_both_apply():
    """NGS 20x20 (np 8, M 16) and LGS 40x40 (np 16, M 32): variant 4 against the specialised
    kernels on the same inputs, within the parity tolerance of either against the oracle."""
    wl = workloads.scao_ngs_20x20()
    g = wl.geometry()
    A = wl.pupil()
    frames = wl.frames(2)
    a, _ = run_plan(g, A, frames, variant=2)
    b, pb = run_plan(g, A, frames, variant=4)
    for k in range(2):
        check_intensity(b[k], a[k].astype(np.float64), g.n_lenses, g.subap_npx)
    wl = workloads.lgs_40x40()
    g = wl.geometry()
    kern = workloads.lgs_kernel(g)
    A = wl.pupil()
    frames = wl.frames(1)
    a, _ = run_plan(g, A, frames, lgs=True, kern=kern)
    b, _ = run_plan(g, A, frames, lgs=True, kern=kern, variant=4)
    check_intensity(b[0], a[0].astype(np.float64), g.n_lenses, g.subap_npx)


# Every distinct SH block of the reference's example and test set-ups (values restated from the
# YAML files; pupil px, pixel pitch, LGS launcher present, SH keywords).
REFERENCE_SH_BLOCKS = workloads.REFERENCE_SH_SETUPS


@pytest.mark.parametrize('block', REFERENCE_SH_BLOCKS, ids=[b[0].split('/')[-1] for b in REFERENCE_SH_BLOCKS])
def test_every_reference_sh_block_gets_a_plan(block):
    """Full-size plans for every SH block of the reference's set-ups: none is refused.  NGS blocks
    also run one flat-wavefront frame through the plan's kernel and check the flux known answer
    sum(I) = S0 * sum(A) * pitch^2 (test_sh.py:18-42).  (LGS blocks stop at the plan: their full
    kernel sets are gigabytes on the host.)"""
    where, pp, pitch, lgs, kw = block
    g = sh_geometry(pp, pitch, **kw)
    print(f'{where}: np {g.ovs_np_sub} N {g.fft_size} cut {g.cutpixels} npx {g.subap_npx} n {g.n_lenses} '
          f'fov_ovs {g.fov_ovs:.3g} rot {g.rot_deg} lgs {lgs}', flush=True)
    if g.ef_ovs != g.n_lenses * g.ovs_np_sub:
        # params_morfeo_scao_plots.yml: fov_ovs 122/7 gives an 8365 px resampled field for 68 lenslets
        # of 123 px (8364): the reference's reshape of each subaperture row (sh.py:373) fails, its
        # divisibility check (sh.py:233-235) using round(fov_ovs) does not catch it; the plan refuses
        assert where.endswith('params_morfeo_scao_plots.yml:94')
        with pytest.raises(RuntimeError, match='ef_size must equal n_lenses'):
            SHPlan(g, lgs=lgs, max_frames=1, device=0)
        return
    plan = SHPlan(g, lgs=lgs, max_frames=1, device=0)
    if lgs:
        return
    A = annular_pupil(pp, 0.15)
    if g.needs_interpolation:
        plan.set_resampling(A)
    out = torch.empty((1, g.ccd_side, g.ccd_side), dtype=torch.float32, device=DEV)
    S0 = 3.0
    plan.forward(torch.tensor(A.astype(np.float32), device=DEV),
                 torch.zeros((1, pp, pp), dtype=torch.float32, device=DEV), out, S0)
    torch.cuda.synchronize()
    tot = float(out.double().sum())
    assert np.isfinite(tot)
    np.testing.assert_allclose(tot, S0 * A.sum() * pitch ** 2, rtol=1e-5)


ROUND_MASK = [
    # (name, ef, pitch, obstruction, LGS, SH keywords with squaremask=False: a round FoV mask that
    # clips the cut window, sh.py:282, so the mask is not separable)
    ('20x20_round', 160, 0.05, 0.14, False,
     dict(wavelengthInNm=750, subap_wanted_fov=2.4, sensor_pxscale=0.3, subap_on_diameter=20, subap_npx=8,
          fov_ovs_coeff=1.0, squaremask=False)),
    ('scao_test_round_fov_ovs2', 64, 0.015625, 0.0, False,
     dict(wavelengthInNm=600, subap_wanted_fov=4.0, sensor_pxscale=0.5, subap_on_diameter=8, subap_npx=8,
          squaremask=False)),
    ('eris_sh_lgs_round', 16, 0.05, 0.0, True,
     dict(subap_wanted_fov=5, sensor_pxscale=0.85, subap_npx=6, wavelengthInNm=589, subap_on_diameter=4,
          squaremask=False)),
]


@pytest.mark.parametrize('case', ROUND_MASK, ids=[c[0] for c in ROUND_MASK])
def test_round_fov_mask_matches_oracle(case):
    """squaremask=False: the plan takes the mask as K separable terms on the any-size path and
    normalises by the image sum; GPU images and slopes against the FP64 oracle."""
    from specula_amd.lib.binning import mask_is_separable
    name, ef, pitch, obs, lgs, kw = case
    g = sh_geometry(ef, pitch, **kw)
    go = O.sh_geometry(ef_size=ef, pixel_pitch=pitch, **kw)
    assert not mask_is_separable(g)
    A = annular_pupil(ef, obs) if obs else O.make_mask(ef, 0.0, 1.0).astype(np.float64)
    scr = von_karman_opd(256, pitch, seed=17)
    frames = np.stack([scr[:ef, k:k + ef] * 0.3 for k in range(2)]).astype(np.float32)
    kern = lgs_kernels(g) if lgs else None
    out, plan = run_plan(g, A, frames, lgs, kern)
    print(f'{name}: np {g.ovs_np_sub} N {g.fft_size} cut {g.cutpixels} npx {g.subap_npx} fov_ovs {g.fov_ovs}: '
          f'{plan.mask_terms} mask terms, parseval {plan.parseval}', flush=True)
    assert plan.mask_terms > 1 and not plan.parseval
    ref0 = None
    for k in range(2):
        ref = O.sh_forward(A, frames[k].astype(np.float64), go, kernels=None if kern is None else kern.kernels)
        st = check_intensity(out[k], ref, g.n_lenses, g.subap_npx)
        print(f'  frame {k}: flux {st["total"]:.1e}, normalised image {st["img"]:.1e}', flush=True)
        ref0 = ref if ref0 is None else ref0
    idxs, _ = detect_subaps(O.sh_forward(A, np.zeros_like(A), go, kernels=None if kern is None else kern.kernels),
                            g.n_lenses, 0.5)
    sp = SlopecPlan(idxs, (g.ccd_side, g.ccd_side), compute_xy_weights(g.subap_npx), max_frames=1, device=0)
    s = torch.empty(2 * idxs.shape[0], dtype=torch.float32, device=DEV)
    fl = torch.empty(idxs.shape[0], dtype=torch.float64, device=DEV)
    tt = torch.empty(2, dtype=torch.float64, device=DEV)
    sp.forward(torch.tensor(out[0], device=DEV), s, fl, tt)
    err = check_slopes(s.cpu().numpy(), O.shslopec(ref0, idxs)['slopes'], g.subap_npx)
    print(f'  slopes: {err / (2 / (g.subap_npx - 1)):.1e} CCD px', flush=True)
