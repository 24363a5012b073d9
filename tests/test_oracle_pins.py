"""Pins of the FP64 oracle: reference fixtures, hand derivations, brute force and invariants.

Each test is chosen so that a plausible slip in the oracle (a transposed operand, a missing
fftshift, a wrong tilt sign, a wrong kernel index, a wrong rebin) fails it.  CPU only.
"""
import math
import os

import numpy as np
import pytest

from oracle import sh_oracle as O
from specula_amd.lib import fits

GOLDEN = os.path.join(os.path.dirname(__file__), 'golden')


# ---------------------------------------------------------------- grids and masks

def test_make_xy_docstring_examples():
    """make_xy.py:93-119: sampling 4 -> -0.75..0.75, sampling 5 -> -0.8..0.8; x along columns."""
    x, y = O.make_xy(4, 1.0)
    np.testing.assert_allclose(x[0], [-0.75, -0.25, 0.25, 0.75])
    np.testing.assert_allclose(y[:, 0], [-0.75, -0.25, 0.25, 0.75])
    x, _ = O.make_xy(5, 1.0)
    np.testing.assert_allclose(x[2], [-0.8, -0.4, 0.0, 0.4, 0.8])


def test_make_mask_closed_forms():
    m = O.make_mask(8, diaratio=0.5, square=True)
    # pixel centres +-(2k+1)/8: |c| <= 0.5 keeps the 4 central rows/cols
    expect = np.zeros((8, 8), np.uint8)
    expect[2:6, 2:6] = 1
    np.testing.assert_array_equal(m, expect)
    c = O.make_mask(16, obsratio=0.25, diaratio=1.0)
    x, y = O.make_xy(16, 1.0)
    r2 = x ** 2 + y ** 2
    np.testing.assert_array_equal(c, ((r2 < 1) & (r2 >= 0.0625)).astype(np.uint8))


# ---------------------------------------------------------------- geometry

def test_geometry_hand_derived_test_sh():
    """test_sh.py geometry (120 px, 20 subaps, 500 nm, 0.5"/px, 6 px, FoV 3").  Derived by hand
    from sh.py:112-237: turbulence px 0.34377", search picks scale 3 (FoV error 0.7%), the real
    FoV is 26 FFT px > 18 turbulence px so fov_ovs = ceil(2) = 2, N = 12*3 = 36, cut 10."""
    g = O.sh_geometry(ef_size=120, pixel_pitch=0.05, wavelengthInNm=500, subap_wanted_fov=3,
                      sensor_pxscale=0.5, subap_on_diameter=20, subap_npx=6)
    assert (g['scale_ovs'], g['fov_ovs'], g['ovs_np_sub'], g['fft_size'], g['cutpixels'], g['cutsize']) == \
        (3, 2.0, 12, 36, 10, 26)


def test_geometry_hand_derived_scao_20x20():
    """20x20 subaps of 8 px, 750 nm, sensor px = lambda/d, FoV 1.05*8 lambda/d, fov_ovs_coeff 1:
    the search keeps scale 4 (smallest with FoV error < 2%), fov_ovs = 1, N = 32, no cut."""
    ld = 750e-9 / 0.4 * O.RAD2ASEC
    g = O.sh_geometry(ef_size=160, pixel_pitch=0.05, wavelengthInNm=750, subap_wanted_fov=1.05 * 8 * ld,
                      sensor_pxscale=ld, subap_on_diameter=20, subap_npx=8, fov_ovs_coeff=1.0)
    assert (g['scale_ovs'], g['fov_ovs'], g['fft_size'], g['cutpixels']) == (4, 1.0, 32, 0)
    assert g['fp_mask'].all()


def test_geometry_pinned_resolution():
    """Pre-set FoV resolution (lambda/d)/8 skips the search (sh.py:130): N = 8 np."""
    pitch = 39 / 640
    ld = 750e-9 / (8 * pitch) * O.RAD2ASEC
    g = O.sh_geometry(ef_size=640, pixel_pitch=pitch, wavelengthInNm=750, subap_wanted_fov=8 * ld,
                      sensor_pxscale=ld / 2, subap_on_diameter=80, subap_npx=16, fov_ovs_coeff=1.0,
                      fov_resolution_arcsec=ld / 8)
    assert (g['scale_ovs'], g['fft_size'], g['cutpixels'], g['ccd_side']) == (8, 64, 0, 1280)


def test_geometry_rejects_indivisible_pupil():
    with pytest.raises(ValueError):
        O.sh_geometry(ef_size=121, pixel_pitch=0.05, wavelengthInNm=500, subap_wanted_fov=3,
                      sensor_pxscale=0.5, subap_on_diameter=20, subap_npx=6)


# ---------------------------------------------------------------- SH forward

def small_geom(**over):
    """ef=16, 2x2 subaps of 8 px, N=16, sensor 6 px of 2 FFT px -> cut 4, 2:1 rebin."""
    pitch = 0.05
    turb = 500e-9 / (8 * pitch) * O.RAD2ASEC
    kw = dict(ef_size=16, pixel_pitch=pitch, wavelengthInNm=500, subap_wanted_fov=0.6 * 6 * turb,
              sensor_pxscale=turb, subap_on_diameter=2, subap_npx=6, fov_ovs_coeff=1.0,
              fov_resolution_arcsec=turb / 2)
    kw.update(over)
    return O.sh_geometry(**kw)


def brute_sh(A, ph, g, S0=1.0, real_kernels=None):
    """Direct O(N^4) evaluation with explicit index arithmetic (no FFT library, no reshape)."""
    n, p, N = g['n_lenses'], g['ovs_np_sub'], g['fft_size']
    c, cs, npx = g['cutpixels'], g['cutsize'], g['subap_npx']
    r = cs // npx
    lo = c // 2
    k = 2 * np.pi / g['wavelengthInNm']
    a = np.arange(p) - math.ceil(p / 2)
    mask = g['fp_mask']
    out = np.zeros((n * npx, n * npx))
    W = np.exp(-2j * np.pi * np.outer(np.arange(N), np.arange(N)) / N)
    for i in range(n):
        for j in range(n):
            E = np.zeros((N, N), complex)
            for y in range(p):
                for x in range(p):
                    E[y, x] = (A[i * p + y, j * p + x] * np.exp(1j * k * ph[i * p + y, j * p + x])
                               * np.exp(-1j * np.pi * (a[x] + a[y]) / N))
            F = W @ E @ W.T
            I = np.abs(F) ** 2
            psf = np.zeros((N, N))
            for u in range(N):
                for v in range(N):
                    if real_kernels is None:     # fftshift: psf[u, v] = I[u - N/2, v - N/2]
                        psf[u, v] = I[(u - N // 2) % N, (v - N // 2) % N]
                    else:                        # circular correlation with the (row i, col j) kernel
                        kk = real_kernels[j * n + i] / real_kernels[j * n + i].sum()
                        psf[u, v] = sum(kk[s, t] * I[(u + s) % N, (v + t) % N]
                                        for s in range(N) for t in range(N))
            psf = psf * mask
            for Y in range(npx):
                for X in range(npx):
                    out[i * npx + Y, j * npx + X] = psf[lo + Y * r: lo + (Y + 1) * r,
                                                        lo + X * r: lo + (X + 1) * r].sum()
    return out * (S0 * g['pixel_pitch'] ** 2 * A.sum() / out.sum())


def test_sh_forward_brute_force_ngs():
    g = small_geom()
    assert (g['fft_size'], g['cutpixels'], g['cutsize']) == (16, 4, 12)
    assert not g['fp_mask'].all()
    rng = np.random.default_rng(3)
    A = (rng.random((16, 16)) > 0.2).astype(float)
    ph = rng.normal(0, 300, (16, 16))
    np.testing.assert_allclose(O.sh_forward(A, ph, g, S0=3.0), brute_sh(A, ph, g, S0=3.0),
                               rtol=1e-10, atol=1e-12)


def test_sh_forward_brute_force_lgs_correlation():
    g = small_geom(subap_wanted_fov=100.0)
    rng = np.random.default_rng(4)
    A = np.ones((16, 16))
    ph = rng.normal(0, 200, (16, 16))
    real = rng.random((4, 16, 16))
    kern = O.process_kernels(real, 2, 2, return_fft=True)
    np.testing.assert_allclose(O.sh_forward(A, ph, g, kernels=kern), brute_sh(A, ph, g, real_kernels=real),
                               rtol=1e-9, atol=1e-12)


def test_flux_known_answer():
    """test_sh.py:18-42: sum of the SH intensity == S0 * masked area."""
    g = small_geom()
    A = np.ones((16, 16))
    A[:3, :5] = 0
    img = O.sh_forward(A, np.zeros((16, 16)), g, S0=100)
    np.testing.assert_allclose(img.sum(), 100 * A.sum() * 0.05 ** 2, rtol=1e-12)


def test_flat_wavefront_zero_slopes_and_tilt_shift():
    """The half-pixel tilt centres a flat-wavefront spot between the 4 central pixels (zero CoG,
    as pinned by scao_sn_n8_th0.5_ref.fits), and a one-CCD-pixel tilt shifts it by one pixel
    (test_sh.py:44-95) giving a CoG slope of 2/(npx-1) (test_sh_slopec.py:109-114)."""
    ef, pitch, pxs, npx, n = 120, 0.05, 0.25, 6, 20
    g = O.sh_geometry(ef_size=ef, pixel_pitch=pitch, wavelengthInNm=500, subap_wanted_fov=npx * pxs,
                      sensor_pxscale=pxs, subap_on_diameter=n, subap_npx=npx, fov_ovs_coeff=1.0)
    flat = O.sh_forward(np.ones((ef, ef)), np.zeros((ef, ef)), g)
    rows, cols = np.meshgrid(np.arange(npx), np.arange(npx), indexing='ij')
    idxs = np.array([((i * npx + rows) * n * npx + j * npx + cols).ravel() for j in range(n) for i in range(n)])
    s0 = O.shslopec(flat, idxs)['slopes']
    assert np.abs(s0).max() < 1e-12
    tilt_value = np.radians(ef * pitch / 3600 * pxs)
    ph = np.broadcast_to(np.linspace(-tilt_value / 2, tilt_value / 2, ef), (ef, ef)) * 1e9
    tilted = O.sh_forward(np.ones((ef, ef)), ph, g)
    shifted = np.roll(flat, (0, 1))
    shifted[:, ::npx] = 0
    t2 = tilted.copy()
    t2[:, ::npx] = 0
    np.testing.assert_array_almost_equal(t2, shifted, decimal=3)


def test_one_pixel_tilt_slope():
    """test_sh_slopec.py:21-114 restated at fov_ovs == 1 (2x2 subapertures of 10 px, 20 CCD px of
    0.08"): x slopes equal 1/(npx/2) within the reference's rtol/atol of 1e-2, y slopes vanish."""
    ef, pitch, pxs, npx, n = 20, 0.05, 0.08, 20, 2
    g = O.sh_geometry(ef_size=ef, pixel_pitch=pitch, wavelengthInNm=500, subap_wanted_fov=npx * pxs,
                      sensor_pxscale=pxs, subap_on_diameter=n, subap_npx=npx, fov_ovs_coeff=1.0)
    tilt_value = np.radians(ef * pitch / 3600 * pxs)
    ph = np.broadcast_to(np.linspace(-tilt_value / 2, tilt_value / 2, ef), (ef, ef)) * 1e9
    img = O.sh_forward(np.ones((ef, ef)), ph, g)
    idxs, _ = O.detect_subaps(img, n, 0.0)
    s = O.shslopec(img, idxs)['slopes']
    np.testing.assert_allclose(s[:4], 1.0 / (npx / 2), rtol=1e-2, atol=1e-2)
    assert np.abs(s[4:]).max() < 1e-12


# ---------------------------------------------------------------- toccd

def test_toccd_integer_ratio_is_block_sum():
    rng = np.random.default_rng(0)
    a = rng.random((12, 12))
    out = O.toccd(a, (4, 4))
    np.testing.assert_allclose(out, a.reshape(4, 3, 4, 3).sum(axis=(1, 3)), rtol=1e-13)


def test_toccd_lcm_matches_replicate_then_mean():
    rng = np.random.default_rng(1)
    a = rng.random((14, 14))
    L = math.lcm(14, 6)
    rep = np.repeat(np.repeat(a, L // 14, axis=0), L // 14, axis=1)       # nearest replicate
    ref = rep.reshape(6, L // 6, 6, L // 6).mean(axis=(1, 3))
    ref = ref / ref.sum() * a.sum()
    np.testing.assert_allclose(O.toccd(a, (6, 6)), ref, rtol=1e-13)


# ---------------------------------------------------------------- ShSlopec

def test_shslopec_hand_computed():
    """2x2 subapertures in a 4x4 frame; x/y weights are -1, +1 per column/row."""
    pix = np.array([[1.0, 3.0, 0.0, 0.0],
                    [0.0, 0.0, 0.0, 0.0],
                    [2.0, 2.0, 5.0, 0.0],
                    [0.0, 0.0, 0.0, 5.0]])
    idxs = np.array([[0, 1, 4, 5], [2, 3, 6, 7], [8, 9, 12, 13], [10, 11, 14, 15]])
    r = O.shslopec(pix, idxs)
    # subap 0: tot 4, sx = (-1*1 + 1*3)/4 = 0.5, sy = (-1*4)/4 = -1
    # subap 1: tot 0 -> factor inf -> clamped to 0
    # subap 2: tot 4, sx 0, sy -1 ; subap 3: tot 10, sx = (-5+5)/10 = 0, sy = (-5+5)/10 = 0
    np.testing.assert_allclose(r['slopes'], [0.5, 0, 0, 0, -1, 0, -1, 0])
    np.testing.assert_allclose(r['flux'], [4, 0, 4, 10])
    assert r['total'] == 18 and r['mean'] == 4.5
    # threshold 1.5: subtract then clamp -> subap 0 keeps 1.5 at x=+1: sx = 1, sy = -1
    r = O.shslopec(pix, idxs, thr_value=1.5)
    np.testing.assert_allclose(r['slopes'][0], 1.0)
    np.testing.assert_allclose(r['flux'], [4, 0, 4, 10])          # flux is pre-threshold
    # pedestal: values below 1.5 zeroed, others kept: subap 0 -> only the 3 survives
    r = O.shslopec(pix, idxs, thr_value=1.5, thr_pedestal=True)
    np.testing.assert_allclose(r['slopes'][[0, 4]], [1.0, -1.0])


def test_shslopec_dim_subaperture_cutoff():
    """factor = 0 where 1/tot > 1/(1e-3 mean) (sh_slopec.py:393-400)."""
    pix = np.zeros((2, 4))
    pix[0, 1] = 1.0           # subap 0: 1.0 at x=+1
    pix[0, 3] = 1e-4          # subap 1: tiny flux, below 1e-3 of the mean
    idxs = np.array([[0, 1, 4, 5], [2, 3, 6, 7]])
    r = O.shslopec(pix, idxs)
    np.testing.assert_allclose(r['slopes'], [1.0, 0.0, -1.0, 0.0])


def test_detect_subaps_reproduces_fixture_layout():
    """scao_subaps_n8_th0.5_ref.fits: 52 subapertures of 8x8 px in a 64x64 frame, idxs stored
    transposed (subap_data.py:62,76).  An image lit exactly on those subapertures must give the
    same index lists and display map, in the same (column-major block) order."""
    hdus = fits.read(os.path.join(GOLDEN, 'scao_subaps_n8_th0.5_ref.fits'))
    idxs_ref = hdus[1][0].T
    dmap_ref = hdus[2][0]
    img = np.zeros(64 * 64)
    img[idxs_ref.ravel()] = 1.0
    idxs, dmap = O.detect_subaps(img.reshape(64, 64), 8, 0.5)
    np.testing.assert_array_equal(idxs, idxs_ref)
    np.testing.assert_array_equal(dmap, dmap_ref)


# ---------------------------------------------------------------- LGS kernels

def test_lgs_map_sh_matches_idl_golden():
    """test_convolution_kernel.py:175-276 parameters and tolerance."""
    dimx = 10
    dimension = 64
    zlayer = np.linspace(80e3, 100e3, 20)
    sigma = 10e3 / (2 * np.sqrt(2 * np.log(2)))
    zprofile = np.exp(-0.5 * ((zlayer - 90e3) / sigma) ** 2)
    zprofile /= zprofile.sum()
    m = O.lgs_map_sh(nsh=dimx, diam=8.0, rl=[5, 5, 0], zb=90e3, dz=zlayer - 90e3, profz=zprofile,
                     fwhmb=1.0, ps=0.1, ssp=dimension, overs=1, theta=[0.0, 0.0])
    k2d = np.zeros((dimx * dimension, dimx * dimension))
    for j in range(dimx):
        for i in range(dimx):
            k2d[j * dimension:(j + 1) * dimension, i * dimension:(i + 1) * dimension] = m[i * dimx + j]
    ref = fits.read(os.path.join(GOLDEN, 'lgs_map_sh_ref.fits'))[0][0].astype(np.float64)
    ref /= ref.sum()
    k2d /= k2d.sum()
    np.testing.assert_allclose(k2d, ref, rtol=1e-5, atol=1e-5)


def test_process_kernels_index_and_normalisation():
    rng = np.random.default_rng(2)
    real = rng.random((9, 8, 8))
    k = O.process_kernels(real, 3, 3)
    for i in range(3):
        for j in range(3):
            np.testing.assert_allclose(k[j * 3 + i], np.fft.ifft2(real[i * 3 + j] / real[i * 3 + j].sum()))
            np.testing.assert_allclose(k[j * 3 + i][0, 0], 1 / 64)


# ---------------------------------------------------------------- field resampling (fov_ovs != 1)

def _golden(name, hdu=0):
    return fits.read(os.path.join(GOLDEN, name))[hdu][0]


@pytest.mark.parametrize('two_pixel,name', [(False, 'extrapol_array_doExt2PixFalse.fits'),
                                            (True, 'extrapol_array_doExt2PixTrue.fits')])
def test_extrapolation_maps_match_idl_golden(two_pixel, name):
    """test_extrapol_mat_define.py: source maps of the 160x160 pupil mask, both variants."""
    s1, s2, _, _ = O.extrapolate_edge_pixel_mat_define(_golden('mask.fits'), do_ext_2_pix=two_pixel)
    np.testing.assert_array_equal(s1, _golden(name, 0).astype(np.int32))
    np.testing.assert_array_equal(s2, _golden(name, 1).astype(np.int32))


@pytest.mark.parametrize('maps,ref', [('extrapol_array_doExt2PixFalse.fits', 'extrapolated1.fits'),
                                      ('extrapol_array_doExt2PixTrue.fits', 'extrapolated2.fits')])
def test_extrapolate_edge_pixel_matches_idl_golden(maps, ref):
    """test_extrapolate_edge_pixel.py (float32 goldens: agreement to their rounding)."""
    m1, m2 = _golden(maps, 0).astype(int), _golden(maps, 1).astype(int)
    i1 = np.where(m1 >= 0)
    i2 = np.where(m2[i1] >= 0)
    out = O.extrapolate_edge_pixel(_golden('phase.fits'), m1, m2, i1, i2)
    want = _golden(ref)
    np.testing.assert_allclose(out, want, rtol=1e-7, atol=0)
    # the extrapolated ring really changed: pixels outside the mask now carry values
    assert np.count_nonzero(out != _golden('phase.fits')) > 100


def test_interp2d_matches_idl_golden():
    """test_interp2d.py:19-43: explicit sampling points from make_xy; rtol 4e-2 with the one
    pixel the reference also exempts."""
    phase = _golden('input_phase.fits')
    ref = _golden('ref_phase.fits')
    xx, yy = O.make_xy(480, 479.84003 / 2.)
    xx1 = np.clip(xx + 240.5 + 0.0930127, 0, phase.shape[1] - 1)
    yy1 = np.clip(yy + 240.5, 0, phase.shape[0] - 1)
    out = O.interp2d(phase, yy1, xx1)
    out[394, 365] = ref[394, 365]
    np.testing.assert_allclose(out, ref, rtol=4e-2)


def test_interp2d_exact_on_affine_and_in_bounds():
    """Bilinear sampling reproduces a + b y + c x exactly (fixes the weights and the corner
    order); the default grid with a rotation never reads outside the input (the reference's
    NaN-padding boundary test, test_interp2d.py:47-62)."""
    y, x = np.mgrid[0:23, 0:31].astype(float)
    f = 1.5 - 0.25 * y + 0.75 * x
    yy, xx = O.interp2d_coords(f.shape, (50, 40), rot_deg=17.0, row_shift=1, col_shift=-2)
    np.testing.assert_allclose(O.interp2d(f, yy, xx), 1.5 - 0.25 * yy + 0.75 * xx, rtol=0, atol=1e-12)
    big = np.full((20, 20), np.nan)
    big[:10] = 1.0
    yy, xx = O.interp2d_coords((10, 20), (5, 5), rot_deg=45)
    assert not np.isnan(O.interp2d(big[:10], yy, xx)).any()
    # default grid without rotation: k (n_in - 1) / n_out along each axis (interp2d.py:61-65)
    yy, xx = O.interp2d_coords((10, 20), (4, 8))
    np.testing.assert_allclose(yy[:, 0], np.arange(4) * 9 / 4)
    np.testing.assert_allclose(xx[0], np.arange(8) * 19 / 8)


def test_interpolated_sh_reproduces_scao_fixtures():
    """params_scao_sh_test*.yml geometry (64 px pupil, 8x8 subapertures, fov_ovs 2, N 32, cut 24):
    the flat-wavefront image selects the 52 subapertures of scao_subaps_n8_th0.5_ref.fits and
    gives the near-zero slopes of scao_sn_n8_th0.5_ref.fits."""
    g = O.sh_geometry(ef_size=64, pixel_pitch=0.015625, wavelengthInNm=600, subap_wanted_fov=4.0,
                      sensor_pxscale=0.5, subap_on_diameter=8, subap_npx=8)
    assert (g['fov_ovs'], g['ovs_np_sub'], g['fft_size'], g['cutpixels'], g['interpolate']) == (2, 16, 32, 24, True)
    A = O.make_mask(64, 0.0, 1.0).astype(float)
    img = O.sh_forward(A, np.zeros((64, 64)), g)
    idxs_ref = fits.read(os.path.join(GOLDEN, 'scao_subaps_n8_th0.5_ref.fits'))[1][0].T
    idxs, _ = O.detect_subaps(img, 8, 0.5)
    np.testing.assert_array_equal(idxs, idxs_ref)
    sn = fits.read(os.path.join(GOLDEN, 'scao_sn_n8_th0.5_ref.fits'))[1][0]
    s = O.shslopec(img, idxs_ref)['slopes']
    np.testing.assert_allclose(s, sn, rtol=0, atol=2e-8)
    # flux normalisation uses the input pupil's masked area (sh.py:426), not the resampled one
    np.testing.assert_allclose(img.sum(), A.sum() * 0.015625 ** 2, rtol=1e-12)


# ---------------------------------------------------------------- modal DM and push-pull calibration

SCAO_SH = dict(wavelengthInNm=600, subap_wanted_fov=4.0, sensor_pxscale=0.5, subap_on_diameter=8, subap_npx=8)


def test_zernike_closed_forms_and_orthonormality():
    """Noll's table (Noll 1976, table 1) at the pixel centres; (1/pi) integral Z_i Z_j = delta_ij."""
    dim = 48
    c = dim / 2
    Y, X = (np.mgrid[0.5:dim + 0.5:1, 0.5:dim + 0.5:1] - c) / (dim / 2)
    r, t = np.hypot(X, Y), np.arctan2(Y, X)
    table = {1: np.ones_like(r), 2: 2 * X, 3: 2 * Y, 4: math.sqrt(3) * (2 * r ** 2 - 1),
             5: math.sqrt(6) * r ** 2 * np.sin(2 * t), 6: math.sqrt(6) * r ** 2 * np.cos(2 * t),
             7: math.sqrt(8) * (3 * r ** 3 - 2 * r) * np.sin(t), 8: math.sqrt(8) * (3 * r ** 3 - 2 * r) * np.cos(t),
             9: math.sqrt(8) * r ** 3 * np.sin(3 * t), 10: math.sqrt(8) * r ** 3 * np.cos(3 * t),
             11: math.sqrt(5) * (6 * r ** 4 - 6 * r ** 2 + 1)}
    for j, z in table.items():
        np.testing.assert_allclose(O.zernike_map(j, dim), z, rtol=0, atol=1e-12, err_msg=f'Z{j}')
    big = 400
    disk = O.make_mask(big, 0.0, 1.0).astype(bool)
    Z = np.array([O.zernike_map(j, big)[disk] for j in range(1, 16)])
    G = Z @ Z.T / disk.sum()
    np.testing.assert_allclose(G, np.eye(15), atol=2e-2)


def test_zernike_ifunc_normalised_modes():
    """compute_zern_ifunc: mean-free, unit std rows spanning the same space as the raw modes."""
    ifunc, mask = O.compute_zern_ifunc(64, 10, obsratio=0.1)
    np.testing.assert_allclose(ifunc.mean(axis=1), 0, atol=1e-12)
    np.testing.assert_allclose(ifunc.std(axis=1), 1, rtol=1e-12)
    idx = np.where(mask)
    raw = np.array([O.zernike_map(j, 64)[idx] for j in range(2, 12)])
    # mode k lies in span(raw modes 1..k, piston): residual of a least-squares fit is ~0
    for k in range(10):
        basis = np.vstack([raw[:k + 1], np.ones(raw.shape[1])]).T
        coef, *_ = np.linalg.lstsq(basis, ifunc[k], rcond=None)
        assert np.abs(basis @ coef - ifunc[k]).max() < 1e-9
    # tip (Z2 = 2x) keeps its sign through the Householder QR on this pupil
    assert np.corrcoef(ifunc[0], raw[0])[0, 1] > 0.999


def scao_pushpull_oracle():
    go = O.sh_geometry(ef_size=64, pixel_pitch=0.015625, **SCAO_SH)
    idxs = fits.read(os.path.join(GOLDEN, 'scao_subaps_n8_th0.5_ref.fits'))[1][0].T
    ifunc, dmask = O.compute_zern_ifunc(64, 40, obsratio=0.1, diaratio=None)
    A = O.make_mask(64, None, 1.0).astype(float) * dmask        # pupilstop x DM layer amplitude
    th = O.pushpull_time_history(40, [50.0] * 40)
    slopes = []
    for cmd in th:
        ph = O.dm_phase(ifunc, dmask, cmd, sign=1)
        # the SH was set up on the not yet propagated (all-ones) field: no edge extrapolation
        img = O.sh_forward(A, ph, go, S0=1.0, ext_mask=np.ones((64, 64)))
        slopes.append(O.shslopec(img, idxs)['slopes'])
    return O.accumulate_intmat(np.array(slopes), th, 40), th, np.array(slopes)


def test_pushpull_calibration_reproduces_scao_rec_fixture():
    """test_sh_calibration.py + params_scao_sh_test{,_rec}.yml: 40 Zernike modes at 50 nm pushed
    and pulled through SH (fov_ovs 2, cut 24) and ShSlopec; pinv of the IM as Recmat.save writes
    it equals scao_rec_n8_th0.5_ref.fits to the test's decimal=3 (|d| < 1.5e-3)."""
    im, _, _ = scao_pushpull_oracle()
    rec = O.recmat_fits_data(im)
    ref = fits.read(os.path.join(GOLDEN, 'scao_rec_n8_th0.5_ref.fits'))[1][0]
    assert rec.shape == ref.shape == (40, 104)
    np.testing.assert_array_almost_equal(rec, ref, decimal=3)
    assert np.abs(rec - ref).max() < 1e-4                 # observed 2.6e-5 (the file is FP32)


# ---------------------------------------------------------------- CCD readout

def test_ccd_oracle_noiseless_binning_and_qe():
    """apply_binning: block sums times binning^2 (the tot/sum factor is 1); apply_qe scales."""
    rng = np.random.default_rng(5)
    img = rng.random((8, 12)).astype(np.float32) * 100
    out = O.ccd_readout(img, binning=2, quantum_eff=0.3, dtype=np.float64)
    blocks = img.astype(np.float64).reshape(4, 2, 6, 2).sum(axis=(1, 3))
    np.testing.assert_allclose(out, blocks * 4 * 0.3, rtol=1e-6)
    np.testing.assert_array_equal(O.ccd_readout(img), img)


def test_ccd_oracle_photon_adu_and_truncated_readout():
    """Poisson counts quantised to 1/ADU_gain; with photon noise the readout noise is added to an
    int64 frame, so only its integer part survives; the clamp at -bias."""
    lam = np.full((200, 200), 25.0, dtype=np.float32)
    out = O.ccd_readout(lam, photon_noise=True, photon_rng=np.random.default_rng(1))
    assert np.all(out == np.round(out))                       # integer counts, ADU gain 8
    assert abs(out.mean() - 25) < 5 * 5 / 200 and abs(out.var() - 25) < 2
    zero = np.zeros((300, 300), dtype=np.float32)
    out = O.ccd_readout(zero, photon_noise=True, readout_noise=True, readout_level=1.5,
                        photon_rng=np.random.default_rng(1), readout_rng=np.random.default_rng(2))
    assert np.all(out == np.trunc(out))
    from scipy.stats import norm
    p0 = norm.cdf(1 / 1.5) - norm.cdf(-1 / 1.5)              # trunc(1.5 Z) == 0
    assert abs(np.mean(out == 0) - p0) < 5 * np.sqrt(p0 * (1 - p0) / out.size)
    big = np.full((4, 4), 0.0, dtype=np.float32)
    out = O.ccd_readout(big, photon_noise=True, readout_noise=True, readout_level=1e4,
                        photon_rng=np.random.default_rng(3), readout_rng=np.random.default_rng(4))
    assert out.min() >= -400 / 8                              # clamp of round(x*8)+400 at 0
    flt = O.ccd_readout(zero, readout_noise=True, readout_level=2.0, readout_rng=np.random.default_rng(2))
    assert flt.dtype == np.float32 and abs(flt.std() - 2.0) < 0.02 and not np.all(flt == np.trunc(flt))
