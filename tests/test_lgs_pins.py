"""Pins of the LGS kernel wiring (CPU): the oracle's convolution_kernels_for_sh and
gaussian_kernels_for_sh, and the product's ConvolutionKernel / GaussianConvolutionKernel, against
what the geometry of the sodium-beacon model fixes, plus the reference's own known answer.

* Known answer (test/test_sh_slopec.py:21-114): Gaussian LLT kernel in the SH path, 2 x 2
  subapertures of 120 CCD px, a one-CCD-pixel tilt gives x slopes of 1/(npx/2) (rtol/atol 1e-2,
  the reference's bar), y slopes unchanged.
* Closed form of a thin layer (convolution_kernel.py:14-108): the layer at zb + dz seen from the
  subaperture at rs, launched from rl, is a Gaussian centred at the angle
      theta_tt + dz (rs - rl) / (zb (zb + dz))                 [rad, per axis]
  with width sigma / (1 + dz/zb) (sigma = fwhm / (2 sqrt(2 ln 2))), sampled on the field grid
  linspace(-N ps/2, N ps/2, N).  theta_tt = +-0.5 ps (positive_shift_tt) + theta
  (convolution_kernel.py:199-204), the layer offsets are taken from zfocus (:207-208), airmass
  multiplies layers and focus (:190-192), and process_kernels stores the kernel of subaperture
  (row, col) at index row * n + col (:258-274).
* Elongation: with a thick profile the spot of each subaperture is elongated along rs - rl.
* The flat-wavefront LGS image is the NGS image moved by 1/(2N) FFT pixel (N: FFT size), because
  the +0.5 ps shift puts the kernel centre at (N^2 - 1)/(2N) instead of N/2.

Each check is also run on kernels built with a planted mistake (wrong lgs_tt sign, transposed
subaperture index, layer offsets not taken from the focus, wrong plate scale) to show it fails.
"""
import numpy as np
import pytest

from oracle import sh_oracle as O
from specula_amd.data_objects.convolution_kernel import ConvolutionKernel, GaussianConvolutionKernel

FWHM2SIG = 1.0 / (2 * np.sqrt(2 * np.log(2)))


def kgeom(n=6, N=64, ps=0.1, D=6.0, ef=120):
    """Geometry dict of the fields convolution_kernels_for_sh reads (pixel scale ps arcsec)."""
    return dict(n_lenses=n, fft_size=N, fp4_pixel_pitch=ps / O.RAD2ASEC, pixel_pitch=D / ef, ef_size=ef)


def real_kernels(kernels):
    """Normalised real kernels back from kernels = ifft2(k) (process_kernels)."""
    return np.fft.fft2(kernels, axes=(1, 2)).real


def moments(k):
    """Centroid (cy, cx) in pixel index and covariance [[vyy, vyx], [vxy, vxx]] of each image."""
    N = k.shape[-1]
    idx = np.arange(N, dtype=np.float64)
    tot = k.sum(axis=(1, 2))
    cy = (k.sum(axis=2) @ idx) / tot
    cx = (k.sum(axis=1) @ idx) / tot
    dy = idx[None, :] - cy[:, None]
    dx = idx[None, :] - cx[:, None]
    vyy = np.einsum('sy,syx->s', dy ** 2, k) / tot
    vxx = np.einsum('sx,syx->s', dx ** 2, k) / tot
    vxy = np.einsum('sy,sx,syx->s', dy, dx, k) / tot
    return cy, cx, np.stack([np.stack([vyy, vxy], -1), np.stack([vxy, vxx], -1)], -2)


def angle_to_index(theta_arcsec, N, ps):
    """Pixel index of angle theta on linspace(-N ps/2, N ps/2, N) (convolution_kernel.py:53-54)."""
    return (theta_arcsec + N * ps / 2) * (N - 1) / (N * ps)


def thin_layer_centres(g, rl, zb, dz, tt):
    """Expected centres (cy, cx) [pixel index] for subaperture s = row * n + col."""
    n, N = g['n_lenses'], g['fft_size']
    ps = g['fp4_pixel_pitch'] * O.RAD2ASEC
    D = g['pixel_pitch'] * g['ef_size']
    lin = np.linspace(-D / 2, D / 2, n)
    rows, cols = np.divmod(np.arange(n * n), n)
    par = dz / (zb * (zb + dz)) * O.RAD2ASEC
    tx = tt[0] + par * (lin[cols] - rl[0])
    ty = tt[1] + par * (lin[rows] - rl[1])
    return angle_to_index(ty, N, ps), angle_to_index(tx, N, ps)


def check_thin_layer(kernels, g, rl, zb, dz, tt, fwhm, tol_px=2e-3):
    k = real_kernels(kernels)
    if not np.all(k.sum(axis=(1, 2)) > 0):          # a spot outside the field: no centre at all
        return False
    cy, cx, cov = moments(k)
    ey, ex = thin_layer_centres(g, rl, zb, dz, tt)
    N, ps = g['fft_size'], g['fp4_pixel_pitch'] * O.RAD2ASEC
    sig_px = fwhm * FWHM2SIG / (1 + dz / zb) * (N - 1) / (N * ps)
    ok_c = np.abs(cy - ey).max() < tol_px and np.abs(cx - ex).max() < tol_px
    ok_w = (np.abs(np.sqrt(cov[:, 0, 0]) / sig_px - 1).max() < 1e-3 and
            np.abs(np.sqrt(cov[:, 1, 1]) / sig_px - 1).max() < 1e-3)
    return ok_c and ok_w


THIN = dict(rl=(7.0, -4.0, 0.0), zfocus=90e3, layer=96e3, spot=0.6, theta=(0.3, -0.2), airmass=1.0)


def oracle_thin(g, positive=True, **over):
    p = dict(THIN, **over)
    return O.convolution_kernels_for_sh(g, list(p['rl']), p['spot'], [p['layer']], [1.0], zfocus=p['zfocus'],
                                        theta=p['theta'], airmass=p['airmass'], positive_shift_tt=positive)


def expected_tt(g, positive, theta):
    ps = g['fp4_pixel_pitch'] * O.RAD2ASEC
    s = 0.5 if positive else -0.5
    return (s * ps + theta[0], s * ps + theta[1])


@pytest.mark.parametrize('positive', [True, False])
@pytest.mark.parametrize('airmass', [1.0, 1.3])
def test_oracle_thin_layer_closed_form(positive, airmass):
    g = kgeom()
    K = oracle_thin(g, positive, airmass=airmass)
    zb = THIN['zfocus'] * airmass
    dz = THIN['layer'] * airmass - zb
    assert check_thin_layer(K, g, THIN['rl'], zb, dz, expected_tt(g, positive, THIN['theta']), THIN['spot'])


def test_thin_layer_check_rejects_planted_mistakes():
    """The closed-form check fails for kernels built with a plausible wiring mistake."""
    g = kgeom()
    zb, dz = THIN['zfocus'], THIN['layer'] - THIN['zfocus']
    tt = expected_tt(g, True, THIN['theta'])
    good = oracle_thin(g)
    assert check_thin_layer(good, g, THIN['rl'], zb, dz, tt, THIN['spot'])
    # lgs_tt sign flipped
    assert not check_thin_layer(oracle_thin(g, positive=False), g, THIN['rl'], zb, dz, tt, THIN['spot'])
    # subaperture index transposed (kernels[col * n + row])
    n = g['n_lenses']
    tr = good.reshape(n, n, *good.shape[1:]).transpose(1, 0, 2, 3).reshape(good.shape)
    assert not check_thin_layer(tr, g, THIN['rl'], zb, dz, tt, THIN['spot'])
    # layer offsets not taken from the focus: dz = layer altitude
    ps = g['fp4_pixel_pitch'] * O.RAD2ASEC
    real = O.lgs_map_sh(n, g['pixel_pitch'] * g['ef_size'], list(THIN['rl']), zb, np.array([THIN['layer']]),
                        [1.0], THIN['spot'], ps, g['fft_size'], overs=1, theta=np.array(tt))
    assert not check_thin_layer(O.process_kernels(real, n, n), g, THIN['rl'], zb, dz, tt, THIN['spot'])
    # wrong plate scale (field sampled at 0.9 ps)
    g9 = dict(g, fp4_pixel_pitch=0.9 * g['fp4_pixel_pitch'])
    assert not check_thin_layer(oracle_thin(g9), g, THIN['rl'], zb, dz, tt, THIN['spot'])


@pytest.mark.parametrize('positive', [True, False])
def test_oracle_gaussian_kernel_centre_and_width(positive):
    """GaussianConvolutionKernel (gaussian_convolution_kernel.py:38-57): on-axis launcher, one
    layer at the focus: centre at (N^2 - 1)/(2N) (positive shift) or (N - 1)^2/(2N) (negative), the
    FWHM the spot size, identical for every subaperture; dimx is raised to at least 2."""
    for n in (1, 4):
        g = kgeom(n=n, N=48, ps=0.12)
        K = O.gaussian_kernels_for_sh(g, 0.8, positive_shift_tt=positive)
        assert K.shape[0] == max(n, 2) ** 2
        cy, cx, cov = moments(real_kernels(K))
        N = 48
        c = (N * N - 1) / (2 * N) if positive else (N - 1) ** 2 / (2 * N)
        np.testing.assert_allclose(cy, c, atol=2e-4)
        np.testing.assert_allclose(cx, c, atol=2e-4)
        sig = 0.8 * FWHM2SIG * (N - 1) / (N * 0.12)
        np.testing.assert_allclose(np.sqrt(cov[:, 0, 0]), sig, rtol=1e-3)
        np.testing.assert_allclose(np.sqrt(cov[:, 1, 1]), sig, rtol=1e-3)


def test_oracle_elongation_along_launcher_baseline():
    """Thick sodium profile, launcher off-axis: each spot's major axis is parallel to rs - rl and
    the elongation grows with |rs - rl|."""
    g = kgeom(n=6, N=96, ps=0.15)
    rl = (9.0, 5.0, 0.0)
    layers = np.linspace(80e3, 100e3, 11)
    prof = np.exp(-0.5 * ((layers - 90e3) / (10e3 * FWHM2SIG)) ** 2)
    K = O.convolution_kernels_for_sh(g, list(rl), 0.6, layers.tolist(), (prof / prof.sum()).tolist())
    _, _, cov = moments(real_kernels(K))
    n = g['n_lenses']
    D = g['pixel_pitch'] * g['ef_size']
    lin = np.linspace(-D / 2, D / 2, n)
    rows, cols = np.divmod(np.arange(n * n), n)
    bx, by = lin[cols] - rl[0], lin[rows] - rl[1]
    w, v = np.linalg.eigh(cov)                       # ascending: v[:, :, 1] major axis (y, x)
    major = v[:, :, 1]
    ang = np.arctan2(major[:, 0], major[:, 1])
    base = np.arctan2(by, bx)
    d = np.abs(((ang - base) + np.pi / 2) % np.pi - np.pi / 2)
    assert d.max() < 2e-3
    elong = np.sqrt(w[:, 1] / w[:, 0])
    order = np.argsort(np.hypot(bx, by))
    assert np.all(np.diff(elong[order]) > -1e-4 * elong[order][1:])   # ties: equal |rs - rl|
    assert elong.min() > 1.05 and elong.max() > 2.0


@pytest.mark.parametrize('positive', [True, False])
def test_product_kernels_pass_the_same_pins(positive):
    """The product's kernel objects reproduce the closed forms (thin layer with airmass and beacon
    offset, Gaussian LLT) and the oracle's kernels."""
    g = kgeom()
    am = 1.2
    k = ConvolutionKernel(g['n_lenses'], g['n_lenses'], g['fp4_pixel_pitch'] * O.RAD2ASEC,
                          g['pixel_pitch'] * g['ef_size'], g['fft_size'], launcher_pos=list(THIN['rl']),
                          launcher_size=THIN['spot'], zfocus=THIN['zfocus'], theta=list(THIN['theta']),
                          airmass=am, positive_shift_tt=positive, target_device_idx=-1, precision=1)
    k.prepare_for_sh(sodium_altitude=[THIN['layer']], sodium_intensity=[1.0])
    zb = THIN['zfocus'] * am
    assert check_thin_layer(k.kernels, g, THIN['rl'], zb, THIN['layer'] * am - zb,
                            expected_tt(g, positive, THIN['theta']), THIN['spot'])
    ko = oracle_thin(g, positive, airmass=am)
    np.testing.assert_allclose(k.kernels, ko, rtol=0, atol=2e-6 * np.abs(ko).max())
    gk = GaussianConvolutionKernel(3, 3, 0.12, 48, 0.8, pupil_size_m=6.0, positive_shift_tt=positive,
                                   target_device_idx=-1, precision=1)
    gk.prepare_for_sh()
    go = O.gaussian_kernels_for_sh(kgeom(n=3, N=48, ps=0.12), 0.8, positive_shift_tt=positive)
    np.testing.assert_allclose(gk.kernels, go, rtol=0, atol=2e-6 * np.abs(go).max())


# --------------------------------------------------------------------------- kernels inside SH

def _reference_slopec_case():
    """test_sh_slopec.py:21-114: 20 px pupil of 1 m, 2 x 2 subapertures, 500 nm, 0.1 arcsec
    pixels, 120 px per subaperture, Gaussian LLT of 0.1 arcsec."""
    pp, pitch, lam, ps, n, npx = 20, 0.05, 500, 0.1, 2, 120
    go = O.sh_geometry(ef_size=pp, pixel_pitch=pitch, wavelengthInNm=lam, subap_wanted_fov=npx * ps,
                       sensor_pxscale=ps, subap_on_diameter=n, subap_npx=npx)
    W = n * npx
    idxs = []
    for i in range(n):                                    # test_sh_slopec.py:41-57
        for j in range(n):
            m = np.zeros((W, W))
            m[i * npx:(i + 1) * npx, j * npx:(j + 1) * npx] = 1
            idxs.append(np.ravel_multi_index(np.where(m == 1), m.shape))
    tv = np.radians(pp * pitch / 3600 * ps)               # test_sh_slopec.py:83-85
    tilt = np.broadcast_to(np.linspace(-tv / 2, tv / 2, pp), (pp, pp)) * 1e9
    return go, np.array(idxs), tilt, np.ones((pp, pp)), npx


def _lgs_flat_offset_slope(go, npx, positive):
    """Slope of the flat-wavefront LGS image: the kernel centre (N^2 - 1)/(2N) moves the image by
    1/(2N) FFT px (positive shift) or 1 - 1/(2N) (negative) from the NGS centre."""
    N = go['fft_size']
    d_fft = 1 / (2 * N) if positive else 1 - 1 / (2 * N)
    return d_fft * npx / go['cutsize'] * 2 / (npx - 1)


def test_reference_lgs_slope_known_answer():
    go, idxs, tilt, A, npx = _reference_slopec_case()
    assert (go['fft_size'], go['cutsize'], go['interpolate']) == (360, 174, True)
    K = O.gaussian_kernels_for_sh(go, 0.1)
    flat = O.shslopec(O.sh_forward(A, np.zeros_like(tilt), go, kernels=K), idxs)['slopes']
    tilted = O.shslopec(O.sh_forward(A, tilt, go, kernels=K), idxs)['slopes']
    n_sub = len(idxs)
    # the reference's own assertion (test_sh_slopec.py:109-114)
    np.testing.assert_allclose(tilted[:n_sub], 1.0 / (npx / 2), rtol=1e-2, atol=1e-2)
    # tighter: the tilt moves the spot by one CCD pixel along x only (slope 2/(npx-1) per pixel;
    # 0.5% of CoG bias from the 174 -> 120 LCM rebin); y slopes do not move
    np.testing.assert_allclose(tilted[:n_sub] - flat[:n_sub], 2 / (npx - 1), rtol=1e-2)
    np.testing.assert_allclose(tilted[n_sub:], flat[n_sub:], atol=1e-9)
    # flat wavefront: the 1/(2N) kernel-centring offset, in both axes
    np.testing.assert_allclose(flat, _lgs_flat_offset_slope(go, npx, True), rtol=5e-3)


def test_lgs_flat_offset_closed_form_and_sign():
    """Flat wavefront, Gaussian LLT, both shift signs, at a cut-free geometry (N = 32): the slopes
    are the kernel-centring offsets; the NGS image is centred (zero slopes)."""
    pitch = 0.05
    turb = 589e-9 / (8 * pitch) * O.RAD2ASEC
    kw = dict(wavelengthInNm=589, subap_wanted_fov=8 * turb, sensor_pxscale=turb / 2, subap_on_diameter=8,
              subap_npx=16, fov_ovs_coeff=1.0)
    go = O.sh_geometry(ef_size=64, pixel_pitch=pitch, fov_resolution_arcsec=turb / 4, **kw)
    assert go['cutpixels'] == 0 and go['fft_size'] == 32
    A = np.ones((64, 64))
    img = O.sh_forward(A, np.zeros((64, 64)), go)
    W, npx = 128, 16
    idxs = np.array([np.ravel_multi_index(np.mgrid[i * npx:(i + 1) * npx, j * npx:(j + 1) * npx].reshape(2, -1),
                                          (W, W)) for j in range(8) for i in range(8)])
    np.testing.assert_allclose(O.shslopec(img, idxs)['slopes'], 0, atol=1e-12)
    # 10%: in this 8-lambda/d field the periodic PSF wings reach the box edge, so part of a shift
    # wraps around and the CoG gain is ~0.93; the planted mistakes below move the slopes 30x more
    for positive in (True, False):
        K = O.gaussian_kernels_for_sh(go, 2.0 * turb, positive_shift_tt=positive)
        s = O.shslopec(O.sh_forward(A, np.zeros((64, 64)), go, kernels=K), idxs)['slopes']
        np.testing.assert_allclose(s, _lgs_flat_offset_slope(go, npx, positive), rtol=0.1)
    # planted mistake: kernels without the half-pixel centring (lgs_tt = 0)
    ps = go['fp4_pixel_pitch'] * O.RAD2ASEC
    real = O.lgs_map_sh(8, 3.2, 0, 90e3, [0], [1.0], 2.0 * turb, ps, 32, overs=1, theta=[0.0, 0.0])
    s0 = O.shslopec(O.sh_forward(A, np.zeros((64, 64)), go, kernels=O.process_kernels(real, 8, 8)), idxs)['slopes']
    assert np.abs(s0 / _lgs_flat_offset_slope(go, npx, True) - 1).min() > 10
