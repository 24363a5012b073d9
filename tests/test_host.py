"""Host-side logic of the product (CPU): geometry, binning operators, FITS, subaperture addressing.

The binning-operator test evaluates the GPU algorithm's mathematics in FP64 NumPy (pruned
M-point spectrum, then P S Q^T, then the Parseval or summed normalisation) and compares it
with the oracle, so the operators the plan uploads are checked without a GPU.
"""
import os
import tempfile

import numpy as np
import pytest

from oracle import sh_oracle as O
from specula_amd import workloads
from specula_amd.lib import fits
from specula_amd.lib.binning import build_operators, spec_size
from specula_amd.lib.geometry import sh_geometry
from specula_amd.lib.phasescreen import annular_pupil, von_karman_opd
from specula_amd.lib.subaps import block_origins, detect_subaps
from specula_amd.processing_objects.sh_slopec import compute_xy_weights

GEOMS = [
    # (ef, pitch, keys)
    (160, 0.05, workloads.scao_ngs_20x20().sh_kwargs),
    (120, 0.05, dict(wavelengthInNm=500, subap_wanted_fov=1.5, sensor_pxscale=0.25, subap_on_diameter=20,
                     subap_npx=6, fov_ovs_coeff=1.0)),
    (120, 0.05, dict(wavelengthInNm=500, subap_wanted_fov=1.5, sensor_pxscale=0.25, subap_on_diameter=20,
                     subap_npx=4, fov_ovs_coeff=1.0)),
    (120, 0.05, dict(wavelengthInNm=500, subap_wanted_fov=1.2, sensor_pxscale=0.2, subap_on_diameter=20,
                     subap_npx=6, fov_ovs_coeff=1.0)),
    (64, 0.1, dict(wavelengthInNm=700, subap_wanted_fov=2.0, sensor_pxscale=0.3, subap_on_diameter=16,
                   subap_npx=5, fov_ovs_coeff=1.0)),
]

SWEEP = [
    dict(wavelengthInNm=500, subap_wanted_fov=3, sensor_pxscale=0.5, subap_on_diameter=20, subap_npx=6),
    dict(wavelengthInNm=600, subap_wanted_fov=4.0, sensor_pxscale=0.5, subap_on_diameter=8, subap_npx=8),
    dict(wavelengthInNm=750, subap_wanted_fov=2.2, sensor_pxscale=0.3, subap_on_diameter=20, subap_npx=8,
         fov_ovs_coeff=1.0),
    dict(wavelengthInNm=750, subap_wanted_fov=2.2, sensor_pxscale=0.3, subap_on_diameter=20, subap_npx=8,
         do_not_double_fov_ovs=True),
    dict(wavelengthInNm=1650, subap_wanted_fov=2.0, sensor_pxscale=2.0, subap_on_diameter=10, subap_npx=4),
    dict(wavelengthInNm=500, subap_wanted_fov=3, sensor_pxscale=0.5, subap_on_diameter=20, subap_npx=6,
         FoVres30mas=True),
]


@pytest.mark.parametrize('kw', SWEEP)
def test_geometry_matches_oracle(kw):
    g = sh_geometry(120 if kw['subap_on_diameter'] != 8 else 160, 0.05, **kw)
    go = O.sh_geometry(ef_size=g.ef_size, pixel_pitch=0.05, **kw)
    assert (g.np_sub, g.ovs_np_sub, g.scale_ovs, g.fft_size, g.cutpixels, g.cutsize) == \
        (go['np_sub'], go['ovs_np_sub'], go['scale_ovs'], go['fft_size'], go['cutpixels'], go['cutsize'])
    assert g.fov_ovs == go['fov_ovs']
    assert g.fov_resolution_arcsec == go['fov_resolution_arcsec']
    np.testing.assert_array_equal(g.fp_mask, go['fp_mask'])


def test_workload_geometries():
    g = workloads.elt_ngs_80x80().geometry()
    assert (g.ovs_np_sub, g.fft_size, g.cutpixels, g.subap_npx, g.ccd_side) == (8, 64, 0, 16, 1280)
    g = workloads.lgs_40x40().geometry()
    assert (g.ovs_np_sub, g.fft_size, g.cutpixels, g.subap_npx, g.ccd_side) == (16, 128, 0, 16, 640)
    g = workloads.scao_ngs_20x20().geometry()
    assert (g.ovs_np_sub, g.fft_size, g.cutpixels, g.subap_npx, g.ccd_side) == (8, 32, 0, 8, 160)


def emulate_gpu_algorithm(A, ph, g, S0=1.0):
    row, col, M, norm = build_operators(g)
    n, p, N, npx = g.n_lenses, g.ovs_np_sub, g.fft_size, g.subap_npx
    a = np.arange(p) + (-p) // 2
    tilt = np.exp(-1j * np.pi * (a[None, :] + a[:, None]) / N)
    k = 2 * np.pi / g.wavelengthInNm
    out = np.zeros((n * npx, n * npx))
    for i in range(n):
        for j in range(n):
            sl = np.s_[i * p:(i + 1) * p, j * p:(j + 1) * p]
            E = A[sl] * np.exp(1j * k * ph[sl]) * tilt
            S = np.abs(np.fft.fft2(E, s=(M, M))) ** 2
            out[i * npx:(i + 1) * npx, j * npx:(j + 1) * npx] = row @ S @ col.T
    phot = S0 * g.pixel_pitch ** 2 * A.sum()
    scale = phot / (norm * (A ** 2).sum()) if norm is not None else phot / out.sum()
    return out * scale, norm


@pytest.mark.parametrize('case', range(len(GEOMS)))
def test_binning_operators_reproduce_oracle(case):
    ef, pitch, kw = GEOMS[case]
    g = sh_geometry(ef, pitch, **kw)
    go = O.sh_geometry(ef_size=ef, pixel_pitch=pitch, **kw)
    A = annular_pupil(ef, 0.2)
    A[3, 5:9] = 0.5                                     # a non-binary amplitude
    ph = von_karman_opd(256, pitch, seed=case)[:ef, :ef]
    out, norm = emulate_gpu_algorithm(A, ph, g, S0=2.0)
    ref = O.sh_forward(A, ph, go, S0=2.0)
    np.testing.assert_allclose(out, ref, rtol=0, atol=1e-12 * ref.max())
    assert (norm is not None) == (g.cutpixels == 0 and g.fp_mask.all())


def lag_row_operator(row, np_sub):
    """P~ = P A: columns [C0, Re C1..C_{np-1}, Im C1..C_{np-1}] of the lag-domain power spectrum."""
    npx, M = row.shape
    out = np.zeros((npx, M))
    ky = np.arange(M)
    out[:, 0] = row.sum(axis=1)
    for d in range(1, np_sub):
        th = 2 * np.pi * ((ky * d) % M) / M
        out[:, d] = 2 * row @ np.cos(th)
        out[:, np_sub - 1 + d] = 2 * row @ np.sin(th)
    return out


def _lag_domain_cases():
    """GEOMS entries the M = 16 lag-domain kernel serves (np_sub 5..8)."""
    out = []
    for c, (ef, pitch, kw) in enumerate(GEOMS):
        p = sh_geometry(ef, pitch, **kw).ovs_np_sub
        if spec_size(p) == 16 and p <= 8:
            out.append(c)
    return out


@pytest.mark.parametrize('case', _lag_domain_cases())
def test_lag_domain_binning_reproduces_oracle(case):
    """MFMA-kernel formulation: FFT_16 along x, autocorrelation along y (2np-1 lag reals), then
    B = (P A) C Q^T with the cosine/sine synthesis A folded into the row operator."""
    ef, pitch, kw = GEOMS[case]
    g = sh_geometry(ef, pitch, **kw)
    row, col, M, norm = build_operators(g)
    p = g.ovs_np_sub
    assert M == 16 and p <= 8
    go = O.sh_geometry(ef_size=ef, pixel_pitch=pitch, **kw)
    A = annular_pupil(ef, 0.2)
    ph = von_karman_opd(256, pitch, seed=case)[:ef, :ef]
    ref = O.sh_forward(A, ph, go, S0=1.0)
    rl = lag_row_operator(row, p)
    n, N, npx = g.n_lenses, g.fft_size, g.subap_npx
    a = np.arange(p) + (-p) // 2
    tilt = np.exp(-1j * np.pi * (a[None, :] + a[:, None]) / N)
    k = 2 * np.pi / g.wavelengthInNm
    out = np.zeros((n * npx, n * npx))
    for i in range(n):
        for j in range(n):
            sl = np.s_[i * p:(i + 1) * p, j * p:(j + 1) * p]
            R = np.fft.fft(A[sl] * np.exp(1j * k * ph[sl]) * tilt, n=M, axis=1)   # R[y][kx]
            C = np.zeros((M, M))
            C[0] = (np.abs(R) ** 2).sum(axis=0)
            for d in range(1, p):
                c = (R[d:] * np.conj(R[:p - d])).sum(axis=0)
                C[d], C[p - 1 + d] = c.real, c.imag
            out[i * npx:(i + 1) * npx, j * npx:(j + 1) * npx] = rl @ C @ col.T
    phot = g.pixel_pitch ** 2 * A.sum()
    out *= phot / (norm * (A ** 2).sum()) if norm is not None else phot / out.sum()
    np.testing.assert_allclose(out, ref, rtol=0, atol=1e-12 * ref.max())


def test_spec_size():
    assert [spec_size(n) for n in (4, 5, 6, 8, 9, 16, 17)] == [8, 16, 16, 16, 32, 32, 64]


ROUND_MASK = [
    # round FoV mask (squaremask=False, sh.py:282) clipping the cut window: not separable
    (120, 0.05, dict(GEOMS[1][2], squaremask=False)),
    (160, 0.05, dict(wavelengthInNm=750, subap_wanted_fov=2.4, sensor_pxscale=0.3, subap_on_diameter=20,
                     subap_npx=8, fov_ovs_coeff=1.0, squaremask=False)),
    (64, 0.015625, dict(wavelengthInNm=600, subap_wanted_fov=4.0, sensor_pxscale=0.5, subap_on_diameter=8,
                        subap_npx=8, fov_ovs_coeff=1.0, squaremask=False)),
]


@pytest.mark.parametrize('case', range(len(ROUND_MASK)))
def test_round_mask_terms_reproduce_oracle(case):
    """A non-separable FoV mask as K separable terms (one per distinct row pattern of the cut
    window): sum_k P_k S Q_k^T, normalised by its own sum, equals the oracle's image; the terms
    add up to the mask exactly; dropping one term (a plausible mistake) breaks the image."""
    from specula_amd.lib.binning import build_operator_terms, mask_is_separable, mask_terms
    ef, pitch, kw = ROUND_MASK[case]
    g = sh_geometry(ef, pitch, **kw)
    go = O.sh_geometry(ef_size=ef, pixel_pitch=pitch, **kw)
    assert not mask_is_separable(g)
    with pytest.raises(NotImplementedError):
        build_operators(g)
    terms = mask_terms(g)
    lo, cs = g.cutpixels // 2, g.cutsize
    W = sum(np.outer(a, b) for a, b in terms)
    np.testing.assert_array_equal(W[lo:lo + cs, lo:lo + cs], g.fp_mask[lo:lo + cs, lo:lo + cs])
    rows, cols, M, norm = build_operator_terms(g)
    assert norm is None and rows.shape[0] == len(terms) > 1
    A = annular_pupil(ef, 0.2)
    ph = von_karman_opd(256, pitch, seed=30 + case)[:ef, :ef]
    ref = O.sh_forward(A, ph, go, S0=1.0)
    n, p, N, npx = g.n_lenses, g.ovs_np_sub, g.fft_size, g.subap_npx
    a = np.arange(p) + (-p) // 2
    tilt = np.exp(-1j * np.pi * (a[None, :] + a[:, None]) / N)
    k = 2 * np.pi / g.wavelengthInNm

    def image(K):
        out = np.zeros((n * npx, n * npx))
        for i in range(n):
            for j in range(n):
                sl = np.s_[i * p:(i + 1) * p, j * p:(j + 1) * p]
                S = np.abs(np.fft.fft2(A[sl] * np.exp(1j * k * ph[sl]) * tilt, s=(M, M))) ** 2
                out[i * npx:(i + 1) * npx, j * npx:(j + 1) * npx] = sum(rows[t] @ S @ cols[t].T for t in K)
        return out * (g.pixel_pitch ** 2 * A.sum() / out.sum())
    out = image(range(rows.shape[0]))
    np.testing.assert_allclose(out, ref, rtol=0, atol=1e-12 * ref.max())
    bad = image(range(1, rows.shape[0]))
    assert np.abs(bad - ref).max() > 1e-4 * ref.max()


def test_fits_round_trip_and_subapdata():
    from specula_amd.data_objects.subap_data import SubapData
    with tempfile.TemporaryDirectory() as d:
        f = os.path.join(d, 'x.fits')
        arrs = [np.arange(6, dtype=np.float32).reshape(2, 3), np.arange(5, dtype=np.int64),
                np.linspace(0, 1, 7)]
        fits.write(f, [(None, {'VERSION': 1, 'NAME': "a'b", 'FLAG': True, 'X': 1.5})] + [(a, {}) for a in arrs])
        h = fits.read(f)
        assert h[0][1]['VERSION'] == 1 and h[0][1]['NAME'] == "a'b" and h[0][1]['FLAG'] is True
        assert h[0][1]['X'] == 1.5
        for a, (b, _) in zip(arrs, h[1:]):
            np.testing.assert_array_equal(a, b)
        idxs = np.arange(12).reshape(3, 4)
        sd = SubapData(idxs, np.array([0, 1, 3]), nx=2, ny=2, energy_th=0.5, target_device_idx=-1)
        sd.save(f)
        sd2 = SubapData.restore(f, target_device_idx=-1)
        np.testing.assert_array_equal(sd2.idxs, idxs)
        np.testing.assert_array_equal(sd2.display_map, [0, 1, 3])
        assert (sd2.nx, sd2.ny, sd2.energy_th, sd2.np_sub) == (2, 2, 0.5, 2)


def test_subapdata_fixture_restore():
    from specula_amd.data_objects.subap_data import SubapData
    sd = SubapData.restore(os.path.join(os.path.dirname(__file__), 'golden', 'scao_subaps_n8_th0.5_ref.fits'),
                           target_device_idx=-1)
    assert sd.n_subaps == 52 and sd.np_sub == 8 and sd.nx == 8
    assert block_origins(sd.idxs, 64) is not None


def test_detect_subaps_matches_oracle():
    rng = np.random.default_rng(0)
    img = rng.random((96, 96))
    img[:30, :30] *= 0.1
    i1, d1 = detect_subaps(img, 12, 0.5)
    i2, d2 = O.detect_subaps(img, 12, 0.5)
    np.testing.assert_array_equal(i1, i2)
    np.testing.assert_array_equal(d1, d2)


def test_block_origins():
    idxs, _ = O.detect_subaps(np.ones((40, 40)), 5, 0.5)
    org = block_origins(idxs, 40)
    np.testing.assert_array_equal(org, idxs[:, 0])
    assert block_origins(idxs[:, ::-1], 40) is None
    # a block that would wrap around the right edge is not a block
    assert block_origins(np.array([[38, 39, 78, 79 + 1]]), 40) is None


@pytest.mark.parametrize('e', [1.0, 1.5])
def test_xy_weights_match_oracle(e):
    a = compute_xy_weights(8, e)
    b = O.compute_xy_weights(8, e)
    for k in ('x', 'y', 'mask_weighted'):
        np.testing.assert_array_equal(a[k], b[k])


def test_von_karman_structure_function():
    """D(r) = 6.88 (r/r0)^(5/3) for r << L0 (phase at 500 nm), checked loosely on one screen."""
    pitch, r0 = 0.02, 0.15
    opd = von_karman_opd(1024, pitch, r0=r0, L0=1e4, seed=3)
    ph = opd * 2 * np.pi / 500.0
    for lag in (4, 8):
        d = np.mean((ph[:, lag:] - ph[:, :-lag]) ** 2)
        expect = 6.88 * (lag * pitch / r0) ** (5 / 3)
        assert 0.6 < d / expect < 1.4


# ---------------------------------------------------------------- field resampling (host side)

def test_edge_extrapolation_maps_match_golden_and_oracle():
    from specula_amd.lib.resample import edge_extrapolation_maps
    golden = os.path.join(os.path.dirname(__file__), 'golden')
    mask = fits.read(os.path.join(golden, 'mask.fits'))[0][0]
    for two, name in ((False, 'extrapol_array_doExt2PixFalse.fits'), (True, 'extrapol_array_doExt2PixTrue.fits')):
        s1, s2 = edge_extrapolation_maps(mask, two_pixel=two)
        h = fits.read(os.path.join(golden, name))
        np.testing.assert_array_equal(s1, h[0][0].astype(np.int32))
        np.testing.assert_array_equal(s2, h[1][0].astype(np.int32))
    rng = np.random.default_rng(3)
    for k in range(10):   # masks touching the borders exercise the periodic neighbour test
        m = (rng.random((23, 29)) < 0.45).astype(float)
        a1, a2, _, _ = O.extrapolate_edge_pixel_mat_define(m, do_ext_2_pix=True)
        b1, b2 = edge_extrapolation_maps(m, two_pixel=True)
        np.testing.assert_array_equal(a1, b1)
        np.testing.assert_array_equal(a2, b2)
    for k in range(6):    # grey amplitudes: (1 - A) > 0 inside the pupil pulls interior pixels in
        m = (rng.random((23, 29)) < 0.5) * rng.choice([0.5, 1.0, 0.25], size=(23, 29))
        a1, a2, _, _ = O.extrapolate_edge_pixel_mat_define(m, do_ext_2_pix=True)
        b1, b2 = edge_extrapolation_maps(m, two_pixel=True)
        np.testing.assert_array_equal(a1, b1)
        np.testing.assert_array_equal(a2, b2)
    half = 0.5 * fits.read(os.path.join(golden, 'mask.fits'))[0][0]
    b1, _ = edge_extrapolation_maps(half, two_pixel=True)
    assert np.count_nonzero((b1 >= 0) & (half > 0)) > 0     # interior pixels of a grey pupil selected


@pytest.mark.parametrize('rot,rs,cs', [(0.0, 0, 0), (12.5, 2, -3), (-40.0, 0, 1)])
def test_resampling_map_matches_oracle_coordinates(rot, rs, cs):
    from specula_amd.lib.resample import resampling_map
    yy, xx = O.interp2d_coords((40, 48), (90, 100), rot, rs, cs)
    c = resampling_map((40, 48), (90, 100), rot, rs, cs)
    Y, X = np.mgrid[0:90, 0:100]
    np.testing.assert_allclose(np.clip(c[0] * Y + c[1] * X + c[2], 0, 39), yy, rtol=0, atol=1e-12)
    np.testing.assert_allclose(np.clip(c[3] * Y + c[4] * X + c[5], 0, 47), xx, rtol=0, atol=1e-12)


def test_geometry_rotation_shift_keys():
    kw = dict(wavelengthInNm=600, subap_wanted_fov=4.0, sensor_pxscale=0.5, subap_on_diameter=8, subap_npx=8)
    g = sh_geometry(64, 0.015625, **kw)
    go = O.sh_geometry(ef_size=64, pixel_pitch=0.015625, **kw)
    assert g.needs_interpolation and go['interpolate'] and g.ef_ovs == go['ef_ovs'] == 128
    kw1 = dict(kw, fov_ovs_coeff=1.0, xShiftPhInPixel=0.6, rotAnglePhInDeg=5.0, aRotAnglePhInDeg=1.0)
    g = sh_geometry(64, 0.015625, **kw1)
    go = O.sh_geometry(ef_size=64, pixel_pitch=0.015625, **kw1)
    assert g.fov_ovs == go['fov_ovs'] == 1 and g.needs_interpolation
    assert g.rot_deg == go['rot_deg'] == 6.0 and tuple(g.xy_shift) == tuple(go['xy_shift']) == (1, 0)
    g = sh_geometry(64, 0.015625, **dict(kw, fov_ovs_coeff=1.0))
    assert not g.needs_interpolation
