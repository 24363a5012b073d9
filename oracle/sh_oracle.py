"""FP64 NumPy restatement of the reference SH / ShSlopec / toccd / LGS-kernel algorithms.

TEST INFRASTRUCTURE ONLY.  Imported by ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` -- never by the product package ``specula_amd``.

Every function cites the reference file:line it follows (paths relative to the
reference checkout).  The restatement is deliberately literal: same loop structure
(per-subaperture-row FFT batches), same FFT library (numpy pocketfft), same rebin
index arithmetic, so that it reproduces the reference's numbers.  Deviations:

* ``cutpixels == 0``: the reference slice ``psf[:, c//2:-c//2]`` is empty and raises
  (sh.py:405,410).  Here the slice end is ``N - ceil(c/2)``, which equals the
  reference's end for every ``c >= 1`` and is ``N`` for ``c == 0``.
* The interpolated path (fov_ovs != 1, rotation or shift: sh.py:350-360, 450-460) is
  restated with FP64 sampling coordinates (the reference builds them in float32 when
  precision is single).

Pinning (what checks each function; see tests/test_oracle_pins.py):
  make_xy / make_mask ....... closed forms + reference docstring examples (make_xy.py:93-119)
  sh_geometry ............... hand-derived geometries (survey cfg0/cfg1/cfg2, test_sh.py
                              geometry, params_scao_sh_test.yml geometry: N=32 cut 24)
  sh_forward (NGS) .......... brute-force O(N^4) DFT on tiny N, Parseval flux known answer
                              (test_sh.py:18-42), 1-px tilt shift (test_sh.py:44-95)
  sh_forward (LGS) .......... direct spatial circular correlation on tiny N, flux preservation
  toccd / rebin2d ........... block sums (integer ratio), brute-force replicate-then-mean
                              (LCM ratio), total conservation
  shslopec .................. hand-computed tiny cases (threshold, clamp, 1e-3*mean cut),
                              1-px tilt slope 2/(npx-1) (test_sh_slopec.py:109-114)
  detect_subaps ............. idx layout / ordering of scao_subaps_n8_th0.5_ref.fits
  lgs_map_sh ................ IDL golden lgs_map_sh_ref.fits (test_convolution_kernel.py:175-276)
  extrapolate_edge_pixel_mat_define / extrapolate_edge_pixel
                              IDL goldens extrapol_array_doExt2Pix{False,True}.fits and
                              extrapolated{1,2}.fits (test_extrapol_mat_define.py,
                              test_extrapolate_edge_pixel.py)
  interp2d .................. IDL golden ref_phase.fits (test_interp2d.py:19-43), exact on
                              affine functions, no reads outside the input
  sh_forward (interpolated) . flat-wavefront slopes of scao_sn_n8_th0.5_ref.fits (fov_ovs 2),
                              reduces to the plain path for an identity resampling
  zernike_degree / zernike_map  Noll's table (closed forms of Z1..Z11), orthonormality on a fine
                              disk
  compute_zern_ifunc, dm_phase, pushpull_time_history, accumulate_intmat, recmat_fits_data
                              the reference's calibration golden scao_rec_n8_th0.5_ref.fits
                              (test_sh_calibration.py, params_scao_sh_test{,_rec}.yml) to
                              2.6e-5 (its bar: decimal 3), which also pins sh_forward at fov_ovs 2
                              with cut 24 and shslopec end to end
  ccd_readout ............... closed forms (block sums x binning^2, QE), Poisson moments, the
                              integer frame after Poisson (readout truncation), the ADU clamp;
                              statistical only for the noise streams (numpy generators)
"""
import math

import numpy as np

ASEC2RAD = np.pi / (3600 * 180)          # specula/__init__.py:21
RAD2ASEC = 1.0 / ASEC2RAD                # specula/__init__.py:22


# --------------------------------------------------------------------------- grids / masks

def make_xy(sampling, ratio, dtype=None):
    """specula/lib/make_xy.py:12-177 (cartesian, not quarter, not fft-ordered)."""
    if sampling <= 1:
        raise ValueError('make_xy: sampling must be larger than 1')
    x0 = (sampling - 1) / 2.0                                   # make_xy.py:157
    x = (np.arange(sampling, dtype=dtype) - x0) / (sampling / 2.0) * ratio   # :161
    x = np.tile(x, (sampling, 1))                               # :172
    y = np.transpose(x).copy()                                  # :173
    return x, y


def make_mask(np_size, obsratio=0.0, diaratio=1.0, xc=0.0, yc=0.0, square=False, inverse=False):
    """specula/lib/make_mask.py:7-147 (no spider, no centeronpixel)."""
    if diaratio is None:
        diaratio = 1.0
    if obsratio is None:
        obsratio = 0.0
    x, y = make_xy(np_size, 1.0)
    ir = obsratio
    if square:                                                  # make_mask.py:60-64
        sel = ((np.abs(x - xc) <= diaratio) & (np.abs(y - yc) <= diaratio) &
               ((np.abs(x - xc) >= diaratio * ir) | (np.abs(y - yc) >= diaratio * ir)))
    else:                                                       # make_mask.py:65-69
        r2 = (x - xc) ** 2 + (y - yc) ** 2
        sel = (r2 < diaratio ** 2) & (r2 >= (diaratio * ir) ** 2)
    mask = np.zeros((np_size, np_size), dtype=np.uint8)
    mask[sel] = 1
    if inverse:
        mask = 1 - mask
    return mask


def lenslet(n_lenses):
    """specula/data_objects/lenslet.py:6-39 -> lenses[i][j] = [x(j), y(i), 2/n]."""
    if n_lenses > 1:
        x, y = make_xy(n_lenses, 1.0)
    else:
        x = np.zeros((1, 1))
        y = np.zeros((1, 1))
    size = 2.0 / n_lenses
    return [[[x[i, j], y[i, j], size] for j in range(n_lenses)] for i in range(n_lenses)]


# --------------------------------------------------------------------------- geometry

def sh_geometry(ef_size, pixel_pitch, wavelengthInNm, subap_wanted_fov, sensor_pxscale,
                subap_on_diameter, subap_npx, FoVres30mas=False, squaremask=True,
                fov_ovs_coeff=0.0, do_not_double_fov_ovs=False, set_fov_res_to_turbpxsc=False,
                fov_resolution_arcsec=None, xShiftPhInPixel=0, yShiftPhInPixel=0,
                aXShiftPhInPixel=0, aYShiftPhInPixel=0, rotAnglePhInDeg=0, aRotAnglePhInDeg=0):
    """Restates SH.__init__ (sh.py:43-110), set_in_ef (sh.py:112-237) and the geometry part of
    calc_trigger_geometry (sh.py:239-282).  Angles in arcsec as in the YAML keys.
    ``fov_resolution_arcsec`` mirrors assigning ``sh._fov_resolution_arcsec`` before setup()."""
    lens = lenslet(subap_on_diameter)
    n_lenses = subap_on_diameter
    dimx = len(lens)
    _subap_wanted_fov = subap_wanted_fov / RAD2ASEC                  # sh.py:68
    _sensor_pxscale = sensor_pxscale / RAD2ASEC                      # sh.py:69
    _fov_resolution_arcsec = 0.03 if FoVres30mas else 0              # sh.py:73
    if fov_resolution_arcsec is not None:
        _fov_resolution_arcsec = fov_resolution_arcsec
    _fov_ovs_coeff = fov_ovs_coeff
    l0 = lens[0][0]

    _np_sub = max([1, round((ef_size * l0[2]) / 2.0)])               # sh.py:118
    if _np_sub * n_lenses > ef_size:                                 # sh.py:119-120
        _np_sub -= 1
    np_sub = (ef_size * l0[2]) / 2.0                                 # sh.py:122
    sensor_pxscale_arcsec = _sensor_pxscale * RAD2ASEC
    dSubApInM = np_sub * pixel_pitch
    turbulence_pxscale = wavelengthInNm * 1e-9 / dSubApInM * RAD2ASEC
    subap_wanted_fov_arcsec = _subap_wanted_fov * RAD2ASEC
    subap_real_fov_arcsec = _sensor_pxscale * subap_npx * RAD2ASEC

    if _fov_resolution_arcsec == 0:                                  # sh.py:130-179
        if set_fov_res_to_turbpxsc:
            if turbulence_pxscale >= sensor_pxscale_arcsec:
                raise ValueError('set_fov_res_to_turbpxsc property should be set to one only if '
                                 'turb. pix. sc. is < sensor pix. sc.')
            _fov_resolution_arcsec = turbulence_pxscale
        elif turbulence_pxscale < sensor_pxscale_arcsec and sensor_pxscale_arcsec / 2.0 > 0.5:
            _fov_resolution_arcsec = turbulence_pxscale * 0.5
        else:
            i = 0
            resTry = turbulence_pxscale / (i + 2)
            while resTry >= sensor_pxscale_arcsec:
                i += 1
                resTry = turbulence_pxscale / (i + 2)
            iMin = i
            nTry = 10
            resTry = np.zeros(nTry)
            scaleTry = np.zeros(nTry)
            fftScaleTry = np.zeros(nTry)
            subapRealTry = np.zeros(nTry)
            mcmxTry = np.zeros(nTry)
            for i in range(nTry):
                resTry[i] = turbulence_pxscale / (iMin + i + 2)
                scaleTry[i] = round(turbulence_pxscale / resTry[i])
                fftScaleTry[i] = (wavelengthInNm / 1e9 * dimx / (ef_size * pixel_pitch * scaleTry[i])
                                  * RAD2ASEC)
                subapRealTry[i] = round(subap_wanted_fov_arcsec / fftScaleTry[i] / 2.0) * 2
                mcmxTry[i] = np.lcm(int(subap_npx), int(subapRealTry[i]))
            FoV = subapRealTry * fftScaleTry
            FoVerror = np.abs(FoV - subap_wanted_fov_arcsec) / subap_wanted_fov_arcsec
            idxGood = np.where(FoVerror < 0.02)[0]
            if len(idxGood) == 0:
                idxGood = np.arange(nTry)
            ratioMcm = mcmxTry[idxGood] / scaleTry[idxGood]
            idxMin = np.argmin(ratioMcm)
            if (idxGood[idxMin] != 0 and mcmxTry[idxGood[0]] / mcmxTry[idxGood[idxMin]]
                    > scaleTry[idxGood[idxMin]] / scaleTry[idxGood[0]]):
                _fov_resolution_arcsec = resTry[idxGood[idxMin]]
            else:
                _fov_resolution_arcsec = resTry[idxGood[0]]

    scale_ovs = round(turbulence_pxscale / _fov_resolution_arcsec)  # sh.py:185
    dTelPaddedInM = ef_size * pixel_pitch * scale_ovs
    dSubApPaddedInM = dTelPaddedInM / dimx
    fft_pxscale_arcsec = wavelengthInNm * 1e-9 / dSubApPaddedInM * RAD2ASEC
    subap_real_fov_pix = round(subap_real_fov_arcsec / fft_pxscale_arcsec / 2.0) * 2.0   # :192
    subap_real_fov_arcsec = subap_real_fov_pix * fft_pxscale_arcsec
    turbulence_fov_pix = int(scale_ovs * np_sub)                     # :196

    fov_ovs = 1
    if turbulence_fov_pix > 2 * subap_real_fov_pix:                  # :199-215
        fov_ovs = 1
        if _fov_ovs_coeff != 0.0:
            fov_ovs = _fov_ovs_coeff
    else:
        ratio = float(subap_real_fov_pix) / float(turbulence_fov_pix)
        np_factor = 1 if abs(np_sub - int(np_sub)) < 1e-3 else int(np_sub)
        if do_not_double_fov_ovs and _fov_ovs_coeff == 0.0:
            _fov_ovs_coeff = 1.0
            fov_ovs = np.ceil(np_factor * ratio / 2.0) * 2.0 / float(np_factor)
        else:
            if _fov_ovs_coeff == 0.0:
                _fov_ovs_coeff = 2.0
            if ratio < 2:
                fov_ovs = np.ceil(np_factor * _fov_ovs_coeff) / float(np_factor)
            else:
                fov_ovs = np.ceil(np_factor * ratio * _fov_ovs_coeff) / float(np_factor)

    _sensor_pxscale = subap_real_fov_arcsec / subap_npx / RAD2ASEC   # :217
    ovs_np_sub = int(ef_size * fov_ovs * l0[2] * 0.5)                # :218
    fft_size = ovs_np_sub * scale_ovs                                # :219

    chk = (ef_size * round(fov_ovs) * l0[2]) / 2.0                   # :234
    if abs(chk - round(chk)) > 1e-4:
        raise ValueError(f'ERROR: interpolated input phase size {ef_size} * {round(fov_ovs)} '
                         f'is not divisible by  {n_lenses} subapertures.')

    # calc_trigger_geometry (sh.py:239-282)
    rot0 = rotAnglePhInDeg + aRotAnglePhInDeg
    shift0 = np.round(np.array([xShiftPhInPixel + aXShiftPhInPixel,
                                yShiftPhInPixel + aYShiftPhInPixel]) * fov_ovs).astype(int)
    resample = fov_ovs != 1 or rot0 != 0 or np.sum(np.abs(shift0)) != 0
    wf1_pitch = pixel_pitch / fov_ovs if resample else pixel_pitch    # :254-259
    fp4_pixel_pitch = wavelengthInNm / 1e9 / (wf1_pitch * fft_size)      # :268
    fov_complete = fft_size * fp4_pixel_pitch
    sensor_subap_fov = _sensor_pxscale * subap_npx
    fov_cut = fov_complete - sensor_subap_fov
    cutpixels = int(np.round(fov_cut / fp4_pixel_pitch) / 2 * 2)         # :274
    cutsize = fft_size - cutpixels
    if cutpixels < 0:
        raise ValueError(f'sensor FoV exceeds the FFT FoV (cutpixels={cutpixels})')
    fp_mask = make_mask(fft_size, diaratio=_subap_wanted_fov / fov_complete, square=squaremask)  # :282

    # phase resampling (calc_trigger_geometry, sh.py:247-254; setup, sh.py:448-452)
    rot = rotAnglePhInDeg + aRotAnglePhInDeg
    xy_shift = np.round(np.array([xShiftPhInPixel + aXShiftPhInPixel,
                                  yShiftPhInPixel + aYShiftPhInPixel]) * fov_ovs).astype(int)
    interpolate = fov_ovs != 1 or rot != 0 or np.sum(np.abs(xy_shift)) != 0

    return dict(ef_size=ef_size, pixel_pitch=pixel_pitch, wavelengthInNm=wavelengthInNm,
                interpolate=bool(interpolate), rot_deg=rot, xy_shift=xy_shift,
                ef_ovs=int(ef_size * fov_ovs), wf1_pitch=wf1_pitch,
                n_lenses=n_lenses, np_sub=_np_sub, ovs_np_sub=ovs_np_sub, fov_ovs=fov_ovs,
                scale_ovs=scale_ovs, fft_size=fft_size, cutpixels=cutpixels, cutsize=cutsize,
                subap_npx=subap_npx, ccd_side=subap_npx * n_lenses, fp_mask=fp_mask,
                fp4_pixel_pitch=fp4_pixel_pitch, fov_complete=fov_complete,
                fov_resolution_arcsec=_fov_resolution_arcsec, sensor_pxscale=_sensor_pxscale,
                turbulence_pxscale=turbulence_pxscale)


def tilt_f(p, c):
    """SH.get_tlt_f, sh.py:469-477: half-pixel tilt exp(-2 pi i (x+y) / (2 (p+c)))."""
    a = np.arange(-p // 2, p // 2)
    xx, yy = np.meshgrid(a, a)
    return np.exp(-2 * np.pi * 1j * (xx + yy) / (2 * (p + c)))


# --------------------------------------------------------------------------- rebin / toccd

def rebin2d(a, shape, sample=False):
    """specula/lib/rebin.py:12-35."""
    if a.shape == tuple(shape):
        return a
    if sample:
        slices = [slice(0, old, float(old) / new) for old, new in zip(a.shape, shape)]
        idx = np.mgrid[slices].astype(np.int32)
        return a[tuple(idx)]
    M, N = a.shape
    m, n = shape
    if m <= M and n <= M:
        if (M // m != M / m) or (N // n != N / n):
            raise ValueError('Resampling by non-integer factors is not supported')
        return a.reshape((m, M // m, n, N // n)).mean(3).mean(1)
    raise ValueError('Upsampling with sample=False is not supported')


def toccd(a, newshape, set_total=None):
    """specula/lib/toccd.py:33-63 (CPU path)."""
    newshape = tuple(newshape)
    if a.shape == newshape:
        return a
    if set_total is None:
        set_total = a.sum()
    mcmx = math.lcm(a.shape[0], newshape[0])
    mcmy = math.lcm(a.shape[1], newshape[1])
    temp = rebin2d(a, (mcmx, a.shape[1]), sample=True)
    temp = rebin2d(temp, (newshape[0], a.shape[1]))
    temp = rebin2d(temp, (newshape[0], mcmy), sample=True)
    rebinned = rebin2d(temp, newshape)
    return rebinned / rebinned.sum() * set_total


# --------------------------------------------------------------------------- SH forward

def sh_forward(A, phase_nm, geom, S0=1.0, kernels=None, return_psfimage=False, ext_mask=None):
    """SH.trigger_code + post_trigger (sh.py:348-428) in FP64.

    ``kernels``: (n_lenses**2, N, N) complex array as produced by ``process_kernels``
    (ConvolutionKernel.kernels), or None for the NGS path.
    ``ext_mask``: with resampling, the amplitude the edge-extrapolation maps were built from
    (in_ef.A at set-up, sh.py:453); None = ``A``.
    Returns the CCD intensity (ccd_side x ccd_side) float64."""
    A = np.asarray(A, dtype=np.float64)
    phase_nm = np.asarray(phase_nm, dtype=np.float64)
    A_in = A
    if geom.get('interpolate', False):                              # sh.py:350-360
        A, phase_nm = resample_field(A, phase_nm, geom, ext_mask)
    n = geom['n_lenses']
    p = geom['ovs_np_sub']
    N = geom['fft_size']
    c = geom['cutpixels']
    cs = geom['cutsize']
    lam = geom['wavelengthInNm']
    tltf = tilt_f(p, N - p)                                          # sh.py:280
    fp_mask = geom['fp_mask'].astype(np.float64)
    wf3 = np.zeros((n, N, N), dtype=np.complex128)                   # sh.py:265
    psfimage = np.zeros((cs * n, cs * n))                            # sh.py:276
    lo, hi = c // 2, N - (c - c // 2)                                # slice fix (see header)
    for i in range(n):                                               # sh.py:364
        sl = slice(i * p, (i + 1) * p)
        phi = phase_nm[sl, :] * ((2 * np.pi) / lam)                  # electric_field.py:65
        ef_row = np.exp(1j * phi) * A[sl, :]                         # electric_field.py:73-74
        cube = ef_row.reshape(p, n, p).swapaxes(0, 1)                # sh.py:373
        wf3[:, :p, :p] = cube * tltf[np.newaxis, :, :]               # sh.py:376
        fp4 = np.fft.fft2(wf3, axes=(1, 2))                          # sh.py:378
        psf_shifted = np.real(fp4 * np.conj(fp4))                    # sh.py:25-27,379
        if kernels is not None:                                      # sh.py:382-394
            kern = kernels[i * n:(i + 1) * n, :, :]
            psf_fft = np.fft.fft2(psf_shifted) * kern
            psf = np.fft.ifft2(psf_fft, norm='forward').real         # in-place scipy ifft2, norm='forward'
        else:
            psf = np.fft.fftshift(psf_shifted, axes=(1, 2))          # sh.py:396
        psf = psf * fp_mask[np.newaxis, :, :]                        # sh.py:399
        cut = psf[:, lo:hi, lo:hi]                                   # sh.py:405 (fixed)
        psfimage[i * cs:(i + 1) * cs, :] = cut.swapaxes(0, 1).reshape(-1, n * cs)   # sh.py:410-413
    ccd_side = geom['ccd_side']
    out = toccd(psfimage, (ccd_side, ccd_side))                      # sh.py:420
    phot = S0 * (geom['pixel_pitch'] ** 2) * np.sum(A_in)            # sh.py:426 (in_ef, not wf1); electric_field.py:87-89
    out = out * (phot / out.sum())                                   # sh.py:427
    if return_psfimage:
        return out, psfimage
    return out


# --------------------------------------------------------------------------- phase resampling

def extrapolate_edge_pixel_mat_define(mask, do_ext_2_pix=False):
    """extrapolate_edge_pixel_mat_define.py:5-140 (loop for loop).  Returns int32 maps of the
    source pixel (linear index) for 1-pixel and 2-pixel linear extrapolation (-1: none), the
    indices of the pixels to fill, and the positions among them with a 2-pixel source."""
    smask = mask.shape
    idx_mask = np.where(mask)
    float_mask = mask.astype(float)
    s1 = np.full(float_mask.shape, -1.0)
    s2 = np.full(float_mask.shape, -1.0)
    lin = np.zeros_like(float_mask)
    lin[idx_mask] = np.ravel_multi_index(idx_mask, smask)
    find1 = (np.roll(float_mask, 1, axis=0) + np.roll(float_mask, -1, axis=0) +     # :26-28
             np.roll(float_mask, 1, axis=1) + np.roll(float_mask, -1, axis=1)) * (1.0 - float_mask)
    find2 = (np.roll(float_mask, 2, axis=0) + np.roll(float_mask, -2, axis=0) +     # :31-33
             np.roll(float_mask, 2, axis=1) + np.roll(float_mask, -2, axis=1)) * (1.0 - float_mask)
    idx1 = np.where(find1 > 0)
    idx2 = np.where((find2 > 0) & (find1 == 0))

    def visit(idx, step):                 # one pixel, the four directions in the reference order
        r, c = idx
        test = -1
        near, far = step, step + 1
        arr = s1 if step == 1 else s2
        # (row offset, col offset) for "y+", "y-", "x+", "x-" (:45-79 / :93-127)
        for dr, dc, ok_near, ok_far in (
                (0, 1, c + near < smask[1] - 1, c + far < smask[1] - 1),
                (0, -1, c - near > 0, c - far > 0),
                (1, 0, r + near < smask[0] - 1, r + far < smask[0] - 1),
                (-1, 0, r - near > 0, r - far > 0)):
            if ok_near and arr[idx] == -1:
                if float_mask[r + dr * near, c + dc * near] > 0:
                    if ok_far and float_mask[r + dr * far, c + dc * far] > 0:
                        s2[idx] = lin[r + dr * far, c + dc * far]
                        s1[idx] = lin[r + dr * near, c + dc * near]
                    else:
                        test = lin[r + dr * near, c + dc * near]
        if s1[idx] == -1 and test >= 0:                                     # :81-82
            s1[idx] = test

    for idx in zip(*idx1):
        visit(idx, 1)
    if do_ext_2_pix:
        for idx in zip(*idx2):
            visit(idx, 2)
    idx_1pix = np.where(s1 >= 0)                                            # :134-135
    idx_2pix = np.where(s2[idx_1pix] >= 0)
    return s1.astype(np.int32), s2.astype(np.int32), idx_1pix, idx_2pix


def extrapolate_edge_pixel(phase, s1, s2, idx_1pix, idx_2pix):
    """extrapolate_edge_pixel.py:4-29: fill the pixels just outside the pupil with the value of
    the nearest pupil pixel, or the linear extrapolation of the two nearest along one axis."""
    out = np.array(phase, dtype=np.float64, copy=True)
    flat = out.ravel().copy()
    v1 = flat[s1[idx_1pix]]
    v2 = flat[s2[idx_1pix]]
    v1[idx_2pix] = 2 * v1[idx_2pix] - v2[idx_2pix]
    out[idx_1pix] = v1
    return out


def interp2d_coords(input_shape, output_shape, rot_deg=0.0, row_shift=0.0, col_shift=0.0):
    """Interp2D.__init__ sampling points, interp2d.py:48-93 (default grid): (yy, xx) in input
    pixel coordinates, clamped to the input."""
    yy, xx = np.mgrid[0:output_shape[0], 0:output_shape[1]].astype(np.float64)
    yy *= (input_shape[0] - 1) / output_shape[0]                            # :64
    xx *= (input_shape[1] - 1) / output_shape[1]
    if rot_deg != 0:                                                         # :73-81
        yc = input_shape[0] / 2 - 0.5
        xc = input_shape[1] / 2 - 0.5
        cos_ = np.cos(rot_deg * 3.1415 / 180.0)
        sin_ = np.sin(rot_deg * 3.1415 / 180.0)
        xxr = (xx - xc) * cos_ - (yy - yc) * sin_
        yyr = (xx - xc) * sin_ + (yy - yc) * cos_
        xx, yy = xxr + xc, yyr + yc
    if row_shift != 0 or col_shift != 0:                                     # :83-85
        yy = yy + row_shift
        xx = xx + col_shift
    yy = np.clip(yy, 0, input_shape[0] - 1)                                  # :87-90
    xx = np.clip(xx, 0, input_shape[1] - 1)
    return yy, xx


def interp2d(value, yy, xx):
    """Bilinear interpolation on the integer grid (Interp2D.interpolate CPU path, interp2d.py:
    117-122: scipy RegularGridInterpolator, method 'linear'); points on the last grid line use
    the last cell with weight 1."""
    value = np.asarray(value, dtype=np.float64)
    n0, n1 = value.shape
    y = np.asarray(yy, dtype=np.float64)
    x = np.asarray(xx, dtype=np.float64)
    i0 = np.clip(np.floor(y).astype(np.int64), 0, n0 - 2)
    j0 = np.clip(np.floor(x).astype(np.int64), 0, n1 - 2)
    ty, tx = y - i0, x - j0
    return ((1 - ty) * (1 - tx) * value[i0, j0] + (1 - ty) * tx * value[i0, j0 + 1] +
            ty * (1 - tx) * value[i0 + 1, j0] + ty * tx * value[i0 + 1, j0 + 1])


def resample_field(A, phase_nm, geom, ext_mask=None):
    """SH.trigger_code interpolation block, sh.py:350-360, with the setup of sh.py:448-458:
    edge extrapolation of the phase (mask = in_ef.A at set-up, 2-pixel extension), then A and
    the extrapolated phase resampled on the ef*fov_ovs grid (rotation and rounded shifts)."""
    ef = A.shape[0]
    mask = A if ext_mask is None else np.asarray(ext_mask, dtype=np.float64)
    s1, s2, i1, i2 = extrapolate_edge_pixel_mat_define(mask, do_ext_2_pix=True)
    ph = extrapolate_edge_pixel(phase_nm, s1, s2, i1, i2)
    out = (geom['ef_ovs'], geom['ef_ovs'])
    yy, xx = interp2d_coords((ef, ef), out, geom['rot_deg'], geom['xy_shift'][0], geom['xy_shift'][1])
    return interp2d(A, yy, xx), interp2d(ph, yy, xx)


# --------------------------------------------------------------------------- ShSlopec

def compute_xy_weights(np_sub, exp_weight=1.0):
    """ShSlopec.computeXYweights, sh_slopec.py:86-132 (quadcell off, weightedPixRad == 0)."""
    x, y = np.meshgrid(np.linspace(-1, 1, np_sub), np.linspace(-1, 1, np_sub))   # :98
    xc, yc = x, y
    x = np.where(x > 0, np.power(x, exp_weight), -np.power(np.abs(x), exp_weight))   # :108
    y = np.where(y > 0, np.power(y, exp_weight), -np.power(np.abs(y), exp_weight))
    xc = np.where(x > 0, xc, -np.abs(xc))
    yc = np.where(y > 0, yc, -np.abs(yc))
    mask_weighted = np.ones((np_sub, np_sub))
    return dict(x=x, y=y, xc=xc, yc=yc, mask_weighted=mask_weighted)


def shslopec(pixels, idxs, thr_value=-1.0, exp_weight=1.0, thr_pedestal=False):
    """ShSlopec.calc_slopes_nofor, sh_slopec.py:308-438, in FP64.
    Returns dict(slopes=[sx..., sy...], flux (n_sub,), total, mean)."""
    pixels = np.asarray(pixels, dtype=np.float64)
    idxs = np.asarray(idxs).astype(np.int64)
    n_sub, npix = idxs.shape
    np_sub = int(math.sqrt(npix))
    w = compute_xy_weights(np_sub, exp_weight)
    rows = idxs // pixels.shape[1]                                    # utils.py:37-49
    cols = idxs - rows * pixels.shape[1]
    p = pixels[rows, cols].T.copy()                                   # :330-331 -> (np^2, n_sub)
    flux = p.sum(axis=0)                                              # :365
    if thr_pedestal or thr_value > 0:                                 # :377-380
        thr = thr_value
    else:
        thr = 0
    if thr_pedestal:                                                  # :382-386
        p = np.where(p < thr, 0, p)
    else:
        p = p - thr
        p = np.where(p < 0, 0, p)
    subap_tot = np.sum(p * w['mask_weighted'].reshape(npix, 1), axis=0)   # :392
    mean_tot = np.mean(subap_tot)
    with np.errstate(divide='ignore'):
        factor = 1.0 / subap_tot                                      # :394
    factor = np.where(factor > 1.0 / (mean_tot * 1e-3), 0, factor)    # :400
    sx = np.sum(p * w['x'].reshape(npix, 1) * factor[np.newaxis, :], axis=0)   # :403
    sy = np.sum(p * w['y'].reshape(npix, 1) * factor[np.newaxis, :], axis=0)   # :404
    return dict(slopes=np.concatenate([sx, sy]), flux=flux, total=flux.sum(), mean=flux.mean())


def detect_subaps(image, n_lenses, energy_th):
    """ShSubapCalibrator._detect_subaps, sh_subap_calibrator.py:46-92.
    Returns (idxs (n_sub, np_sub**2) int64, display_map (n_sub,) int64)."""
    npx = image.shape[0]
    lens = lenslet(n_lenses)
    spot = np.zeros((n_lenses, n_lenses))
    xs = np.zeros((n_lenses, n_lenses))
    ys = np.zeros((n_lenses, n_lenses))
    np_sub = 0
    masks = {}
    for i in range(n_lenses):
        for j in range(n_lenses):
            l = lens[i][j]
            xs[i, j] = npx / 2.0 * (1 + l[0])
            ys[i, j] = npx / 2.0 * (1 + l[1])
            np_sub = round(npx / 2.0 * l[2])
            r0, r1 = int(np.round(xs[i, j] - np_sub / 2)), int(np.round(xs[i, j] + np_sub / 2))
            c0, c1 = int(np.round(ys[i, j] - np_sub / 2)), int(np.round(ys[i, j] + np_sub / 2))
            masks[i, j] = (r0, r1, c0, c1)
            spot[i, j] = image[r0:r1, c0:c1].sum()
    idxs, dmap = [], []
    for i in range(n_lenses):
        for j in range(n_lenses):
            if spot[i, j] > energy_th * spot.max():                   # :72
                r0, r1, c0, c1 = masks[i, j]
                rr, cc = np.meshgrid(np.arange(r0, r1), np.arange(c0, c1), indexing='ij')
                idxs.append(np.ravel_multi_index((rr.ravel(), cc.ravel()), image.shape))
                dmap.append(j * n_lenses + i)                         # :77
    if not idxs:
        raise ValueError('Error: no subapertures selected')
    return np.array(idxs, dtype=np.int64), np.array(dmap, dtype=np.int64)


# --------------------------------------------------------------------------- LGS kernels

def lgs_map_sh(nsh, diam, rl, zb, dz, profz, fwhmb, ps, ssp, overs=2, theta=(0.0, 0.0),
               rprof_type=0, dtype=np.float32):
    """convolution_kernel.py:14-108 (doCube=True, no pupil mask).  Computed in ``dtype``
    (float32 by default, as the reference)."""
    theta = np.asarray(theta, dtype=np.float64)
    ossp = ssp * overs
    xsh, ysh = np.meshgrid(np.linspace(-diam / 2, diam / 2, nsh, dtype=dtype),
                           np.linspace(-diam / 2, diam / 2, nsh, dtype=dtype))
    xfov, yfov = np.meshgrid(np.linspace(-ssp * ps / 2, ssp * ps / 2, ossp, dtype=dtype),
                             np.linspace(-ssp * ps / 2, ssp * ps / 2, ossp, dtype=dtype))
    sigma = (fwhmb * ASEC2RAD * zb) / (2 * np.sqrt(2 * np.log(2)))
    one_over_sigma2 = 1.0 / sigma ** 2
    exp_sigma = -0.5 * one_over_sigma2
    rb = np.array([theta[0] * ASEC2RAD * zb, theta[1] * ASEC2RAD * zb, 0], dtype=dtype)
    kv = np.array([0, 0, 1], dtype=dtype)
    BL = zb * kv + rb - np.array(rl, dtype=dtype)
    el = BL / BL[2]
    rs_x, rs_y = xsh, ysh
    rf_x = np.tile(xfov * ASEC2RAD * zb, (nsh, nsh)).reshape(ossp * nsh, ossp * nsh)
    rf_y = np.tile(yfov * ASEC2RAD * zb, (nsh, nsh)).reshape(ossp * nsh, ossp * nsh)
    es_x = (rf_x - np.repeat(np.repeat(rs_x, ossp, axis=0), ossp, axis=1)) / zb
    es_y = (rf_y - np.repeat(np.repeat(rs_y, ossp, axis=0), ossp, axis=1)) / zb
    es_z = 1.0
    fmap = np.zeros((nsh * ossp, nsh * ossp), dtype=dtype)
    if rprof_type == 0:
        gnorm = 1.0 / (sigma * np.pi * np.sqrt(2.0))
    elif rprof_type == 1:
        gnorm = 1.0 / (np.pi / 4 * (fwhmb * ASEC2RAD * zb) ** 2)
    else:
        raise ValueError('Unsupported radial profile type')
    for iz in range(len(dz)):
        if profz[iz] > 0:
            d2 = ((rf_x + dz[iz] * es_x - (rb[0] + dz[iz] * el[0])) ** 2 +
                  (rf_y + dz[iz] * es_y - (rb[1] + dz[iz] * el[1])) ** 2 +
                  (0.0 + dz[iz] * es_z - (rb[2] + dz[iz] * el[2])) ** 2)
            if rprof_type == 0:
                fmap += (gnorm * profz[iz]) * np.exp(d2 * exp_sigma, dtype=dtype)
            else:
                fmap += (gnorm * profz[iz]) * ((d2 * one_over_sigma2) <= 1.0)
    ccd = rebin2d(fmap, (ssp * nsh, ssp * nsh), sample=True)
    ccd = ccd.reshape(nsh, ssp, nsh, ssp).transpose((2, 0, 1, 3)).reshape(nsh * nsh, ssp, ssp)
    return ccd


def process_kernels(real_kernels, dimx, dimy, return_fft=True):
    """ConvolutionKernel.process_kernels, convolution_kernel.py:258-274 (FP64 output)."""
    if np.any(~np.isfinite(real_kernels)):
        raise ValueError('Kernel contains non-finite values!')
    out = np.zeros(real_kernels.shape, dtype=np.complex128 if return_fft else np.float64)
    for i in range(dimx):
        for j in range(dimy):
            k = np.array(real_kernels[i * dimx + j, :, :])
            total = np.sum(k)
            if total > 0:
                k = k / total
            out[j * dimx + i] = np.fft.ifft2(k) if return_fft else k
    return out


def convolution_kernels_for_sh(geom, launcher_pos, spot_size, zlayer, zprofile, zfocus=90e3,
                               theta=(0.0, 0.0), seeing=0.0, airmass=1.0, positive_shift_tt=True):
    """ConvolutionKernel as built by SH.calc_trigger_geometry (sh.py:307-320) and
    calculate_lgs_map (convolution_kernel.py:179-225) -> ``kernels`` (n^2, N, N) complex."""
    n = geom['n_lenses']
    pxscale = geom['fp4_pixel_pitch'] * RAD2ASEC
    zf = zfocus if zfocus != -1 else np.sum(np.array(zlayer) * np.array(zprofile)) / np.sum(zprofile)
    lay = np.array(zlayer) * airmass
    zf = zf * airmass
    spot = np.sqrt(seeing ** 2 + spot_size ** 2)
    if spot <= 0:
        raise ValueError('Spot size must be greater than zero')
    lgs_tt = np.array([0.5, 0.5] if positive_shift_tt else [-0.5, -0.5]) * pxscale + np.asarray(theta)
    real = lgs_map_sh(n, geom['pixel_pitch'] * geom['ef_size'], launcher_pos, zf, lay - zf,
                      zprofile, spot, pxscale, geom['fft_size'], overs=1, theta=lgs_tt)
    return process_kernels(real, n, n, return_fft=True)


def gaussian_kernels_for_sh(geom, spot_size, positive_shift_tt=True):
    """GaussianConvolutionKernel.calculate_lgs_map, gaussian_convolution_kernel.py:38-57."""
    n = max(geom['n_lenses'], 2)
    pxscale = geom['fp4_pixel_pitch'] * RAD2ASEC
    lgs_tt = [x * pxscale for x in ([0.5, 0.5] if positive_shift_tt else [-0.5, -0.5])]
    real = lgs_map_sh(n, geom['pixel_pitch'] * geom['ef_size'], 0, 90e3, [0], profz=[1.0],
                      fwhmb=spot_size, ps=pxscale, ssp=geom['fft_size'], overs=1, theta=lgs_tt)
    return process_kernels(real, n, n, return_fft=True)


# --------------------------------------------------------------------------- modal DM, push-pull IM

def zernike_degree(j):
    """ZernikeGenerator.degree, zernike_generator.py:141-151 (Noll order)."""
    n = int(np.ceil(0.5 * (np.sqrt(8 * j + 1) - 3)))                      # :404-421
    cn = n * (n + 1) / 2 + 1
    if n % 2 == 0:
        m = int(j - cn + 1) // 2 * 2
    else:
        m = int(j - cn) // 2 * 2 + 1
    return n, m


def zernike_map(j, dim):
    """ZernikeGenerator(dim).getZernike(j), zernike_generator.py:83-99,153-190,192-246, FP64:
    pixel-centre coordinates of a unit disk of diameter ``dim``, the factorial sum for R_n^m."""
    sz = math.ceil(dim)
    c = sz / 2
    Y, X = (np.mgrid[0.5:sz + 0.5:1, 0.5:sz + 0.5:1] - c) / (dim / 2)
    rho = np.sqrt(X ** 2 + Y ** 2)
    theta = np.arctan2(Y, X)
    n, m = zernike_degree(j)
    if n == 0 and m == 0:
        R = np.ones_like(rho)
    else:
        R = np.zeros_like(rho)
        for s in range((n - abs(m)) // 2 + 1):
            R = R + ((-1) ** s * math.factorial(n - s) /
                     (math.factorial(s) * math.factorial((n + abs(m)) // 2 - s) *
                      math.factorial((n - abs(m)) // 2 - s))) * rho ** (n - 2 * s)
    NC = np.sqrt(2 * (n + 1))
    if m == 0:
        return np.sqrt(0.5) * NC * R
    return NC * R * (np.cos(m * theta) if j % 2 == 0 else np.sin(m * theta))


def compute_zern_ifunc(dim, nzern, obsratio=0.0, diaratio=1.0):
    """compute_zern_ifunc.py:6-29 with make_orto_modes (utils.py:51-83): Zernike 2..nzern+1 on
    the make_mask pixels, Householder QR (numpy.linalg.qr), mean removed, unit std.
    Returns (ifunc (nzern, n_px), mask (dim, dim) float)."""
    mask = make_mask(dim, obsratio, diaratio).astype(float)
    idx = np.where(mask)
    z = np.array([zernike_map(j, dim)[idx] for j in range(2, nzern + 2)])
    q, _ = np.linalg.qr(z.T)
    q = q.T
    q = q - np.mean(q, axis=1, keepdims=True)
    q = q / np.std(q, axis=1, keepdims=True)
    return q, mask


def dm_phase(ifunc, mask, cmd, sign=-1):
    """DM.trigger_code, dm.py:72-86 (no m2c): phase[idx] = (sign * cmd) @ ifunc, zero elsewhere."""
    out = np.zeros(mask.shape)
    out[np.where(mask)] = (sign * np.asarray(cmd, dtype=np.float64)) @ ifunc
    return out


def pushpull_time_history(n_modes, vect_amplitude):
    """modal_pushpull_signal.py:35-65 (ncycles 1, nsamples 1): rows +a_m, -a_m for each mode."""
    th = np.zeros((2 * n_modes, n_modes))
    for i in range(n_modes):
        th[2 * i:2 * i + 2, i] = [vect_amplitude[i], -vect_amplitude[i]]
    return th


def accumulate_intmat(slopes_hist, commands_hist, n_modes):
    """ImRecCalibrator.trigger_code, im_rec_calibrator.py:42-71: IM[mode] += slopes / command of
    the non-zero command of each step."""
    im = np.zeros((n_modes, slopes_hist.shape[1]))
    for s, c in zip(slopes_hist, commands_hist):
        nz = np.nonzero(c)[0]
        if len(nz) > 0 and nz[0] < n_modes:
            im[nz[0]] += s / c[nz[0]]
    return im


def recmat_fits_data(im):
    """Intmat.generate_rec (intmat.py:74-87: pinv) as Recmat.save writes it (recmat.py:40-50:
    the transpose), i.e. the (n_modes, n_slopes) array of the file's extension 1."""
    return np.linalg.pinv(im).T


# --------------------------------------------------------------------------- CCD

def ccd_readout(integrated, binning=1, quantum_eff=1.0, photon_noise=False, readout_noise=False,
                excess_noise=False, darkcurrent_noise=False, background_noise=False, cic_noise=False,
                readout_level=0.0, darkcurrent_level=0.0, background_level=0.0, cic_level=0.0,
                ADU_gain=8, ADU_bias=400, emccd_gain=1, excess_delta=1.0, pixelGains=None,
                keep_ADU_bias=False, bg_remove_average=False, do_not_remove_dark=False,
                photon_rng=None, readout_rng=None, cic_rng=None, excess_rng=None, dtype=np.float32,
                cte_noise=False, cte_mat=None):
    """CCD.apply_binning + apply_qe + apply_noise, ccd.py:220-273, with numpy generators and the
    reference's dtype flow (float32 frame; int64 after Poisson, so an added readout noise is
    truncated by ``astype``; float64 after the ADU conversion).  ``excess_noise`` uses
    numpy's gamma in place of scipy.stats.gamma.rvs (same distribution)."""
    if excess_noise:                                                   # ccd.py:104-106
        emccd_gain = 400
        ADU_gain = 1 / 20
    frame = np.array(integrated, dtype=dtype)
    if binning > 1:                                                    # :264-268
        tot = np.sum(frame)
        h, w = frame.shape[0] // binning, frame.shape[1] // binning
        frame = frame.reshape(h, binning, w, binning).sum(axis=(1, 3))
        frame = frame * binning ** 2 * (tot / np.sum(frame))
    if quantum_eff != 1:                                               # :276-277
        frame *= dtype(quantum_eff) if dtype is np.float32 else quantum_eff
    if background_noise or darkcurrent_noise:                          # :220-221
        frame += (background_level + darkcurrent_level)
    if cte_noise:                                                      # :223-224
        m = np.asarray(cte_mat, dtype=frame.dtype)
        frame = np.dot(np.dot(m[:, :, 0], frame), m[:, :, 1])
    if cic_noise:                                                      # :228-229
        frame = frame + cic_rng.binomial(1, cic_level, frame.shape)
    if photon_noise:                                                   # :231-232
        frame = photon_rng.poisson(frame)
    if excess_noise:                                                   # :234-235
        frame = 1.0 / excess_delta * excess_rng.gamma(excess_delta * frame, emccd_gain)
    if readout_noise:                                                  # :237-239
        ron = readout_rng.standard_normal(size=frame.size)
        frame += (ron.reshape(frame.shape) * readout_level).astype(frame.dtype)
    if pixelGains is not None:                                         # :241-242 (in place there,
        frame = frame * pixelGains                                     # which numpy refuses for int64)
    if photon_noise:                                                   # :249-262
        frame = np.round(frame * ADU_gain) + ADU_bias
        frame = np.where(frame < 0, 0, frame)
        if not keep_ADU_bias:
            frame = frame - ADU_bias
        frame = frame / ADU_gain
        if excess_noise:
            frame = frame / emccd_gain
        if darkcurrent_noise and not do_not_remove_dark:
            frame = frame - darkcurrent_level
        if bg_remove_average and not do_not_remove_dark:
            frame = frame - background_level
    return frame
