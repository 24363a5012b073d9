"""Shared comparison helpers for the parity tests (tolerances stated in one place).

Tolerances (FP32 GPU path vs FP64 oracle on identical inputs):
  total flux                         1e-5 relative
  per-subaperture flux               1e-5 relative (+ 1e-7 of the brightest subaperture as an
                                     absolute floor for almost-dark edge subapertures)
  per-subaperture-normalised image   1e-4 of the subaperture's peak (max-abs)
  slopes                             1e-3 CCD px = 1e-3 * 2/(npx-1) slope units
"""
import numpy as np

from oracle import sh_oracle as O

TOL_FLUX = 1e-5
TOL_IMG = 1e-4
TOL_SLOPE_PX = 1e-3


def oracle_geometry(wl):
    return O.sh_geometry(ef_size=wl.ef_size, pixel_pitch=wl.pixel_pitch,
                         fov_resolution_arcsec=wl.fov_resolution_arcsec or None, **wl.sh_kwargs)


def subap_tiles(img, n, npx):
    return img.reshape(n, npx, n, npx).transpose(0, 2, 1, 3).reshape(n * n, npx, npx)


def check_intensity(gpu, ref, n, npx):
    gpu = np.asarray(gpu, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    assert np.all(np.isfinite(gpu))
    tot_err = abs(gpu.sum() - ref.sum()) / abs(ref.sum())
    assert tot_err <= TOL_FLUX, f'total flux rel err {tot_err:.3e}'
    tg, tr = subap_tiles(gpu, n, npx), subap_tiles(ref, n, npx)
    fg, fr = tg.sum(axis=(1, 2)), tr.sum(axis=(1, 2))
    ferr = np.abs(fg - fr)
    bound = TOL_FLUX * np.abs(fr) + 1e-7 * fr.max()
    assert np.all(ferr <= bound), f'subap flux: worst excess {np.max(ferr / bound):.3f}x tolerance'
    lit = fr > 0
    ng = tg[lit] / fg[lit, None, None]
    nr = tr[lit] / fr[lit, None, None]
    rel = np.abs(ng - nr).max(axis=(1, 2)) / nr.max(axis=(1, 2))
    assert rel.max() <= TOL_IMG, f'normalised image err {rel.max():.3e}'
    return dict(total=tot_err, img=float(rel.max()))


def check_slopes(gpu, ref, npx):
    tol = TOL_SLOPE_PX * 2.0 / (npx - 1)
    err = np.abs(np.asarray(gpu, np.float64) - np.asarray(ref, np.float64))
    assert err.max() <= tol, f'slope err {err.max():.3e} > {tol:.3e}'
    return float(err.max())
