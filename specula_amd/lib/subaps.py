"""Subaperture selection and pixel addressing on the host.

``detect_subaps`` restates ShSubapCalibrator._detect_subaps
(specula/processing_objects/sh_subap_calibrator.py:46-92): a subaperture is valid when its flux
is strictly above energy_th * max; subapertures are emitted in column-major block order with
row-major pixel indices inside each block, display_map = row * n + col.
"""
import numpy as np

from specula_amd.lib.grids import make_xy


def detect_subaps(image, n_lenses, energy_th):
    image = np.asarray(image)
    npx = image.shape[0]
    if n_lenses > 1:
        lx, ly = make_xy(n_lenses, 1.0)
    else:
        lx = ly = np.zeros((1, 1))
    size = 2.0 / n_lenses
    np_sub = round(npx / 2.0 * size)
    boxes = {}
    flux = np.zeros((n_lenses, n_lenses))
    for i in range(n_lenses):
        for j in range(n_lenses):
            cx = npx / 2.0 * (1 + lx[i, j])
            cy = npx / 2.0 * (1 + ly[i, j])
            r0, r1 = int(np.round(cx - np_sub / 2)), int(np.round(cx + np_sub / 2))
            c0, c1 = int(np.round(cy - np_sub / 2)), int(np.round(cy + np_sub / 2))
            boxes[i, j] = (r0, r1, c0, c1)
            flux[i, j] = image[r0:r1, c0:c1].sum()
    idxs, dmap = [], []
    thr = energy_th * flux.max()
    for i in range(n_lenses):
        for j in range(n_lenses):
            if flux[i, j] > thr:
                r0, r1, c0, c1 = boxes[i, j]
                rows = np.arange(r0, r1)[:, None]
                cols = np.arange(c0, c1)[None, :]
                idxs.append((rows * image.shape[1] + cols).ravel())
                dmap.append(j * n_lenses + i)
    if not idxs:
        raise ValueError('Error: no subapertures selected')
    return np.array(idxs, dtype=np.int64), np.array(dmap, dtype=np.int64)


def block_origins(idxs, frame_w):
    """If every subaperture is a square block of contiguous rows (idx = origin + r*W + c),
    return the int32 origins, else None."""
    idxs = np.asarray(idxs, dtype=np.int64)
    n_sub, npp = idxs.shape
    np_sub = int(round(np.sqrt(npp)))
    if np_sub * np_sub != npp:
        return None
    r, c = np.divmod(np.arange(npp), np_sub)
    pattern = r * frame_w + c
    origins = idxs[:, 0]
    if not np.array_equal(idxs, origins[:, None] + pattern[None, :]):
        return None
    if origins.max() >= 2 ** 31 or np.any(origins % frame_w + np_sub > frame_w):
        return None
    return origins.astype(np.int32)
