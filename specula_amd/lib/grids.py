"""Coordinate grids and masks: specula/lib/make_xy.py:12-177 and make_mask.py:7-147."""
import numpy as np


def make_xy(sampling, ratio, dtype=None):
    """Pixel-centre coordinates on [-ratio, ratio]; x varies along columns (make_xy.py:141-177)."""
    if sampling <= 1:
        raise ValueError('make_xy: sampling must be larger than 1')
    centre = (sampling - 1) / 2.0
    v = (np.arange(sampling, dtype=dtype) - centre) / (sampling / 2.0) * ratio
    x = np.broadcast_to(v, (sampling, sampling)).copy()
    return x, x.T.copy()


def make_mask(np_size, obsratio=0.0, diaratio=1.0, xc=0.0, yc=0.0, square=False, inverse=False):
    """uint8 pupil/FoV mask (make_mask.py:7-147 without spiders / centeronpixel)."""
    diaratio = 1.0 if diaratio is None else diaratio
    obsratio = 0.0 if obsratio is None else obsratio
    x, y = make_xy(np_size, 1.0)
    dx, dy = np.abs(x - xc), np.abs(y - yc)
    if square:
        inside = (dx <= diaratio) & (dy <= diaratio) & ((dx >= diaratio * obsratio) | (dy >= diaratio * obsratio))
    else:
        r2 = (x - xc) ** 2 + (y - yc) ** 2
        inside = (r2 < diaratio ** 2) & (r2 >= (diaratio * obsratio) ** 2)
    mask = inside.astype(np.uint8)
    return (1 - mask) if inverse else mask
