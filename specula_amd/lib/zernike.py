"""Noll-ordered Zernike polynomials and the orthonormalised Zernike influence functions of a
modal DM (host-side set-up, FP64).

Conventions of lib/zernike_generator.py:73-270: a pupil of ``dim`` pixels is the unit disk
sampled at pixel centres, X along columns and Y along rows measured from the array centre
(``dim / 2``), ``theta = atan2(Y, X)``; index j >= 1 in Noll's order with Noll's normalisation
(the mean of Z_j^2 over the unit disk is 1), even j -> cos(m theta), odd j -> sin(m theta).
Values outside the disk are the polynomial continued (the reference's masked array keeps them).

``zernike_ifunc`` follows lib/compute_zern_ifunc.py:6-29: modes j = 2 .. nzern + 1 sampled on
the pupil mask pixels (row-major), orthonormalised by a Householder QR (numpy / LAPACK, the
same factorisation as lib/utils.py:51-83, so the column signs agree), then each mode made
mean-free and scaled to unit standard deviation over the mask.
"""
import math

import numpy as np

from specula_amd.lib.grids import make_mask


def noll_degree(j):
    """(n, m) of Noll index j >= 1: n the radial order, m >= 0 the azimuthal frequency."""
    n = int(math.ceil((math.sqrt(8 * j + 1) - 3) / 2))
    first = n * (n + 1) // 2 + 1            # first index of radial order n
    if n % 2 == 0:
        m = 2 * ((j - first + 1) // 2)
    else:
        m = 2 * ((j - first) // 2) + 1
    return n, m


def radial_poly(n, m, rho):
    """R_n^m(rho) as the terminating hypergeometric sum with integer coefficients."""
    if (n - m) % 2:
        raise ValueError(f'n - m must be even (n={n}, m={m})')
    out = np.zeros_like(rho, dtype=np.float64)
    for s in range((n - m) // 2 + 1):
        c = ((-1) ** s) * math.factorial(n - s) // (
            math.factorial(s) * math.factorial((n + m) // 2 - s) * math.factorial((n - m) // 2 - s))
        out += c * rho ** (n - 2 * s)
    return out


def pupil_polar(dim):
    """rho, theta of the pixel centres of a ``dim``-pixel unit disk."""
    sz = int(math.ceil(dim))
    c = sz / 2.0
    r = dim / 2.0
    yy, xx = np.mgrid[0:sz, 0:sz].astype(np.float64)
    Y = (yy + 0.5 - c) / r
    X = (xx + 0.5 - c) / r
    return np.hypot(X, Y), np.arctan2(Y, X)


def zernike(j, dim, polar=None):
    """Z_j on the ``dim`` x ``dim`` grid (FP64)."""
    if j < 1:
        raise ValueError(f'invalid Zernike index {j}')
    rho, theta = polar if polar is not None else pupil_polar(dim)
    n, m = noll_degree(j)
    rad = radial_poly(n, m, rho)
    if m == 0:
        return math.sqrt(n + 1) * rad
    norm = math.sqrt(2 * (n + 1))
    return norm * rad * (np.cos(m * theta) if j % 2 == 0 else np.sin(m * theta))


def zernike_ifunc(dim, nzern, obsratio=0.0, diaratio=1.0, start_mode=0, mask=None, dtype=np.float32):
    """(modes (nzern - start_mode, n_mask_px), mask (dim, dim)) as compute_zern_ifunc returns them."""
    if mask is None:
        mask = make_mask(dim, obsratio, diaratio).astype(np.float64)
    else:
        mask = np.asarray(mask, dtype=np.float64)
    idx = np.nonzero(mask)
    polar = pupil_polar(dim)
    z = np.stack([zernike(j, dim, polar)[idx] for j in range(2, nzern + 2)])[start_mode:]
    if z.shape[1] > z.shape[0]:
        q = np.linalg.qr(z.T)[0].T                # K x n_px, Householder (LAPACK) signs
    else:
        q = np.linalg.qr(z)[0]
    q = q - q.mean(axis=1, keepdims=True)
    q = q / q.std(axis=1, keepdims=True)
    return q.astype(dtype), mask.astype(dtype)
