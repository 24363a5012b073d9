"""Binning operators of the spectral SH formulation (host, FP64).

For one subaperture with field E (np x np, zero padded to N), every step of the reference
after ``|FFT_N(E)|^2`` -- fftshift (sh.py:396), FoV mask (sh.py:399), FoV cut (sh.py:405),
reassembly and CCD rebin (lib/toccd.py:33-63) -- is a fixed linear map of the N x N power
image.  When the mask is separable (square masks always are) the map factorises per axis:

    B[Y, X] = sum_{u, v} w_row[Y, u] w_col[X, v] |F_N[u, v]|^2        (u, v: FFT-order indices)

|F_N|^2 is a trigonometric polynomial of degree np-1 per axis: |F_N[u]|^2 = sum_l C[l] W_N^{u l}
with C the autocorrelation of E on lags |l| <= np-1.  Any grid M >= 2 np - 1 samples it without
aliasing, C[l] = 1/M sum_j S_M[j] W_M^{-j l} with S_M = |FFT_M(E)|^2, hence

    B = P S_M Q^T,   P[Y, j] = 1/M sum_{|l|<np} (sum_u w_row[Y, u] W_N^{u l}) W_M^{-j l}   (real)

For the LGS path (sh.py:382-394) the psf is Re IFFT_N(FFT_N(I) K) with K the kernel spectrum;
FFT_N(I)[m] = N^2 C[-m] on the lag grid, so

    B = Re( Pt D Qt^T ),  D[m] = N^2 C[-m] K[m],  Pt[Y, m] = sum_u w_row[Y, u] W_N^{-u m}
"""
import math

import numpy as np


def spec_size(np_sub):
    """Smallest power of two >= 2 np - 1."""
    m = 1
    while m < 2 * np_sub - 1:
        m *= 2
    return m


def _rebin_map_1d(n_in, n_out):
    """Rows of ``toccd`` along one axis as a (n_out x n_in) weight matrix.

    toccd (toccd.py:55-61) replicates each input sample onto the LCM grid with
    ``rebin2d(sample=True)`` (rebin.py:18-21: index = int(k * old/new), float step as numpy's
    mgrid) and then block-averages.  The weights below count those replicas (the constant
    1/dx_out and the final renormalisation cancel in the flux normalisation)."""
    if n_in == n_out:
        return np.eye(n_in)
    mcm = math.lcm(n_in, n_out)
    if mcm == n_in:
        src = np.arange(n_in)
    else:
        src = np.mgrid[slice(0, n_in, float(n_in) / mcm)].astype(np.int32)
        if src.size != mcm:
            raise ValueError(f'rebin {n_in}->{n_out}: LCM sampling yields {src.size} != {mcm} samples')
    dout = mcm // n_out
    w = np.zeros((n_out, n_in))
    for m, s in enumerate(src):
        w[m // dout, s] += 1.0
    return w


def axis_weights(geom, lgs, axis, m1=None):
    """w[Y, u] (subap_npx x N) for one axis of one subaperture.

    axis 0: rows (y), axis 1: columns (x).  ``m1``: the mask factor of this axis over the
    shifted psf index (length N); None takes it from a separable FoV mask and raises if the mask
    is not separable.  Raises if the global rebin is not periodic per subaperture."""
    N = geom.fft_size
    n = geom.n_lenses
    cs = geom.cutsize
    npx = geom.subap_npx
    if cs <= 0 or geom.cutpixels < 0:
        raise ValueError(f'unsupported FoV cut: cutpixels={geom.cutpixels}, cutsize={cs}')
    if m1 is None:
        mask = geom.fp_mask.astype(np.float64)
        m_rows = mask.max(axis=1)
        m_cols = mask.max(axis=0)
        if not np.array_equal(mask, np.outer(m_rows, m_cols)):
            raise NotImplementedError('FoV mask is not separable (squaremask=False with a clipping FoV)')
        m1 = m_rows if axis == 0 else m_cols
    # global rebin of the reassembled image, then check per-subaperture periodicity
    wg = _rebin_map_1d(cs * n, npx * n)
    rho = wg[:npx, :cs]
    for s in range(n):
        blk = wg[s * npx:(s + 1) * npx, s * cs:(s + 1) * cs]
        if not np.array_equal(blk, rho) or np.count_nonzero(wg[s * npx:(s + 1) * npx]) != np.count_nonzero(blk):
            raise NotImplementedError('CCD rebin is not periodic per subaperture for this geometry')
    lo = geom.cutpixels // 2
    w = np.zeros((npx, N))
    for t in range(cs):
        p = lo + t                                    # index in the (shifted) psf
        u = p if lgs else (p - N // 2) % N            # index in FFT order (fftshift for NGS)
        w[:, u] += rho[:, t] * m1[p]
    return w


def mask_is_separable(geom):
    mask = geom.fp_mask.astype(np.float64)
    return bool(np.array_equal(mask, np.outer(mask.max(axis=1), mask.max(axis=0))))


def mask_terms(geom):
    """Exact separable decomposition of the FoV mask (sh.py:282, make_mask with square=False) on
    the window the FoV cut keeps: W = sum_k a_k b_k^T, one term per distinct non-zero row pattern
    of the window (a_k: indicator of the rows with that pattern, b_k: the pattern).  A disc of
    diameter D gives about D / 2 terms.  Vectors over the shifted psf index (length N)."""
    N = geom.fft_size
    lo, cs = geom.cutpixels // 2, geom.cutsize
    mask = geom.fp_mask.astype(np.float64)
    win = mask[lo:lo + cs, lo:lo + cs]
    groups = {}
    for t in range(cs):
        row = win[t]
        if not row.any():
            continue
        groups.setdefault(row.tobytes(), (row, []))[1].append(lo + t)
    terms = []
    for row, rows in groups.values():
        a = np.zeros(N)
        a[rows] = 1.0
        b = np.zeros(N)
        b[lo:lo + cs] = row
        terms.append((a, b))
    return terms


def ngs_operator(w, N, np_sub, M):
    """P (npx x M, real) from axis weights w (npx x N)."""
    L = np.arange(-(np_sub - 1), np_sub)
    u = np.arange(N)
    P = w @ np.exp(-2j * np.pi * np.outer(u, L) / N)
    j = np.arange(M)
    Pp = (P @ np.exp(2j * np.pi * np.outer(L, j) / M)) / M
    if np.abs(Pp.imag).max() > 1e-9 * max(1.0, np.abs(Pp.real).max()):
        raise AssertionError('binning operator is not real')
    return np.ascontiguousarray(Pp.real)


def lgs_operator(w, N, np_sub):
    """Pt (npx x (2np-1), complex) from axis weights w (npx x N)."""
    L = np.arange(-(np_sub - 1), np_sub)
    u = np.arange(N)
    return np.ascontiguousarray(w @ np.exp(2j * np.pi * np.outer(u, L) / N))


def parseval_norm(w_row, w_col, N):
    """If every FFT sample carries the same total weight on both axes, the sum of the binned
    image over a subaperture equals norm * sum|E|^2 = norm * sum(A^2) (Parseval); return norm,
    else None."""
    cr = w_row.sum(axis=0)
    cc = w_col.sum(axis=0)
    if cr.min() <= 0 or cc.min() <= 0:
        return None
    if np.ptp(cr) > 1e-12 * cr.max() or np.ptp(cc) > 1e-12 * cc.max():
        return None
    return float(cr.mean() * cc.mean() * N * N)


def build_operators(geom, lgs=False):
    """Everything the C plan needs: (row_op, col_op, M, parseval_norm_or_None)."""
    np_sub = geom.ovs_np_sub
    M = spec_size(np_sub)
    N = geom.fft_size
    wr = axis_weights(geom, lgs, 0)
    wc = axis_weights(geom, lgs, 1)
    norm = parseval_norm(wr, wc, N)
    if lgs:
        return lgs_operator(wr, N, np_sub), lgs_operator(wc, N, np_sub), M, norm
    return ngs_operator(wr, N, np_sub, M), ngs_operator(wc, N, np_sub, M), M, norm


def build_operator_terms(geom, lgs=False):
    """Operators of a possibly non-separable FoV mask: (row_ops, col_ops, M, norm_or_None) with
    row_ops, col_ops stacked over the K separable terms of ``mask_terms`` (K = 1 and the
    operators of build_operators when the mask is separable).  B = sum_k P_k S Q_k^T (NGS) or
    Re sum_k Pt_k D Qt_k^T (LGS); with K > 1 the image is normalised by its own sum."""
    if mask_is_separable(geom):
        r, c, M, norm = build_operators(geom, lgs=lgs)
        return r[None], c[None], M, norm
    np_sub = geom.ovs_np_sub
    M = spec_size(np_sub)
    N = geom.fft_size
    rows, cols = [], []
    for a, b in mask_terms(geom):
        wr = axis_weights(geom, lgs, 0, m1=a)
        wc = axis_weights(geom, lgs, 1, m1=b)
        if lgs:
            rows.append(lgs_operator(wr, N, np_sub))
            cols.append(lgs_operator(wc, N, np_sub))
        else:
            rows.append(ngs_operator(wr, N, np_sub, M))
            cols.append(ngs_operator(wc, N, np_sub, M))
    return np.stack(rows), np.stack(cols), M, None
