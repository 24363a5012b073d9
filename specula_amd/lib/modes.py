"""Seeded smooth modal bases on a pupil (synthetic inputs for interaction matrices).

``fourier_modes`` returns cosine/sine pupil modes ordered by spatial frequency (a KL-like
low-to-high-order basis): mean-free and normalised to unit RMS over the pupil, in nm per unit
command, like the orthonormalised influence functions of lib/compute_zern_ifunc.py:6-29.
Built with torch so that large bases (1000 modes x 640^2) can be generated on the device.
"""
import math

import numpy as np
import torch


def _frequencies(n_modes):
    """(kx, ky, parity) triples sorted by |k|, excluding piston."""
    out = []
    r = 1
    while len(out) < n_modes:
        cand = []
        for ky in range(-r, r + 1):
            for kx in range(0, r + 1):
                if kx == 0 and ky <= 0:
                    continue
                k2 = kx * kx + ky * ky
                if (r - 1) ** 2 < k2 <= r * r:
                    cand.append((k2, kx, ky))
        for k2, kx, ky in sorted(cand):
            out.append((kx, ky, 0))
            out.append((kx, ky, 1))
        r += 1
    return out[:n_modes]


def fourier_modes(pupil, n_modes, device='cpu', dtype=torch.float32, seed=None):
    """(n_modes, n, n) modes over the binary/graded ``pupil`` (n x n).  ``seed`` adds a seeded
    random phase offset to every mode (different but reproducible bases)."""
    pupil = torch.as_tensor(np.asarray(pupil), dtype=torch.float64, device=device)
    n = pupil.shape[0]
    w = pupil > 0
    npx = float(w.sum())
    coords = (torch.arange(n, dtype=torch.float64, device=device) - (n - 1) / 2) / n
    yy, xx = torch.meshgrid(coords, coords, indexing='ij')
    freqs = _frequencies(n_modes)
    rng = np.random.default_rng(seed) if seed is not None else None
    out = torch.empty((n_modes, n, n), dtype=dtype, device=device)
    for m, (kx, ky, par) in enumerate(freqs):
        off = 0.0 if rng is None else float(rng.uniform(0, 2 * math.pi))
        arg = 2 * math.pi * (kx * xx + ky * yy) + off
        mode = torch.cos(arg) if par == 0 else torch.sin(arg)
        mode = torch.where(w, mode, torch.zeros_like(mode))
        mean = mode.sum() / npx
        mode = torch.where(w, mode - mean, torch.zeros_like(mode))
        rms = torch.sqrt((mode * mode).sum() / npx)
        out[m] = (mode / rms).to(dtype)
    return out
