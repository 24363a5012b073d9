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
    offs = [0.0 if rng is None else float(rng.uniform(0, 2 * math.pi)) for _ in freqs]
    out = torch.empty((n_modes, n, n), dtype=dtype, device=device)
    wf = w.to(torch.float64)
    chunk = 100                                          # modes per batch of tensor operations
    for m0 in range(0, n_modes, chunk):
        f = freqs[m0:m0 + chunk]
        kx = torch.tensor([v[0] for v in f], dtype=torch.float64, device=device).view(-1, 1, 1)
        ky = torch.tensor([v[1] for v in f], dtype=torch.float64, device=device).view(-1, 1, 1)
        par = torch.tensor([v[2] for v in f], device=device).view(-1, 1, 1)
        off = torch.tensor(offs[m0:m0 + chunk], dtype=torch.float64, device=device).view(-1, 1, 1)
        arg = 2 * math.pi * (kx * xx + ky * yy) + off
        mode = torch.where(par == 0, torch.cos(arg), torch.sin(arg)) * wf
        mean = mode.sum(dim=(1, 2), keepdim=True) / npx
        mode = (mode - mean) * wf
        rms = torch.sqrt((mode * mode).sum(dim=(1, 2), keepdim=True) / npx)
        out[m0:m0 + len(f)] = (mode / rms).to(dtype)
    return out
