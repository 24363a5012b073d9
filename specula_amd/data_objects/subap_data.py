import math

import numpy as np
import torch

from specula_amd.base import BaseDataObj


class SubapData(BaseDataObj):
    """specula/data_objects/subap_data.py:10-79.

    idxs: (n_subaps, np_sub**2) linear pixel indices of each subaperture (host int64);
    display_map: (n_subaps,) subaperture position on the nx*ny grid."""

    def __init__(self, idxs, display_map, nx: int, ny: int, energy_th: float = 0,
                 target_device_idx: int = None, precision: int = None):
        super().__init__(target_device_idx=target_device_idx, precision=precision)
        self.idxs = np.asarray(idxs.cpu() if isinstance(idxs, torch.Tensor) else idxs).astype(np.int64)
        self.display_map = np.asarray(display_map.cpu() if isinstance(display_map, torch.Tensor)
                                      else display_map).astype(np.int64)
        self.nx = int(nx)
        self.ny = int(ny)
        self.energy_th = float(energy_th)

    @property
    def n_subaps(self):
        return self.idxs.shape[0]

    @property
    def np_sub(self):
        return int(math.sqrt(self.idxs.shape[1]))

    def single_mask(self):
        f = torch.zeros(self.nx * self.ny, dtype=self.dtype, device=self.device)
        f[torch.as_tensor(self.display_map, device=self.device)] = 1
        return f.reshape((self.nx, self.ny))

    def subap_idx(self, n):
        return self.idxs[n, :]

    def display_map_idx(self, n):
        return self.display_map[n]

    def save(self, filename):
        """FITS layout of subap_data.py:53-63 (idxs stored transposed)."""
        from specula_amd.lib import fits
        hdr = {'VERSION': 1, 'ENRGYTH': self.energy_th, 'NP_SUB': self.np_sub, 'NX': self.nx, 'NY': self.ny}
        fits.write(filename, [(np.zeros(2), hdr), (self.idxs.T.copy(), {}), (self.display_map, {})])

    @classmethod
    def restore(cls, filename, target_device_idx=None):
        """subap_data.py:65-79."""
        from specula_amd.lib import fits
        hdus = fits.read(filename)
        hdr = hdus[0][1]
        if hdr.get('VERSION') != 1:
            raise ValueError(f"Unknown version {hdr.get('VERSION')} in file {filename}")
        return cls(idxs=hdus[1][0].T, display_map=hdus[2][0], nx=hdr.get('NX'), ny=hdr.get('NY'),
                   energy_th=hdr.get('ENRGYTH'), target_device_idx=target_device_idx)
