import numpy as np
import torch

from specula_amd.base import BaseDataObj


class Recmat(BaseDataObj):
    """Reconstruction matrix, specula/data_objects/recmat.py:8-70.  ``recmat`` is (n_slopes,
    n_modes) as pinv(intmat) returns it; the FITS file holds its transpose in extension 1 under a
    primary header with VERSION and NORMFACT, plus an optional modes2recLayer extension."""

    def __init__(self, recmat, modes2recLayer=None, norm_factor: float = 0,
                 target_device_idx: int = None, precision: int = None):
        super().__init__(target_device_idx=target_device_idx, precision=precision)
        self.recmat = torch.as_tensor(np.asarray(recmat) if not torch.is_tensor(recmat) else recmat,
                                      device=self.device)
        self.norm_factor = norm_factor
        self.proj_list = []
        self.modes2recLayer = None if modes2recLayer is None else np.asarray(modes2recLayer)

    def reduce_size(self, nModesToBeDiscarded):
        nmodes = self.recmat.shape[1]
        if nModesToBeDiscarded >= nmodes:
            raise ValueError(f'nModesToBeDiscarded should be less than nmodes (<{nmodes})')
        self.recmat = self.recmat[:, :nmodes - nModesToBeDiscarded]

    def save(self, filename):
        from specula_amd.lib import fits
        filename = filename if filename.endswith('.fits') else filename + '.fits'
        hdus = [(np.zeros(2), {'VERSION': 1, 'NORMFACT': self.norm_factor}),
                (self.recmat.cpu().numpy().T.copy(), {})]
        if self.modes2recLayer is not None:
            hdus.append((self.modes2recLayer, {}))
        fits.write(filename, hdus)

    @staticmethod
    def restore(filename, target_device_idx=None):
        """recmat.py:52-68.  As in the reference, the array read back is the stored
        (n_modes, n_slopes) transpose."""
        from specula_amd.lib import fits
        hdus = fits.read(filename)
        hdr = hdus[0][1]
        if int(hdr.get('VERSION', -1)) != 1:
            raise ValueError(f"Error: unknown version {hdr.get('VERSION')} in file {filename}")
        m2r = hdus[2][0] if len(hdus) >= 3 else None
        return Recmat(hdus[1][0], m2r, float(hdr['NORMFACT']), target_device_idx=target_device_idx)
