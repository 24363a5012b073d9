import numpy as np
import torch

from specula_amd.base import BaseDataObj
from specula_amd.data_objects.recmat import Recmat


def _fname(filename):
    return filename if filename.endswith('.fits') else filename + '.fits'


class Intmat(BaseDataObj):
    """Interaction matrix (n_modes, n_slopes), specula/data_objects/intmat.py:7-111: FITS layout
    (primary header VERSION/PUP_TAG/NORMFACT over a 2-element dummy, the matrix in extension 1,
    optional slope min/max and rms extensions), ``generate_rec`` = pseudo-inverse."""

    def __init__(self, intmat, slope_mm=None, slope_rms=None, pupdata_tag: str = '', norm_factor: float = 0.0,
                 target_device_idx: int = None, precision: int = None):
        super().__init__(target_device_idx=target_device_idx, precision=precision)
        self.intmat = torch.as_tensor(np.asarray(intmat) if not torch.is_tensor(intmat) else intmat,
                                      device=self.device)
        self.slope_mm = slope_mm
        self.slope_rms = slope_rms
        self.pupdata_tag = pupdata_tag
        self.norm_factor = norm_factor

    def reduce_size(self, n_modes_to_be_discarded):
        nmodes = self.intmat.shape[0]
        if n_modes_to_be_discarded >= nmodes:
            raise ValueError(f'nModesToBeDiscarded should be less than nmodes (<{nmodes})')
        self.intmat = self.intmat[:nmodes - n_modes_to_be_discarded, :]

    def reduce_slopes(self, n_slopes_to_be_discarded):
        nslopes = self.intmat.shape[1]
        if n_slopes_to_be_discarded >= nslopes:
            raise ValueError(f'nSlopesToBeDiscarded should be less than nslopes (<{nslopes})')
        self.intmat = self.intmat[:, :nslopes - n_slopes_to_be_discarded]

    def set_start_mode(self, start_mode):
        nmodes = self.intmat.shape[0]
        if start_mode >= nmodes:
            raise ValueError(f'start_mode should be less than nmodes (<{nmodes})')
        self.intmat = self.intmat[start_mode:, :]

    def generate_rec(self, nmodes=None, cut_modes=0, w_vec=None, interactive=False):
        """intmat.py:74-87: Recmat of pinv(intmat[:nmodes]) (n_slopes, n_modes).  Calibration-time
        work: the SVD runs on the host in FP64 and the result returns to the object's device."""
        im = self.intmat[:nmodes, :] if nmodes is not None else self.intmat
        rec = torch.linalg.pinv(im.detach().double().cpu()).to(device=im.device, dtype=im.dtype)
        out = Recmat(rec, target_device_idx=self.target_device_idx, precision=self.precision)
        out.im_tag = self.norm_factor
        return out

    def save(self, filename):
        from specula_amd.lib import fits
        hdus = [(np.zeros(2), {'VERSION': 1, 'PUP_TAG': self.pupdata_tag or '', 'NORMFACT': float(self.norm_factor)}),
                (self.intmat.cpu().numpy(), {})]
        if self.slope_mm is not None:
            hdus.append((np.asarray(self.slope_mm), {}))
        if self.slope_rms is not None:
            hdus.append((np.asarray(self.slope_rms), {}))
        fits.write(_fname(filename), hdus)

    @staticmethod
    def restore(filename, target_device_idx=None):
        from specula_amd.lib import fits
        hdus = fits.read(filename)
        hdr = hdus[0][1]
        mm = rms = None
        if len(hdus) >= 4:
            mm, rms = hdus[2][0], hdus[3][0]
        return Intmat(hdus[1][0], mm, rms, hdr.get('PUP_TAG', ''), float(hdr.get('NORMFACT', 0.0)),
                      target_device_idx=target_device_idx)
