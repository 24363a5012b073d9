"""Slopec base (specula/processing_objects/slopec.py:11-104)."""
import torch

from specula_amd.base import BaseProcessingObj, BaseValue, InputValue
from specula_amd.data_objects.pixels import Pixels
from specula_amd.data_objects.slopes import Slopes


class Slopec(BaseProcessingObj):
    def __init__(self, sn: Slopes = None, use_sn: bool = False, accumulate: bool = False,
                 weight_from_accumulated: bool = False, recmat=None, filt_intmat=None,
                 filt_recmat=None, filtmat=None, accumulation_dt: float = 0,
                 accumulated_pixels: tuple = (0, 0), target_device_idx: int = None,
                 precision: int = None):
        super().__init__(target_device_idx=target_device_idx, precision=precision)
        if use_sn and not sn:
            raise ValueError('Slopes null are not valid')
        if weight_from_accumulated and accumulate:
            raise ValueError('weightFromAccumulated and accumulate must not be set together')
        self.slopes = Slopes(2, target_device_idx=self.target_device_idx, precision=self.precision)
        self.sn = sn
        self.flux_per_subaperture_vector = BaseValue(target_device_idx=self.target_device_idx)
        self.max_flux_per_subaperture_vector = BaseValue(target_device_idx=self.target_device_idx)
        self.use_sn = use_sn
        self.accumulate = accumulate
        self.weight_from_accumulated = weight_from_accumulated
        self.recmat = recmat
        if filtmat is not None:
            if filt_intmat:
                raise ValueError('filt_intmat must not be set if "filtmat" is set')
            if filt_recmat:
                raise ValueError('filt_recmat must not be set if "filtmat" is set')
            self.filt_intmat = filtmat[0]      # Intmat(filtmat[0]), slopec.py:49
            self.filt_recmat = filtmat[1]      # Recmat(filtmat[1]), slopec.py:50
        else:
            if bool(filt_intmat is not None) != bool(filt_recmat is not None):
                raise ValueError('Both filt_intmat and filt_recmat must be set for slopes filtering')
            self.filt_intmat = filt_intmat
            self.filt_recmat = filt_recmat
        self.accumulation_dt = self.seconds_to_t(accumulation_dt)
        self.accumulated_slopes = Slopes(2, target_device_idx=self.target_device_idx, precision=self.precision)
        self.inputs['in_pixels'] = InputValue(type=Pixels)
        self.outputs['out_slopes'] = self.slopes

    def trigger_code(self):
        raise NotImplementedError(f'{self.__class__.__name__}: please implement trigger_code() in your derived class!')

    @staticmethod
    def _mat(m):
        return getattr(m, 'recmat', getattr(m, 'intmat', m))

    def post_trigger(self):
        """Optional reconstruction / modal filtering (slopec.py:92-104) as library GEMMs."""
        super().post_trigger()
        if self.recmat is not None:
            R = torch.as_tensor(self._mat(self.recmat), device=self.device, dtype=self.slopes.slopes.dtype)
            self.slopes.slopes = self.slopes.slopes @ R
        if self.filt_intmat is not None and self.filt_recmat is not None:
            IM = torch.as_tensor(self._mat(self.filt_intmat), device=self.device, dtype=self.slopes.slopes.dtype)
            RM = torch.as_tensor(self._mat(self.filt_recmat), device=self.device, dtype=self.slopes.slopes.dtype)
            m = self.slopes.slopes @ RM
            self.slopes.slopes -= m @ IM.T
