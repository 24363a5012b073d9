"""Slopec base (specula/processing_objects/slopec.py:11-104)."""
import numpy as np
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

    def _device_mat(self, key, m, dtype):
        """Matrix of a Recmat / Intmat as a contiguous device tensor of ``dtype``.  A conversion
        (dtype change or a copy of a sliced view) is made once per source tensor; reassigning the
        object's matrix (as Recmat.reduce_size does) triggers a new one."""
        src = self._mat(m)
        cache = self.__dict__.setdefault('_mat_cache', {})
        hit = cache.get(key)
        if hit is None or hit[0] is not src or hit[1].dtype != dtype:
            t = torch.as_tensor(src if torch.is_tensor(src) else np.asarray(src), device=self.device, dtype=dtype)
            hit = (src, t.contiguous())
            cache[key] = hit
        return hit[1]

    @staticmethod
    def _product(a, b, out, alpha=1.0, beta=0.0, trans_b=False):
        """out = alpha a @ op(b) + beta out: the library's matrix-core GEMM for FP32 on the GPU,
        torch for FP64 and on the CPU."""
        if out.is_cuda and out.dtype == torch.float32:
            from specula_amd import _native
            return _native.gemm(a, b, out, alpha=alpha, beta=beta, trans_b=trans_b)
        bb = b.T if trans_b else b
        return torch.addmm(out, a, bb, beta=beta, alpha=alpha, out=out) if beta != 0 else torch.mm(a, bb, out=out)

    def post_trigger(self):
        """Optional reconstruction / modal filtering (slopec.py:92-104) as device GEMMs."""
        super().post_trigger()
        s = self.slopes.slopes
        if self.recmat is not None:
            R = self._device_mat('recmat', self.recmat, s.dtype)
            out = torch.empty(R.shape[1], dtype=s.dtype, device=self.device)
            self._product(s.reshape(1, -1).contiguous(), R, out.view(1, -1))
            self.slopes.slopes = out
        if self.filt_intmat is not None and self.filt_recmat is not None:
            s = self.slopes.slopes
            IM = self._device_mat('filt_intmat', self.filt_intmat, s.dtype)
            RM = self._device_mat('filt_recmat', self.filt_recmat, s.dtype)
            m = torch.empty(RM.shape[1], dtype=s.dtype, device=self.device)
            self._product(s.reshape(1, -1).contiguous(), RM, m.view(1, -1))
            # slopes -= m @ IM^T, in place
            self._product(m.view(1, -1), IM, s.view(1, -1), alpha=-1.0, beta=1.0, trans_b=True)
