import torch

from specula_amd.base import BaseDataObj


class Pixels(BaseDataObj):
    """specula/data_objects/pixels.py:6-95 (pixel values stored in the float dtype, as the reference)."""

    def __init__(self, dimx: int, dimy: int, bits: int = 16, signed: int = 0,
                 target_device_idx: int = None, precision: int = None):
        super().__init__(target_device_idx=target_device_idx, precision=precision)
        if bits > 64:
            raise ValueError('Cannot create pixel object with more than 64 bits per pixel')
        self.signed = signed
        self.pixels = torch.zeros((dimx, dimy), dtype=self.dtype, device=self.device)
        self.bpp = bits
        self.bytespp = (bits + 7) // 8

    def set_value(self, v):
        self.pixels[:] = torch.as_tensor(v, dtype=self.dtype, device=self.device)

    @property
    def size(self):
        return tuple(self.pixels.shape)

    def multiply(self, factor):
        self.pixels *= factor

    def set_size(self, size):
        self.pixels = torch.zeros(size, dtype=self.dtype, device=self.device)
