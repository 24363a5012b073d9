import numpy as np

from specula_amd.data_objects.electric_field import ElectricField


class Layer(ElectricField):
    """specula/data_objects/layer.py:7-24: an ElectricField with a conjugation height and the
    geometric transforms (shift, rotation, magnification) a propagation step reads."""

    def __init__(self, dimx: int, dimy: int, pixel_pitch: float, height: float,
                 shiftXYinPixel: tuple = (0.0, 0.0), rotInDeg: float = 0.0, magnification: float = 1.0,
                 target_device_idx: int = None, precision: int = None):
        super().__init__(dimx, dimy, pixel_pitch, target_device_idx=target_device_idx, precision=precision)
        self.height = height
        self.shiftXYinPixel = np.asarray(shiftXYinPixel, dtype=np.float64)
        self.rotInDeg = rotInDeg
        self.magnification = magnification
