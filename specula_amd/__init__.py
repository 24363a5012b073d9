"""specula_amd -- MI355X-native Shack-Hartmann wavefront-sensor hot path.

A drop-in for SPECULA's ``SH`` and ``ShSlopec`` processing objects (same constructor
keywords = YAML keys, same ports and lifecycle) whose per-frame work runs in hand-written
gfx950 HIP kernels called through a ctypes C ABI (``include/shwfs.h``).  Device buffers are
torch tensors; torch is used for device memory and streams only.

``init(device_idx, precision)`` mirrors ``specula.init`` (specula/__init__.py:36-94):
device_idx -1 selects the CPU for data objects; the SH/ShSlopec compute path itself exists
only on the GPU and raises if asked to run elsewhere.
"""
import numpy as np

__version__ = '0.1.0'

ASEC2RAD = np.pi / (3600 * 180)
RAD2ASEC = 1.0 / ASEC2RAD

default_target_device_idx = 0
global_precision = 1


def init(device_idx=0, precision=1):
    """Set the default target device and precision (0 = FP64, 1 = FP32), as specula.init."""
    global default_target_device_idx, global_precision
    if precision not in (0, 1):
        raise ValueError('precision must be 0 (double) or 1 (single)')
    default_target_device_idx = int(device_idx)
    global_precision = int(precision)
