"""The C-ABI library: it loads, exports every symbol include/shwfs.h declares, and rejects bad
arguments with a message before touching the GPU.  CPU only (no kernels are launched)."""
import ctypes
import os
import re

import numpy as np

from specula_amd import _native

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), 'include', 'shwfs.h')


def declared():
    text = open(HEADER).read()
    text = re.sub(r'/\*.*?\*/', '', text, flags=re.S)
    return set(re.findall(r'\b((?:shwfs|shslopec|shccd|shgemm)_[a-z0-9_]+)\s*\(', text))


def test_library_exports_header_symbols():
    lib = _native.load_library()
    names = declared()
    assert names == set(_native.EXPORTS)
    for n in names:
        assert hasattr(lib, n), n
    assert lib.shwfs_version().startswith(b'shwfs-gfx950')


def test_geom_struct_layout():
    # int32 x 10 then 5 doubles, no padding surprises
    assert ctypes.sizeof(_native.ShwfsGeom) == 10 * 4 + 5 * 8
    # int32 x 6, 9 doubles, 4 uint64
    assert ctypes.sizeof(_native.ShccdParams) == 6 * 4 + 9 * 8 + 4 * 8


def test_ccd_argument_errors_are_reported():
    lib = _native.load_library()
    p = _native.ShccdParams(in_rows=8, in_cols=8, out_rows=4, out_cols=4, binning=3, flags=0, qe=1.0)
    rc = lib.shccd_readout(ctypes.byref(p), 1, None, 0, ctypes.c_void_p(16), 64, 1.0, None, 0,
                           ctypes.c_void_p(16), 16, None)
    assert rc == -1 and b'in = out * binning' in lib.shwfs_last_error()
    p.binning = 2
    p.flags = _native.SHCCD_PHOTON
    p.adu_gain = 0.0
    rc = lib.shccd_readout(ctypes.byref(p), 1, None, 0, ctypes.c_void_p(16), 64, 1.0, None, 0,
                           ctypes.c_void_p(16), 16, None)
    assert rc == -1 and b'ADU gain' in lib.shwfs_last_error()
    rc = lib.shccd_integrate(1, None, 0, 1.0, None, 0, 16, None)
    assert rc == -1
    p.flags = 0
    rc = lib.shccd_readout_cte(ctypes.byref(p), 1, None, 0, ctypes.c_void_p(16), 64, 1.0, None, None,
                               ctypes.c_void_p(16), None, 0, ctypes.c_void_p(16), 16, None)
    assert rc == -1 and b'CTE matrices' in lib.shwfs_last_error()
    rc = lib.shccd_readout_cte(ctypes.byref(p), 0, None, 0, ctypes.c_void_p(16), 64, 1.0, ctypes.c_void_p(16),
                               ctypes.c_void_p(16), ctypes.c_void_p(16), None, 0, ctypes.c_void_p(16), 16, None)
    assert rc == -1 and b'n_frames' in lib.shwfs_last_error()
    assert lib.shccd_philox(-1, ctypes.c_void_p(16), ctypes.c_void_p(16), None) == -1
    assert lib.shccd_philox(4, None, ctypes.c_void_p(16), None) == -1
    assert lib.shccd_philox(0, None, None, None) == 0


def test_argument_errors_are_reported():
    lib = _native.load_library()
    h = ctypes.c_void_p()
    rc = lib.shwfs_plan_create(None, None, None, 0, ctypes.byref(h))
    assert rc == -1 and b'null' in lib.shwfs_last_error()
    g = _native.ShwfsGeom(n_lenses=4, ef_size=30, np_sub=8, spec_size=16, subap_npx=8, max_frames=1)
    w = np.zeros((8, 16))
    rc = lib.shwfs_plan_create(ctypes.byref(g), w.ctypes.data_as(ctypes.c_void_p),
                               w.ctypes.data_as(ctypes.c_void_p), 0, ctypes.byref(h))
    assert rc == -1 and b'ef_size' in lib.shwfs_last_error()
    g.ef_size = 32
    g.spec_size = 8
    rc = lib.shwfs_plan_create(ctypes.byref(g), w.ctypes.data_as(ctypes.c_void_p),
                               w.ctypes.data_as(ctypes.c_void_p), 0, ctypes.byref(h))
    assert rc == -1 and b'spec_size' in lib.shwfs_last_error()
    idx = np.array([[0, 1, 2, 99]], dtype=np.int32)
    wt = np.zeros(4)
    vp = lambda a: a.ctypes.data_as(ctypes.c_void_p)   # noqa: E731
    rc = lib.shslopec_plan_create(1, 2, 4, 4, None, vp(idx), vp(wt), vp(wt), vp(wt), 1, 0, ctypes.byref(h))
    assert rc == -1 and b'out of range' in lib.shwfs_last_error()
    org = np.array([3], dtype=np.int32)
    rc = lib.shslopec_plan_create(1, 2, 4, 4, vp(org), None, vp(wt), vp(wt), vp(wt), 1, 0, ctypes.byref(h))
    assert rc == -1 and b'outside' in lib.shwfs_last_error()
    rc = lib.shwfs_plan_set_timing_events(None, None, None)
    assert rc == -1


def test_grid_limits_are_rejected():
    lib = _native.load_library()
    h = ctypes.c_void_p()
    g = _native.ShwfsGeom(n_lenses=4, ef_size=32, np_sub=8, fft_size=32, spec_size=16, subap_npx=8,
                          max_frames=70000, parseval=1, norm_const=1.0, pixel_area=1.0)
    w = np.zeros((8, 16))
    rc = lib.shwfs_plan_create(ctypes.byref(g), w.ctypes.data_as(ctypes.c_void_p), w.ctypes.data_as(ctypes.c_void_p),
                               0, ctypes.byref(h))
    assert rc == -1 and b'65535' in lib.shwfs_last_error()
    p = _native.ShccdParams(in_rows=8, in_cols=8, out_rows=8, out_cols=8, binning=1, flags=0, qe=1.0)
    rc = lib.shccd_readout(ctypes.byref(p), 70000, None, 0, ctypes.c_void_p(16), 64, 1.0, None, 0,
                           ctypes.c_void_p(16), 64, None)
    assert rc == -1 and b'65535' in lib.shwfs_last_error()


def test_sensor_vjp_rejects_mismatched_frames_before_any_launch():
    """SHSensor.vjp: a per-frame phase must hold one frame per cotangent (the VJP kernel reads
    phase frame c for cotangent c); checked on the host, so no device buffer is touched."""
    import pytest
    import torch
    from types import SimpleNamespace
    from specula_amd.sensor import SHSensor
    s = SHSensor.__new__(SHSensor)
    s.geom = SimpleNamespace(ef_size=16, ccd_side=8)
    s.max_frames = 4
    s.n_sub = 5
    with pytest.raises(ValueError, match='2 phase frames for 3 cotangents'):
        s.vjp(torch.zeros((2, 16, 16)), torch.zeros((3, 10)))
    with pytest.raises(ValueError, match='grad_slopes must be'):
        s.vjp(torch.zeros((16, 16)), torch.zeros((3, 9)))
    with pytest.raises(ValueError, match='phase must be'):
        s.vjp(torch.zeros((3, 16, 15)), torch.zeros((3, 10)))
    with pytest.raises(ValueError, match='cotangents > max_frames'):
        s.vjp(torch.zeros((16, 16)), torch.zeros((5, 10)))


def test_noalloc_im_rejects_sizes_beyond_its_scratch_before_any_launch():
    """shwfs_im_from_jacobian_noalloc: n_valid * 4 + 8 n_lenses^2 bytes must fit the 16 MiB scratch;
    the size check precedes any HIP call (aligned dummy pointers are never dereferenced)."""
    lib = _native.load_library()
    p = ctypes.c_void_p(1 << 20)
    rc = lib.shwfs_im_from_jacobian_noalloc(80, 8, 640, p, p, 10, p, 640 * 640, p, 5_000_000, 2.0, p, None)
    assert rc == -1 and b'16 MiB scratch' in lib.shwfs_last_error()
