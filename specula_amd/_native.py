"""ctypes binding of libshwfs.so (declarations in include/shwfs.h).

torch is imported before the library is loaded: torch ships its own HIP runtime whose
SONAME (libamdhip64.so.7) then satisfies libshwfs.so's dependency, so torch streams and
allocations and the library's launches share one runtime.  There is no CPU fallback: if the
library is missing every entry point raises.
"""
import ctypes
import os

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'lib', 'libshwfs.so')

# every symbol declared in include/shwfs.h
SHWFS_OK, SHWFS_ERR_ARG, SHWFS_ERR_HIP, SHWFS_ERR_UNSUPPORTED = 0, -1, -2, -3   # include/shwfs.h

EXPORTS = (
    'shwfs_version', 'shwfs_last_error',
    'shwfs_plan_create', 'shwfs_plan_create_masked', 'shwfs_plan_destroy', 'shwfs_plan_set_timing_events', 'shwfs_plan_set_frames_per_block',
    'shwfs_set_kernels', 'shwfs_forward',
    'shslopec_plan_create', 'shslopec_plan_destroy', 'shslopec_forward',
    'shwfs_plan_bind_slopec', 'shwfs_forward_slopes', 'shwfs_plan_set_pupil', 'shwfs_plan_pupil_registered',
    'shwfs_plan_set_resampling',
    'shslopec_vjp', 'shwfs_vjp', 'shwfs_vjp_sensor_grid', 'shwfs_resample', 'shwfs_pushpull_fields', 'shwfs_pushpull_rows', 'shwfs_im_from_jacobian', 'shwfs_im_from_jacobian_noalloc',
    'shwfs_batch_create', 'shwfs_batch_forward', 'shwfs_batch_set_slopes', 'shwfs_batch_latencies', 'shwfs_batch_info',
    'shwfs_batch_destroy',
    'shccd_integrate', 'shccd_readout', 'shccd_readout_cte', 'shccd_philox',
    'shwfs_graph_begin', 'shwfs_graph_end', 'shwfs_graph_launch', 'shwfs_graph_destroy',
    'shwfs_graph_node_count',
    'shgemm_f32',
)

SHCCD_PHOTON, SHCCD_READOUT, SHCCD_EXCESS, SHCCD_DARK, SHCCD_BACKGROUND, SHCCD_CIC = 1, 2, 4, 8, 16, 32
SHCCD_KEEP_BIAS, SHCCD_BG_REMOVE_AVG, SHCCD_KEEP_DARK, SHCCD_OFFSET_APPLIED = 64, 128, 256, 512


class ShccdParams(ctypes.Structure):
    """Mirror of ``shccd_params`` in include/shwfs.h."""
    _fields_ = [
        ('in_rows', ctypes.c_int32), ('in_cols', ctypes.c_int32), ('out_rows', ctypes.c_int32),
        ('out_cols', ctypes.c_int32), ('binning', ctypes.c_int32), ('flags', ctypes.c_int32),
        ('qe', ctypes.c_double), ('background_level', ctypes.c_double), ('darkcurrent_level', ctypes.c_double),
        ('readout_level', ctypes.c_double), ('cic_level', ctypes.c_double), ('adu_gain', ctypes.c_double),
        ('adu_bias', ctypes.c_double), ('emccd_gain', ctypes.c_double), ('excess_delta', ctypes.c_double),
        ('photon_seed', ctypes.c_uint64), ('readout_seed', ctypes.c_uint64), ('excess_seed', ctypes.c_uint64),
        ('cic_seed', ctypes.c_uint64),
    ]


class ShwfsGeom(ctypes.Structure):
    """Mirror of ``shwfs_geom`` in include/shwfs.h."""
    _fields_ = [
        ('n_lenses', ctypes.c_int32), ('ef_size', ctypes.c_int32), ('np_sub', ctypes.c_int32),
        ('fft_size', ctypes.c_int32), ('spec_size', ctypes.c_int32), ('subap_npx', ctypes.c_int32),
        ('lgs', ctypes.c_int32), ('parseval', ctypes.c_int32), ('max_frames', ctypes.c_int32),
        ('variant', ctypes.c_int32),
        ('k_rad_per_nm', ctypes.c_double), ('tilt_rad', ctypes.c_double), ('tilt_a0', ctypes.c_double),
        ('norm_const', ctypes.c_double), ('pixel_area', ctypes.c_double),
    ]


_lib = None

_P = ctypes.c_void_p
_I = ctypes.c_int
_I64 = ctypes.c_int64
_D = ctypes.c_double
_F = ctypes.c_float


def _declare(lib):
    sig = {
        'shwfs_version': (ctypes.c_char_p, []),
        'shwfs_last_error': (ctypes.c_char_p, []),
        'shwfs_plan_create': (_I, [ctypes.POINTER(ShwfsGeom), _P, _P, _I, ctypes.POINTER(_P)]),
        'shwfs_plan_create_masked': (_I, [ctypes.POINTER(ShwfsGeom), _I, _P, _P, _I, ctypes.POINTER(_P)]),
        'shwfs_plan_destroy': (_I, [_P]),
        'shwfs_plan_set_timing_events': (_I, [_P, _P, _P]),
        'shwfs_plan_set_frames_per_block': (_I, [_P, _I]),
        'shwfs_set_kernels': (_I, [_P, _P, _P]),
        'shwfs_forward': (_I, [_P, _I, _P, _I64, _P, _I64, _D, _P, _I64, _P]),
        'shslopec_plan_create': (_I, [_I, _I, _I, _I, _P, _P, _P, _P, _P, _I, _I, ctypes.POINTER(_P)]),
        'shslopec_plan_destroy': (_I, [_P]),
        'shslopec_forward': (_I, [_P, _I, _P, _I64, _F, _I, _P, _I64, _P, _P, _P]),
        'shwfs_plan_bind_slopec': (_I, [_P, _P]),
        'shwfs_plan_set_pupil': (_I, [_P, _P, _I64, _I, _D, _P]),
        'shwfs_plan_pupil_registered': (_I, [_P, _P, _I64, _D]),
        'shwfs_plan_set_resampling': (_I, [_P, _I, _P, _P, _P]),
        'shwfs_forward_slopes': (_I, [_P, _I, _P, _I64, _P, _I64, _D, _P, _I64, _F, _I, _P, _I64, _P, _P, _P]),
        'shslopec_vjp': (_I, [_P, _I, _P, _I64, _F, _I, _P, _I64, _P, _P]),
        'shwfs_vjp': (_I, [_P, _I, _P, _I64, _P, _I64, _D, _P, _I64, _P, _I64, _P]),
        'shwfs_vjp_sensor_grid': (_I, [_P, _I, _P, _I64, _P, _I64, _D, _P, _I64, _P, _I64, _P]),
        'shwfs_resample': (_I, [_P, _I, _P, _I64, _P, _I64, _P]),
        'shwfs_pushpull_fields': (_I, [_I, _P, _I64, _P, _I64, _F, _P, _I64, _P]),
        'shwfs_pushpull_rows': (_I, [_I, _P, _I64, _I, _F, _P, _I64, _P]),
        'shwfs_im_from_jacobian': (_I, [_I, _I, _I, _P, _P, _I, _P, _I64, _P, _I, _F, _P, _P, _P]),
        'shwfs_im_from_jacobian_noalloc': (_I, [_I, _I, _I, _P, _P, _I, _P, _I64, _P, _I, _F, _P, _P]),
        'shwfs_batch_create': (_I, [_I, _P, _P, _P, _P, _P, _P, _I, _I, _I, ctypes.POINTER(_P)]),
        'shwfs_batch_forward': (_I, [_P, _P]),
        'shwfs_batch_set_slopes': (_I, [_P, _I, _F, _I, _P, _I64, _P, _P]),
        'shwfs_batch_latencies': (_I, [_P, _P, _P]),
        'shwfs_batch_info': (_I, [_P, _P, _P]),
        'shwfs_batch_destroy': (_I, [_P]),
        'shccd_integrate': (_I, [_I, _P, _I64, _D, _P, _I64, _I64, _P]),
        'shccd_philox': (_I, [_I, _P, _P, _P]),
        'shccd_readout': (_I, [ctypes.POINTER(ShccdParams), _I, _P, _I64, _P, _I64, _D, _P, ctypes.c_uint64, _P,
                               _I64, _P]),
        'shccd_readout_cte': (_I, [ctypes.POINTER(ShccdParams), _I, _P, _I64, _P, _I64, _D, _P, _P, _P, _P,
                                   ctypes.c_uint64, _P, _I64, _P]),
    }
    sig.update({
        'shwfs_graph_begin': (_I, [_P]),
        'shwfs_graph_end': (_I, [_P, ctypes.POINTER(_P)]),
        'shwfs_graph_launch': (_I, [_P, _P]),
        'shwfs_graph_destroy': (_I, [_P]),
        'shwfs_graph_node_count': (_I, [_P, ctypes.POINTER(_I64)]),
        'shgemm_f32': (_I, [_I, _I, _I, _F, _P, _I64, _P, _I64, _I, _F, _P, _P, _I64, _P]),
    })
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args


def load_library(path=LIB_PATH):
    """Load libshwfs.so (after torch, see module docstring).  Raises if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    import torch  # noqa: F401  -- must precede the dlopen (shared HIP runtime)
    if not os.path.exists(path):
        raise RuntimeError(f'{path} not found: build it with `python -m specula_amd.build`')
    lib = ctypes.CDLL(path)
    _declare(lib)
    _lib = lib
    return lib


def lib():
    return load_library()


def check(rc, what=''):
    if rc != 0:
        msg = lib().shwfs_last_error().decode(errors='replace')
        raise RuntimeError(f'{what} failed ({rc}): {msg}')


def stream_handle(stream=None):
    """hipStream_t of a torch stream (default: the current stream of the current device) as a
    ctypes pointer.  The current stream is read as a raw handle (building a torch Stream object
    for it costs microseconds per call)."""
    import torch
    if stream is None:
        return ctypes.c_void_p(torch._C._cuda_getCurrentRawStream(torch.cuda.current_device()))
    return ctypes.c_void_p(stream.cuda_stream)


def ptr(t):
    """Device pointer of a torch tensor."""
    return ctypes.c_void_p(t.data_ptr())


def gemm(A, B, out, alpha=1.0, beta=0.0, bias=None, trans_b=False, stream=None):
    """out = alpha * A @ op(B) + beta * out + bias (shgemm_f32, FP32 matrix cores).  A (m, k), B (k, n)
    or, with trans_b, (n, k), out (m, n): float32 device tensors with unit column stride; bias (n,)
    or None."""
    import torch
    if A.dim() != 2 or B.dim() != 2 or out.dim() != 2:
        raise ValueError('gemm takes 2-D tensors')
    m, k = A.shape
    n = B.shape[0] if trans_b else B.shape[1]
    if (B.shape[1] if trans_b else B.shape[0]) != k or tuple(out.shape) != (m, n):
        raise ValueError(f'gemm shapes do not match: A {tuple(A.shape)}, B {tuple(B.shape)}, out {tuple(out.shape)}')
    for t in (A, B, out) + ((bias,) if bias is not None else ()):
        if t.dtype != torch.float32 or not t.is_cuda or t.stride(-1) != 1:
            raise ValueError('gemm operands must be float32 device tensors with unit column stride')
    if bias is not None and (bias.dim() != 1 or bias.shape[0] != n or not bias.is_contiguous()):
        raise ValueError(f'bias must be a contiguous ({n},) tensor')
    check(lib().shgemm_f32(m, n, k, float(alpha), ptr(A), A.stride(0), ptr(B), B.stride(0), 1 if trans_b else 0,
                           float(beta), ptr(bias) if bias is not None else None, ptr(out), out.stride(0),
                           stream_handle(stream)), 'shgemm_f32')
    return out


class HipGraph:
    """A HIP graph of library calls (shwfs_graph_*): ``capture(fn, stream)`` records the calls
    that ``fn`` issues on ``stream`` (a non-default torch stream, made current while ``fn`` runs);
    ``replay(stream)`` launches them on ``stream`` (default: the current stream)."""

    def __init__(self):
        self._h = None

    def capture(self, fn, stream):
        import torch
        h = stream_handle(stream)
        check(lib().shwfs_graph_begin(h), 'shwfs_graph_begin')
        out = ctypes.c_void_p()
        try:
            with torch.cuda.stream(stream):
                fn()
        finally:
            rc = lib().shwfs_graph_end(h, ctypes.byref(out))
        check(rc, 'shwfs_graph_end')
        self.destroy()
        self._h = out
        return self

    def replay(self, stream=None):
        if self._h is None:
            raise RuntimeError('no graph captured')
        check(lib().shwfs_graph_launch(self._h, stream_handle(stream)), 'shwfs_graph_launch')

    def node_count(self):
        """Kernel launches (graph nodes) that one replay issues."""
        if self._h is None:
            raise RuntimeError('no graph captured')
        n = ctypes.c_int64()
        check(lib().shwfs_graph_node_count(self._h, ctypes.byref(n)), 'shwfs_graph_node_count')
        return n.value

    def destroy(self):
        if self._h is not None and self._h.value:
            lib().shwfs_graph_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass
