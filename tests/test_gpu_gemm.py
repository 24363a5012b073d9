"""shgemm_f32 (the matrix-core FP32 GEMM behind the DM projection and the Slopec reconstruction)
against float64 numpy on the same inputs: GEMV rows (m = 1, the per-step DM trigger), ragged edges
of the 64 x 64 block tile and the 16-wide K chunk, op(B) = B^T, bias, alpha / beta accumulation,
and row pitches wider than the rows.  FP32 products accumulated in FP32: the bound is a few ulp
times sqrt(k) of the sum of |terms|, written below per case."""
import numpy as np
import pytest
import torch

from specula_amd import _native

pytestmark = pytest.mark.gpu

SHAPES = [
    (1, 4096, 40),      # DM trigger: one command vector, 64 x 64 grid, 40 modes
    (1, 25600, 5),      # K below one chunk
    (37, 1000, 129),    # every edge ragged
    (64, 64, 16),       # exactly one tile and one chunk
    (200, 3, 517),      # narrow output, long K
    (130, 257, 1),      # K = 1
]


def _case(m, n, k, trans_b, seed):
    g = np.random.default_rng(seed)
    A = g.standard_normal((m, k)).astype(np.float32)
    B = g.standard_normal((n, k) if trans_b else (k, n)).astype(np.float32)
    return A, B


def _check(got, A, B, trans_b, alpha=1.0, beta=0.0, C0=None, bias=None):
    Bo = B.astype(np.float64).T if trans_b else B.astype(np.float64)
    want = alpha * (A.astype(np.float64) @ Bo)
    mag = abs(alpha) * (np.abs(A.astype(np.float64)) @ np.abs(Bo))
    if C0 is not None:
        want += beta * C0.astype(np.float64)
        mag += abs(beta) * np.abs(C0.astype(np.float64))
    if bias is not None:
        want += bias.astype(np.float64)[None, :]
        mag += np.abs(bias.astype(np.float64))[None, :]
    k = A.shape[1]
    tol = 4 * np.finfo(np.float32).eps * np.sqrt(k + 2) * mag + 1e-30
    err = np.abs(got.astype(np.float64) - want)
    assert np.all(err <= tol), float((err / tol).max())


@pytest.mark.parametrize('trans_b', [False, True])
@pytest.mark.parametrize('shape', SHAPES)
def test_gemm_matches_float64(shape, trans_b):
    m, n, k = shape
    A, B = _case(m, n, k, trans_b, seed=m * 7 + n + k)
    dev = 'cuda:0'
    C = torch.full((m, n), np.nan, dtype=torch.float32, device=dev)
    _native.gemm(torch.from_numpy(A).to(dev), torch.from_numpy(B).to(dev), C, trans_b=trans_b)
    torch.cuda.synchronize()
    _check(C.cpu().numpy(), A, B, trans_b)


@pytest.mark.parametrize('trans_b', [False, True])
def test_gemm_alpha_beta_bias_and_pitch(trans_b):
    m, n, k = 45, 300, 77
    A, B = _case(m, n, k, trans_b, seed=3)
    g = np.random.default_rng(4)
    C0 = g.standard_normal((m, n)).astype(np.float32)
    bias = g.standard_normal(n).astype(np.float32)
    dev = 'cuda:0'
    # operands and output as column slices of wider arrays: leading dimension > row length
    Aw = torch.zeros((m, k + 9), dtype=torch.float32, device=dev)
    Aw[:, 3:3 + k] = torch.from_numpy(A).to(dev)
    Bw = torch.zeros((B.shape[0], B.shape[1] + 5), dtype=torch.float32, device=dev)
    Bw[:, :B.shape[1]] = torch.from_numpy(B).to(dev)
    Cw = torch.full((m, n + 11), 7.0, dtype=torch.float32, device=dev)
    Cw[:, 2:2 + n] = torch.from_numpy(C0).to(dev)
    _native.gemm(Aw[:, 3:3 + k], Bw[:, :B.shape[1]], Cw[:, 2:2 + n], alpha=-0.5, beta=1.5,
                 bias=torch.from_numpy(bias).to(dev), trans_b=trans_b)
    torch.cuda.synchronize()
    out = Cw.cpu().numpy()
    _check(out[:, 2:2 + n], A, B, trans_b, alpha=-0.5, beta=1.5, C0=C0, bias=bias)
    # the columns outside the output view are untouched
    assert np.all(out[:, :2] == 7.0) and np.all(out[:, 2 + n:] == 7.0)


def test_gemm_beta_zero_ignores_nan_output():
    # beta = 0 overwrites: NaN already in C must not leak through 0 * NaN
    m, n, k = 5, 70, 20
    A, B = _case(m, n, k, False, seed=9)
    dev = 'cuda:0'
    C = torch.full((m, n), float('nan'), dtype=torch.float32, device=dev)
    _native.gemm(torch.from_numpy(A).to(dev), torch.from_numpy(B).to(dev), C, beta=0.0)
    torch.cuda.synchronize()
    _check(C.cpu().numpy(), A, B, False)


def test_dm_trigger_and_project_products():
    """The DM trigger (a GEMV) and project() (a GEMM with the base phase as bias) agree with the
    influence-function product in float64."""
    from specula_amd.data_objects.ifunc import IFunc
    from specula_amd.processing_objects.dm import DM
    from specula_amd.base import BaseValue
    ifunc = IFunc(type_str='zernike', npixels=48, nmodes=30, target_device_idx=0, precision=1)
    dm = DM(ifunc=ifunc, pixel_pitch=0.05, sign=-1, target_device_idx=0, precision=1)
    cmd = np.linspace(-50, 50, 30).astype(np.float32)
    v = BaseValue(value=torch.from_numpy(cmd).to('cuda:0'), target_device_idx=0)
    v.generation_time = 1
    dm.inputs['in_command'].set(v)
    dm.setup()
    assert dm.check_ready(1)
    dm.trigger()
    dm.post_trigger()
    torch.cuda.synchronize()
    grid = dm._ifunc_grid.double().cpu().numpy()
    want = (-cmd.astype(np.float64)) @ grid
    got = dm.layer.phaseInNm.double().cpu().numpy().ravel()
    mag = np.abs(cmd.astype(np.float64)) @ np.abs(grid)
    assert np.all(np.abs(got - want) <= 4 * np.finfo(np.float32).eps * np.sqrt(32) * mag + 1e-30)
    F = np.random.default_rng(0).standard_normal((7, 30)).astype(np.float32)
    base = np.random.default_rng(1).standard_normal((48, 48)).astype(np.float32)
    ph = dm.project(torch.from_numpy(F).to('cuda:0'), base=torch.from_numpy(base).to('cuda:0'))
    torch.cuda.synchronize()
    want = base.reshape(1, -1).astype(np.float64) + (-F.astype(np.float64)) @ grid
    mag = np.abs(base.reshape(1, -1)) + np.abs(F.astype(np.float64)) @ np.abs(grid)
    got = ph.double().cpu().numpy().reshape(7, -1)
    assert np.all(np.abs(got - want) <= 4 * np.finfo(np.float32).eps * np.sqrt(32) * mag + 1e-30)
