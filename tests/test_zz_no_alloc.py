"""shwfs_im_from_jacobian_noalloc (library-owned scratch, no allocation) against
shwfs_im_from_jacobian with a workspace, through the C ABI: two calls in a row, the second on a
smaller sensor and fewer modes, each equal bit for bit.  The first call (1000 modes on the ELT
80x80 sensor) goes through the scratch in four chunks of modes."""
import ctypes

import numpy as np
import pytest
import torch

from specula_amd import _native, workloads
from specula_amd.lib.modes import fourier_modes
from specula_amd.sensor import SHSensor

pytestmark = pytest.mark.gpu
DEV = 'cuda:0'


def _both(sen, ph, modes):
    g = sen.geom
    jx, jy = sen.jacobian_maps(ph)
    K = modes.shape[0]
    ef = g.ef_size
    slot = sen.slot_map()
    lib = _native.lib()
    ref = torch.full((K, 2 * sen.n_sub), float('nan'), dtype=torch.float32, device=DEV)
    work = torch.empty(K * 2 * g.n_lenses ** 2 + sen.n_sub, dtype=torch.float32, device=DEV)
    st = _native.stream_handle()
    _native.check(lib.shwfs_im_from_jacobian(g.n_lenses, g.ovs_np_sub, ef, _native.ptr(jx), _native.ptr(jy), K,
                                             _native.ptr(modes), ef * ef, _native.ptr(slot), sen.n_sub, 2.0,
                                             _native.ptr(ref), _native.ptr(work), st), 'im with workspace')
    out = torch.full_like(ref, float('nan'))
    _native.check(lib.shwfs_im_from_jacobian_noalloc(g.n_lenses, g.ovs_np_sub, ef, _native.ptr(jx),
                                                     _native.ptr(jy), K, _native.ptr(modes), ef * ef,
                                                     _native.ptr(slot), sen.n_sub, 2.0, _native.ptr(out), st),
                  'im without allocation')
    torch.cuda.synchronize()
    return ref, out


def test_noalloc_im_equals_workspace_im_on_two_calls():
    print('no-alloc IM: call 1, elt_ngs_80x80, 1000 modes', flush=True)
    wl = workloads.elt_ngs_80x80()
    sen = SHSensor(wl.geometry(), wl.pupil(), max_frames=2, device=0)
    ph = torch.tensor(wl.frames(1)[0].astype(np.float32), device=DEV)
    ref, out = _both(sen, ph, fourier_modes(wl.pupil(), 1000, device=DEV))
    assert torch.isfinite(ref).all() and torch.equal(out, ref)
    print('no-alloc IM: call 2, scao_ngs_20x20, 37 modes', flush=True)
    wl2 = workloads.scao_ngs_20x20()
    sen2 = SHSensor(wl2.geometry(), wl2.pupil(), max_frames=2, device=0)
    ph2 = torch.tensor(wl2.frames(1)[0].astype(np.float32), device=DEV)
    ref2, out2 = _both(sen2, ph2, fourier_modes(wl2.pupil(), 37, device=DEV))
    assert torch.isfinite(ref2).all() and torch.equal(out2, ref2)
