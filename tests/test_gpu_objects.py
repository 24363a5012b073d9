"""The drop-in SH and ShSlopec processing objects: lifecycle, ports, graph replay, errors.

Mirrors the reference's unit-test pattern (test/test_sh.py:34-40, test/test_sh_slopec.py:96-107):
``inputs[...].set(obj); setup(); check_ready(t); trigger(); post_trigger()``.
"""
import numpy as np
import pytest
import torch

import specula_amd
from helpers import check_intensity, check_slopes, oracle_geometry
from oracle import sh_oracle as O
from specula_amd import workloads
from specula_amd.data_objects.electric_field import ElectricField
from specula_amd.data_objects.pixels import Pixels
from specula_amd.data_objects.subap_data import SubapData
from specula_amd.lib.subaps import detect_subaps
from specula_amd.processing_objects.sh import SH
from specula_amd.processing_objects.sh_slopec import ShSlopec

pytestmark = pytest.mark.gpu
specula_amd.init(0, precision=1)


def make_sh(wl):
    sh = SH(**wl.sh_kwargs, target_device_idx=0)
    if wl.fov_resolution_arcsec:
        sh._fov_resolution_arcsec = wl.fov_resolution_arcsec
    return sh


def test_sh_and_slopec_lifecycle_matches_oracle():
    wl = workloads.scao_ngs_20x20()
    frames = wl.frames(3)
    A = wl.pupil()
    ef = ElectricField(wl.ef_size, wl.ef_size, wl.pixel_pitch, S0=10.0, target_device_idx=0)
    ef.A[:] = torch.tensor(A, device='cuda:0')
    sh = make_sh(wl)
    sh.inputs['in_ef'].set(ef)
    sh.setup()
    go = oracle_geometry(wl)
    flat = O.sh_forward(A, np.zeros_like(A), go)
    idxs, dmap = detect_subaps(flat, 20, 0.5)
    subapdata = SubapData(idxs=idxs, display_map=dmap, nx=20, ny=20, target_device_idx=0)
    slopec = ShSlopec(subapdata, target_device_idx=0)
    pix = Pixels(*sh.outputs['out_i'].i.shape, target_device_idx=0)
    slopec.inputs['in_pixels'].set(pix)
    for t, ph in enumerate(frames, start=1):
        ef.phaseInNm[:] = torch.tensor(ph, device='cuda:0')
        ef.generation_time = t
        assert sh.check_ready(t)
        sh.trigger()
        sh.post_trigger()
        out = sh.outputs['out_i']
        assert out.generation_time == t
        img = out.i.cpu().numpy()
        ref = O.sh_forward(A, ph.astype(np.float32).astype(np.float64), go, S0=10.0)
        check_intensity(img, ref, 20, 8)
        pix.pixels = out.i
        pix.generation_time = t
        slopec.check_ready(t)
        slopec.trigger()
        slopec.post_trigger()
        sref = O.shslopec(img.astype(np.float64), idxs)
        check_slopes(slopec.outputs['out_slopes'].slopes.cpu().numpy(), sref['slopes'], 8)
        np.testing.assert_allclose(slopec.flux_per_subaperture_vector.value.cpu().numpy(), sref['flux'], rtol=1e-12)
        np.testing.assert_allclose(float(slopec.total_counts.value), sref['total'], rtol=1e-12)
        assert slopec.outputs['out_subapdata'] is subapdata
    # no new input -> not ready
    assert not sh.check_ready(len(frames) + 1)


def test_graph_replay_matches_eager():
    """Replays of the captured SH graph, with eager work and syncs in between, equal eager runs."""
    wl = workloads.scao_ngs_20x20()
    frames = wl.frames(4)
    ef = ElectricField(wl.ef_size, wl.ef_size, wl.pixel_pitch, S0=1.0, target_device_idx=0)
    ef.A[:] = torch.tensor(wl.pupil(), device='cuda:0')
    eager, graph = make_sh(wl), make_sh(wl)
    eager.inputs['in_ef'].set(ef)
    graph.inputs['in_ef'].set(ef)
    eager.setup()
    graph.setup()
    graph.use_graph = True
    graph.graph_min_nodes = 1                       # keep the graph whatever its size
    for t, ph in enumerate(frames, start=1):
        ef.phaseInNm[:] = torch.tensor(ph, device='cuda:0')
        ef.generation_time = t
        for obj in (graph, eager):
            obj.check_ready(t)
            obj.trigger()
            obj.post_trigger()
        torch.cuda.synchronize()
        _ = (ef.phaseInNm * 2).sum().item()          # unrelated eager work between replays
        assert torch.equal(graph.outputs['out_i'].i, eager.outputs['out_i'].i)
    assert graph.cuda_graph is not None
    # an in-place change of the pupil: both refresh the registered statistics and subaperture
    # order (the graph outside its replay) and stay equal
    g0 = graph.cuda_graph
    ef.A[: wl.ef_size // 3] = 0.0
    for t, ph in enumerate(frames[:2], start=len(frames) + 1):
        ef.phaseInNm[:] = torch.tensor(ph, device='cuda:0')
        ef.generation_time = t
        for obj in (graph, eager):
            obj.check_ready(t)
            obj.trigger()
            obj.post_trigger()
        assert torch.equal(graph.outputs['out_i'].i, eager.outputs['out_i'].i)
    assert graph.cuda_graph is g0                    # same storage: no re-capture
    img = graph.outputs['out_i'].i.cpu().numpy()
    assert np.all(img[: img.shape[0] // 3 - 8] == 0)        # rows of subapertures under the dark band


def test_small_graph_declined():
    """use_graph on a one-launch trigger: the capture holds one node, fewer than graph_min_nodes,
    so the object keeps no graph and runs its trigger directly, with the same image."""
    wl = workloads.scao_ngs_20x20()
    ef = ElectricField(wl.ef_size, wl.ef_size, wl.pixel_pitch, S0=1.0, target_device_idx=0)
    ef.A[:] = torch.tensor(wl.pupil(), device='cuda:0')
    eager, graph = make_sh(wl), make_sh(wl)
    for obj in (eager, graph):
        obj.inputs['in_ef'].set(ef)
        obj.setup()
    graph.use_graph = True
    for t, ph in enumerate(wl.frames(2), start=1):
        ef.phaseInNm[:] = torch.tensor(ph, device='cuda:0')
        ef.generation_time = t
        for obj in (graph, eager):
            obj.check_ready(t)
            obj.trigger()
            obj.post_trigger()
        assert torch.equal(graph.outputs['out_i'].i, eager.outputs['out_i'].i)
    assert graph.graph_nodes == 1 and graph.graph_min_nodes > 1
    assert graph.cuda_graph is None


def test_setup_errors():
    wl = workloads.scao_ngs_20x20()
    sh = make_sh(wl)
    with pytest.raises(ValueError):
        sh.setup()                                     # in_ef not set (base_processing_obj.py:175-177)
    sh2 = SH(wavelengthInNm=500, subap_wanted_fov=3, sensor_pxscale=0.5, subap_on_diameter=20,
             subap_npx=6, target_device_idx=0)
    sh2.inputs['in_ef'].set(ElectricField(120, 100, 0.05, S0=1, target_device_idx=0))
    with pytest.raises(ValueError):
        sh2.setup()                                    # non-square field
    with pytest.raises(ValueError):
        sh2.inputs['in_ef'].set(Pixels(4, 4, target_device_idx=0))


def test_subapdata_restore_block_mode_and_thresholds():
    wl = workloads.scao_ngs_20x20()
    go = oracle_geometry(wl)
    A = wl.pupil()
    img = O.sh_forward(A, wl.frames(1)[0], go)
    idxs, dmap = detect_subaps(img, 20, 0.5)
    sd = SubapData(idxs=idxs, display_map=dmap, nx=20, ny=20, target_device_idx=0)
    pix = Pixels(*img.shape, target_device_idx=0)
    pix.pixels = torch.tensor(img, dtype=torch.float32, device='cuda:0')
    pix.generation_time = 1
    for thr, ped in ((-1, False), (3e-7, False), (3e-7, True)):
        sl = ShSlopec(sd, thr_value=thr, target_device_idx=0)
        sl.thr_pedestal = ped
        sl.inputs['in_pixels'].set(pix)
        sl.check_ready(1)
        sl.trigger()
        sl.post_trigger()
        ref = O.shslopec(img.astype(np.float32).astype(np.float64), idxs, thr_value=thr, thr_pedestal=ped)
        check_slopes(sl.slopes.slopes.cpu().numpy(), ref['slopes'], 8)


def test_slopec_post_trigger_recmat_and_filtering():
    """Slopec.post_trigger (slopec.py:92-104): slopes @ recmat, and the modal filtering
    s - (s @ filt_recmat) @ filt_intmat^T, as device GEMMs on the ShSlopec output."""
    from specula_amd.data_objects.intmat import Intmat
    from specula_amd.data_objects.recmat import Recmat
    wl = workloads.scao_ngs_20x20()
    go = oracle_geometry(wl)
    img = O.sh_forward(wl.pupil(), wl.frames(1)[0], go)
    idxs, dmap = detect_subaps(img, 20, 0.5)
    sd = SubapData(idxs=idxs, display_map=dmap, nx=20, ny=20, target_device_idx=0)
    pix = Pixels(*img.shape, target_device_idx=0)
    pix.pixels = torch.tensor(img, dtype=torch.float32, device='cuda:0')
    pix.generation_time = 1
    n_sl = 2 * idxs.shape[0]
    rng = np.random.default_rng(4)
    R = rng.standard_normal((n_sl, 12)).astype(np.float32)
    s_ref = O.shslopec(img.astype(np.float32).astype(np.float64), idxs)['slopes']
    sl = ShSlopec(sd, target_device_idx=0)
    sl.recmat = Recmat(R, target_device_idx=0)
    sl.inputs['in_pixels'].set(pix)
    sl.check_ready(1); sl.trigger(); sl.post_trigger()
    np.testing.assert_allclose(sl.slopes.slopes.cpu().numpy(), s_ref @ R, rtol=0, atol=1e-3 * np.abs(s_ref @ R).max())
    # the reference's formula m @ filt_intmat.T with m = s @ filt_recmat (n_modes) takes the
    # filtering interaction matrix as (n_slopes, n_modes)
    IM = rng.standard_normal((n_sl, 5)).astype(np.float32)
    RM = rng.standard_normal((n_sl, 5)).astype(np.float32) * 1e-2
    sl2 = ShSlopec(sd, filtmat=(Intmat(IM, target_device_idx=0), Recmat(RM, target_device_idx=0)), target_device_idx=0)
    sl2.inputs['in_pixels'].set(pix)
    sl2.check_ready(1); sl2.trigger(); sl2.post_trigger()
    expect = s_ref - (s_ref @ RM) @ IM.T
    np.testing.assert_allclose(sl2.slopes.slopes.cpu().numpy(), expect, rtol=0, atol=2e-3 * 2 / 7)
