"""Uneven multi-sensor persistent batch: equality with per-sensor launches, graph replays,
latency read-back, autotuner."""
import numpy as np
import pytest
import torch

from specula_amd import autotune, workloads
from specula_amd.batch import MultiSensorBatch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module')
def entries():
    return workloads.multi_sensor_entries(n_lgs=2, n_pushpull=40)


def reference_outputs(entries):
    refs = []
    for e in entries:
        out = torch.empty_like(e.ccd)
        e.plan.forward(e.pupil, e.phase, out, e.S0, n_frames=e.n_frames)
        refs.append(out)
    torch.cuda.synchronize()
    return refs


@pytest.mark.parametrize('grid,order', [(0, 0), (256, 1), (64, 2)])
def test_batch_equals_per_sensor_launches(entries, grid, order):
    refs = reference_outputs(entries)
    b = MultiSensorBatch(entries, grid_blocks=grid, order=order)
    for e in entries:
        e.ccd.fill_(-1.0)
    b.forward()
    torch.cuda.synchronize()
    for e, r in zip(entries, refs):
        assert torch.equal(e.ccd, r), e.name
    lat = b.latencies_us()
    assert np.all(lat > 0) and np.all(lat < 1e6)


def test_batch_graph_replays(entries):
    b = MultiSensorBatch(entries, grid_blocks=512, order=0)
    b.forward()
    torch.cuda.synchronize()
    eager = [(e.ccd.clone(), e.slopes.clone()) for e in entries]
    b.capture()
    for k in range(3):
        for e in entries:
            e.ccd.zero_()
            e.slopes.zero_()
        torch.cuda.synchronize()
        b.replay()
        torch.cuda.synchronize()
        _ = sum(float(e.ccd.sum()) for e in entries)          # eager work between replays
        for e, (c, s) in zip(entries, eager):
            assert torch.equal(e.ccd, c) and torch.equal(e.slopes, s), (k, e.name)


def test_autotuner(entries, tmp_path, monkeypatch):
    monkeypatch.setenv('SPECULA_AMD_TUNE_CACHE', str(tmp_path / 'tune.json'))
    b, rep = autotune.tune_batch(entries, grids=[0, 256], orders=(0, 2), reps=2)
    assert rep['grid'] in (0, 256) and rep['order'] in (0, 2) and len(rep['candidates']) == 4
    b2, rep2 = autotune.tune_batch(entries, grids=[0, 256], orders=(0, 2), reps=2)
    assert rep2['grid'] == rep['grid']                             # served from the cache
    wl = workloads.elt_ngs_80x80()
    best, times = autotune.tune_plan_variant(wl.geometry(), wl.pupil(), n_frames=8)
    assert best in times and set(times) == {1, 2, 3}
