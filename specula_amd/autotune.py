"""Measurement-driven choice of launch shapes.

* ``tune_plan_variant``: register-GEMV, FP32 MFMA or bf16-split MFMA binning, or the any-size LDS FFT
  path, for an NGS plan (the matrix cores are used only when they measure faster).
* ``tune_batch``: persistent grid size and work-list order of a MultiSensorBatch.

Results are cached in a JSON file keyed by device name and workload signature
(``$SPECULA_AMD_TUNE_CACHE`` or ``~/.cache/specula_amd/autotune.json``).
"""
import json
import os
import time

import torch

from specula_amd.batch import MultiSensorBatch
from specula_amd.processing_objects.sh import SHPlan


def _cache_path():
    return os.environ.get('SPECULA_AMD_TUNE_CACHE',
                          os.path.join(os.path.expanduser('~'), '.cache', 'specula_amd', 'autotune.json'))


def _load():
    try:
        with open(_cache_path()) as fh:
            return json.load(fh)
    except (OSError, ValueError):
        return {}


def _store(key, value):
    data = _load()
    data[key] = value
    path = _cache_path()
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        with open(path, 'w') as fh:
            json.dump(data, fh, indent=1)
    except OSError:
        pass


def _time(fn, reps=10, warm=2):
    for _ in range(warm):
        fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / reps


def device_name(device=0):
    return torch.cuda.get_device_properties(device).name


def tune_plan_variant(geom, pupil, n_frames=32, device=0, use_cache=True, variants=(1, 2, 3, 4)):
    """Time the kernel variants of an NGS plan -- register-GEMV binning (1), FP32 MFMA binning (2),
    bf16-split MFMA binning (3), full M-point LDS FFTs of the any-size path (4) -- and return
    (best_variant, {variant: seconds/frame}); variants the geometry does not support are skipped."""
    key = f'variant|{device_name(device)}|{geom.n_lenses}|{geom.ovs_np_sub}|{geom.fft_size}|{geom.subap_npx}|{n_frames}'
    if use_cache:
        hit = _load().get(key)
        if hit:
            return hit['best'], {int(k): v for k, v in hit['times'].items()}
    dev = torch.device(f'cuda:{device}')
    A = torch.as_tensor(pupil, dtype=torch.float32, device=dev)
    ph = torch.zeros((n_frames, geom.ef_size, geom.ef_size), dtype=torch.float32, device=dev)
    out = torch.empty((n_frames, geom.ccd_side, geom.ccd_side), dtype=torch.float32, device=dev)
    times = {}
    for v in variants:
        try:
            plan = SHPlan(geom, max_frames=n_frames, device=device, variant=v)
        except RuntimeError:
            continue
        times[v] = _time(lambda: plan.forward(A, ph, out, 1.0, n_frames=n_frames)) / n_frames
    best = min(times, key=times.get)
    _store(key, {'best': best, 'times': times})
    return best, times


def tune_batch(entries, device=0, grids=None, orders=(0, 1, 2), reps=5, use_cache=True, signature=None):
    """Pick (grid_blocks, order) for a MultiSensorBatch by timing graph replays.
    Returns (batch, report) with the batch built at the best shape."""
    cus = torch.cuda.get_device_properties(device).multi_processor_count
    grids = grids or [0, cus, 2 * cus, 4 * cus, 8 * cus]
    sig = signature or '|'.join(f'{e.plan.geom.n_lenses}x{e.plan.geom.ovs_np_sub}x{int(e.plan.lgs)}x{e.n_frames}'
                                for e in entries)
    key = f'batch|{device_name(device)}|{sig}'
    if use_cache:
        hit = _load().get(key)
        if hit:
            b = MultiSensorBatch(entries, grid_blocks=hit['grid'], order=hit['order'], device=device)
            return b, hit
    results = []
    for g in grids:
        for o in orders:
            b = MultiSensorBatch(entries, grid_blocks=g, order=o, device=device)
            t = _time(lambda: b.forward(slopes=False), reps=reps, warm=1)
            results.append({'grid': g, 'order': o, 'seconds': t})
            del b
    best = min(results, key=lambda r: r['seconds'])
    report = {'grid': best['grid'], 'order': best['order'], 'seconds': best['seconds'], 'candidates': results}
    _store(key, report)
    return MultiSensorBatch(entries, grid_blocks=best['grid'], order=best['order'], device=device), report
