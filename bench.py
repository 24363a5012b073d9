"""Benchmark: SH frames/s on the ELT-class NGS sensor (80x80 subapertures, N=64), one MI355X.

One step = the hot path over one batch of 100 sequential frames of a seeded von Karman screen
moving at 10 m/s (1 ms per frame): SH forward (field, spectra, binning, normalisation) plus the
ShSlopec centre of gravity, with all inputs resident in HBM before the timed region.

Prints one JSON line (rank 0).  The `roofline` object is for the dominant kernel (the fused SH
forward): algorithmic bytes per launch / its average duration measured with HIP events recorded
on its own stream.  `cpu_baseline` times the FP64 NumPy oracle (a restatement of the reference
CPU path) on a bounded sample of the same frames, single-threaded.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HBM_PEAK_GBS = 8000.0
# HBM bytes per launch of the main kernel, measured with rocprofv3 --pmc FETCH_SIZE and then
# WRITE_SIZE (KiB per dispatch) and scaled by what the same counters read for a 256 MiB device copy
# (fetch x2.0, write x1.0); command and raw figures: profiles/pmc_traffic_elt_ngs_80x80.md.
# Keyed by (CoG fused, frames per launch, default frames per block).  The static pupil is served
# from L2 after its first read and phase rows under a dark pupil row are not read, so the traffic
# is below the algorithmic bytes (pupil and phase counted in full per frame).
MEASURED_TRAFFIC = {(True, 100, True): 835_400_000}
# Executed floating-point work per frame of the main kernel, from rocprofv3 counters of the same
# launches (profiles/pmc_flops_elt_ngs_80x80.md): SQ_INSTS_VALU_MFMA_MOPS_F32 x 512 (FP32 MFMA) +
# (SQ_INSTS_VALU_FLOPS_FP32 + SQ_INSTS_VALU_FLOPS_FP64) x 64 lanes, divided by the 100 frames.
MEASURED_FLOP_PER_FRAME = {(True, True): 111_692_800}
# LGS kernel (lgs_40x40, 50-frame launches): measured HBM bytes per frame (FETCH_SIZE x2.0 +
# WRITE_SIZE x1.0, profiles/pmc_traffic_lgs_40x40.md); the kernel spectra come from the caches
# after the first frame.
MEASURED_LGS_TRAFFIC_PER_FRAME = 3_308_000
FP32_PEAK_TFLOPS = 157.3
# Instruction issue per frame, from rocprofv3 counters of the same launches (SQ_INSTS_VALU and
# SQ_INSTS_MFMA per dispatch / frames per launch): profiles/pmc_issue_elt_ngs_80x80.md,
# pmc_issue_lgs.md and pmc_issue_ccd.md.  A wave64 VALU instruction holds its SIMD 4 cycles, an
# FP32 16x16x4 MFMA 32 cycles (FP32 MFMA does not co-issue with VALU on this part: the counters
# show zero co-issue cycles), so (4 VALU + 32 MFMA) SIMD cycles per frame against 1024 SIMDs at
# the 2.4 GHz peak engine clock is the issue roofline that bounds these kernels.
MEASURED_ISSUE_PER_FRAME = {
    'sh_ngs_mfma_kernel<8, true, 16, 3, false>': (464_520, 38_400),     # (VALU wave instr, MFMA)
    'sh_lgs_forward_kernel<16, 32, 8, 0>': (974_000, 40_000),
    'ccd_photon_lean_kernel': (6_960_000, 0),
}
N_SIMDS = 1024
PEAK_CLOCK_HZ = 2.4e9


def issue_roofline(kernel, seconds_per_frame):
    """Fraction of the SIMDs' peak issue rate that the kernel's counter-measured instructions
    per frame take at the measured time per frame."""
    per = MEASURED_ISSUE_PER_FRAME.get(kernel)
    if per is None:
        return None
    cycles = 4 * per[0] + 32 * per[1]
    frac = cycles / (N_SIMDS * PEAK_CLOCK_HZ * seconds_per_frame)
    return {'bound': 'issue', 'valu_wave_instr_per_frame': per[0], 'mfma_per_frame': per[1],
            'simd_cycles_per_frame': cycles, 'peak_clock_ghz': PEAK_CLOCK_HZ / 1e9, 'frac': round(frac, 4)}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=50)
    ap.add_argument('--warmup', type=int, default=10)
    ap.add_argument('--frames', type=int, default=100)
    ap.add_argument('--cpu-seconds', type=float, default=15.0,
                    help='budget for the CPU baseline sample (0 disables it)')
    ap.add_argument('--unfused', action='store_true',
                    help='run ShSlopec as its own pass instead of in the SH kernel epilogue')
    ap.add_argument('--fpb', type=int, default=0, help='frames per block of the SH kernel (0: plan default)')
    ap.add_argument('--copy-probe', action='store_true',
                    help='also run two 256 MiB device copies (reference traffic for rocprofv3 counters)')
    ap.add_argument('--extras', default='all',
                    help="extra measurements: 'all' (or 1), 'none' (or 0), or a comma list of "
                         "bf16, ccd, lgs, im, multi, hbm, variants, objects")
    args = ap.parse_args()
    names = ('bf16', 'ccd', 'lgs', 'im', 'multi', 'hbm', 'variants', 'objects', 'setups')
    e = str(args.extras).strip().lower()
    args.extras = set(names) if e in ('all', '1') else set() if e in ('none', '0', '') else \
        {x.strip() for x in e.split(',') if x.strip()}
    unknown = args.extras - set(names)
    if unknown:
        ap.error(f'unknown extras {sorted(unknown)}')
    return args


def executed_flops(fused, default_fpb, frames, kern_ms):
    """Executed FP32 + FP64 arithmetic of the main kernel (counter-measured per frame) per second,
    against MI355X's FP32 peak."""
    per = MEASURED_FLOP_PER_FRAME.get((bool(fused), bool(default_fpb)))
    if per is None:
        return None
    tflops = per * frames / (kern_ms * 1e-3) / 1e12
    return {'flop_per_frame': per, 'tflops': round(tflops, 2), 'fp32_peak_tflops': FP32_PEAK_TFLOPS,
            'fp32_peak_frac': round(tflops / FP32_PEAK_TFLOPS, 4)}


def progress(msg):
    """One line per phase on stderr (stdout carries only the JSON result)."""
    print(f'bench: {msg}', file=sys.stderr, flush=True)


def cpu_baseline(wl, A, frames, idxs, budget_s):
    """Oracle SH + ShSlopec per frame on one host core (bounded sample)."""
    os.environ.setdefault('OMP_NUM_THREADS', '1')
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from oracle import sh_oracle as O
    go = O.sh_geometry(ef_size=wl.ef_size, pixel_pitch=wl.pixel_pitch,
                       fov_resolution_arcsec=wl.fov_resolution_arcsec or None, **wl.sh_kwargs)
    A64 = A.astype(np.float64)
    times = []
    t_start = time.perf_counter()
    k = 0
    while True:
        t0 = time.perf_counter()
        img = O.sh_forward(A64, frames[k % len(frames)].astype(np.float64), go)
        O.shslopec(img, idxs)
        times.append(time.perf_counter() - t0)
        k += 1
        if time.perf_counter() - t_start > budget_s or k >= 50:
            break
    warm = times[1:] if len(times) > 1 else times
    per_frame = float(np.median(warm))
    return {'value': 1.0 / per_frame, 'unit': 'frames/s', 'cores': 1, 'kind': 'port',
            'sample': f'{len(times)} frames of elt_ngs_80x80 (FP64 NumPy oracle SH + ShSlopec, '
                      f'median of {len(warm)} after 1 warm-up)',
            'cpu': _cpu_model(), 'cpu_count': os.cpu_count(),
            'affinity': len(os.sched_getaffinity(0))}


def _hip():
    import ctypes
    import torch  # noqa: F401  -- the HIP runtime torch loaded is the one the library uses
    return ctypes.CDLL('libamdhip64.so.7')


def hip_event():
    """A raw hipEvent_t (timing enabled) on the current device."""
    import ctypes
    ev = ctypes.c_void_p()
    if _hip().hipEventCreate(ctypes.byref(ev)) != 0:
        raise RuntimeError('hipEventCreate failed')
    return ev


def hip_elapsed_ms(a, b):
    import ctypes
    ms = ctypes.c_float()
    h = _hip()
    if h.hipEventSynchronize(b) != 0 or h.hipEventElapsedTime(ctypes.byref(ms), a, b) != 0:
        raise RuntimeError('hipEventElapsedTime failed')
    return ms.value


def timed(fn, reps, sync):
    fn()
    sync()
    t0 = time.perf_counter()
    for _ in range(reps):
        fn()
    sync()
    return (time.perf_counter() - t0) / reps


def bench_lgs(torch, dev):
    """lgs_40x40: 40x40 subapertures of 16 px, N=128, elongated Na kernels; 50-frame batches
    (measured: 2.97 us per frame at 50 frames per launch, 3.15 at 20, 3.02 at 100)."""
    from specula_amd import _native, workloads
    from specula_amd.processing_objects.sh import SHPlan
    wl = workloads.lgs_40x40()
    g = wl.geometry()
    kern = workloads.lgs_kernel(g)
    F = 50
    plan = SHPlan(g, lgs=True, max_frames=F, device=0, kernels_normalised=kern.kernels_normalised())
    plan.set_kernels(torch.tensor(kern.lag_spectra(g.ovs_np_sub), device=dev))
    A = torch.tensor(wl.pupil().astype(np.float32), device=dev)
    ph = torch.tensor(wl.frames(F).astype(np.float32), device=dev)
    out = torch.empty((F, g.ccd_side, g.ccd_side), dtype=torch.float32, device=dev)
    plan.set_static_pupil(A, 1.0)      # as the SH object does for an unchanged telescope pupil
    dt = timed(lambda: plan.forward(A, ph, out, 1.0, n_frames=F), 10, torch.cuda.synchronize)
    # the kernel's own duration: events attached to its dispatch (plan timing hook)
    ev = [(hip_event(), hip_event()) for _ in range(10)]
    for k in range(10):
        _native.check(_native.lib().shwfs_plan_set_timing_events(plan._h, ev[k][0], ev[k][1]), 'set_timing_events')
        plan.forward(A, ph, out, 1.0, n_frames=F)
    _native.check(_native.lib().shwfs_plan_set_timing_events(plan._h, None, None), 'set_timing_events')
    torch.cuda.synchronize()
    kt = float(np.mean([hip_elapsed_ms(a, b) for a, b in ev])) * 1e-3
    # SH + ShSlopec of the same frames: the CoG in the kernel's epilogue against its own pass
    from specula_amd.lib.subaps import detect_subaps
    from specula_amd.processing_objects.sh_slopec import SlopecPlan, compute_xy_weights
    plan.forward(A, torch.zeros((1, g.ef_size, g.ef_size), dtype=torch.float32, device=dev), out[:1], 1.0)
    idxs, _ = detect_subaps(out[0].cpu().numpy(), g.n_lenses, 0.5)
    sc = SlopecPlan(idxs, (g.ccd_side, g.ccd_side), compute_xy_weights(g.subap_npx), max_frames=F, device=0)
    n = idxs.shape[0]
    sl = torch.empty((F, 2 * n), dtype=torch.float32, device=dev)
    fl = torch.empty((F, n), dtype=torch.float64, device=dev)
    tt = torch.empty((F, 2), dtype=torch.float64, device=dev)

    def two_pass():
        plan.forward(A, ph, out, 1.0, n_frames=F)
        sc.forward(out, sl, fl, tt, n_frames=F)
    dt_two = timed(two_pass, 10, torch.cuda.synchronize)
    if not plan.bind_slopec(sc):
        raise RuntimeError('lgs_40x40: the ShSlopec could not be fused')
    dt_fused = timed(lambda: plan.forward_slopes(A, ph, out, 1.0, sl, fl, tt, n_frames=F), 10, torch.cuda.synchronize)
    plan.bind_slopec(None)
    kbytes = g.n_lenses ** 2 * (2 * g.ovs_np_sub - 1) * g.ovs_np_sub * 8     # Hermitian half of the lag grid read
    frame_bytes = 4 * (2 * g.ef_size ** 2 + g.ccd_side ** 2) + kbytes          # pupil + phase + CCD + spectra
    gbps = frame_bytes * F / kt / 1e9
    return {'workload': 'lgs_40x40', 'frames_per_s': round(F / dt, 1), 'ms_per_frame': round(dt * 1e3 / F, 5),
            'kernel': 'sh_lgs_forward_kernel<16, 32, 8, 0>', 'kernel_ms_per_launch': round(kt * 1e3, 5),
            # algorithmic bytes count the kernel spectra on the Hermitian half of the lag grid
            # (|m| <= np - 1, m_x >= 0: all the correlation uses), not the full N x N spectra;
            # the HBM fraction is per kernel duration, frames_per_s per call (host launch included)
            'kernel_spectra_MB_per_frame': round(kbytes / 1e6, 2), 'algorithmic_MB_per_frame': round(frame_bytes / 1e6, 2),
            'algorithmic_GBps': round(gbps, 1), 'hbm_frac': round(gbps / HBM_PEAK_GBS, 4),
            'measured_traffic_MB_per_frame': round(MEASURED_LGS_TRAFFIC_PER_FRAME / 1e6, 3),
            'issue_roofline': issue_roofline('sh_lgs_forward_kernel<16, 32, 8, 0>', kt / F),
            'sh_slopes_frames_per_s': {'fused_cog': round(F / dt_fused, 1), 'two_pass': round(F / dt_two, 1),
                                       'valid_subaps': int(n)},
            'fft_size': g.fft_size, 'np_sub': g.ovs_np_sub}


def bench_im(torch, dev, n_modes=1000):
    """Interaction matrix of n_modes smooth modes on elt_ngs_80x80: two VJPs + the Jacobian
    contraction, against batched push-pull (2 n_modes sensor frames) on the same GPU."""
    from specula_amd import workloads
    from specula_amd.lib.modes import fourier_modes
    from specula_amd.sensor import SHSensor
    wl = workloads.elt_ngs_80x80()
    sen = SHSensor(wl.geometry(), wl.pupil(), max_frames=200, device=0)
    ph = torch.tensor(wl.frames(1)[0].astype(np.float32), device=dev)
    modes = fourier_modes(wl.pupil(), n_modes, device=dev)
    sync = torch.cuda.synchronize
    box = {}

    def adj():
        box['adj'] = sen.interaction_matrix(ph, modes)

    def pp():
        box['pp'] = sen.pushpull_interaction_matrix(ph, modes, amp=50.0, chunk=100)

    t_adj = timed(adj, 10, sync)
    t_pp = timed(pp, 3, sync)
    im, ref = box['adj'], box['pp']
    rel = float((im - ref).norm() / ref.norm())
    rel200 = float((im[:200] - ref[:200]).norm() / ref[:200].norm())
    pp5 = sen.pushpull_interaction_matrix(ph, modes, amp=5.0, chunk=100)
    rel5 = float((im - pp5).norm() / pp5.norm())
    # Richardson extrapolation of push-pull at 50 and 25 nm, (4 IM(a/2) - IM(a)) / 3: cancels its
    # O(a^2) term, leaving the derivative that the adjoint computes
    pp25 = sen.pushpull_interaction_matrix(ph, modes, amp=25.0, chunk=100)
    rich = (4.0 * pp25 - ref) / 3.0
    rel_rich = float((im - rich).norm() / rich.norm())
    return {'workload': 'elt_ngs_80x80', 'modes': n_modes, 'slopes': 2 * sen.n_sub,
            'adjoint_ms': round(t_adj * 1e3, 3), 'pushpull_ms': round(t_pp * 1e3, 3),
            'speedup': round(t_pp / t_adj, 2),
            # Frobenius relative differences to push-pull: the 200 lowest modes (the parity case),
            # all modes at 50 nm RMS, all modes at 5 nm (push-pull's O(a^2) term shrinks 100x), and
            # all modes against the Richardson extrapolation of the 50 and 25 nm push-pulls
            'rel_diff_pushpull_50nm_200modes': rel200, 'rel_diff_pushpull_50nm': rel,
            'rel_diff_pushpull_5nm': rel5, 'rel_diff_pushpull_richardson_50_25nm': rel_rich}


def bench_multi(torch, dev, reps=20):
    """Uneven batch: 6 LGS 40x40 (N 128), NGS 10x10 / 20x20 (N 32) and 80x80 (N 64), plus 2000
    push-pull fields of the 80x80 sensor, in one persistent SH launch + CoG launches captured in one
    HIP graph; grid/order from the autotuner."""
    from specula_amd import autotune, workloads
    entries = workloads.multi_sensor_entries(n_lgs=6, n_pushpull=2000)
    cus = torch.cuda.get_device_properties(0).multi_processor_count
    batch, rep = autotune.tune_batch(entries, grids=[0, 2 * cus, 8 * cus], orders=(0, 2), reps=3, use_cache=False)
    batch.capture()
    lat = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        batch.replay()
        lat.append(batch.latencies_us())       # synchronises
    dt = (time.perf_counter() - t0) / reps
    lat = np.array(lat)
    frames = sum(e.n_frames for e in entries)
    subaps = sum(e.n_frames * e.plan.geom.n_lenses ** 2 for e in entries)
    per = {e.name: {'p50_us': round(float(np.percentile(lat[:, k], 50)), 1),
                    'p99_us': round(float(np.percentile(lat[:, k], 99)), 1)} for k, e in enumerate(entries)}
    return {'sensors': len(entries), 'frames_per_replay': frames, 'ms_per_replay': round(dt * 1e3, 3),
            'frames_per_s': round(frames / dt, 1), 'subap_ffts_per_s': round(subaps / dt, 1),
            'autotune': {'grid': rep['grid'], 'order': rep['order'],
                         'candidates_ms': {f"{c['grid']}/{c['order']}": round(c['seconds'] * 1e3, 3)
                                           for c in rep['candidates']}},
            'forward_completion_us': per}


def bench_objects(torch, dev, steps=200):
    """One simulation-loop step of the drop-in objects at elt_ngs_80x80: ElectricField phase
    update, SH (check_ready / trigger / post_trigger) and ShSlopec on the SH pixels, one frame per
    step, host overhead included: eager, with ``use_graph`` (the object keeps a graph only when it
    holds at least ``graph_min_nodes`` launches; the SH trigger is one launch, so it runs direct)
    and with the one-launch graph kept and replayed (``graph_min_nodes`` 1)."""
    from specula_amd import workloads
    from specula_amd.data_objects.electric_field import ElectricField
    from specula_amd.data_objects.pixels import Pixels
    from specula_amd.data_objects.subap_data import SubapData
    from specula_amd.lib.subaps import detect_subaps
    from specula_amd.processing_objects.sh import SH
    from specula_amd.processing_objects.sh_slopec import ShSlopec
    wl = workloads.elt_ngs_80x80()
    g = wl.geometry()
    ef = ElectricField(wl.ef_size, wl.ef_size, wl.pixel_pitch, S0=1.0, target_device_idx=0)
    ef.A[:] = torch.tensor(wl.pupil(), device=dev)
    frames = torch.tensor(wl.frames(8).astype(np.float32), device=dev)
    out = {}
    for mode in ('eager', 'use_graph', 'graph_forced'):
        use_graph = mode != 'eager'
        sh = SH(**wl.sh_kwargs, target_device_idx=0)
        if mode == 'graph_forced':
            sh.graph_min_nodes = 1
        if wl.fov_resolution_arcsec:
            sh._fov_resolution_arcsec = wl.fov_resolution_arcsec
        sh.inputs['in_ef'].set(ef)
        sh.setup()
        sh.use_graph = use_graph
        ef.phaseInNm[:] = 0.0
        ef.generation_time = 1
        sh.check_ready(1)
        sh.trigger()
        sh.post_trigger()
        idxs, dmap = detect_subaps(sh.outputs['out_i'].i.cpu().numpy(), g.n_lenses, 0.5)
        sl = ShSlopec(SubapData(idxs=idxs, display_map=dmap, nx=g.n_lenses, ny=g.n_lenses, target_device_idx=0),
                      target_device_idx=0)
        pix = Pixels(*sh.outputs['out_i'].i.shape, target_device_idx=0)
        sl.inputs['in_pixels'].set(pix)

        def step(t):
            ef.phaseInNm.copy_(frames[t % 8])
            ef.generation_time = t
            sh.check_ready(t)
            sh.trigger()
            sh.post_trigger()
            pix.pixels = sh.outputs['out_i'].i
            pix.generation_time = t
            sl.check_ready(t)
            sl.trigger()
            sl.post_trigger()
        for t in range(2, 12):
            step(t)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for t in range(12, 12 + steps):
            step(t)
        torch.cuda.synchronize()
        out[mode + '_us'] = round((time.perf_counter() - t0) / steps * 1e6, 1)
        if use_graph:
            out[mode + '_graph_kept'] = sh.cuda_graph is not None
            out['sh_graph_nodes'] = sh.graph_nodes
    return {'workload': 'elt_ngs_80x80, 1 frame per step', **out}


def bench_setups(torch, dev, frames=8):
    """Every SH block of the reference's own set-ups (ERIS, MORFEO, SCAO; workloads.REFERENCE_SH_SETUPS)
    at full size on this GPU: frames/s of the plan's forward over ``frames`` seeded screens per call
    (host launch, resampling and the rescale pass included), and the time per frame of its
    vector-Jacobian product (one cotangent image per frame).  LGS blocks use a Gaussian kernel of
    2 FFT pixels FWHM in every subaperture (their sodium-profile kernel sets are gigabytes; the
    kernel values do not change the work)."""
    from specula_amd import workloads
    from specula_amd.lib.geometry import sh_geometry
    from specula_amd.lib.phasescreen import annular_pupil, von_karman_opd
    from specula_amd.processing_objects.sh import SHPlan
    out = {}
    for where, pp, pitch, lgs, kw in workloads.REFERENCE_SH_SETUPS:
        g = sh_geometry(pp, pitch, **kw)
        key = where.split('/')[-1]
        rec = {'lgs': lgs, 'subaps': g.n_lenses, 'np_sub': g.ovs_np_sub, 'fft_size': g.fft_size, 'cut': g.cutpixels,
               'ccd_px_per_subap': g.subap_npx, 'fov_ovs': round(float(g.fov_ovs), 4)}
        try:
            plan = SHPlan(g, lgs=lgs, max_frames=frames, device=0)
        except RuntimeError as e:
            rec['refused'] = str(e).split(':')[-1].strip()
            out[key] = rec
            continue
        A = annular_pupil(pp, 0.15)
        if g.needs_interpolation:
            plan.set_resampling(A)
        if lgs:
            plan.set_kernels(torch.tensor(workloads.gaussian_lag_spectra(g), device=dev))
        scr = von_karman_opd(1 << int(np.ceil(np.log2(pp + frames + 8))), pitch, seed=5)
        ph = torch.tensor(np.stack([scr[:pp, k:k + pp] for k in range(frames)]).astype(np.float32), device=dev)
        At = torch.tensor(A.astype(np.float32), device=dev)
        ccd = torch.empty((frames, g.ccd_side, g.ccd_side), dtype=torch.float32, device=dev)
        plan.set_static_pupil(At, 1.0)
        dt = timed(lambda: plan.forward(At, ph, ccd, 1.0, n_frames=frames), 5, torch.cuda.synchronize)
        rec.update(frames_per_s=round(frames / dt, 1), ms_per_frame=round(dt * 1e3 / frames, 4),
                   parseval=bool(plan.parseval), mask_terms=int(plan.mask_terms))
        # adjoint: one cotangent image per frame (non-Parseval plans rerun the forward inside)
        gccd = torch.randn((frames, g.ccd_side, g.ccd_side), dtype=torch.float32, device=dev)
        gph = torch.empty_like(ph)
        dv = timed(lambda: plan.vjp(At, ph, gccd, gph, 1.0, n_frames=frames), 3, torch.cuda.synchronize)
        rec['vjp_ms_per_frame'] = round(dv * 1e3 / frames, 4)
        out[key] = rec
        del plan, ph, ccd, gccd, gph
    return out


def cpu_baseline_setups(setups, seconds):
    """The FP64 oracle (one core) on one flat-plus-screen frame of the NGS set-ups of
    ``bench_setups`` whose full frame it finishes quickly, within a total budget of ``seconds``."""
    from oracle import sh_oracle as O
    from specula_amd import workloads
    from specula_amd.lib.phasescreen import annular_pupil, von_karman_opd
    t_all = time.perf_counter()
    for where, pp, pitch, lgs, kw in workloads.REFERENCE_SH_SETUPS:
        key = where.split('/')[-1]
        rec = setups.get(key)
        if lgs or rec is None or 'refused' in rec or time.perf_counter() - t_all > seconds:
            continue
        go = O.sh_geometry(ef_size=pp, pixel_pitch=pitch, **kw)
        if go['ef_ovs'] > 1024:          # resampled fields of 8k px: minutes per frame in FP64
            continue
        A = annular_pupil(pp, 0.15)
        ph = von_karman_opd(1 << int(np.ceil(np.log2(pp + 8))), pitch, seed=5)[:pp, :pp]
        t0 = time.perf_counter()
        O.sh_forward(A, ph, go)
        dt = time.perf_counter() - t0
        rec['cpu_oracle_frames_per_s'] = round(1.0 / dt, 3)
        rec['gpu_over_cpu'] = round(rec['frames_per_s'] * dt, 1)


def bench_variants(torch, dev):
    """The plan's binning variants on elt_ngs_80x80 as the autotuner times them (100 frames, flat
    phase, per-call pupil statistics): register-GEMV binning on the VALU after the FFTs, FP32
    MFMA binning, bf16-split MFMA binning, and the any-size path (four-step FFTs on the M grid)."""
    from specula_amd import autotune, workloads
    wl = workloads.elt_ngs_80x80()
    best, times = autotune.tune_plan_variant(wl.geometry(), wl.pupil(), n_frames=100, use_cache=False)
    names = {1: 'register_gemv_valu', 2: 'fp32_mfma', 3: 'bf16_split_mfma', 4: 'general_lds_fft'}
    return {'best': names[best], 'us_per_frame': {names[v]: round(t * 1e6, 3) for v, t in times.items()}}


def bench_hbm(torch, dev):
    """Practical HBM ceilings of this GPU, for reading the roofline fraction: a 1 GiB write
    (torch fill) and a 1 GiB device-to-device copy (read + write), median of 20 timed runs."""
    n = 256 * 1024 * 1024
    a = torch.empty(n, dtype=torch.float32, device=dev)
    b = torch.empty(n, dtype=torch.float32, device=dev)
    a.fill_(1.0)
    b.fill_(2.0)

    def med(fn):
        for _ in range(3):
            fn()
        torch.cuda.synchronize()
        ts = []
        for _ in range(20):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            fn()
            e1.record()
            torch.cuda.synchronize()
            ts.append(e0.elapsed_time(e1) * 1e-3)
        return float(np.median(ts))
    w = med(lambda: a.fill_(0.5))
    c = med(lambda: b.copy_(a))
    del a, b
    return {'write_GBps': round(4 * n / w / 1e9, 1), 'copy_GBps': round(8 * n / c / 1e9, 1),
            'bytes': 4 * n, 'note': 'copy counts read + write bytes'}


def bench_ccd(torch, dev, plan, sl_plan, At, ph, ccd, slopes, flux, totals, F, g):
    """Detector post-step at the ELT frame (1280^2): photon + readout noise (1 e-), QE 0.3, on the
    SH images scaled to ~1000 photons per subaperture-average pixel; and the noisy loop step
    SH -> CCD -> ShSlopec (the CoG must see the noisy pixels, so it runs as its own pass)."""
    from specula_amd.data_objects.simul_params import SimulParams
    from specula_amd.processing_objects.ccd import CCD, readout_frames
    det = CCD(SimulParams(time_step=0.001), size=[g.ccd_side, g.ccd_side], dt=0.001, bandw=1.0,
              photon_noise=True, readout_noise=True, readout_level=1.0, quantum_eff=0.3, target_device_idx=0)
    p = det.params()
    plan.forward(At, ph, ccd, 1.0, n_frames=F)
    factor = float(1000.0 / ccd[0].mean())
    pix = torch.empty_like(ccd)
    box = {'n': 0}

    def ro():
        readout_frames(p, ccd, pix, factor=factor, counter=box['n'])
        box['n'] += 1
    t_ro = timed(lambda: [ro() for _ in range(5)], 3, torch.cuda.synchronize) / 5
    nbytes = 2 * 4 * g.ccd_side ** 2 * F                       # intensity read + pixels written

    def loop():
        plan.forward(At, ph, ccd, 1.0, n_frames=F)
        ro()
        sl_plan.forward(pix, slopes, flux, totals, n_frames=F)
    t_loop = timed(lambda: [loop() for _ in range(5)], 3, torch.cuda.synchronize) / 5
    return {'frame': f'{g.ccd_side}x{g.ccd_side}', 'noise': 'photon + readout 1 e-, QE 0.3, ADU 8',
            'mean_photons_per_px': 300.0, 'readout_frames_per_s': round(F / t_ro, 1),
            'readout_GBps': round(nbytes / t_ro / 1e9, 1), 'readout_hbm_frac': round(nbytes / t_ro / 1e9 / HBM_PEAK_GBS, 4),
            # the photon sampler is bound by instruction issue, not by its 13 MB per frame
            'issue_roofline': issue_roofline('ccd_photon_lean_kernel', t_ro / F),
            'noisy_loop_frames_per_s': round(F / t_loop, 1), 'noisy_loop_ms_per_frame': round(t_loop * 1e3 / F, 5)}


def _cpu_model():
    try:
        for line in open('/proc/cpuinfo'):
            if line.startswith('model name'):
                return line.split(':', 1)[1].strip()
    except OSError:
        pass
    return 'unknown'


def main():
    args = parse()
    if args.gpus != 1:
        print(f'bench.py: this build runs on one GPU only; --gpus {args.gpus} is not supported',
              file=sys.stderr)
        sys.exit(2)
    import torch
    from specula_amd import _native, workloads
    from specula_amd.lib.subaps import detect_subaps
    from specula_amd.processing_objects.sh import SHPlan
    from specula_amd.processing_objects.sh_slopec import SlopecPlan, compute_xy_weights

    dev = torch.device('cuda:0')
    wl = workloads.elt_ngs_80x80()
    g = wl.geometry()
    F = args.frames
    A = wl.pupil().astype(np.float32)
    frames = wl.frames(F).astype(np.float32)

    plan = SHPlan(g, max_frames=F, device=0)
    if args.fpb > 0:
        plan.set_frames_per_block(args.fpb)
    At = torch.tensor(A, device=dev)
    ph = torch.tensor(frames, device=dev)
    ccd = torch.empty((F, g.ccd_side, g.ccd_side), dtype=torch.float32, device=dev)

    # valid subapertures: flux > 0.5 max on a flat wavefront (ShSubapCalibrator convention)
    flat = torch.zeros((1, g.ef_size, g.ef_size), dtype=torch.float32, device=dev)
    flat_img = torch.empty((1, g.ccd_side, g.ccd_side), dtype=torch.float32, device=dev)
    plan.forward(At, flat, flat_img, 1.0, n_frames=1)
    idxs, _ = detect_subaps(flat_img[0].cpu().numpy(), g.n_lenses, 0.5)
    n_sub = idxs.shape[0]
    sl_plan = SlopecPlan(idxs, (g.ccd_side, g.ccd_side), compute_xy_weights(g.subap_npx), max_frames=F, device=0)
    slopes = torch.empty((F, 2 * n_sub), dtype=torch.float32, device=dev)
    flux = torch.empty((F, n_sub), dtype=torch.float64, device=dev)
    totals = torch.empty((F, 2), dtype=torch.float64, device=dev)

    stream = torch.cuda.current_stream()
    ev = [(hip_event(), hip_event()) for _ in range(args.steps)]
    # CoG fused into the SH kernel epilogue (one pass; the CCD frames are still written)
    fused = (not args.unfused) and plan.bind_slopec(sl_plan)
    # the telescope pupil is static: its statistics (masked area, Parseval scale) once
    plan.set_static_pupil(At, 1.0)

    def step(k=None):
        if k is not None:
            _native.check(_native.lib().shwfs_plan_set_timing_events(plan._h, ev[k][0], ev[k][1]),
                          'set_timing_events')
        if fused:
            plan.forward_slopes(At, ph, ccd, 1.0, slopes, flux, totals, n_frames=F, stream=stream)
        else:
            plan.forward(At, ph, ccd, 1.0, n_frames=F, stream=stream)
        if k is not None:
            _native.check(_native.lib().shwfs_plan_set_timing_events(plan._h, None, None), 'set_timing_events')
        if not fused:
            sl_plan.forward(ccd, slopes, flux, totals, n_frames=F, stream=stream)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        step(k)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0

    if args.copy_probe:
        src = torch.ones(64 * 1024 * 1024, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        for _ in range(2):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    kern_ms = float(np.mean([hip_elapsed_ms(a, b) for a, b in ev]))
    ms_per_step = elapsed * 1e3 / args.steps
    fps = F * args.steps / elapsed
    px = g.ef_size * g.ef_size
    bytes_per_frame = 4 * (2 * px + g.ccd_side * g.ccd_side)      # A + phase read, CCD written
    if fused:
        bytes_per_frame += 32 * n_sub                              # FP64 CoG moments written
    achieved = bytes_per_frame * F / (kern_ms * 1e-3) / 1e9
    N = g.fft_size
    fft_flop = 5 * N * N * np.log2(N * N) * g.n_lenses ** 2       # 5 N^2 log2(N^2) per subaperture
    result = {
        'metric': 'SH frames/s (subap-FFTs/s, % HBM roofline); adjoint vs push-pull IM time',
        'value': round(fps, 1),
        'unit': 'frames/s',
        'n_gpus': 1,
        'steps': args.steps,
        'warmup': args.warmup,
        'ms_per_step': round(ms_per_step, 4),
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': 'f32 (f64 flux/CoG accumulation)',
        'data': 'synthetic: seeded von Karman screen (r0=0.15 m, L0=25 m) moving at 10 m/s',
        'config': {'workload': 'elt_ngs_80x80', 'pupil_px': g.ef_size, 'subaps': f'{g.n_lenses}x{g.n_lenses}',
                   'np_sub': g.ovs_np_sub, 'fft_size': N, 'ccd_px_per_subap': g.subap_npx,
                   'frames_per_step': F, 'valid_subaps': int(n_sub), 'parallelism': 'single GPU'},
        'subap_ffts_per_s': round(fps * g.n_lenses ** 2, 1),
        'sh_kernel_ms_per_launch': round(kern_ms, 5),
        'sh_kernel_frames_per_s': round(F / (kern_ms * 1e-3), 1),
        # the reference's FFT work (N = 64 padded transforms, 5 N^2 log2 N^2 per subaperture) per
        # second: an equivalent rate, not executed arithmetic (the kernel never runs those FFTs)
        'fft_equivalent_gflops': round(fft_flop * F / (kern_ms * 1e-3) / 1e9, 1),
        'executed': executed_flops(fused, args.fpb in (0, 4), F, kern_ms),
        # the bound the counters show for this kernel: instruction issue (HBM is the contract's roofline)
        'issue_roofline': issue_roofline(f'sh_ngs_mfma_kernel<8, true, 16, {3 if fused else 0}, false>'
                                         if args.fpb in (0, 4) else None, kern_ms * 1e-3 / F),
        'roofline': {'bound': 'hbm', 'achieved': round(achieved, 1), 'peak': HBM_PEAK_GBS, 'unit': 'GB/s',
                     'frac': round(achieved / HBM_PEAK_GBS, 4),
                     'traffic': MEASURED_TRAFFIC.get((bool(fused), F, args.fpb in (0, 4))),
                     'kernel': f'sh_ngs_mfma_kernel<8, true, 16, {3 if fused else 0}, false>',
                     'fused_cog': bool(fused), 'frames_per_block': plan.fpb,
                     'bytes_per_launch': bytes_per_frame * F},
    }
    progress(f'main step: {fps:.0f} frames/s')
    if 'bf16' in args.extras and fused:
        progress('bf16-split variant')
        # same step with the bf16-split MFMA binning (variant 3): x = hi + lo in bf16, all four cross
        # products on the matrix cores; ~1e-5 relative image error (FP32 path: ~4e-7)
        plan3 = SHPlan(g, max_frames=F, device=0, variant=3)
        plan3.bind_slopec(sl_plan)
        plan3.set_static_pupil(At, 1.0)

        def step3():
            plan3.forward_slopes(At, ph, ccd, 1.0, slopes, flux, totals, n_frames=F, stream=stream)
        dt3 = timed(lambda: [step3() for _ in range(10)], 3, torch.cuda.synchronize) / 10
        result['bf16_split_variant'] = {'frames_per_s': round(F / dt3, 1), 'ms_per_step': round(dt3 * 1e3, 4),
                                        'dtype': 'bf16 hi+lo pieces, f32 accumulation'}
        plan.bind_slopec(sl_plan)
    if 'ccd' in args.extras:
        progress('ccd readout')
        result['ccd'] = bench_ccd(torch, dev, plan, sl_plan, At, ph, ccd, slopes, flux, totals, F, g)
    if 'lgs' in args.extras:
        progress('lgs sensor')
        result['lgs'] = bench_lgs(torch, dev)
    if 'im' in args.extras:
        progress('interaction matrix')
        result['interaction_matrix'] = bench_im(torch, dev)
    if 'multi' in args.extras:
        progress('multi-sensor batch')
        result['multi_sensor'] = bench_multi(torch, dev)
    if 'objects' in args.extras:
        progress('drop-in objects')
        result['object_step'] = bench_objects(torch, dev)
    if 'variants' in args.extras:
        progress('binning variants')
        result['binning_variants'] = bench_variants(torch, dev)
    if 'hbm' in args.extras:
        progress('hbm ceilings')
        result['hbm_ceiling'] = bench_hbm(torch, dev)
    if 'setups' in args.extras:
        progress('reference set-ups')
        result['reference_setups'] = bench_setups(torch, dev)
    if args.cpu_seconds > 0:
        progress('cpu baseline')
        result['cpu_baseline'] = cpu_baseline(wl, A, frames, idxs, args.cpu_seconds)
        if 'reference_setups' in result:
            cpu_baseline_setups(result['reference_setups'], args.cpu_seconds)
    print(json.dumps(result), flush=True)


if __name__ == '__main__':
    main()
