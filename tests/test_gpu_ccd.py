"""CCD readout on the GPU (shccd_readout / the drop-in CCD object) against the oracle's restatement
of ccd.py:198-273: exact without noise, distributional with noise (the generators differ: the
kernel's streams are counter-based, the oracle's are numpy's default_rng as in the reference).
Every sample set is fixed by its seed and counter, so the statistical checks are deterministic."""
import numpy as np
import pytest
import torch
from scipy import stats

from oracle import sh_oracle as O
from specula_amd import _native
from specula_amd.processing_objects.ccd import CCD, readout_frames

pytestmark = pytest.mark.gpu
DEV = 'cuda:0'


def params(rows, cols, binning=1, **kw):
    from specula_amd.data_objects.simul_params import SimulParams
    ccd = CCD(SimulParams(time_step=0.001), size=[rows, cols], dt=0.001, bandw=1.0, binning=binning,
              target_device_idx=0, **kw)
    return ccd.params()


def run(p, frames, counter=0, factor=1.0, gains=None):
    x = torch.as_tensor(np.ascontiguousarray(frames, dtype=np.float32), device=DEV)
    if x.dim() == 2:
        x = x[None]
    out = torch.empty((x.shape[0], p.out_rows, p.out_cols), dtype=torch.float32, device=DEV)
    readout_frames(p, x.contiguous(), out, factor=factor, counter=counter, pixel_gains=gains)
    torch.cuda.synchronize()
    return out.cpu().numpy()


def chi2_two_sample(a, b, min_count=20):
    """Chi-square homogeneity test of two integer-valued samples (bins with enough counts)."""
    lo, hi = int(min(a.min(), b.min())), int(max(a.max(), b.max()))
    ha = np.bincount((a - lo).astype(np.int64), minlength=hi - lo + 1)
    hb = np.bincount((b - lo).astype(np.int64), minlength=hi - lo + 1)
    keep = (ha + hb) >= min_count
    tab = np.vstack([np.append(ha[keep], ha[~keep].sum()), np.append(hb[keep], hb[~keep].sum())])
    tab = tab[:, tab.sum(axis=0) > 0]
    return stats.chi2_contingency(tab)[1]


# Known-answer vectors of Philox4x32-10 published with the Random123 library (kat_vectors):
# counter (4 words), key (2 words) -> output (4 words).
PHILOX_KAT = [
    ((0x00000000, 0x00000000, 0x00000000, 0x00000000, 0x00000000, 0x00000000),
     (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
    ((0xffffffff, 0xffffffff, 0xffffffff, 0xffffffff, 0xffffffff, 0xffffffff),
     (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
    ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344, 0xa4093822, 0x299f31d0),
     (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)),
]


def test_noise_generator_known_answers():
    """Every noise draw of the readout kernels comes from this generator (shccd_philox runs the
    same device function): the Random123 known answers pin it."""
    inp = np.array([c for c, _ in PHILOX_KAT], dtype=np.uint32)
    exp = np.array([o for _, o in PHILOX_KAT], dtype=np.uint32)
    src = torch.as_tensor(inp.view(np.int32), device=DEV)
    out = torch.zeros((len(PHILOX_KAT), 4), dtype=torch.int32, device=DEV)
    _native.check(_native.lib().shccd_philox(len(PHILOX_KAT), _native.ptr(src), _native.ptr(out), None),
                  'shccd_philox')
    torch.cuda.synchronize()
    got = out.cpu().numpy().view(np.uint32)
    np.testing.assert_array_equal(got, exp)


@pytest.mark.parametrize('binning', [1, 2])
def test_noiseless_readout_matches_oracle(binning):
    rng = np.random.default_rng(0)
    img = (rng.random((64, 96)) * 1e3).astype(np.float32)
    p = params(64, 96, binning=binning, quantum_eff=0.3)
    out = run(p, img, factor=0.25)[0]
    ref = O.ccd_readout((img * np.float32(0.25)).astype(np.float32), binning=binning, quantum_eff=0.3)
    np.testing.assert_allclose(out, ref, rtol=2e-6)


@pytest.mark.parametrize('lam', [0.3, 2.5, 9.9, 10.0, 37.0, 850.0, 2.5e5])
def test_poisson_counts_follow_poisson_law(lam):
    print(f'poisson lambda={lam}', flush=True)
    p = params(512, 512, photon_noise=True)
    out = run(p, np.full((512, 512), lam, np.float32))[0].ravel()
    n = out.size
    assert np.all(out == np.round(out)) and out.min() >= 0
    assert abs(out.mean() - lam) < 5 * np.sqrt(lam / n)
    assert abs(out.var() / lam - 1) < 5 * np.sqrt(2.0 / n) + 1e-3
    # goodness of fit against the Poisson pmf over the central bins
    k = out.astype(np.int64)
    lo, hi = int(stats.poisson.ppf(1e-4, lam)), int(stats.poisson.ppf(1 - 1e-4, lam))
    edges = np.unique(np.linspace(lo, hi + 1, min(hi - lo + 2, 60)).astype(np.int64))
    obs = np.histogram(k, bins=edges)[0]
    cdf = stats.poisson.cdf(edges - 1, lam)
    exp = np.diff(cdf) * n
    sel = exp > 20
    chi2 = ((obs[sel] - exp[sel]) ** 2 / exp[sel]).sum()
    pval = stats.chi2.sf(chi2, sel.sum() - 1)
    # same law as numpy's sampler (the reference's generator)
    ref = np.random.default_rng(1).poisson(lam, size=n)
    p2 = chi2_two_sample(k, ref)
    print(f'  mean {out.mean():.6g} var {out.var():.6g}; GOF p={pval:.3g}; vs numpy p={p2:.3g}', flush=True)
    assert pval > 1e-4 and p2 > 1e-4


def test_readout_noise_float_and_truncated():
    p = params(512, 512, readout_noise=True, readout_level=2.0)
    out = run(p, np.zeros((512, 512), np.float32))[0].ravel()
    ks = stats.kstest(out / 2.0, 'norm')
    print(f'readout: mean {out.mean():.3g} std {out.std():.5g} KS p={ks.pvalue:.3g}', flush=True)
    assert ks.pvalue > 1e-4
    # with photon noise the frame is integer and the readout noise is truncated toward zero
    p = params(512, 512, photon_noise=True, readout_noise=True, readout_level=1.5)
    out = run(p, np.zeros((512, 512), np.float32))[0].ravel()
    assert np.all(out == np.trunc(out))
    ref = O.ccd_readout(np.zeros((512, 512), np.float32), photon_noise=True, readout_noise=True,
                        readout_level=1.5, photon_rng=np.random.default_rng(1),
                        readout_rng=np.random.default_rng(2)).ravel()
    pv = chi2_two_sample(out, ref)
    print(f'truncated readout: P(0) {np.mean(out == 0):.4f} vs oracle {np.mean(ref == 0):.4f}, p={pv:.3g}', flush=True)
    assert pv > 1e-4


def test_excess_cic_dark_and_adu_paths_match_oracle_distribution():
    lam = np.full((400, 400), 6.0, np.float32)
    kw = dict(photon_noise=True, excess_noise=True, readout_noise=True, readout_level=0.5)
    out = run(params(400, 400, **kw), lam)[0].ravel()
    ref = O.ccd_readout(lam, **kw, photon_rng=np.random.default_rng(1), readout_rng=np.random.default_rng(2),
                        excess_rng=np.random.default_rng(3)).ravel()
    q = 20 / 400            # EMCCD output quantum: 1/(ADU_gain * emccd_gain)
    pv = chi2_two_sample(np.round(out / q), np.round(ref / q))
    print(f'excess: mean {out.mean():.4g} vs {ref.mean():.4g}, p={pv:.3g}', flush=True)
    assert pv > 1e-4
    kw = dict(photon_noise=True, cic_noise=True, cic_level=0.2, darkcurrent_noise=True, darkcurrent_level=3.0,
              background_noise=True, background_level=1.0)
    out = run(params(400, 400, **kw), np.full((400, 400), 4.0, np.float32))[0].ravel()
    ref = O.ccd_readout(np.full((400, 400), 4.0, np.float32), **kw, photon_rng=np.random.default_rng(1),
                        cic_rng=np.random.default_rng(5)).ravel()
    pv = chi2_two_sample(np.round(out * 8), np.round(ref * 8))
    print(f'cic+dark+bg: mean {out.mean():.4g} vs {ref.mean():.4g} (expect {4 + 0.2 + 3 + 1 - 3:.4g}), p={pv:.3g}',
          flush=True)
    assert pv > 1e-4 and abs(out.mean() - 5.2) < 0.05


def test_counter_reproducibility_and_batch_independence():
    p = params(128, 128, photon_noise=True, readout_noise=True, readout_level=1.0)
    lam = np.full((3, 128, 128), 50.0, np.float32)
    a = run(p, lam, counter=7)
    b = run(p, lam, counter=7)
    c = run(p, lam, counter=8)
    np.testing.assert_array_equal(a, b)
    assert np.mean(a == c) < 0.2
    assert np.mean(a[0] == a[1]) < 0.2                    # frames of a batch draw different numbers



def test_lean_and_full_photon_kernels_agree():
    """The common-case photon kernel and the full one (forced here by unit pixel gains, which
    also carries binning, CIC and EM gain) draw from the same generator keyed by (pixel, counter,
    try): the same values, bit for bit, on a ragged frame mixing dark pixels, CDF inversion
    (signal below 10) and PTRS."""
    rng = np.random.default_rng(5)
    rows, cols = 97, 211                                      # 20467 px: ragged last chunk
    lam = rng.gamma(0.6, 60.0, size=(2, rows, cols)).astype(np.float32)
    lam[:, :, :40] = 0.0                                      # dark columns
    lam[:, 10:30] *= 0.05                                     # many pixels below 10
    p = params(rows, cols, photon_noise=True, readout_noise=True, readout_level=1.5, quantum_eff=0.8)
    lean = run(p, lam, counter=3)
    full = run(p, lam, counter=3, gains=torch.ones((rows, cols), dtype=torch.float32, device=DEV))
    frac = [np.mean(lam == 0), np.mean((lam > 0) & (lam * 0.8 < 10)), np.mean(lam * 0.8 >= 10)]
    assert min(frac) > 0.05, frac
    np.testing.assert_array_equal(lean, full)

def _readout_raw(p, integrated, in_i, factor, counter, gains=None):
    """shccd_readout with both sources: the integration buffer (read, then zeroed) and the fused
    last step in_i * factor."""
    import ctypes
    F = in_i.shape[0]
    n_in, n_out = p.in_rows * p.in_cols, p.out_rows * p.out_cols
    out = torch.empty((F, p.out_rows, p.out_cols), dtype=torch.float32, device=DEV)
    _native.check(_native.lib().shccd_readout(
        ctypes.byref(p), F, _native.ptr(integrated), n_in, _native.ptr(in_i), n_in, float(factor),
        _native.ptr(gains) if gains is not None else None, ctypes.c_uint64(counter), _native.ptr(out), n_out,
        None), 'shccd_readout')
    torch.cuda.synchronize()
    return out.cpu().numpy()


@pytest.mark.parametrize('readout', [False, True])
def test_lean_and_full_photon_kernels_agree_with_integration_buffer(readout):
    """Both signal sources at once (integration buffer + fused last step), a ragged frame of
    several tiles, with and without readout noise (without it, pixels without signal draw
    nothing): the lean kernel equals the full one bit for bit, and the buffer is zeroed."""
    rng = np.random.default_rng(8)
    F, rows, cols = 3, 61, 173                                  # 10553 px per frame: 11 tiles, ragged
    integ = rng.gamma(0.5, 30.0, size=(F, rows, cols)).astype(np.float32)
    integ[:, :, :25] = 0.0
    cur = rng.gamma(0.5, 30.0, size=(F, rows, cols)).astype(np.float32)
    cur[:, :, :25] = 0.0
    kw = dict(photon_noise=True, quantum_eff=0.7)
    if readout:
        kw.update(readout_noise=True, readout_level=2.0)
    p = params(rows, cols, **kw)
    gains = torch.ones((rows, cols), dtype=torch.float32, device=DEV)
    cur_t = torch.tensor(cur, device=DEV)
    outs = []
    for g in (None, gains):
        buf = torch.tensor(integ, device=DEV)
        outs.append(_readout_raw(p, buf, cur_t, 0.5, counter=11, gains=g))
        assert float(buf.abs().max()) == 0.0
    np.testing.assert_array_equal(outs[0], outs[1])
    lam = (integ + cur * np.float32(0.5)) * np.float32(0.7)
    assert np.mean(lam == 0) > 0.1 and np.mean((lam > 0) & (lam < 10)) > 0.05 and np.mean(lam >= 10) > 0.05
    if not readout:
        assert np.all(outs[0][lam == 0] == 0.0)


def test_ccd_object_integrates_over_dt_and_resets():
    from specula_amd.data_objects.intensity import Intensity
    from specula_amd.data_objects.simul_params import SimulParams
    ccd = CCD(SimulParams(time_step=0.001), size=[32, 32], dt=0.003, bandw=300, quantum_eff=0.5,
              target_device_idx=0)
    inp = Intensity(32, 32, target_device_idx=0)
    ccd.inputs['in_i'].set(inp)
    ccd.setup()
    rng = np.random.default_rng(3)
    frames = [rng.random((32, 32)).astype(np.float32) for _ in range(6)]
    step = ccd.seconds_to_t(0.001)
    outs = []
    for k, fr in enumerate(frames):
        t = (k + 1) * step
        inp.i[:] = torch.tensor(fr, device=DEV)
        inp.generation_time = t
        ccd.check_ready(t); ccd.trigger(); ccd.post_trigger()
        outs.append((ccd.outputs['out_pixels'].generation_time == t, ccd.outputs['out_pixels'].pixels.cpu().numpy().copy()))
    # ccd.py:207: readout when (t + loop_dt - dt) % dt == 0 -> t = 2, 5 ms (steps 2 and 5, 1-based... )
    read_steps = [k for k, (fresh, _) in enumerate(outs) if fresh]
    expect = [k for k in range(6) if ((k + 1) * step + step - 3 * step) % (3 * step) == 0]
    assert read_steps == expect
    f32 = np.float32(0.001 * 300)
    last = 0
    for k in read_steps:
        acc = np.zeros((32, 32), np.float32)
        for j in range(last, k + 1):
            acc += frames[j] * f32
        last = k + 1
        np.testing.assert_allclose(outs[k][1], acc * np.float32(0.5), rtol=2e-6)
    assert float(ccd.outputs['integrated_i'].i.abs().max()) == 0.0 or read_steps[-1] != 5


def test_noisy_loop_slopes_match_oracle_statistics():
    """SH -> CCD (photon + readout noise) -> ShSlopec with the drop-in objects over 40 frames of the
    same field: per-subaperture slope mean and spread agree with the oracle chain run with numpy's
    generators (ccd.py dtype flow), to their sampling errors."""
    from oracle import sh_oracle as O2
    from specula_amd import workloads
    from specula_amd.data_objects.electric_field import ElectricField
    from specula_amd.data_objects.simul_params import SimulParams
    from specula_amd.data_objects.subap_data import SubapData
    from specula_amd.lib.subaps import detect_subaps
    from specula_amd.processing_objects.sh import SH
    from specula_amd.processing_objects.sh_slopec import ShSlopec
    wl = workloads.scao_ngs_20x20()
    A = wl.pupil()
    ph = wl.frames(1)[0].astype(np.float32)
    go = O2.sh_geometry(ef_size=wl.ef_size, pixel_pitch=wl.pixel_pitch, **wl.sh_kwargs)
    clean = O2.sh_forward(A, ph.astype(np.float64), go, S0=50.0)
    idxs, dmap = detect_subaps(O2.sh_forward(A, np.zeros_like(A), go), 20, 0.5)
    sh = SH(**wl.sh_kwargs, target_device_idx=0)
    if wl.fov_resolution_arcsec:
        sh._fov_resolution_arcsec = wl.fov_resolution_arcsec
    ef = ElectricField(wl.ef_size, wl.ef_size, wl.pixel_pitch, S0=50.0, target_device_idx=0)
    ef.A[:] = torch.tensor(A.astype(np.float32), device=DEV)
    ef.phaseInNm[:] = torch.tensor(ph, device=DEV)
    sh.inputs['in_ef'].set(ef)
    det = CCD(SimulParams(time_step=0.001), size=[160, 160], dt=0.001, bandw=300, photon_noise=True,
              readout_noise=True, readout_level=2.0, quantum_eff=0.3, target_device_idx=0)
    det.inputs['in_i'].set(sh.outputs['out_i'])
    slopec = ShSlopec(SubapData(idxs=idxs, display_map=dmap, nx=20, ny=20, target_device_idx=0), target_device_idx=0)
    slopec.inputs['in_pixels'].set(det.outputs['out_pixels'])
    for o in (sh, det, slopec):
        o.setup()
    step = det.seconds_to_t(0.001)
    gpu = []
    for k in range(1, 41):
        t = k * step
        ef.generation_time = t
        for o in (sh, det, slopec):
            o.check_ready(t); o.trigger(); o.post_trigger()
        gpu.append(slopec.outputs['out_slopes'].slopes.cpu().numpy().copy())
    gpu = np.array(gpu)
    rp, rr = np.random.default_rng(1), np.random.default_rng(2)
    ref = []
    for k in range(40):
        pix = O2.ccd_readout((clean * np.float32(0.001 * 300)).astype(np.float32), quantum_eff=0.3, photon_noise=True,
                             readout_noise=True, readout_level=2.0, photon_rng=rp, readout_rng=rr)
        ref.append(O2.shslopec(np.asarray(pix, np.float64), idxs)['slopes'])
    ref = np.array(ref)
    n = len(gpu)
    diff = np.abs(gpu.mean(0) - ref.mean(0))
    se = np.sqrt(gpu.var(0) / n + ref.var(0) / n) + 1e-12
    z = diff / se
    ratio = np.median(gpu.std(0) / (ref.std(0) + 1e-12))
    print(f'noisy loop: max z of slope means {z.max():.2f}, median noise ratio {ratio:.3f}, '
          f'noise rms {ref.std(0).mean():.2e}', flush=True)
    assert np.mean(z < 3.5) > 0.99 and z.max() < 5.5
    assert 0.85 < ratio < 1.15


def _cte_pair(n, seed):
    """Charge-transfer matrices of an n-pixel square frame: mostly diagonal, a few per cent of the
    charge moved to the neighbouring rows / columns, plus a small dense positive part."""
    rng = np.random.default_rng(seed)
    mats = []
    for k in range(2):
        m = np.eye(n) * 0.97 + np.eye(n, k=1) * 0.02 + np.eye(n, k=-1) * 0.005 + rng.random((n, n)) * 1e-3
        mats.append(m.astype(np.float32))
    return np.stack(mats, axis=2)


def test_cte_noiseless_matches_oracle():
    """cte_noise (ccd.py:223-224) between the background/dark offset and the noise chain: binning 2,
    QE, background + dark, cte_mat[:, :, 0] . frame . cte_mat[:, :, 1], against the oracle in FP64;
    the FP32 products are bounded by a few ulp x sqrt(64) of the positive terms."""
    rng = np.random.default_rng(5)
    img = (rng.random((128, 128)) * 500).astype(np.float32)
    cte = _cte_pair(64, 7)
    p = params(128, 128, binning=2, quantum_eff=0.5, background_noise=True, darkcurrent_noise=True,
               background_level=3.0, darkcurrent_level=2.0)
    rows = torch.tensor(np.ascontiguousarray(cte[:, :, 0]), device=DEV)
    cols = torch.tensor(np.ascontiguousarray(cte[:, :, 1]), device=DEV)
    x = torch.tensor(img, device=DEV)[None].contiguous()
    out = torch.empty((1, 64, 64), dtype=torch.float32, device=DEV)
    readout_frames(p, x, out, cte=(rows, cols))
    torch.cuda.synchronize()
    ref = O.ccd_readout(img, binning=2, quantum_eff=0.5, background_noise=True, darkcurrent_noise=True,
                        background_level=3.0, darkcurrent_level=2.0, cte_noise=True, cte_mat=cte.astype(np.float64),
                        dtype=np.float64)
    got = out[0].cpu().numpy().astype(np.float64)
    err = np.abs(got - ref).max() / np.abs(ref).max()
    print(f'cte noiseless: max rel err {err:.2e}', flush=True)
    assert err < 2e-6 * np.sqrt(64)
    # planted mistake on the oracle side: the transposed matrices give a different frame
    wrong = O.ccd_readout(img, binning=2, quantum_eff=0.5, background_noise=True, darkcurrent_noise=True,
                          background_level=3.0, darkcurrent_level=2.0, cte_noise=True,
                          cte_mat=cte.transpose(1, 0, 2).astype(np.float64), dtype=np.float64)
    assert np.abs(got - wrong).max() / np.abs(ref).max() > 1e-3


def test_cte_with_photon_noise_and_dark_subtraction():
    """With photon noise the CTE output is Poisson around cte_rows . (x QE + dark) . cte_cols, ADU
    converted, with the dark current subtracted once: mean and variance over 256 frames per pixel."""
    rng = np.random.default_rng(9)
    n, F = 32, 256
    img = (rng.random((n, n)) * 200 + 20).astype(np.float32)
    cte = _cte_pair(n, 11)
    p = params(n, n, quantum_eff=0.8, photon_noise=True, darkcurrent_noise=True, darkcurrent_level=4.0)
    rows = torch.tensor(np.ascontiguousarray(cte[:, :, 0]), device=DEV)
    cols = torch.tensor(np.ascontiguousarray(cte[:, :, 1]), device=DEV)
    x = torch.tensor(np.broadcast_to(img, (F, n, n)).copy(), device=DEV)
    out = torch.empty((F, n, n), dtype=torch.float32, device=DEV)
    readout_frames(p, x, out, counter=3, cte=(rows, cols))
    torch.cuda.synchronize()
    o = out.cpu().numpy().astype(np.float64) + 4.0          # undo the subtraction of the dark level
    lam = cte[:, :, 0].astype(np.float64) @ (img.astype(np.float64) * np.float32(0.8) + 4.0) @ cte[:, :, 1]
    assert np.all(np.abs(o * 8 - np.round(o * 8)) < 1e-3)    # ADU gain 8: multiples of 1/8
    z = (o.mean(axis=0) - lam) / np.sqrt(lam / F)
    ratio = o.var(axis=0).mean() / lam.mean()
    print(f'cte + photon: mean z {z.mean():.3f}, rms z {np.sqrt((z ** 2).mean()):.3f}, var/mean {ratio:.4f}',
          flush=True)
    assert abs(z.mean()) < 5 / np.sqrt(n * n) * 1.2 and abs(np.sqrt((z ** 2).mean()) - 1) < 0.1
    assert abs(ratio - 1) < 0.05


def test_ccd_object_cte_and_auto_levels():
    """The drop-in object with cte_noise reads out as shccd_readout_cte with its own parameters
    (bit-equal), and 'auto' read-out / dark levels come from the detector model at 1/dt."""
    from specula_amd.data_objects.intensity import Intensity
    from specula_amd.data_objects.simul_params import SimulParams
    from specula_amd.lib.detector_noise import detector_noise
    n = 48
    cte = _cte_pair(n, 13)
    ccd = CCD(SimulParams(time_step=0.001), size=[n, n], dt=0.001, bandw=300, name='OCAM2k', photon_noise=True,
              readout_noise=True, darkcurrent_noise=True, readout_level='auto', darkcurrent_level='auto',
              background_level=1.0, cte_noise=True, cte_mat=cte, target_device_idx=0)
    ron, dark, _ = detector_noise(1000.0, 'OCAM2k', 1)
    assert ccd._readout_level == ron and ccd._darkcurrent_level == dark
    inp = Intensity(n, n, target_device_idx=0)
    ccd.inputs['in_i'].set(inp)
    ccd.setup()
    img = (np.random.default_rng(2).random((n, n)) * 2).astype(np.float32)
    inp.i[:] = torch.tensor(img, device=DEV)
    t = ccd.seconds_to_t(0.001)
    inp.generation_time = t
    ccd.check_ready(t); ccd.trigger(); ccd.post_trigger()
    torch.cuda.synchronize()
    got = ccd.outputs['out_pixels'].pixels.cpu().numpy()
    out = torch.empty((1, n, n), dtype=torch.float32, device=DEV)
    readout_frames(ccd.params(), torch.tensor(img, device=DEV)[None].contiguous(), out, factor=0.001 * 300,
                   counter=0, cte=(torch.tensor(np.ascontiguousarray(cte[:, :, 0]), device=DEV),
                                   torch.tensor(np.ascontiguousarray(cte[:, :, 1]), device=DEV)))
    torch.cuda.synchronize()
    assert np.array_equal(got, out[0].cpu().numpy())
    with pytest.raises(ValueError, match='cte_mat'):
        CCD(SimulParams(time_step=0.001), size=[n, n], dt=0.001, bandw=300, cte_noise=True,
            cte_mat=np.zeros((n, n + 1, 2), np.float32), target_device_idx=0)
