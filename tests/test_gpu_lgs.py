"""GPU parity of the LGS (elongated-spot) SH path against the FP64 oracle."""
import numpy as np
import pytest
import torch

import specula_amd
from helpers import check_intensity, oracle_geometry
from oracle import sh_oracle as O
from specula_amd import workloads
from specula_amd.base import BaseValue
from specula_amd.data_objects.convolution_kernel import ConvolutionKernel
from specula_amd.data_objects.electric_field import ElectricField
from specula_amd.data_objects.laser_launch_telescope import LaserLaunchTelescope
from specula_amd.lib.geometry import sh_geometry
from specula_amd.lib.phasescreen import annular_pupil, von_karman_opd
from specula_amd.processing_objects.sh import SH, SHPlan

pytestmark = pytest.mark.gpu
DEV = 'cuda:0'
specula_amd.init(0, precision=1)


def lag_spectra(kernels, np_sub):
    N = kernels.shape[-1]
    lags = np.arange(-(np_sub - 1), np_sub) % N
    return np.ascontiguousarray(kernels[:, lags][:, :, lags].astype(np.complex64))


def run_lgs(g, A, frames, kernels, normalised):
    plan = SHPlan(g, lgs=True, max_frames=len(frames), device=0, kernels_normalised=normalised)
    plan.set_kernels(torch.tensor(lag_spectra(kernels, g.ovs_np_sub), device=DEV))
    out = torch.empty((len(frames), g.ccd_side, g.ccd_side), dtype=torch.float32, device=DEV)
    plan.forward(torch.tensor(A, dtype=torch.float32, device=DEV),
                 torch.tensor(np.asarray(frames, np.float32), device=DEV), out, 1.0, n_frames=len(frames))
    torch.cuda.synchronize()
    return out.cpu().numpy(), plan


def small_geom(cut=True):
    pitch = 0.05
    turb = 589e-9 / (8 * pitch) * O.RAD2ASEC
    kw = dict(wavelengthInNm=589, subap_wanted_fov=4 * turb * (0.8 if cut else 2), sensor_pxscale=turb / 2 if cut
              else turb, subap_on_diameter=8, subap_npx=8, fov_ovs_coeff=1.0)
    return kw, pitch, turb / 4


@pytest.mark.parametrize('normalised,cut', [(True, False), (False, True), (True, True)])
def test_small_lgs_parity(normalised, cut):
    kw, pitch, fr = small_geom(cut)
    g = sh_geometry(64, pitch, fov_resolution_arcsec=fr, **kw)
    go = O.sh_geometry(ef_size=64, pixel_pitch=pitch, fov_resolution_arcsec=fr, **kw)
    print(f'lgs small N={g.fft_size} cut={g.cutpixels} normalised={normalised}', flush=True)
    rng = np.random.default_rng(7)
    real = rng.random((64, 32, 32)) ** 6
    K = O.process_kernels(real, 8, 8)
    if not normalised:
        K = K * 1.7
    A = annular_pupil(64, 0.2)
    scr = von_karman_opd(128, pitch, seed=3)
    frames = np.stack([scr[:64, k:k + 64] for k in range(2)]).astype(np.float32)
    out, plan = run_lgs(g, A, frames, K, normalised)
    assert plan.parseval == (normalised and not cut)
    for k in range(2):
        ref = O.sh_forward(A, frames[k].astype(np.float64), go, kernels=K)
        check_intensity(out[k], ref, 8, 8)



@pytest.mark.parametrize('ef,npx,fovmul', [(48, 6, 1.6), (96, 12, 2), (96, 8, 1.6), (128, 12, 2)])
def test_lgs_np_variants_parity(ef, npx, fovmul):
    """np_sub 6, 12 and 16 (8 x 8 lenslets), npx below 16 and cut or not: the tile shapes
    (8 or 16 subapertures per block, one or two subapertures per 16-column matrix block)."""
    np_sub = ef // 8
    pitch = 0.05
    turb = 589e-9 / (np_sub * pitch) * O.RAD2ASEC
    kw = dict(wavelengthInNm=589, subap_wanted_fov=npx * turb * fovmul / 2, sensor_pxscale=turb * fovmul / 2,
              subap_on_diameter=8, subap_npx=npx, fov_ovs_coeff=1.0)
    g = sh_geometry(ef, pitch, fov_resolution_arcsec=turb / 4, **kw)
    go = O.sh_geometry(ef_size=ef, pixel_pitch=pitch, fov_resolution_arcsec=turb / 4, **kw)
    assert (g.ovs_np_sub, g.subap_npx) == (np_sub, npx)
    print(f'lgs np={np_sub} N={g.fft_size} cut={g.cutpixels} npx={npx}', flush=True)
    rng = np.random.default_rng(11)
    K = O.process_kernels(rng.random((64, g.fft_size, g.fft_size)) ** 6, 8, 8)
    A = annular_pupil(ef, 0.2)
    scr = von_karman_opd(2 * ef, pitch, seed=5)
    frames = np.stack([scr[:ef, k:k + ef] for k in range(2)]).astype(np.float32)
    out, _ = run_lgs(g, A, frames, K, True)
    for k in range(2):
        ref = O.sh_forward(A, frames[k].astype(np.float64), go, kernels=K)
        check_intensity(out[k], ref, 8, npx)

def test_lgs_40x40_parity():
    wl = workloads.lgs_40x40()
    g = wl.geometry()
    assert (g.ovs_np_sub, g.fft_size, g.cutpixels, g.subap_npx) == (16, 128, 0, 16)
    kern = workloads.lgs_kernel(g)
    assert kern.kernels_normalised()
    A = wl.pupil()
    frames = wl.frames(1)
    out, plan = run_lgs(g, A, frames, kern.kernels, True)
    assert plan.parseval
    ref = O.sh_forward(A, frames[0].astype(np.float32).astype(np.float64), oracle_geometry(wl),
                       kernels=kern.kernels)
    check_intensity(out[0], ref, 40, 16)
    # a registered static pupil regroups the tiles lit-first (1248 lit of 1600 subapertures;
    # dark tiles store zeros): bit-identical image
    At = torch.tensor(A, dtype=torch.float32, device=DEV)
    ph = torch.tensor(np.asarray(frames, np.float32), device=DEV)
    plan.set_static_pupil(At, 1.0)
    out2 = torch.empty((1, g.ccd_side, g.ccd_side), dtype=torch.float32, device=DEV)
    plan.forward(At, ph, out2, 1.0, n_frames=1)
    torch.cuda.synchronize()
    assert np.array_equal(out2.cpu().numpy(), out)


def test_sh_object_gaussian_lgs_and_sodium_inputs():
    """SH with a LaserLaunchTelescope (test_sh_slopec.py:62-71 pattern): Gaussian kernel when
    tel_position is empty, sodium-profile kernel (BaseValue inputs) otherwise."""
    kw, pitch, fr = small_geom(cut=False)
    A = annular_pupil(64, 0.2)
    ph = von_karman_opd(128, pitch, seed=4)[:64, :64].astype(np.float32)
    go = O.sh_geometry(ef_size=64, pixel_pitch=pitch, fov_resolution_arcsec=fr, **kw)
    ef = ElectricField(64, 64, pitch, S0=3.0, target_device_idx=0)
    ef.A[:] = torch.tensor(A, device=DEV)
    ef.phaseInNm[:] = torch.tensor(ph, device=DEV)
    ef.generation_time = 1
    for tel_pos in ([], [5.0, -3.0, 0.0]):
        llt = LaserLaunchTelescope(spot_size=0.8, tel_position=tel_pos)
        sh = SH(**kw, laser_launch_tel=llt, target_device_idx=0)
        sh._fov_resolution_arcsec = fr
        sh.inputs['in_ef'].set(ef)
        zl, zp = np.linspace(85e3, 95e3, 5), np.ones(5) / 5
        if tel_pos:
            sh.inputs['sodium_altitude'].set(BaseValue(value=zl, target_device_idx=-1))
            sh.inputs['sodium_intensity'].set(BaseValue(value=zp, target_device_idx=-1))
        sh.setup()
        sh.check_ready(1)
        sh.trigger()
        sh.post_trigger()
        img = sh.outputs['out_i'].i.cpu().numpy()
        if tel_pos:
            K = O.convolution_kernels_for_sh(go, tel_pos, 0.8, zl.tolist(), zp.tolist())
        else:
            K = O.gaussian_kernels_for_sh(go, 0.8)
        ref = O.sh_forward(A, ph.astype(np.float64), go, S0=3.0, kernels=K)
        check_intensity(img, ref, 8, 8)
