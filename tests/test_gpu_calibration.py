"""Modal push-pull calibration on the GPU against the reference's scao_rec_n8_th0.5_ref.fits
(test_sh_calibration.py, params_scao_sh_test{,_rec}.yml): 40 Zernike modes at 50 nm through the
DM projection GEMM, SH with field resampling (fov_ovs 2, N 32, cut 24) and ShSlopec."""
import os

import numpy as np
import pytest
import torch

from oracle import sh_oracle as O
from specula_amd.lib import fits
from specula_amd.lib.geometry import sh_geometry
from specula_amd.lib.grids import make_mask

pytestmark = pytest.mark.gpu
DEV = 'cuda:0'
GOLDEN = os.path.join(os.path.dirname(__file__), 'golden')
SCAO_SH = dict(wavelengthInNm=600, subap_wanted_fov=4.0, sensor_pxscale=0.5, subap_on_diameter=8, subap_npx=8)
PITCH = 0.015625


def _fixture_rec():
    return fits.read(os.path.join(GOLDEN, 'scao_rec_n8_th0.5_ref.fits'))[1][0]


def _idxs():
    return fits.read(os.path.join(GOLDEN, 'scao_subaps_n8_th0.5_ref.fits'))[1][0].T


def _dm(device):
    from specula_amd.processing_objects.dm import DM
    return DM(pixel_pitch=PITCH, height=0, type_str='zernike', nmodes=40, npixels=64, obsratio=0.1, sign=1,
              target_device_idx=device)


def test_batched_pushpull_reproduces_scao_rec():
    """80 push-pull fields in one batch: DM GEMM -> SH forward -> CoG; IM rows accumulated as
    ImRecCalibrator does; pinv written as Recmat.save writes it."""
    from specula_amd.data_objects.intmat import Intmat
    from specula_amd.sensor import SHSensor
    print('calibration: batched push-pull, 80 fields', flush=True)
    g = sh_geometry(64, PITCH, **SCAO_SH)
    dm = _dm(0)
    A = make_mask(64, None, 1.0).astype(np.float32) * dm.layer.A.cpu().numpy()
    sen = SHSensor(g, A, idxs=_idxs(), max_frames=80, device=0, extrapolation_mask=np.ones((64, 64)))
    th = torch.tensor(O.pushpull_time_history(40, [50.0] * 40), dtype=torch.float32, device=DEV)
    phases = dm.project(th)
    s = sen.forward(phases).clone()
    torch.cuda.synchronize()
    amp = th.abs().amax(dim=1)
    sign = torch.sign(th.sum(dim=1))
    im = torch.zeros((40, s.shape[1]), dtype=torch.float32, device=DEV)
    im.index_add_(0, th.abs().argmax(dim=1), s / (sign * amp)[:, None])
    rec = Intmat(im, target_device_idx=0).generate_rec().recmat.T.cpu().numpy()
    ref = _fixture_rec()
    err = np.abs(rec - ref).max()
    im_o, _, _ = __import__('test_oracle_pins').scao_pushpull_oracle()
    im_err = np.abs(im.cpu().numpy() - im_o).max() / np.abs(im_o).max()
    print(f'  rec max |gpu - fixture| {err:.2e} (decimal 3: 1.5e-3); IM vs oracle {im_err:.2e} of max', flush=True)
    np.testing.assert_array_almost_equal(rec, ref, decimal=3)
    assert im_err < 1e-4
    # the same IM from the library's push-pull fields and rows (shwfs_pushpull_fields / _rows)
    modes = dm.project(torch.eye(40, device=DEV))
    im2 = sen.pushpull_interaction_matrix(torch.zeros((64, 64), device=DEV), modes, 50.0)
    d = float((im2.double() - im.double()).norm() / im.double().norm())
    print(f'  shwfs_pushpull IM vs calibrator accumulation: rel {d:.2e}', flush=True)
    assert d < 1e-5


def test_drop_in_objects_calibration_loop(tmp_path):
    """DM -> field -> SH -> CCD -> ShSlopec -> ImRecCalibrator objects stepped as LoopControl does
    (SH set up on the still all-ones field), rec file written at finalize equals the fixture."""
    from specula_amd.base import BaseValue
    from specula_amd.data_objects.electric_field import ElectricField
    from specula_amd.data_objects.subap_data import SubapData
    from specula_amd.processing_objects.im_rec_calibrator import ImRecCalibrator
    from specula_amd.processing_objects.sh import SH
    from specula_amd.processing_objects.sh_slopec import ShSlopec
    print('calibration: drop-in objects loop, 80 steps', flush=True)
    dm = _dm(0)
    pup = torch.tensor(make_mask(64, None, 1.0).astype(np.float32), device=DEV)
    cmd = BaseValue(value=torch.zeros(40, device=DEV), target_device_idx=0)
    dm.inputs['in_command'].set(cmd)
    ef = ElectricField(64, 64, PITCH, S0=1.0, target_device_idx=0)
    sh = SH(target_device_idx=0, **SCAO_SH)
    sh.inputs['in_ef'].set(ef)
    from specula_amd.data_objects.simul_params import SimulParams
    from specula_amd.processing_objects.ccd import CCD
    # detector of params_scao_sh_test.yml with the _rec override (no photon / readout noise)
    ccd = CCD(SimulParams(pixel_pupil=64, pixel_pitch=PITCH, time_step=0.001), size=[64, 64], dt=0.001, bandw=300,
              photon_noise=False, readout_noise=False, readout_level=1.0, quantum_eff=0.3, target_device_idx=0)
    ccd.inputs['in_i'].set(sh.outputs['out_i'])
    pix = ccd.outputs['out_pixels']
    sd = SubapData.restore(os.path.join(GOLDEN, 'scao_subaps_n8_th0.5_ref.fits'), target_device_idx=0)
    slopec = ShSlopec(sd, target_device_idx=0)
    slopec.inputs['in_pixels'].set(pix)
    cal = ImRecCalibrator(nmodes=40, data_dir=str(tmp_path), rec_tag='rec', target_device_idx=0)
    cal.inputs['in_slopes'].set(slopec.outputs['out_slopes'])
    cal.inputs['in_commands'].set(cmd)
    for obj in (dm, sh, ccd, slopec, cal):
        obj.setup()
    th = O.pushpull_time_history(40, [50.0] * 40)
    step = ccd.seconds_to_t(0.001)                   # LoopControl advances by the time step in ns
    for k, c in enumerate(th, start=1):
        t = k * step
        cmd.value = torch.tensor(c, dtype=torch.float32, device=DEV)
        cmd.generation_time = t
        dm.check_ready(t); dm.trigger(); dm.post_trigger()
        ef.A.copy_(pup * dm.layer.A)                 # pupilstop x DM layer, as the propagation step
        ef.phaseInNm.copy_(dm.layer.phaseInNm)
        ef.generation_time = t
        sh.check_ready(t); sh.trigger(); sh.post_trigger()
        ccd.check_ready(t); ccd.trigger(); ccd.post_trigger()
        assert pix.generation_time == t
        slopec.check_ready(t); slopec.trigger(); slopec.post_trigger()
        cal.check_ready(t); cal.trigger(); cal.post_trigger()
    cal.finalize()
    rec = fits.read(str(tmp_path / 'rec.fits'))[1][0]
    err = np.abs(rec - _fixture_rec()).max()
    print(f'  rec max |objects - fixture| {err:.2e}', flush=True)
    np.testing.assert_array_almost_equal(rec, _fixture_rec(), decimal=3)
