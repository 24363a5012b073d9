"""Host side of the modal push-pull calibration: influence functions, DM projection, Intmat /
Recmat FITS formats and the ImRecCalibrator accumulation, against the oracle and the reference's
scao_rec_n8_th0.5_ref.fits.  CPU only (data objects on target_device_idx -1)."""
import os

import numpy as np
import pytest
import torch

from oracle import sh_oracle as O
from specula_amd.data_objects.ifunc import IFunc
from specula_amd.data_objects.intmat import Intmat
from specula_amd.data_objects.recmat import Recmat
from specula_amd.lib import fits
from specula_amd.lib.zernike import noll_degree, zernike, zernike_ifunc

GOLDEN = os.path.join(os.path.dirname(__file__), 'golden')


def test_noll_degree_and_maps_match_oracle():
    for j in range(1, 60):
        assert noll_degree(j) == O.zernike_degree(j)
    for j in (1, 2, 3, 7, 12, 23, 44):
        np.testing.assert_allclose(zernike(j, 33), O.zernike_map(j, 33), rtol=0, atol=1e-11)


def test_zernike_ifunc_matches_oracle():
    q, m = zernike_ifunc(64, 40, obsratio=0.1, dtype=np.float64)
    qo, mo = O.compute_zern_ifunc(64, 40, obsratio=0.1)
    np.testing.assert_array_equal(m, mo)
    np.testing.assert_allclose(q, qo, rtol=0, atol=1e-9)


def test_ifunc_cut_keeps_reference_slice_semantics():
    f = IFunc(type_str='zernike', npixels=32, nmodes=10, obsratio=0.0, target_device_idx=-1)
    assert f.size[0] == 10
    full = f.influence_function.clone()
    f.cut(start_mode=2, nmodes=6)          # ifunc.py:154-157: rows start_mode:nmodes
    torch.testing.assert_close(f.influence_function, full[2:6])
    g = IFunc(ifunc=full.numpy(), mask=f.mask_inf_func.numpy(), idx_modes=[0, 3, 5], target_device_idx=-1)
    torch.testing.assert_close(g.influence_function, full[[0, 3, 5]])
    with pytest.raises(NotImplementedError):
        IFunc(type_str='zonal', npixels=32, target_device_idx=-1)
    with pytest.raises(ValueError):
        IFunc(type_str='bogus', npixels=32, target_device_idx=-1)


def _dm(sign=-1, precision=0):
    from specula_amd.processing_objects.dm import DM
    return DM(pixel_pitch=0.015625, height=0, type_str='zernike', nmodes=40, npixels=64, obsratio=0.1,
              sign=sign, target_device_idx=-1, precision=precision)


def test_dm_trigger_and_batched_projection_match_oracle():
    from specula_amd.base import BaseValue
    dm = _dm(sign=1)
    ifunc, dmask = O.compute_zern_ifunc(64, 40, obsratio=0.1)
    rng = np.random.default_rng(1)
    cmds = rng.standard_normal((5, 40)) * 30
    cmd_in = BaseValue(value=torch.tensor(cmds[2]), target_device_idx=-1)
    cmd_in.generation_time = 1
    dm.inputs['in_command'].set(cmd_in)
    dm.setup()
    dm.check_ready(1)
    dm.trigger()
    np.testing.assert_allclose(dm.layer.phaseInNm.numpy(), O.dm_phase(ifunc, dmask, cmds[2], sign=1), atol=1e-9)
    np.testing.assert_array_equal(dm.layer.A.numpy(), dmask)
    base = torch.tensor(rng.standard_normal((64, 64)))
    ph = dm.project(torch.tensor(cmds), base=base).numpy()
    for k in range(5):
        np.testing.assert_allclose(ph[k], base.numpy() + O.dm_phase(ifunc, dmask, cmds[k], sign=1), atol=1e-9)


def test_intmat_recmat_fits_round_trip_and_fixture_layout(tmp_path):
    rng = np.random.default_rng(2)
    im = Intmat(rng.standard_normal((6, 20)), pupdata_tag='tag', norm_factor=2.5, target_device_idx=-1, precision=0)
    im.save(str(tmp_path / 'im'))
    back = Intmat.restore(str(tmp_path / 'im.fits'), target_device_idx=-1)
    np.testing.assert_array_equal(back.intmat.numpy(), im.intmat.numpy())
    assert (back.pupdata_tag, back.norm_factor) == ('tag', 2.5)
    rec = im.generate_rec()
    assert tuple(rec.recmat.shape) == (20, 6)
    np.testing.assert_allclose(rec.recmat.numpy(), np.linalg.pinv(im.intmat.numpy()), atol=1e-12)
    rec.save(str(tmp_path / 'rec'))
    raw = fits.read(str(tmp_path / 'rec.fits'))
    np.testing.assert_array_equal(raw[1][0], rec.recmat.numpy().T)   # stored transposed (recmat.py:47)
    ref = Recmat.restore(os.path.join(GOLDEN, 'scao_rec_n8_th0.5_ref.fits'), target_device_idx=-1)
    assert tuple(ref.recmat.shape) == (40, 104) and ref.norm_factor == 0.0


def test_im_rec_calibrator_accumulates_and_writes_fixture_rec(tmp_path):
    """The drop-in calibrator fed the oracle's push-pull slopes writes a Recmat file equal to
    scao_rec_n8_th0.5_ref.fits (decimal 3)."""
    from test_oracle_pins import scao_pushpull_oracle
    from specula_amd.base import BaseValue
    from specula_amd.data_objects.slopes import Slopes
    from specula_amd.processing_objects.im_rec_calibrator import ImRecCalibrator
    im_o, th, slopes = scao_pushpull_oracle()
    cal = ImRecCalibrator(nmodes=40, data_dir=str(tmp_path), rec_tag='rec', im_tag='im', target_device_idx=-1,
                          precision=0)
    sl = Slopes(length=104, target_device_idx=-1, precision=0)
    cmd = BaseValue(value=torch.zeros(40, dtype=torch.float64), target_device_idx=-1)
    cal.inputs['in_slopes'].set(sl)
    cal.inputs['in_commands'].set(cmd)
    cal.setup()
    for t, (s, c) in enumerate(zip(slopes, th), start=1):
        sl.slopes[:] = torch.tensor(s)
        sl.generation_time = t
        cmd.value = torch.tensor(c)
        cmd.generation_time = t
        assert cal.check_ready(t)
        cal.trigger()
        cal.post_trigger()
    np.testing.assert_allclose(cal.im.numpy(), im_o, rtol=1e-12, atol=1e-15)
    cal.finalize()
    rec = fits.read(str(tmp_path / 'rec.fits'))[1][0]
    ref = fits.read(os.path.join(GOLDEN, 'scao_rec_n8_th0.5_ref.fits'))[1][0]
    np.testing.assert_array_almost_equal(rec, ref, decimal=3)
    np.testing.assert_allclose(Intmat.restore(str(tmp_path / 'im.fits'), target_device_idx=-1).intmat.numpy(), im_o,
                               rtol=1e-12)
    with pytest.raises(FileExistsError):
        again = ImRecCalibrator(nmodes=40, data_dir=str(tmp_path), rec_tag='rec', target_device_idx=-1)
        again.inputs['in_slopes'].set(sl)
        again.inputs['in_commands'].set(cmd)
        again.setup()


def test_slopes_fits_restore_fixture_and_round_trip(tmp_path):
    from specula_amd.data_objects.slopes import Slopes
    sn = Slopes.restore(os.path.join(GOLDEN, 'scao_sn_n8_th0.5_ref.fits'), target_device_idx=-1)
    raw = fits.read(os.path.join(GOLDEN, 'scao_sn_n8_th0.5_ref.fits'))[1][0]
    np.testing.assert_array_equal(sn.slopes.numpy(), raw)
    assert sn.size == 104 and not sn.interleave and sn.pupdata_tag == ''
    s = Slopes(slopes=np.arange(10, dtype=np.float32), interleave=True, target_device_idx=-1)
    s.pupdata_tag = 'pup'
    s.save(str(tmp_path / 's.fits'))
    b = Slopes.restore(str(tmp_path / 's.fits'), target_device_idx=-1)
    np.testing.assert_array_equal(b.slopes.numpy(), s.slopes.numpy())
    assert b.interleave and b.pupdata_tag == 'pup'
    torch.testing.assert_close(b.xslopes, torch.tensor([0., 2, 4, 6, 8]))


def test_convolution_kernel_fits_cache_round_trip(tmp_path):
    from specula_amd.data_objects.convolution_kernel import ConvolutionKernel
    kw = dict(dimx=4, dimy=4, pxscale=0.3, pupil_size_m=8.0, dimension=32, launcher_pos=[5.0, 5.0, 0.0],
              launcher_size=1.0, target_device_idx=-1)
    k = ConvolutionKernel(**kw)
    k.cache_dir = str(tmp_path)
    k.prepare_for_sh(sodium_altitude=[88e3, 90e3, 92e3], sodium_intensity=[0.2, 0.6, 0.2])
    files = list(tmp_path.glob('ConvolutionKernel*.fits'))
    assert len(files) == 1
    hdr = fits.read(str(files[0]))[0][1]
    assert hdr['VERSION'] == 1.1 and hdr['DIM'] == 32 and hdr['POSTT'] is True
    k2 = ConvolutionKernel(**kw)
    k2.cache_dir = str(tmp_path)
    k2.prepare_for_sh(sodium_altitude=[88e3, 90e3, 92e3], sodium_intensity=[0.2, 0.6, 0.2])   # loads the file
    np.testing.assert_allclose(k2.kernels, k.kernels, rtol=0, atol=1e-7)   # FP32 file
    k3 = ConvolutionKernel.restore(str(files[0]), target_device_idx=-1, return_fft=True)
    np.testing.assert_allclose(k3.kernels, k.kernels, rtol=0, atol=1e-7)
