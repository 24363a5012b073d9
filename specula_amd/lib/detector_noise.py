"""Detector read-out noise and dark current per detector model, for the CCD's 'auto' noise levels
(specula/lib/calc_detector_noise.py:3-122, used at processing_objects/ccd.py:67-74).

The numbers are the detectors' specifications as the reference tabulates them (read-out noise in
electrons, dark current in electrons per pixel per frame, read-out time in seconds); they are data
that must stay alike.  Each model is a function of the frame rate ``fs`` [Hz] and the binning.
"""
import math

__all__ = ['detector_noise', 'MODELS']


def _ocam_dark(fs, b):
    return (0.444 / fs + 0.00556) * b ** 2


def _ccd220_dark(fs, b):
    # dark current per frame of the CCD220 by frame-rate band (calc_detector_noise.py:78-88)
    for fmin, rate in ((900, 3.6), (500, 1.8), (360, 1.4), (180, 1.08)):
        if fs >= fmin:
            return rate * b ** 2 / fs
    return 0.84 * b ** 2 / fs


def _ccd39(fs, b, speed):
    speeds, rons = (0.95e-3, 1.56e-3, 0.89e-3, 0.68e-3), (10.5, 4.2, 4.5, 4.6)
    if not 1 <= b <= 4:
        raise ValueError(f'CCD39 has binning modes 1-4, not {b}')
    return rons[b - 1], 0.0, speed or speeds[b - 1]


def _ocam2k(fs, b, speed):
    return 150.0 / 400.0, _ocam_dark(fs, b), speed or 0.5e-3


def _ocam2k_sw(fs, b, speed):
    return 150.0 / 400.0, _ocam_dark(fs, b), speed or 0.24e-3


def _ocam2k_sw_spec(fs, b, speed):
    rons = {1: 0.40, 2: 0.38, 3: 0.45, 4: 0.51}
    if b not in rons:
        raise ValueError(f'OCAM2kSWspec has binning modes 1-4, not {b}')
    return rons[b], 1.38 / fs * b ** 2, speed or 0.24e-3


def _ocam2(fs, b, speed):
    if fs > 1.5e3:
        return 150.0 / 400.0, _ocam_dark(fs, b), speed or 0.5e-3
    return 50.0 / 400.0, _ocam_dark(fs, b), speed or 0.667e-3


def _hnu(fs, b, speed):
    # the reference fixes the read-out time of this model whatever CCD_speed says
    return 0.1, (0.0006 / fs + 0.005) * b ** 2, 0.995e-3


def _ccd220(fs, b, speed):
    return 80.0 / 400.0 * math.sqrt(b), _ccd220_dark(fs, b), speed or 0.667e-3


def _ccd220_good_bin(fs, b, speed):
    return 80.0 / 400.0, _ccd220_dark(fs, b), speed or 0.667e-3


def _ccd220_eso(fs, b, speed):
    return 1.0 * math.sqrt(b), 1.2 * b ** 2 / fs, speed or 0.667e-3


def _c_red(fs, b, speed):
    return 24.0 / 30.0, 100.0 / fs, speed or 0.5e-3


def _ccid75(fs, b, speed):
    # read-out modes (time, noise), slowest first; the first mode whose read-out time fits in the
    # frame period is taken (calc_detector_noise.py:104-118)
    f = max(fs) if isinstance(fs, (list, tuple)) else fs
    if b > 1:
        modes, period = ((0.588e-3, 1.7), (0.294e-3, 2.1)), 1.0 / f
    else:
        modes = ((2.0e-3, 1.7), (1.0e-3, 2.1), (0.588e-3, 2.6), (0.435e-3, 3.6), (0.294e-3, 4.4))
        period = 1.0 / round(f)
    for t, ron in modes:
        if period >= t:
            return ron, 26.0 * b ** 2 / f, speed or t
    raise ValueError(f'CCID75: no read-out mode fits a frame rate of {f} Hz')


MODELS = {
    'CCD39': _ccd39, 'OCAM2k': _ocam2k, 'OCAM2kSW': _ocam2k_sw, 'OCAM2kSWspec': _ocam2k_sw_spec,
    'OCAM2': _ocam2, 'HnU': _hnu, 'CCD220': _ccd220, 'CCD220goodBin': _ccd220_good_bin,
    'CCD220_ESO': _ccd220_eso, 'C-RED': _c_red, 'CCID75': _ccid75,
}


def detector_noise(fs, ccd_type, binning=1, ccd_speed=None):
    """(read-out noise [e-], dark current [e-/px/frame], read-out time [s]) of detector ``ccd_type``
    at frame rate ``fs`` [Hz] and ``binning``; ``ccd_speed`` overrides the model's read-out time
    (except for 'HnU').  Unknown models raise ValueError, as the reference does."""
    model = MODELS.get(ccd_type)
    if model is None:
        raise ValueError(f'Unknown CCD type: {ccd_type}')
    return model(fs, int(binning), ccd_speed or 0.0)
