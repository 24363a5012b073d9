"""'auto' detector noise levels (calc_detector_noise.py:3-122 through ccd.py:67-74), CPU only.

Known answers worked out by hand from the detector specifications the reference tabulates (read-out
noise [e-], dark current [e-/px/frame] = dark rate x binning^2 / fs, read-out time [s]), at frame
rates that select each band or read-out mode; plus the resolution rule of the CCD object (only
when readout, dark and background levels are all given; the background never resolved; a string
level is an error only where the noise chain uses it)."""
import math

import pytest

from specula_amd.lib.detector_noise import MODELS, detector_noise
from specula_amd.processing_objects.ccd import resolve_noise_levels

CASES = [
    # (model, fs, binning) -> (RON, dark, speed)
    (('CCD39', 1000.0, 1), (10.5, 0.0, 0.95e-3)),
    (('CCD39', 1000.0, 3), (4.5, 0.0, 0.89e-3)),
    (('OCAM2k', 1000.0, 1), (0.375, 0.444e-3 + 0.00556, 0.5e-3)),
    (('OCAM2k', 500.0, 2), (0.375, 4 * (0.444 / 500 + 0.00556), 0.5e-3)),
    (('OCAM2kSW', 2000.0, 1), (0.375, 0.444 / 2000 + 0.00556, 0.24e-3)),
    (('OCAM2kSWspec', 1000.0, 3), (0.45, 9 * 1.38e-3, 0.24e-3)),
    (('OCAM2', 2000.0, 1), (0.375, 0.444 / 2000 + 0.00556, 0.5e-3)),       # fast mode above 1.5 kHz
    (('OCAM2', 1000.0, 1), (0.125, 0.444e-3 + 0.00556, 0.667e-3)),        # slow mode
    (('HnU', 100.0, 2), (0.1, 4 * (0.0006 / 100 + 0.005), 0.995e-3)),
    (('CCD220', 1000.0, 4), (0.2 * 2.0, 16 * 3.6 / 1000, 0.667e-3)),     # band fs >= 900
    (('CCD220', 600.0, 1), (0.2, 1.8 / 600, 0.667e-3)),                  # band 500 .. 900
    (('CCD220', 400.0, 1), (0.2, 1.4 / 400, 0.667e-3)),                  # band 360 .. 500
    (('CCD220', 200.0, 1), (0.2, 1.08 / 200, 0.667e-3)),                 # band 180 .. 360
    (('CCD220', 100.0, 1), (0.2, 0.84 / 100, 0.667e-3)),                 # below 180
    (('CCD220goodBin', 1000.0, 4), (0.2, 16 * 3.6 / 1000, 0.667e-3)),
    (('CCD220_ESO', 500.0, 4), (2.0, 16 * 1.2 / 500, 0.667e-3)),
    (('C-RED', 1000.0, 1), (0.8, 0.1, 0.5e-3)),
    (('CCID75', 400.0, 1), (1.7, 26 / 400, 2.0e-3)),        # 2.5 ms period: slowest mode fits
    (('CCID75', 1000.0, 1), (2.1, 26 / 1000, 1.0e-3)),
    (('CCID75', 2000.0, 1), (3.6, 26 / 2000, 0.435e-3)),     # 0.5 ms: first mode <= 0.5 ms is 0.435 ms
    (('CCID75', 3000.0, 2), (2.1, 4 * 26 / 3000, 0.294e-3)), # binned modes: 0.333 ms fits only 0.294 ms
]


@pytest.mark.parametrize('args,want', CASES, ids=[f'{a[0]}-{a[1]:g}Hz-b{a[2]}' for a, _ in CASES])
def test_detector_models(args, want):
    model, fs, b = args
    got = detector_noise(fs, model, b)
    for g, w in zip(got, want):
        assert math.isclose(g, w, rel_tol=1e-12, abs_tol=1e-15)


def test_every_model_is_covered_and_unknown_models_raise():
    assert {a[0] for a, _ in CASES} == set(MODELS)
    with pytest.raises(ValueError, match='Unknown CCD type'):
        detector_noise(1000.0, 'NoSuchCam')
    with pytest.raises(ValueError):
        detector_noise(1000.0, 'CCD39', 5)


def test_speed_override():
    assert detector_noise(1000.0, 'OCAM2k', 1, ccd_speed=1e-3)[2] == 1e-3
    assert detector_noise(1000.0, 'HnU', 1, ccd_speed=1e-3)[2] == 0.995e-3      # fixed by the model


def test_resolution_rule_of_the_ccd_object():
    ron, dark, _ = detector_noise(1.0 / 0.002, 'OCAM2k', 2)
    # all three given: the 'auto' ones resolved at fs = 1/dt and the given binning
    assert resolve_noise_levels('auto', 'auto', 3.0, 'OCAM2k', 0.002, 2, True, True) == (ron, dark, 3.0)
    assert resolve_noise_levels(1.5, 'auto', 3.0, 'OCAM2k', 0.002, 2, True, True) == (1.5, dark, 3.0)
    # background 'auto' is never resolved: fine while unused, an error where the offset is added
    assert resolve_noise_levels('auto', 'auto', 'auto', 'OCAM2k', 0.002, 2, True, False) == (ron, dark, 0.0)
    with pytest.raises(ValueError, match='background_level'):
        resolve_noise_levels('auto', 'auto', 'auto', 'OCAM2k', 0.002, 2, True, True)
    # one level missing: nothing resolved (ccd.py:66), an 'auto' read-out noise in use is an error
    with pytest.raises(ValueError, match='readout_level'):
        resolve_noise_levels('auto', '', 3.0, 'OCAM2k', 0.002, 2, True, False)
    assert resolve_noise_levels('auto', '', 3.0, 'OCAM2k', 0.002, 2, False, False) == (0.0, 0.0, 3.0)
    # plain numbers pass through
    assert resolve_noise_levels(1.0, 2.0, 3.0, '', 0.001, 1, True, True) == (1.0, 2.0, 3.0)
