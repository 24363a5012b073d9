from specula_amd.base import BaseDataObj


class LaserLaunchTelescope(BaseDataObj):
    """specula/data_objects/laser_launch_telescope.py:3-33."""

    def __init__(self, spot_size: float = 0.0, tel_position: list = [], beacon_focus: float = 90e3,
                 beacon_tt: list = [0.0, 0.0], target_device_idx: int = None, precision: int = None):
        super().__init__(target_device_idx=target_device_idx, precision=precision)
        self.spot_size = spot_size
        self.tel_pos = list(tel_position)
        self.beacon_focus = beacon_focus
        self.beacon_tt = list(beacon_tt)
