from dataclasses import dataclass


@dataclass
class SimulParams:
    """Simulation-wide parameters read by DM and CCD (specula/data_objects/simul_params.py:3-32)."""
    pixel_pupil: int = None
    pixel_pitch: float = None
    root_dir: str = '.'
    total_time: float = 0.1
    time_step: float = 0.001
    zenithAngleInDeg: float = 0
    display_server: bool = False
