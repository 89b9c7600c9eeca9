"""bsx: BallaSim's per-pixel continuous-time event-sensor path on MI355X (gfx950).

The compute runs in ballasim_amd/lib/libbsx.so (hand-written HIP kernels behind the C ABI of
include/bsx/bsx_sensor.h); this package holds the Python surface over it.
"""
from .sensor import Rig, SensorModel, SensorParams, device_count, noise_tables, preset, preset_names  # noqa: F401
