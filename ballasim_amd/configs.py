"""The sensor parameters of the benchmark workloads, as the reference's SensorParams (the mapping of
each workload's wording onto the reference's keys):

- config 1 (346x260 disk, 100 ms): the cascade with a 200-Hz photoreceptor, mapped onto the
  reference's three-stage cascade as tau_sf = 1 / (2 pi 200 Hz) = 795.8 us and tau_o = tau_sf / 4
  (the prophesee family's ratio), k = 1e16 (k I >> 2 pi 200 Hz at 100 lux); thresholds 0.2 +- 0.02,
  refractory 100 us with the physical reset, noise, leak and hot pixels off;
- config 2 (1280x720, 64 rectangles, 500 ms): thresholds 0.15 +- 0.03, refractory 50 us, leak
  0.05 Hz/px; the instantaneous response with the library's rate-specified shot process at
  0.1 Hz/px, or the cascade (noise off, or the reference's in-path noise in place of the shot
  process);
- config 3 (the 4-camera rig): config 1's thresholds and refractory, noise off.
Irradiance is in lux with illuminance_scale = LUX_TO_FA fA per lux.
"""
import math

from .scene import LUX_TO_FA
from .sensor import preset


def config1_params(photoreceptor="cascade"):
    p = preset("prophesee_gen3")
    p.photoreceptor = photoreceptor
    p.noise = False
    p.theta_on = p.theta_off = 0.2
    p.sigma_theta_on = p.sigma_theta_off = 0.02
    p.tau_sf = 1.0 / (2.0 * math.pi * 200.0)
    p.tau_o = p.tau_sf / 4.0
    p.k = 1e16
    p.refractory_s = 100e-6
    p.reset_mode = "physical"
    p.leak_rate_hz = 0.0
    p.hot_fraction = 0.0
    p.shot_rate_hz = 0.0
    p.illuminance_scale = LUX_TO_FA
    return p


def config2_params(photoreceptor="instantaneous", noise=False):
    """noise: the cascade's in-path noise (cascade only); the instantaneous response carries the shot
    process instead."""
    p = preset("prophesee_gen3")
    p.photoreceptor = photoreceptor
    p.noise = bool(noise) and photoreceptor == "cascade"
    p.theta_on = p.theta_off = 0.15
    p.sigma_theta_on = p.sigma_theta_off = 0.03
    p.refractory_s = 50e-6
    p.leak_rate_hz = 0.05
    p.shot_rate_hz = 0.1 if photoreceptor == "instantaneous" else 0.0
    p.illuminance_scale = LUX_TO_FA
    return p


def config3_params(photoreceptor="instantaneous"):
    p = preset("prophesee_gen3")
    p.photoreceptor = photoreceptor
    p.noise = False
    p.theta_on = p.theta_off = 0.2
    p.sigma_theta_on = p.sigma_theta_off = 0.02
    p.refractory_s = 100e-6
    p.leak_rate_hz = 0.0
    p.shot_rate_hz = 0.0
    p.illuminance_scale = LUX_TO_FA
    return p
