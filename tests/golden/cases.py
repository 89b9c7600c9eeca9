"""The golden cases: parameters (a preset plus overrides) and seeded inputs. make_golden.py stores the
inputs with the reference's outputs, so the fixtures do not depend on this generator staying stable."""
import ctypes as C

import numpy as np

from ballasim_amd import _abi
from tests.helpers import random_runs

WINDOW = 2 ** 23

MAP_CASES = {
    "prophesee_gen3": {"preset": "prophesee_gen3", "n": 4000, "seed": 123456789, "camera": 2},
    "samsung_gen3": {"preset": "samsung_gen3", "n": 2000, "seed": 7, "camera": 0},
    "prophesee_corr_symmetric": {"preset": "prophesee_gen3", "n": 2000, "seed": 99, "camera": 5,
                                 "overrides": {"theta_onoff_corr": -0.4, "initial_phase": 1, "leak_rate_hz": 0.3,
                                               "temperature_K": 310.0}},
}

CASES = {
    "instant_ideal": {"preset": "ideal", "width": 24, "height": 16, "groups": 1, "seed": 11, "camera": 0,
                      "t0": -1000, "calls": 3, "input_seed": 1},
    "instant_physical_leak": {"preset": "ideal", "width": 24, "height": 16, "groups": 1, "seed": 12, "camera": 1,
                              "t0": 0, "calls": 3, "input_seed": 2,
                              "overrides": {"reset_mode": 0, "refractory_s": 1e-4, "leak_rate_hz": 2.0,
                                            "sigma_theta_on": 0.03, "sigma_theta_off": 0.02, "dark_sigma_ln": 0.3,
                                            "leak_sigma_log10": 0.1}},
    "instant_groups2_readout": {"preset": "ideal", "width": 16, "height": 16, "groups": 2, "seed": 13, "camera": 3,
                                "t0": 5 * WINDOW - 77, "calls": 3, "input_seed": 3,
                                "overrides": {"reset_mode": 0, "refractory_s": 0.0, "eps_ci": 0.002,
                                              "initial_phase": 1, "readout_quantum_s": 1e-6,
                                              "readout_delay_s": 3e-6, "clock_offset_ns": 1234,
                                              "clock_drift_ppm": 2.5}},
    "cascade_prophesee_physical": {"preset": "prophesee_gen3", "width": 16, "height": 12, "groups": 1, "seed": 21,
                                   "camera": 0, "t0": 3 * WINDOW + 5, "calls": 3, "input_seed": 4,
                                   "overrides": {"noise": 0}},
    "cascade_samsung_ideal_leak": {"preset": "samsung_gen3", "width": 16, "height": 12, "groups": 2, "seed": 22,
                                   "camera": 6, "t0": -2 * WINDOW, "calls": 3, "input_seed": 5,
                                   "overrides": {"noise": 0, "reset_mode": 1, "refractory_s": 2e-5,
                                                 "leak_rate_hz": 1.5}},
    # the reference's default sensor: the cascade with its in-path noise, unmodified
    "cascade_prophesee_noise": {"preset": "prophesee_gen3", "width": 16, "height": 12, "groups": 1, "seed": 23,
                                "camera": 1, "t0": 2 * WINDOW - 40_000, "calls": 3, "input_seed": 6},
    "cascade_noise_ideal_levels3": {"preset": "prophesee_gen3", "width": 16, "height": 12, "groups": 1, "seed": 24,
                                    "camera": 4, "t0": -3 * 65536, "calls": 3, "input_seed": 7,
                                    "overrides": {"reset_mode": 1, "refractory_s": 1e-5, "noise_leaf_level": 3,
                                                  "sigma_o": 0.12, "leak_rate_hz": 0.5}},
}
# Fields of the saved state each case stores (the noise fields only where the noise is on).
STATE_FIELDS = ("z", "x1", "s2", "s3", "i_scale", "tv", "s3_ref", "n3_ref", "dr_ref", "t_ref", "blind_until",
                "t_search", "flags")
NOISE_STATE_FIELDS = ("nk", "nk1", "ln_a1", "cell", "have_k1")


def noise_case(spec):
    return spec["preset"] == "prophesee_gen3" and spec.get("overrides", {}).get("noise", 1) != 0


def case_params(spec):
    lib = _abi.load_library()
    p = _abi.Params()
    _abi.check(lib.bsx_params_preset(spec["preset"].encode(), C.byref(p)))
    for k, v in spec.get("overrides", {}).items():
        setattr(p, k, v)
    return p


def case_inputs(spec):
    """irr0 and the calls (ids, t, irr, offsets, t_end), seeded by spec['input_seed']."""
    rng = np.random.default_rng(spec["input_seed"])
    P, G = spec["width"] * spec["height"], spec["groups"]
    irr0 = np.exp(rng.normal(3, 1, size=(P, G))).astype(np.float32)
    calls = []
    t = spec["t0"]
    for _ in range(spec["calls"]):
        t1 = t + int(rng.integers(1, 2 * WINDOW))
        ids, ts, ir, offs = random_runs(rng, P, t, t1, G=G)
        calls.append((ids, ts, ir, offs, t1))
        t = t1
    return irr0, calls
