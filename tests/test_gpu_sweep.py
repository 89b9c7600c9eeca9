"""GPU parity of the instantaneous response on randomized configurations: resets, refractory
periods, leak from none to 40 Hz, threshold spreads, one to three light groups, hot pixels and the
shot process, each over four calls of ragged runs with up to 200 samples per pixel (up to 5.6 M
events per call: a call with more than 2^20 events overflows the event buffer's first size and the
walk re-runs, which those trials assert). Events, reported times and the sub-step and piece counters
must equal the oracle's bit for bit."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import events_from_dict, first_difference, random_runs

pytestmark = pytest.mark.gpu


def _params(rng):
    import ballasim_amd as bsx
    p = bsx.preset("ideal")
    p.reset_mode = "physical" if rng.random() < 0.5 else "ideal"
    p.refractory_s = float(rng.choice([0.0, 1e-5, 5e-5, 2e-4]))
    p.leak_rate_hz = float(rng.choice([0.0, 0.05, 0.5, 5.0, 40.0]))
    p.leak_sigma_log10 = float(rng.uniform(0, 0.3))
    p.theta_on = float(rng.uniform(0.08, 0.3))
    p.theta_off = float(rng.uniform(0.08, 0.3))
    p.sigma_theta_on = float(rng.uniform(0, 0.03))
    p.sigma_theta_off = float(rng.uniform(0, 0.03))
    p.dark_sigma_ln = float(rng.uniform(0, 0.4))
    p.eps_ci = float(rng.uniform(-0.003, 0.003))
    p.initial_phase = ["uniform", "symmetric", "zero"][int(rng.integers(0, 3))]
    if rng.random() < 0.4:
        p.hot_fraction = float(rng.uniform(0.01, 0.2))
        p.hot_rate_law = int(rng.integers(0, 2))
        p.hot_rate_min_hz = 20.0
        p.hot_rate_max_hz = min(3000.0, 0.5 / max(p.refractory_s, 1e-9))
    if rng.random() < 0.4:
        p.shot_rate_hz = float(rng.choice([10.0, 100.0, 400.0]))
    return p


@pytest.mark.parametrize("trial", range(24))
def test_random_configuration_matches_oracle(trial):
    import ballasim_amd as bsx
    rng = np.random.default_rng(5000 + trial)
    p = _params(rng)
    G = int(rng.integers(1, 4))
    W, H = 96, 64
    seed = int(rng.integers(0, 2 ** 40))
    gpu = bsx.SensorModel(p, W, H, seed, ["constant"] * G, 0, device="cuda")
    ref = O.OracleModel(p.c, W, H, seed, G, 0)
    irr0 = np.exp(rng.normal(3, 1, size=(W * H, G))).astype(np.float32)
    gpu.init(-1000, irr0)
    ref.init(-1000, irr0)
    t = -1000
    most = 0
    for call in range(4):
        t1 = t + int(rng.integers(2 ** 20, 3 * 2 ** 23))
        ids, ts, ir, offs = random_runs(rng, W * H, t, t1, G=G, max_n=int(rng.choice([12, 60, 200])))
        ev = gpu.advance(ids, ts, ir, t1, offsets=offs)
        want, rep = ref.advance_runs(ids, ts, ir, offs, t1)
        got = events_from_dict(ev, W)
        i, a, b = first_difference(got, want)
        assert len(got) == len(want) and i == len(got), f"call {call}: first difference at {i}: gpu {a} oracle {b}"
        assert np.array_equal(ev["t_reported_ns"], rep)
        most = max(most, len(want))
        t = t1
    ks, ko = gpu.stats(), ref.stats()
    if most > 1 << 20:
        assert ks["overflow_reruns"] >= 1, most
    for f in ("substeps", "pieces"):
        assert ks[f] == ko[f], (f, ks[f], ko[f])
