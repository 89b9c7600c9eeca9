"""The oracle against the compiled reference itself (oracle/_ref, built from the reference's sources
where they are present; skipped elsewhere): randomized parameter sets and inputs, bit-exact events,
reported times, maps and states; the presets."""
import ctypes as C

import numpy as np
import pytest

from ballasim_amd import _abi
from oracle import oracle as O
from tests.helpers import random_runs

pytestmark = pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref (the compiled reference) is not built")


def preset(name):
    lib = _abi.load_library()
    p = _abi.Params()
    _abi.check(lib.bsx_params_preset(name.encode(), C.byref(p)))
    return p


@pytest.mark.parametrize("name", ["prophesee_gen3", "samsung_gen3", "ideal"])
def test_presets_equal_reference(name):
    a = preset(name)
    b = _abi.Params()
    assert O.ref_lib().ref_preset(name.encode(), C.byref(b)) == 0
    for f, _ in _abi.Params._fields_:
        assert getattr(a, f) == getattr(b, f), f


def _random_instant_params(rng):
    p = preset("ideal")
    p.reset_mode = int(rng.integers(0, 2))
    p.refractory_s = float(rng.choice([0.0, 1e-5, 1e-4, 3e-4]))
    p.leak_rate_hz = float(rng.choice([0.0, 0.1, 3.0, 30.0]))
    p.leak_sigma_log10 = float(rng.uniform(0, 0.3))
    p.theta_on = float(rng.uniform(0.08, 0.4))
    p.theta_off = float(rng.uniform(0.08, 0.4))
    p.sigma_theta_on = float(rng.uniform(0, 0.04))
    p.sigma_theta_off = float(rng.uniform(0, 0.04))
    p.theta_onoff_corr = float(rng.uniform(-1, 1))
    p.dark_current_fA = float(np.exp(rng.uniform(-4, 2)))
    p.dark_sigma_ln = float(rng.uniform(0, 0.5))
    p.initial_phase = int(rng.integers(0, 3))
    p.eps_ci = float(rng.uniform(-0.004, 0.004))
    p.label_share_cut = float(rng.uniform(0.3, 0.7))
    p.illuminance_scale = float(np.exp(rng.uniform(-2, 3)))
    return p


@pytest.mark.parametrize("trial", range(6))
def test_instant_random_sweep(trial):
    rng = np.random.default_rng(1000 + trial)
    p = _random_instant_params(rng)
    W, H, G = 20, 12, int(rng.integers(1, 4))
    seed = int(rng.integers(0, 2 ** 63))
    a = O.OracleModel(p, W, H, seed, G, 3, 2)
    b = O.RefModel(p, W, H, seed, G, 3, 2, True)
    irr0 = np.exp(rng.normal(2, 1.5, (W * H, G))).astype(np.float32)
    t0 = int(rng.integers(-10 ** 9, 10 ** 9))
    a.init(t0, irr0)
    b.init(t0, irr0)
    t = t0
    for _ in range(3):
        t1 = t + int(rng.integers(1, 3 * 2 ** 23))
        ids, ts, ir, offs = random_runs(rng, W * H, t, t1, G=G, log_sigma=2.0)
        ea, ra = a.advance_runs(ids, ts, ir, offs, t1)
        eb, rb = b.advance_runs(ids, ts, ir, offs, t1)
        assert np.array_equal(ea, eb) and np.array_equal(ra, rb)
        t = t1
    sa, sb = a.states(), b.states()
    for f in ("s3_ref", "dr_ref", "t_ref", "t_search", "blind_until", "i_scale", "flags"):
        assert np.array_equal(sa[f], sb[f]), f


def _random_cascade_params(rng, family):
    p = preset(family)
    p.noise = 0
    p.reset_mode = int(rng.integers(0, 2))
    p.refractory_s = float(rng.choice([0.0, 2e-5, 1e-4]))
    p.leak_rate_hz = float(rng.choice([0.0, 0.1, 3.0]))
    p.theta_on = float(rng.uniform(0.1, 0.35))
    p.theta_off = float(rng.uniform(0.1, 0.35))
    p.eps_ci = float(rng.uniform(-0.004, 0.004))
    p.illuminance_scale = float(np.exp(rng.uniform(-1, 3)))
    return p


@pytest.mark.parametrize("trial", range(6))
def test_cascade_random_sweep(trial):
    """The cascade without noise, bit for bit: events, reported times, the full state and the work
    counters (sub-steps, attempts, settled steps, leaf intervals, pieces)."""
    rng = np.random.default_rng(3000 + trial)
    p = _random_cascade_params(rng, ["prophesee_gen3", "samsung_gen3"][trial % 2])
    W, H, G = 14, 9, int(rng.integers(1, 3))
    seed = int(rng.integers(0, 2 ** 63))
    a = O.OracleModel(p, W, H, seed, G, 4, 2)
    b = O.RefModel(p, W, H, seed, G, 4, 2, True)
    irr0 = np.exp(rng.normal(2, 1.5, (W * H, G))).astype(np.float32)
    t0 = int(rng.integers(-10 ** 9, 10 ** 9))
    a.init(t0, irr0)
    b.init(t0, irr0)
    t = t0
    for _ in range(3):
        t1 = t + int(rng.integers(1, 3 * 2 ** 23))
        ids, ts, ir, offs = random_runs(rng, W * H, t, t1, G=G, log_sigma=2.0)
        ea, ra = a.advance_runs(ids, ts, ir, offs, t1)
        eb, rb = b.advance_runs(ids, ts, ir, offs, t1)
        assert np.array_equal(ea, eb) and np.array_equal(ra, rb)
        t = t1
    sa, sb = a.states(), b.states()
    for f in ("z", "x1", "s2", "s3", "tv", "i_scale", "s3_ref", "dr_ref", "t_ref", "t_search", "blind_until", "flags"):
        assert np.array_equal(sa[f], sb[f]), f
    ka, kb = a.stats(), b.stats()
    for f in ("substeps", "substep_attempts", "settled_substeps", "leaf_intervals", "fallbacks", "quarantines",
              "pieces"):
        assert ka[f] == kb[f], f
    assert ka["substeps"] > 0


@pytest.mark.parametrize("name", ["prophesee_gen3", "samsung_gen3"])
def test_maps_random_seeds(name):
    rng = np.random.default_rng(5)
    for _ in range(3):
        p = preset(name)
        p.theta_onoff_corr = float(rng.uniform(-1, 1))
        p.initial_phase = int(rng.integers(0, 3))
        p.temperature_K = float(rng.uniform(280, 320))
        seed, cam = int(rng.integers(0, 2 ** 63)), int(rng.integers(0, 256))
        ours = O.draw_maps(p, seed, cam, 3000)
        p.noise = 0
        ref = O.RefModel(p, 3000, 1, seed, 1, cam, 1, True).maps()
        for k in ours:
            assert np.array_equal(ours[k], ref[k]), k


def _random_hot(rng, p):
    """Hot pixels on (a large fraction, so that every law and reset shows up in small frames)."""
    p.noise = 0
    p.hot_fraction = float(rng.uniform(0.05, 0.4))
    p.hot_rate_law = int(rng.integers(0, 2))
    p.hot_rate_min_hz = float(np.exp(rng.uniform(0, 3)))
    p.hot_rate_max_hz = p.hot_rate_min_hz * float(np.exp(rng.uniform(0.5, 4)))
    p.hot_rate_median_hz = float(np.exp(rng.uniform(1, 5)))
    p.hot_rate_sigma_ln = float(rng.uniform(0, 2))
    p.hot_leak_fraction = float(rng.uniform(0, 1))
    if p.hot_rate_max_hz * p.refractory_s >= 1.0:
        p.refractory_s = 0.5 / p.hot_rate_max_hz
    return p


@pytest.mark.parametrize("trial", range(4))
def test_hot_maps_random_seeds(trial):
    """Hot pixels (maps.cpp:106-124, without in-path noise: all leak-type): the flags, target rates and
    the raised leaks equal the reference's, under both rate laws and both resets."""
    rng = np.random.default_rng(7000 + trial)
    p = _random_hot(rng, preset(["prophesee_gen3", "ideal"][trial % 2]))
    p.reset_mode = trial // 2
    seed, cam = int(rng.integers(0, 2 ** 63)), int(rng.integers(0, 256))
    ours = O.draw_maps(p, seed, cam, 3000)
    ref = O.RefModel(p, 3000, 1, seed, 1, cam, 1, True).maps()
    for k in ours:
        assert np.array_equal(ours[k], ref[k]), k
    assert 0 < ours["hot"].sum() < 3000
    assert np.all((ours["hot_rate_hz"] > 0) == (ours["hot"] == 1))


@pytest.mark.parametrize("trial", range(4))
def test_hot_instant_random_sweep(trial):
    """Events of hot pixels: HOT labels (unless SIGNAL) and the hot-pixel flag, bit for bit."""
    rng = np.random.default_rng(8000 + trial)
    p = _random_hot(rng, _random_instant_params(rng))
    W, H, G = 16, 10, 1
    seed = int(rng.integers(0, 2 ** 63))
    a = O.OracleModel(p, W, H, seed, G, 3, 2)
    b = O.RefModel(p, W, H, seed, G, 3, 2, True)
    irr0 = np.exp(rng.normal(2, 1.0, (W * H, G))).astype(np.float32)
    t0 = int(rng.integers(-10 ** 9, 10 ** 9))
    a.init(t0, irr0)
    b.init(t0, irr0)
    t = t0
    n_hot = 0
    for _ in range(3):
        t1 = t + int(rng.integers(2 ** 22, 3 * 2 ** 23))
        ids, ts, ir, offs = random_runs(rng, W * H, t, t1, G=G, log_sigma=1.0)
        ea, ra = a.advance_runs(ids, ts, ir, offs, t1)
        eb, rb = b.advance_runs(ids, ts, ir, offs, t1)
        assert np.array_equal(ea, eb) and np.array_equal(ra, rb)
        n_hot += int(np.count_nonzero(ea["flags"] & (1 << 5)))
        t = t1
    assert n_hot > 0
