"""GPU parity of the photoreceptor cascade (noise off) and a GPU replay of every golden fixture.

- The HIP cascade kernel through the C ABI against the oracle (itself bit-exact to the compiled
  reference: test_oracle_vs_ref.py) on seeded random runs: events (pixel, polarity, int64 t_cross,
  label, flags, fp32 increments), reported times, the full fp32 state and the work counters.
- The fixtures written from the reference's own CPU path (tests/golden) replayed on the GPU: events,
  reported times and states must equal the reference's bit for bit (instantaneous and cascade)."""
import json
import os

import numpy as np
import pytest

from oracle import oracle as O
from tests.golden.cases import CASES, NOISE_STATE_FIELDS, STATE_FIELDS, noise_case
from tests.helpers import events_from_dict, first_difference, random_runs

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
MANIFEST = json.load(open(os.path.join(GOLDEN, "manifest.json")))

CASCADE_CASES = {
    "prophesee_physical": ("prophesee_gen3", dict(noise=False)),
    "samsung_ideal_leak": ("samsung_gen3", dict(noise=False, reset_mode="ideal", refractory_s=2e-5, leak_rate_hz=2.0)),
    "prophesee_zero_refractory": ("prophesee_gen3", dict(noise=False, refractory_s=0.0, eps_ci=0.003,
                                                        leak_rate_hz=0.5, theta_on=0.12, theta_off=0.14)),
    "samsung_fast_light_groups2": ("samsung_gen3", dict(noise=False, illuminance_scale=40.0)),
    # leak-type hot pixels through the cascade's comparator (HOT labels and the hot-pixel flag)
    "prophesee_hot": ("prophesee_gen3", dict(noise=False, hot_fraction=0.25, hot_rate_law=1, hot_rate_median_hz=300.0,
                                             hot_rate_min_hz=20.0, hot_rate_max_hz=2000.0)),
}


def _params(case):
    import ballasim_amd as bsx
    family, over = CASCADE_CASES[case]
    p = bsx.preset(family)
    for k, v in over.items():
        setattr(p, k, v)
    return p


@pytest.mark.parametrize("case", sorted(CASCADE_CASES))
def test_cascade_matches_oracle(case):
    import ballasim_amd as bsx
    G = 2 if "groups2" in case else 1
    W, H, seed = 40, 24, 17
    rng = np.random.default_rng(sum(case.encode()) + 1)
    p = _params(case)
    gpu = bsx.SensorModel(p, W, H, seed, ["constant"] * G, 2, device="cuda")
    ref = O.OracleModel(p.c, W, H, seed, G, 2)
    irr0 = np.exp(rng.normal(2.5, 1.2, size=(W * H, G))).astype(np.float32)
    t = 3 * 2 ** 23 - 12345
    gpu.init(t, irr0)
    ref.init(t, irr0)
    total = n_hot = 0
    for call in range(3):
        t1 = t + int(rng.integers(1, 2 * 2 ** 23))
        ids, ts, ir, offs = random_runs(rng, W * H, t, t1, G=G, log_sigma=1.8)
        ev = gpu.advance(ids, ts, ir, t1, offsets=offs)
        want, rep = ref.advance_runs(ids, ts, ir, offs, t1)
        got = events_from_dict(ev, W)
        i, a, b = first_difference(got, want)
        assert len(got) == len(want) and i == len(got), f"call {call}: first difference at {i}: gpu {a} oracle {b}"
        assert np.array_equal(ev["t_reported_ns"], rep)
        total += len(got)
        n_hot += int(np.count_nonzero(want["flags"] & (1 << 5)))
        t = t1
    assert total > 200
    if "hot" in case:
        assert n_hot > 50
    st = np.frombuffer(gpu.save_state()[56:], O.REF_STATE_DTYPE)
    os_ = ref.states()
    for f in ("z", "x1", "s2", "s3", "tv", "i_scale", "s3_ref", "dr_ref", "t_ref", "t_search", "blind_until",
              "flags"):
        assert np.array_equal(st[f], os_[f]), f
    ks, ko = gpu.stats(), ref.stats()
    for f in ("substeps", "substep_attempts", "settled_substeps", "leaf_intervals", "fallbacks", "quarantines",
              "pieces"):
        assert ks[f] == ko[f], (f, ks[f], ko[f])


@pytest.mark.parametrize("name", sorted(CASES))
def test_golden_fixture_replayed_on_gpu(name):
    import ballasim_amd as bsx
    from ballasim_amd import _abi
    spec = MANIFEST["cases"][name]
    g = np.load(os.path.join(GOLDEN, f"case_{name}.npz"))
    c = _abi.Params()
    for f, _ in _abi.Params._fields_:
        v = spec["params"][f]
        setattr(c, f, v.encode() if isinstance(v, str) else v)
    p = bsx.SensorParams(c)
    W, H, G = spec["width"], spec["height"], spec["groups"]
    gpu = bsx.SensorModel(p, W, H, spec["seed"], ["constant"] * G, spec["camera"], device="cuda")
    gpu.init(spec["t0"], g["irr0"])
    for k in range(spec["calls"]):
        ev = gpu.advance(g[f"ids_{k}"], g[f"t_{k}"], g[f"irr_{k}"], int(g[f"t_end_{k}"]), offsets=g[f"offsets_{k}"])
        got = events_from_dict(ev, W)
        want = g[f"events_{k}"]
        i, a, b = first_difference(got, want)
        assert len(got) == len(want) and i == len(got), f"call {k}: first difference at {i}: gpu {a} reference {b}"
        assert np.array_equal(ev["t_reported_ns"], g[f"reported_{k}"])
    st = np.frombuffer(gpu.save_state()[56:], O.REF_STATE_DTYPE)
    for f in STATE_FIELDS + (NOISE_STATE_FIELDS if noise_case(CASES[name]) else ()):
        assert np.array_equal(st[f], g[f"state_{f}"]), f
    got = gpu.stats()
    instant = p.photoreceptor == "instantaneous"
    for k, v in zip(g["stat_names"], g["stats"]):
        if instant and str(k) == "leaf_intervals":
            continue  # the instantaneous path counts its own searches' evaluations
        assert got[str(k)] == int(v), (k, got[str(k)], int(v))
