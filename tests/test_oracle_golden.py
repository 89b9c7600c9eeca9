"""The oracle (C restatement) and the product's host-side maps against the golden fixtures written by
tests/golden/make_golden.py from the reference's own CPU path (fp32): bit-exact events, reported
times, final states and maps."""
import ctypes as C
import json
import os

import numpy as np
import pytest

from ballasim_amd import _abi
from oracle import oracle as O
from tests.golden.cases import CASES, MAP_CASES, NOISE_STATE_FIELDS, STATE_FIELDS, noise_case

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
MANIFEST = json.load(open(os.path.join(GOLDEN, "manifest.json")))


def params_from(d):
    p = _abi.Params()
    for f, _ in _abi.Params._fields_:
        v = d[f]
        setattr(p, f, v.encode() if isinstance(v, str) else v)
    return p


@pytest.mark.parametrize("name", sorted(CASES))
def test_oracle_matches_reference_fixture(name):
    spec = MANIFEST["cases"][name]
    g = np.load(os.path.join(GOLDEN, f"case_{name}.npz"))
    p = params_from(spec["params"])
    m = O.OracleModel(p, spec["width"], spec["height"], spec["seed"], spec["groups"], spec["camera"], 2)
    m.init(spec["t0"], g["irr0"])
    for k in range(spec["calls"]):
        ev, rep = m.advance_runs(g[f"ids_{k}"], g[f"t_{k}"], g[f"irr_{k}"], g[f"offsets_{k}"], int(g[f"t_end_{k}"]))
        want = g[f"events_{k}"]
        assert len(ev) == len(want)
        assert np.array_equal(ev, want), f"call {k}"
        assert np.array_equal(rep, g[f"reported_{k}"])
    st = m.states()
    for f in STATE_FIELDS + (NOISE_STATE_FIELDS if noise_case(CASES[name]) else ()):
        assert np.array_equal(st[f], g[f"state_{f}"]), f
    got = m.stats()
    for k, v in zip(g["stat_names"], g["stats"]):
        assert got[str(k)] == int(v), (k, got[str(k)], int(v))


@pytest.mark.parametrize("name", sorted(MAP_CASES))
def test_maps_match_reference_fixture(name):
    """Both the oracle's and the product's (bsx_draw_maps, host) maps equal the reference's bit for bit."""
    spec = MANIFEST["maps"][name]
    g = np.load(os.path.join(GOLDEN, f"maps_{name}.npz"))
    p = params_from(spec["params"])
    n = spec["n"]
    lib = _abi.load_library()
    arrs = [np.zeros(n) for _ in range(5)]
    _abi.check(lib.bsx_draw_maps(C.byref(p), spec["seed"], spec["camera"], n,
                                 *[a.ctypes.data_as(C.c_void_p) for a in arrs]))
    for key, a in zip(("theta_on", "theta_off", "leak", "dark_fA", "phi0"), arrs):
        assert np.array_equal(a, g[key]), key
    om = O.draw_maps(p, spec["seed"], spec["camera"], n)
    for key in ("theta_on", "theta_off", "leak", "dark_fA", "phi0"):
        assert np.array_equal(om[key], g[key]), key


def test_fixture_maps_are_nontrivial():
    g = np.load(os.path.join(GOLDEN, "maps_prophesee_gen3.npz"))
    assert 0.15 < g["theta_on"].mean() < 0.35 and g["theta_on"].std() > 0.02
    assert np.all(g["phi0"] >= 0) and np.all(g["phi0"] < g["theta_on"])
    assert np.all(g["leak"] > 0)
