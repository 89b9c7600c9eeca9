"""The device's noise lane program (ballasim_amd/csrc/sensor/noise.h: cells, lattice draws, quiet
cells, the path-memory walk of the bridge tree, n3 at releases, init inside a cell), compiled for the
host under the device build's floating-point policy, against the oracle: events in every field,
the fp32 state (noise nodes, rate, cell included) and the work counters, bit for bit, over calls
that start and end inside cells."""
import numpy as np
import pytest

from oracle import oracle as O
from tests import noise_host as NH
from tests.helpers import first_difference, random_runs

CASES = {
    "default": dict(),
    "ideal": dict(reset_mode="ideal", refractory_s=1e-5),
    "leak_loud": dict(leak_rate_hz=2.0, refractory_s=5e-5, sigma_o=0.2),
    "levels0": dict(noise_leaf_level=0),
    "levels8_eps": dict(noise_leaf_level=8, eps_ci=0.002, refractory_s=2e-5),
    "levels10": dict(noise_leaf_level=10, sigma_o=0.15),
}


@pytest.fixture(scope="module")
def lane(tmp_path_factory):
    L = NH.lib(str(tmp_path_factory.mktemp("noise_lane")))
    if L is None:
        pytest.skip("g++ not found")
    return L


def _params(**kw):
    import ballasim_amd as bsx
    p = bsx.preset("prophesee_gen3")
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def _static_runs(rng, P, t0, t1, irr0, n=6):
    ids = np.arange(P, dtype=np.uint32)
    ts, irs, offs = [], [], [0]
    for q in ids:
        tt = np.unique(rng.integers(t0 + 1, t1 + 1, size=n)).astype(np.int64)
        v = np.repeat(irr0[q:q + 1], len(tt), axis=0) if q < P // 2 else \
            np.exp(rng.normal(3, 1.0, size=(len(tt), 1))).astype(np.float32)
        ts.append(tt)
        irs.append(v.astype(np.float32))
        offs.append(offs[-1] + len(tt))
    return ids, np.concatenate(ts), np.concatenate(irs), np.array(offs, np.int64)


@pytest.mark.parametrize("case", sorted(CASES) + ["quiet_cells", "init_on_node"])
def test_noise_lane_equals_oracle(lane, case):
    W, H, seed = 18, 12, 11 + len(case)
    p = _params(**CASES.get(case, {}))
    rng = np.random.default_rng(100 + len(case))
    host = NH.HostCamera(lane, p, W, H, seed)
    o = O.OracleModel(p.c, W, H, seed, 1, 0, 2)
    irr0 = np.exp(rng.normal(3, 1, size=(W * H, 1))).astype(np.float32)
    t0 = 3 * 65536 if case == "init_on_node" else -70_001
    o.init(t0, irr0)
    host.init(t0, irr0)
    NH.compare_states(host, o.states(), f"{case} init")
    t = t0
    events = 0
    for w in range(3):
        t1 = (w + 1) * (2 ** 21) + (0 if case == "quiet_cells" else 777 * w + 13)
        if case == "quiet_cells":
            ids, ts, irr, offs = _static_runs(rng, W * H, t, t1, irr0)
        else:
            ids, ts, irr, offs = random_runs(rng, W * H, t, t1, max_n=10)
        a, _ = o.advance_runs(ids, ts, irr, offs, t1)
        b = host.advance_runs(ids, ts, irr, offs, t1)
        i, x, y = first_difference(a, b)
        assert len(a) == len(b) and x is None, f"{case} window {w}: events differ at {i} of {len(a)}/{len(b)}: {x} / {y}"
        NH.compare_states(host, o.states(), f"{case} window {w}")
        so = o.stats()
        sh = host.stats()
        for k in NH.COUNTERS:
            assert sh[k] == so[k], f"{case} window {w}: counter {k} {sh[k]} vs {so[k]}"
        events += len(a)
        t = t1
    assert events > 30
