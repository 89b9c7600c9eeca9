"""GPU parity of the cascade with its in-path noise (the reference's default sensor, prophesee_gen3):
the HIP kernels through the C ABI against the oracle, which is itself bit-exact to the compiled
reference (test_noise_oracle.py) and to the golden fixtures. Events in every field, reported times,
the saved state (noise nodes, the cell's rate and the cell included) and the work counters (lattice
cells and bridge refinements included) must be identical, over calls that start and end inside
noise cells, through host runs, device runs, a rig and a saved-and-restored state."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import events_from_dict, first_difference, random_runs

pytestmark = pytest.mark.gpu

CASES = {
    "default": dict(),
    "ideal": dict(reset_mode="ideal", refractory_s=1e-5),
    "leak_loud": dict(leak_rate_hz=2.0, refractory_s=5e-5, sigma_o=0.2),
    "levels0": dict(noise_leaf_level=0),
    "levels8_eps": dict(noise_leaf_level=8, eps_ci=0.002, refractory_s=2e-5),
    "levels10_groups2": dict(noise_leaf_level=10, sigma_o=0.15),
}
STATE = ("z", "x1", "s2", "s3", "tv", "i_scale", "s3_ref", "n3_ref", "dr_ref", "t_ref", "t_search", "blind_until",
         "flags", "nk", "nk1", "ln_a1", "cell", "have_k1")
COUNTERS = ("substeps", "substep_attempts", "settled_substeps", "cells", "refinements", "leaf_intervals",
            "fallbacks", "quarantines", "pieces")


def _params(**kw):
    import ballasim_amd as bsx
    p = bsx.preset("prophesee_gen3")
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def _check(gpu, ref, ev, want, rep, W, where):
    got = events_from_dict(ev, W)
    i, a, b = first_difference(got, want)
    assert len(got) == len(want) and i == len(got), f"{where}: first difference at {i}: gpu {a} oracle {b}"
    assert np.array_equal(ev["t_reported_ns"], rep), where


def _check_state(gpu, ref, where):
    st = np.frombuffer(gpu.save_state()[56:], O.REF_STATE_DTYPE)
    so = ref.states()
    for f in STATE:
        assert np.array_equal(st[f], so[f]), f"{where}: state {f}"
    ks, ko = gpu.stats(), ref.stats()
    for f in COUNTERS:
        assert ks[f] == ko[f], (where, f, ks[f], ko[f])


@pytest.mark.parametrize("case", sorted(CASES))
def test_noise_matches_oracle(case):
    import ballasim_amd as bsx
    G = 2 if "groups2" in case else 1
    W, H, seed = 40, 24, 31 + len(case)
    p = _params(**CASES[case])
    rng = np.random.default_rng(sum(case.encode()))
    print(f"noise case {case}: {W}x{H}, seed {seed}", flush=True)
    gpu = bsx.SensorModel(p, W, H, seed, ["constant"] * G, 1, device="cuda")
    ref = O.OracleModel(p.c, W, H, seed, G, 1)
    irr0 = np.exp(rng.normal(3, 1.2, size=(W * H, G))).astype(np.float32)
    t = 2 ** 23 - 50_001
    gpu.init(t, irr0)
    ref.init(t, irr0)
    _check_state(gpu, ref, f"{case} init")
    total = 0
    for call in range(3):
        t1 = t + int(rng.integers(1, 2 ** 22))
        ids, ts, ir, offs = random_runs(rng, W * H, t, t1, G=G, log_sigma=1.5)
        ev = gpu.advance(ids, ts, ir, t1, offsets=offs)
        want, rep = ref.advance_runs(ids, ts, ir, offs, t1)
        _check(gpu, ref, ev, want, rep, W, f"{case} call {call}")
        _check_state(gpu, ref, f"{case} call {call}")
        total += len(want)
        t = t1
    assert total > 100
    assert ref.stats()["refinements"] > 0 or case == "levels0"


def test_noise_quiet_cells_and_device_runs():
    """Pixels whose light never changes (settled cells advanced in one step by the quiet-cell path),
    unlisted pixels (noise only) and runs given as device arrays."""
    import torch

    import ballasim_amd as bsx
    W, H, seed = 64, 32, 5
    P = W * H
    p = _params()
    rng = np.random.default_rng(9)
    gpu = bsx.SensorModel(p, W, H, seed, ["constant"], 0, device="cuda")
    ref = O.OracleModel(p.c, W, H, seed, 1, 0)
    irr0 = np.exp(rng.normal(3, 1, size=(P, 1))).astype(np.float32)
    gpu.init(0, irr0)
    ref.init(0, irr0)
    t = 0
    for call in range(3):
        t1 = t + 3 * 2 ** 21 + 999
        # a third listed with the init light repeated, a third with random light, a third unlisted
        ids = rng.permutation(P)[: 2 * P // 3].astype(np.uint32)
        ts, irs, offs = [], [], [0]
        for q in ids:
            tt = np.unique(rng.integers(t + 1, t1 + 1, size=5)).astype(np.int64)
            v = (np.repeat(irr0[q:q + 1], len(tt), axis=0) if q % 2 == 0
                 else np.exp(rng.normal(3, 1.0, size=(len(tt), 1))).astype(np.float32))
            ts.append(tt)
            irs.append(v.astype(np.float32))
            offs.append(offs[-1] + len(tt))
        tt, ir, offs = np.concatenate(ts), np.concatenate(irs), np.array(offs, np.int64)
        dev = (torch.from_numpy(ids.astype(np.int32)).cuda(), torch.from_numpy(offs[:-1].copy()).cuda(),
               torch.from_numpy(np.diff(offs).astype(np.int32)).cuda(), torch.from_numpy(tt).cuda(),
               torch.from_numpy(ir.reshape(-1).copy()).cuda())
        ptr, count = gpu.advance_device(*dev, t1)
        ev = gpu.download_events(ptr, count)
        want, rep = ref.advance_runs(ids, tt, ir, offs, t1)
        _check(gpu, ref, ev, want, rep, W, f"call {call}")
        _check_state(gpu, ref, f"call {call}")
        t = t1
    st = ref.stats()
    assert st["settled_substeps"] > 0.2 * st["cells"]


def test_noise_state_resume_mid_cell():
    """A state saved inside a noise cell and loaded into a fresh model continues exactly as the model
    that saved it (the cell's level-0 hull is rebuilt from the loaded nodes and rate)."""
    import ballasim_amd as bsx
    W, H, seed = 32, 16, 77
    p = _params(leak_rate_hz=0.5)
    rng = np.random.default_rng(4)
    a = bsx.SensorModel(p, W, H, seed, ["constant"], 0, device="cuda")
    irr0 = np.exp(rng.normal(3, 1, size=(W * H, 1))).astype(np.float32)
    a.init(1000, irr0)
    t1 = 1000 + 2 ** 21 + 12_345
    ids, ts, ir, offs = random_runs(rng, W * H, 1000, t1)
    a.advance(ids, ts, ir, t1, offsets=offs)
    blob = a.save_state()
    b = bsx.SensorModel(p, W, H, seed, ["constant"], 0, device="cuda")
    b.load_state(blob)
    t2 = t1 + 2 ** 22 + 4321
    ids, ts, ir, offs = random_runs(rng, W * H, t1, t2)
    ea = a.advance(ids, ts, ir, t2, offsets=offs)
    eb = b.advance(ids, ts, ir, t2, offsets=offs)
    for k in ea:
        assert np.array_equal(ea[k], eb[k]), k
    assert len(ea["t_cross_ns"]) > 20
    assert a.save_state() == b.save_state()


def test_noise_rig_equals_solo():
    """A rig of noise-on cameras (one launch) equals each camera advanced alone."""
    import ballasim_amd as bsx
    from ballasim_amd.sensor import Rig
    p = _params()
    shapes = [(40, 24), (24, 16), (16, 9)]
    rng = np.random.default_rng(12)
    cams, solo, irr0s = [], [], []
    for c, (W, H) in enumerate(shapes):
        irr0 = np.exp(rng.normal(3, 1, size=(W * H, 1))).astype(np.float32)
        irr0s.append(irr0)
        for lst in (cams, solo):
            m = bsx.SensorModel(p, W, H, 3 + c, ["constant"], c, device="cuda")
            m.init(-12_345, irr0)
            lst.append(m)
    import torch
    rig = Rig(cams)
    t = -12_345
    for call in range(2):
        t1 = t + 2 ** 21 + 77
        runs = [random_runs(rng, W * H, t, t1) for (W, H) in shapes]
        dev = [(torch.from_numpy(ids.astype(np.int32)).cuda(), torch.from_numpy(offs[:-1].copy()).cuda(),
                torch.from_numpy(np.diff(offs).astype(np.int32)).cuda(), torch.from_numpy(ts).cuda(),
                torch.from_numpy(ir.reshape(-1).copy()).cuda()) for ids, ts, ir, offs in runs]
        out = rig.advance_device(dev, t1)
        for c, (ids, ts, ir, offs) in enumerate(runs):
            got = cams[c].download_events(*out[c])
            want = solo[c].advance(ids, ts, ir, t1, offsets=offs)
            for k in want:
                assert np.array_equal(got[k], want[k]), (call, c, k)
        t = t1
    for c in range(len(shapes)):
        assert cams[c].save_state() == solo[c].save_state()
