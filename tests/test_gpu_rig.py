"""GPU tests of the multi-camera rig: one launch chain (phase A over every camera's segments, one
heavy-first order of the joint pixel space, one walk launch, captured into a hipGraph) must give
every camera exactly what advancing it alone gives: the same events bit for bit, the same saved state
and the same work counters. Cameras of mixed resolutions, ragged runs with heavy-tailed sample counts,
cameras at different model times, event-buffer overflow, graph on and off, both photoreceptor kinds,
the autotuner (it changes no camera) and the refusals (before any camera changes)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

WORK = ("substeps", "substep_attempts", "settled_substeps", "leaf_intervals", "fallbacks", "quarantines", "pieces",
        "events", "input_clamps")


def _params(kind, **over):
    import ballasim_amd as bsx
    if kind == "instant":
        p = bsx.preset("prophesee_gen3")
        p.photoreceptor = "instantaneous"
        p.noise = False
        p.leak_rate_hz = 0.5
        p.refractory_s = 5e-5
    else:
        p = bsx.preset("samsung_gen3")
        p.noise = False
    for k, v in over.items():
        setattr(p, k, v)
    return p


def heavy_runs(rng, P, t0, t1, G=1, mu=np.log(6.0), sigma=1.5, cap=300, frac=0.85):
    """Ragged runs with log-normal sample counts (clipped to [1, cap]) for a random subset of pixels in
    random order: strictly increasing times in (t0, t1], log-normal irradiance drifting per pixel."""
    ids = np.nonzero(rng.random(P) < frac)[0].astype(np.uint32)
    rng.shuffle(ids)
    counts = np.clip(np.round(np.exp(rng.normal(mu, sigma, size=len(ids)))), 1, cap).astype(np.int64)
    counts = np.minimum(counts, t1 - t0)
    ts, irs = [], []
    for n in counts:
        tt = np.sort(rng.choice(np.arange(t0 + 1, t1 + 1, max(1, (t1 - t0) // 4096)), size=int(n), replace=False))
        ts.append(tt.astype(np.int64))
        walk = np.cumsum(rng.normal(0.0, 0.35, size=(int(n), G)), axis=0) + rng.normal(3.0, 1.0)
        irs.append(np.exp(walk).astype(np.float32))
    offs = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
    t = np.concatenate(ts) if ts else np.zeros(0, np.int64)
    irr = np.concatenate(irs).reshape(-1, G) if irs else np.zeros((0, G), np.float32)
    return ids, t, irr, offs


def to_device(ids, t, irr, offs):
    import torch
    dev = torch.device("cuda:0")
    return (torch.from_numpy(ids.astype(np.int32)).to(dev), torch.from_numpy(offs[:-1].copy()).to(dev),
            torch.from_numpy(np.diff(offs).astype(np.int32)).to(dev), torch.from_numpy(t).to(dev),
            torch.from_numpy(np.ascontiguousarray(irr.reshape(-1))).to(dev))


SHAPES = [(40, 24), (23, 17), (64, 9), (31, 31)]


def make_cameras(kind, shapes, seed=5, G=1, **over):
    import ballasim_amd as bsx
    cams = []
    for c, (W, H) in enumerate(shapes):
        extra = dict(over)
        if kind == "instant" and c == 1:
            extra.setdefault("shot_rate_hz", 40.0)  # one camera with the shot process
        cams.append(bsx.SensorModel(_params(kind, **extra), W, H, seed + c, ["constant"] * G, c, device="cuda"))
    return cams


def run_pair(kind, shapes, calls=3, graph=True, G=1, t0s=None, caps=None, tuning=None, seed=11):
    """Advances a rig and an identical set of cameras advanced one by one on the same device runs;
    asserts equality after every call."""
    import torch

    import ballasim_amd as bsx
    rng = np.random.default_rng(seed)
    rig_cams = make_cameras(kind, shapes, G=G)
    solo = make_cameras(kind, shapes, G=G)
    t0s = t0s or [0] * len(shapes)
    for c, (W, H) in enumerate(shapes):
        irr0 = np.exp(rng.normal(3, 1, (W * H, G))).astype(np.float32)
        rig_cams[c].init(t0s[c], irr0)
        solo[c].init(t0s[c], irr0)
        if caps is not None and caps[c]:
            rig_cams[c].set_launch_policy(event_capacity=caps[c])
            solo[c].set_launch_policy(event_capacity=caps[c])
    rig = bsx.Rig(rig_cams)
    rig.set_graph(graph)
    if tuning:
        rig.set_tuning(**tuning)
    t_end = max(t0s)
    for call in range(calls):
        t_end += int(rng.integers(2 ** 21, 2 * 2 ** 23))
        runs, keep = [], []
        for c, (W, H) in enumerate(shapes):
            t_from = rig_cams[c].time_ns
            r = heavy_runs(rng, W * H, t_from, t_end, G=G)
            d = to_device(*r)
            keep.append(d)
            runs.append(d)
        torch.cuda.synchronize()
        got = rig.advance_device(runs, t_end)
        for c in range(len(shapes)):
            ptr, count = solo[c].advance_device(*runs[c], t_end)
            want = solo[c].download_events(ptr, count)
            have = rig_cams[c].download_events(*got[c])
            assert got[c][1] == count, f"call {call} camera {c}: {got[c][1]} events, alone {count}"
            for k in want:
                assert np.array_equal(have[k], want[k]), f"call {call} camera {c}: {k}"
            assert rig_cams[c].save_state() == solo[c].save_state(), f"call {call} camera {c}: state"
    for c in range(len(shapes)):
        a, b = rig_cams[c].stats(), solo[c].stats()
        for k in WORK:
            assert a[k] == b[k], (c, k, a[k], b[k])
    return rig, rig_cams


def test_rig_instant_equals_cameras_alone_graph():
    rig, cams = run_pair("instant", SHAPES)
    st = rig.stats()
    assert st["launches"] == 3 and st["graph_launches"] == 3 and st["captures"] >= 1
    assert sum(c.stats()["events"] for c in cams) > 500


def test_rig_instant_without_graph_and_persistent_queue():
    rig, _ = run_pair("instant", SHAPES, graph=False, tuning=dict(queue_chunk=2, block_size=128, walk_budget=16,
                                                                  search_batch=4))
    assert rig.stats()["graph_launches"] == 0


@pytest.mark.parametrize("graph,tuning", [
    (True, dict(head_lanes=16, head_per_mille=250)),
    (False, dict(head_lanes=8, head_per_mille=500, queue_chunk=2, block_size=128)),
])
def test_rig_instant_head_groups(graph, tuning):
    """The heaviest pixels of the order walked in groups of fewer than 64 (idle lanes in those waves)
    give every camera the events, state and counters of advancing it alone in 64-pixel groups."""
    rig, _ = run_pair("instant", SHAPES, graph=graph, tuning=tuning)
    got = rig.tuning()
    assert got["head_lanes"] == tuning["head_lanes"] and got["head_per_mille"] == tuning["head_per_mille"]
    rig.set_tuning(head_per_mille=-1)
    assert rig.tuning()["head_per_mille"] == 0


def test_rig_cascade_head_groups():
    run_pair("cascade", [(24, 16), (17, 9)], calls=2, tuning=dict(head_lanes=32, head_per_mille=300))


def test_rig_cameras_at_different_times_and_overflow():
    rig, _ = run_pair("instant", SHAPES[:3], t0s=[0, 2 ** 23 + 777, 3 * 2 ** 22 - 5], caps=[32, 0, 16],
                      tuning=dict(queue_chunk=1))
    assert rig.stats()["walk_reruns"] >= 1


def test_rig_two_light_groups():
    run_pair("instant", SHAPES[:2], G=2, calls=2)


def test_rig_cascade_equals_cameras_alone():
    rig, cams = run_pair("cascade", [(24, 16), (17, 9)], calls=2)
    assert sum(c.stats()["events"] for c in cams) > 50


def test_rig_autotune_changes_no_camera():
    import torch

    import ballasim_amd as bsx
    rng = np.random.default_rng(3)
    cams = make_cameras("instant", SHAPES)
    for c, (W, H) in enumerate(SHAPES):
        cams[c].init(0, np.exp(rng.normal(3, 1, (W * H, 1))).astype(np.float32))
    rig = bsx.Rig(cams)
    runs = [to_device(*heavy_runs(rng, W * H, 0, 2 ** 23)) for (W, H) in SHAPES]
    torch.cuda.synchronize()
    before = [c.save_state() for c in cams]
    trials = rig.autotune(runs, 2 ** 23, reps=1)
    assert len(trials) >= 10
    best = min(ms for _, ms in trials)
    tu = rig.tuning()
    assert any(t == tu and ms == best for t, ms in trials)
    assert [c.save_state() for c in cams] == before
    assert all(c.time_ns == 0 for c in cams)
    # the tuned rig still equals the cameras alone
    solo = make_cameras("instant", SHAPES)
    rng = np.random.default_rng(3)
    for c, (W, H) in enumerate(SHAPES):
        solo[c].init(0, np.exp(rng.normal(3, 1, (W * H, 1))).astype(np.float32))
    got = rig.advance_device(runs, 2 ** 23)
    for c in range(len(SHAPES)):
        ptr, count = solo[c].advance_device(*runs[c], 2 ** 23)
        want = solo[c].download_events(ptr, count)
        have = cams[c].download_events(*got[c])
        for k in want:
            assert np.array_equal(have[k], want[k]), (c, k)
        assert cams[c].save_state() == solo[c].save_state()


def test_rig_refusals_change_no_camera():
    import torch

    import ballasim_amd as bsx
    rng = np.random.default_rng(4)
    cams = make_cameras("instant", SHAPES[:3])
    for c, (W, H) in enumerate(SHAPES[:3]):
        cams[c].init(0, np.exp(rng.normal(3, 1, (W * H, 1))).astype(np.float32))
    rig = bsx.Rig(cams)
    runs = [heavy_runs(rng, W * H, 0, 2 ** 23) for (W, H) in SHAPES[:3]]
    bad = list(runs[1])
    bad[0] = bad[0].copy()
    bad[0][0] = SHAPES[1][0] * SHAPES[1][1] + 5  # pixel id out of range in camera 1
    dev = [to_device(*runs[0]), to_device(*bad), to_device(*runs[2])]
    torch.cuda.synchronize()
    before = [c.save_state() for c in cams]
    with pytest.raises(ValueError, match="camera 1"):
        rig.advance_device(dev, 2 ** 23)
    assert [c.save_state() for c in cams] == before
    nan = list(runs[2])
    nan[2] = nan[2].copy()
    nan[2][len(nan[2]) // 2] = np.nan  # a non-finite sample in camera 2: found by phase A, after the launch
    dev = [to_device(*runs[0]), to_device(*runs[1]), to_device(*nan)]
    torch.cuda.synchronize()
    with pytest.raises(ValueError, match="camera 2.*non-finite"):
        rig.advance_device(dev, 2 ** 23)
    assert [c.save_state() for c in cams] == before and all(c.initialized for c in cams)
    dev = [to_device(*r) for r in runs]
    rig.advance_device(dev, 2 ** 23)
    assert all(c.time_ns == 2 ** 23 for c in cams)
    with pytest.raises(ValueError):
        bsx.Rig([cams[0], cams[0]])
    other = bsx.SensorModel(_params("cascade"), 8, 8, 1, device="cuda")
    with pytest.raises(ValueError):
        bsx.Rig([cams[0], other])
    with pytest.raises(ValueError):
        rig.set_tuning(block_size=96)
