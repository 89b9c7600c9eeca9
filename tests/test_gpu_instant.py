"""GPU parity of the instantaneous log response: the HIP kernel through the C ABI against the oracle
(the C restatement, itself pinned bit for bit to the compiled reference by test_oracle_vs_ref.py),
on the same seeded inputs. Integer and event outputs must be bit-exact: event count, pixel,
polarity, t_cross (int64 ns), label, flags, the fp32 increments and the reported times."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import events_from_dict, first_difference, random_runs

pytestmark = pytest.mark.gpu

CASES = {
    "ideal": dict(),
    "physical_leak": dict(reset_mode="physical", refractory_s=1e-4, leak_rate_hz=0.5, sigma_theta_on=0.03,
                          sigma_theta_off=0.02, dark_sigma_ln=0.3),
    "groups3_fast_leak": dict(leak_rate_hz=5.0, leak_sigma_log10=0.2, refractory_s=3e-5, theta_on=0.1,
                              theta_off=0.12, eps_ci=0.001),
    "symmetric_readout": dict(reset_mode="physical", refractory_s=0.0, initial_phase="symmetric", eps_ci=-0.002,
                              readout_quantum_s=1e-6, readout_delay_s=5e-6, clock_drift_ppm=3.0),
    # the rate-specified shot-noise process, at rates far above the physical ones so that a few calls
    # hold thousands of arrivals (inside segments, at segment ends, in dead time)
    "shot_physical": dict(reset_mode="physical", refractory_s=2e-5, leak_rate_hz=0.5, shot_rate_hz=400.0),
    "shot_ideal_leak": dict(refractory_s=1e-5, leak_rate_hz=3.0, shot_rate_hz=150.0, sigma_theta_on=0.02),
    # hot pixels (leak-type, raised leak): HOT labels and the hot-pixel flag, with shot arrivals too
    "hot_physical": dict(reset_mode="physical", refractory_s=5e-5, leak_rate_hz=0.2, hot_fraction=0.15,
                         hot_rate_law=1, hot_rate_min_hz=20.0, hot_rate_max_hz=5000.0, hot_rate_median_hz=300.0,
                         hot_rate_sigma_ln=1.0, shot_rate_hz=50.0),
    "hot_ideal_loguniform": dict(refractory_s=1e-5, hot_fraction=0.3, hot_rate_min_hz=50.0, hot_rate_max_hz=4000.0),
}


def _params(case):
    import ballasim_amd as bsx
    p = bsx.preset("ideal")
    for k, v in CASES[case].items():
        setattr(p, k, v)
    return p


@pytest.mark.parametrize("case", sorted(CASES))
def test_instant_matches_oracle(case):
    import ballasim_amd as bsx
    G = 3 if case == "groups3_fast_leak" else 1
    W, H, seed = 48, 32, 11
    rng = np.random.default_rng(sum(case.encode()))
    p = _params(case)
    gpu = bsx.SensorModel(p, W, H, seed, ["constant"] * G, 0, device="cuda")
    ref = O.OracleModel(p.c, W, H, seed, G, 0)
    maps = gpu.sensor_maps()
    om = ref.maps()
    assert np.array_equal(maps["theta_on"].reshape(-1), om["theta_on"])
    assert np.array_equal(maps["phi0"].reshape(-1), om["phi0"])
    assert np.array_equal(maps["leak"].reshape(-1), om["leak"])
    assert np.array_equal(maps["hot"].reshape(-1), om["hot"])
    irr0 = np.exp(rng.normal(3, 1, size=(W * H, G))).astype(np.float32)
    gpu.init(-1000, irr0)
    ref.init(-1000, irr0)
    t = -1000
    total = n_hot = 0
    for call in range(4):
        t1 = t + int(rng.integers(1, 3 * 2 ** 23))
        ids, ts, ir, offs = random_runs(rng, W * H, t, t1, G=G)
        ev = gpu.advance(ids, ts, ir, t1, offsets=offs)
        want, rep = ref.advance_runs(ids, ts, ir, offs, t1)
        got = events_from_dict(ev, W)
        i, a, b = first_difference(got, want)
        assert len(got) == len(want) and i == len(got), f"call {call}: first difference at {i}: gpu {a} oracle {b}"
        assert np.array_equal(ev["t_reported_ns"], rep)
        total += len(got)
        n_hot += int(np.count_nonzero(want["flags"] & (1 << 5)))
        t = t1
    assert total > 1000
    assert gpu.time_ns == t
    if CASES[case].get("hot_fraction", 0) > 0:
        assert n_hot > 100
    # the work counters the reference defines for this path (the leaf intervals are this backend's own)
    ks, ko = gpu.stats(), ref.stats()
    for f in ("substeps", "pieces"):
        assert ks[f] == ko[f], (f, ks[f], ko[f])


def test_padded_counts_and_unlisted_pixels():
    """The padded form with counts, a call with no listed pixels (held input, leak fires), and a
    zero-length call."""
    import ballasim_amd as bsx
    W, H = 40, 20
    rng = np.random.default_rng(5)
    p = _params("physical_leak")
    p.leak_rate_hz = 20.0
    gpu = bsx.SensorModel(p, W, H, 2, device="cuda")
    ref = O.OracleModel(p.c, W, H, 2, 1, 0)
    irr0 = np.full((W * H, 1), 50.0, np.float32)
    gpu.init(0, irr0)
    ref.init(0, irr0)
    n, m = 300, 9
    ids = rng.choice(W * H, n, replace=False).astype(np.uint32)
    tt = np.sort(rng.choice(np.arange(0, 2 ** 22), size=(n, m), replace=False), axis=1).astype(np.int64)
    irr = np.exp(rng.normal(4, 1, (n, m, 1))).astype(np.float32)
    counts = rng.integers(0, m + 1, n).astype(np.int32)
    ev = gpu.advance(ids, tt, irr, 2 ** 22, counts=counts)
    want, _ = ref.advance_padded(ids, tt, irr, 2 ** 22, counts=counts)
    assert np.array_equal(events_from_dict(ev, W), want)
    empty = np.zeros(0, np.uint32)
    for t_end in (2 ** 22, 3 * 2 ** 24):
        ev = gpu.advance(empty, np.zeros((0, 0), np.int64), np.zeros((0, 0, 1), np.float32), t_end)
        want, _ = ref.advance_padded(empty, np.zeros((0, 0), np.int64), np.zeros((0, 0, 1), np.float32), t_end)
        assert np.array_equal(events_from_dict(ev, W), want)
    assert len(want) > 0  # leak-only events of held pixels


def test_state_roundtrip_and_slicing_invariance():
    """save_state/load_state resume bit for bit; the saved fp32 state equals the oracle's."""
    import ballasim_amd as bsx
    W, H = 32, 16
    rng = np.random.default_rng(9)
    p = _params("physical_leak")
    p.shot_rate_hz = 300.0
    a = bsx.SensorModel(p, W, H, 4, device="cuda")
    b = bsx.SensorModel(p, W, H, 4, device="cuda")
    ref = O.OracleModel(p.c, W, H, 4, 1, 0)
    irr0 = np.exp(rng.normal(3, 1, (W * H, 1))).astype(np.float32)
    a.init(0, irr0)
    ref.init(0, irr0)
    ids, ts, ir, offs = random_runs(rng, W * H, 0, 2 ** 23)
    a.advance(ids, ts, ir, 2 ** 23, offsets=offs)
    ref.advance_runs(ids, ts, ir, offs, 2 ** 23)
    blob = a.save_state()
    st = np.frombuffer(blob[56:], O.REF_STATE_DTYPE)
    os_ = ref.states()
    for f in ("s3_ref", "dr_ref", "i_scale", "t_ref", "t_search", "blind_until", "flags"):
        assert np.array_equal(st[f], os_[f]), f
    # the shot-noise process's next arrival and index ride in the cell / nk[0] fields
    assert np.array_equal(st["cell"], os_["shot_next"])
    assert np.array_equal(st["nk"][:, 0], os_["shot_index"].astype(np.float64))
    b.load_state(blob)
    ids2, ts2, ir2, offs2 = random_runs(rng, W * H, 2 ** 23, 2 ** 24)
    ea = a.advance(ids2, ts2, ir2, 2 ** 24, offsets=offs2)
    eb = b.advance(ids2, ts2, ir2, 2 ** 24, offsets=offs2)
    for k in ea:
        assert np.array_equal(ea[k], eb[k]), k


def test_input_refusal_before_state_change():
    import ballasim_amd as bsx
    W, H = 8, 8
    m = bsx.SensorModel("ideal", W, H, 1, device="cuda")
    with pytest.raises(RuntimeError):
        m.advance(np.zeros(1, np.uint32), np.zeros((1, 1), np.int64), np.ones((1, 1, 1), np.float32), 10)
    m.init(0, np.ones((W * H, 1), np.float32))
    bad = [
        (np.array([64], np.uint32), np.array([[1]], np.int64), np.ones((1, 1, 1), np.float32), 10),
        (np.array([1, 1], np.uint32), np.array([[1], [2]], np.int64), np.ones((2, 1, 1), np.float32), 10),
        (np.array([1], np.uint32), np.array([[3, 3]], np.int64), np.ones((1, 2, 1), np.float32), 10),
        (np.array([1], np.uint32), np.array([[11]], np.int64), np.ones((1, 1, 1), np.float32), 10),
        (np.array([1], np.uint32), np.array([[5]], np.int64), np.full((1, 1, 1), np.nan, np.float32), 10),
    ]
    for ids, t, irr, t_end in bad:
        with pytest.raises(ValueError):
            m.advance(ids, t, irr, t_end)
        assert m.initialized and m.time_ns == 0
    with pytest.raises(ValueError):
        m.advance(np.zeros(0, np.uint32), np.zeros((0, 0), np.int64), np.zeros((0, 0, 1), np.float32), -5)


def _kind_params(kind):
    import ballasim_amd as bsx
    if kind == "instantaneous":
        return bsx.preset("ideal")
    p = bsx.preset("samsung_gen3")
    p.noise = False
    return p


@pytest.mark.parametrize("kind", ["instantaneous", "cascade"])
@pytest.mark.parametrize("offset", [0, 1])
def test_device_sample_checks_large_runs(offset, kind):
    """The device check of the runs' irradiance, made by the chain's phase A (the instantaneous
    response's screen, the cascade's cost pass), on arrays of 163,840 values, aligned (offset 0) and
    starting one float past an aligned address (offset 1): a NaN deep inside and an infinity in the
    last value are refused before any state changes, and values above the largest accepted sample are
    counted as clamps, exactly."""
    import torch

    import ballasim_amd as bsx
    W, H, n = 64, 64, 40
    P = W * H
    m = bsx.SensorModel(_kind_params(kind), W, H, 1, device="cuda")
    m.init(0, np.ones((P, 1), np.float32))
    dev = "cuda"
    pixel = torch.arange(P, dtype=torch.int32, device=dev)
    count = torch.full((P,), n, dtype=torch.int32, device=dev)
    start = torch.arange(P, dtype=torch.int64, device=dev) * n
    t = (torch.arange(n, dtype=torch.int64, device=dev) * 1000 + 1000).repeat(P)
    base = torch.full((P * n + offset,), 5.0, dtype=torch.float32, device=dev)
    irr = base[offset:]
    assert (irr.data_ptr() % 16 == 0) == (offset == 0)
    t_end = int(n * 1000 + 1000)
    for where, value in ((100_003, float("nan")), (P * n - 1, float("inf"))):
        irr[where] = value
        with pytest.raises(ValueError, match="non-finite"):
            m.advance_device(pixel, start, count, t, irr, t_end)
        assert m.initialized and m.time_ns == 0
        irr[where] = 5.0
    big = m.irr_max * 4.0
    hot = torch.tensor([0, 7, 99_999, 150_001, P * n - 3, P * n - 1], device=dev)
    irr[hot] = big
    c0 = m.stats()["input_clamps"]
    m.advance_device(pixel, start, count, t, irr, t_end)
    torch.cuda.synchronize()
    assert m.stats()["input_clamps"] - c0 == len(hot)
    assert m.time_ns == t_end


@pytest.mark.parametrize("kind", ["instantaneous", "cascade"])
def test_device_check_of_a_first_sample_at_the_model_time(kind):
    """A run whose first sample lies at the model time ends no segment, but its value is checked all
    the same: a NaN there refuses the call with the model unchanged (the next call then runs), and an
    over-range value there counts as one clamp."""
    import torch

    import ballasim_amd as bsx
    W, H, n = 16, 8, 6
    P = W * H
    m = bsx.SensorModel(_kind_params(kind), W, H, 3, device="cuda")
    m.init(5000, np.full((P, 1), 20.0, np.float32))
    dev = "cuda"
    pixel = torch.arange(P, dtype=torch.int32, device=dev)
    count = torch.full((P,), n, dtype=torch.int32, device=dev)
    start = torch.arange(P, dtype=torch.int64, device=dev) * n
    t = (torch.arange(n, dtype=torch.int64, device=dev) * 1000 + 5000).repeat(P)
    irr = torch.full((P * n,), 20.0, dtype=torch.float32, device=dev)
    t_end = 5000 + n * 1000
    before = m.save_state()
    irr[37 * n] = float("nan")  # pixel 37's first sample, at the model time
    with pytest.raises(ValueError, match="1 non-finite"):
        m.advance_device(pixel, start, count, t, irr, t_end)
    assert m.initialized and m.time_ns == 5000 and m.save_state() == before
    irr[37 * n] = m.irr_max * 2.0
    c0 = m.stats()["input_clamps"]
    m.advance_device(pixel, start, count, t, irr, t_end)
    torch.cuda.synchronize()
    assert m.stats()["input_clamps"] - c0 == 1 and m.time_ns == t_end


@pytest.mark.parametrize("kind", ["instantaneous", "cascade"])
def test_short_runs_many_runs_per_phase_a_block(kind):
    """Runs of zero to two samples over a 192x128 camera, so one phase-A block of 2048 consecutive
    positions spans far more than the 256 runs it stages in LDS and its threads find their runs by
    bisection in global memory: events, reported times and the state equal the oracle's."""
    import ballasim_amd as bsx
    W, H = 192, 128
    P = W * H
    rng = np.random.default_rng(31)
    p = _kind_params(kind)
    p.leak_rate_hz = 2.0
    gpu = bsx.SensorModel(p, W, H, 9, device="cuda")
    ref = O.OracleModel(p.c, W, H, 9, 1, 0)
    irr0 = np.exp(rng.normal(3, 1, size=(P, 1))).astype(np.float32)
    gpu.init(0, irr0)
    ref.init(0, irr0)
    t = 0
    for call in range(3):
        t1 = t + 3_000_000
        ids, ts, ir, offs = random_runs(rng, P, t, t1, max_n=3, frac=0.95, log_sigma=2.5)
        assert len(ids) > 2048 * 4 and len(ts) < 2 * len(ids)
        ev = gpu.advance(ids, ts, ir, t1, offsets=offs)
        want, rep = ref.advance_runs(ids, ts, ir, offs, t1)
        got = events_from_dict(ev, W)
        i, a, b = first_difference(got, want)
        assert len(got) == len(want) and i == len(got), f"call {call}: first difference at {i}: gpu {a} oracle {b}"
        assert np.array_equal(ev["t_reported_ns"], rep)
        t = t1
    st = np.frombuffer(gpu.save_state()[56:], O.REF_STATE_DTYPE)
    so = ref.states()
    for f in ("z", "s3_ref", "t_ref", "t_search", "blind_until", "flags", "held"):
        assert np.array_equal(st[f], so[f]), f


@pytest.mark.parametrize("kind", ["instantaneous", "cascade"])
def test_sparse_runs_among_many_empty_ones(kind):
    """Every pixel of a 1024x512 camera listed, all but a few hundred with no samples (a static scene's
    pixels take none), so phase A's threads cross stretches of thousands of empty runs, which share one
    position: events, reported times and the state equal the oracle's."""
    import ballasim_amd as bsx
    W, H = 1024, 512
    P = W * H
    rng = np.random.default_rng(57)
    p = _kind_params(kind)
    gpu = bsx.SensorModel(p, W, H, 4, device="cuda")
    ref = O.OracleModel(p.c, W, H, 4, 1, 0)
    irr0 = np.exp(rng.normal(3, 1, size=(P, 1))).astype(np.float32)
    gpu.init(0, irr0)
    ref.init(0, irr0)
    ids = np.arange(P, dtype=np.uint32)
    counts = np.zeros(P, np.int64)
    busy = np.sort(rng.choice(P, size=300, replace=False))
    counts[busy] = rng.integers(1, 40, size=len(busy))
    offs = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
    t = np.zeros(offs[-1], np.int64)
    for k, pix in enumerate(busy):
        s0, c = offs[pix], counts[pix]
        t[s0:s0 + c] = np.sort(rng.choice(np.arange(1, 2 ** 23), size=c, replace=False))
    irr = np.exp(rng.normal(3, 1.5, size=(offs[-1], 1))).astype(np.float32)
    ev = gpu.advance(ids, t, irr, 2 ** 23, offsets=offs)
    want, rep = ref.advance_runs(ids, t, irr, offs, 2 ** 23)
    got = events_from_dict(ev, W)
    i, a, b = first_difference(got, want)
    assert len(got) == len(want) and i == len(got) and len(want) > 100, f"first difference at {i}: gpu {a} oracle {b}"
    assert np.array_equal(ev["t_reported_ns"], rep)
    st = np.frombuffer(gpu.save_state()[56:], O.REF_STATE_DTYPE)
    so = ref.states()
    for f in ("z", "s3_ref", "t_ref", "t_search", "blind_until", "flags", "held"):
        assert np.array_equal(st[f], so[f]), f


def test_device_runs_padding_not_checked():
    """Device runs that leave samples unreferenced: the padding between runs holds NaN, infinities
    and over-range values. Only the referenced samples are checked, as the reference's
    check_inputs_kernel does (sensor_cuda.cu:415-440), so the call succeeds, and input_clamps counts
    only the referenced values the clamp changes, equal to the oracle's count; events equal too."""
    import torch

    import ballasim_amd as bsx
    W, H, n, pad = 32, 24, 9, 3
    P = W * H
    rng = np.random.default_rng(77)
    p = _params("physical_leak")
    gpu = bsx.SensorModel(p, W, H, 5, device="cuda")
    ref = O.OracleModel(p.c, W, H, 5, 1, 0)
    irr0 = np.exp(rng.normal(3, 1, size=(P, 1))).astype(np.float32)
    gpu.init(0, irr0)
    ref.init(0, irr0)
    ids = rng.permutation(P).astype(np.uint32)[: P - 40]
    m = len(ids)
    start = np.arange(m, dtype=np.int64) * (n + pad) + pad
    count = rng.integers(0, n + 1, size=m).astype(np.int32)
    S = m * (n + pad) + pad
    t = np.zeros(S, np.int64)
    irr = np.full(S, np.nan, np.float32)
    irr[1::7] = np.inf
    irr[2::7] = gpu.irr_max * 8.0
    for i in range(m):
        s, c = int(start[i]), int(count[i])
        t[s:s + c] = np.sort(rng.choice(np.arange(1, 2 ** 23), size=c, replace=False))
        irr[s:s + c] = np.exp(rng.normal(3, 1.5, size=c))
    big = rng.choice(np.flatnonzero(np.repeat(count > 0, 1)), size=25, replace=False)
    irr[start[big]] = gpu.irr_max * 4.0
    want_clamps = sum(int(np.count_nonzero(irr[s:s + c] > gpu.irr_max)) for s, c in zip(start, count))
    assert want_clamps >= 25
    want, rep = ref._advance(ids, t, irr.reshape(S, 1), start, count, 2 ** 23)
    assert ref.stats()["input_clamps"] == want_clamps
    dev = "cuda"
    c0 = gpu.stats()["input_clamps"]
    ptr, cnt = gpu.advance_device(torch.from_numpy(ids.astype(np.int32)).to(dev), torch.from_numpy(start).to(dev),
                                  torch.from_numpy(count).to(dev), torch.from_numpy(t).to(dev),
                                  torch.from_numpy(irr).to(dev), 2 ** 23)
    got = events_from_dict(gpu.download_events(ptr, cnt), W)
    assert gpu.stats()["input_clamps"] - c0 == want_clamps
    assert len(got) == len(want) and len(want) > 0
    i, a, b = first_difference(got, want)
    assert a is None, f"event {i}: gpu {a} oracle {b}"


def test_device_runs_tensor_checks():
    """Tensors the C side cannot check are refused by the wrapper: host tensors, strided views,
    wrong dtypes and mismatched lengths."""
    import torch

    import ballasim_amd as bsx
    m = bsx.SensorModel("ideal", 8, 8, 1, device="cuda")
    m.init(0, np.ones((64, 1), np.float32))
    dev = "cuda"
    good = dict(pixel=torch.arange(4, dtype=torch.int32, device=dev),
                row_start=torch.arange(4, dtype=torch.int64, device=dev) * 2,
                row_count=torch.full((4,), 2, dtype=torch.int32, device=dev),
                t=torch.arange(1, 9, dtype=torch.int64, device=dev) * 100,
                irr=torch.ones(8, dtype=torch.float32, device=dev))
    bad = [("pixel", good["pixel"].cpu(), "on the GPU"),
           ("row_start", good["row_start"].to(torch.int32), "dtype"),
           ("row_count", good["row_count"].to(torch.int64), "dtype"),
           ("t", (torch.arange(1, 17, dtype=torch.int64, device=dev) * 100)[::2], "contiguous"),
           ("irr", torch.ones(7, dtype=torch.float32, device=dev), "irr must hold"),
           ("row_count", torch.full((3,), 2, dtype=torch.int32, device=dev), "same length")]
    for name, value, msg in bad:
        args = dict(good, **{name: value})
        with pytest.raises(ValueError, match=msg):
            m.advance_device(args["pixel"], args["row_start"], args["row_count"], args["t"], args["irr"], 10_000)
        assert m.time_ns == 0
    m.advance_device(good["pixel"], good["row_start"], good["row_count"], good["t"], good["irr"], 10_000)
    assert m.time_ns == 10_000


def test_event_buffer_overflow_reruns():
    """A forced tiny event capacity: the call re-runs from its input state with a larger buffer and
    gives the same events."""
    import ballasim_amd as bsx
    W, H = 32, 32
    rng = np.random.default_rng(3)
    p = _params("ideal")
    a = bsx.SensorModel(p, W, H, 8, device="cuda")
    b = bsx.SensorModel(p, W, H, 8, device="cuda")
    a.set_launch_policy(event_capacity=7)
    irr0 = np.exp(rng.normal(3, 1, (W * H, 1))).astype(np.float32)
    a.init(0, irr0)
    b.init(0, irr0)
    ids, ts, ir, offs = random_runs(rng, W * H, 0, 2 ** 23)
    ea = a.advance(ids, ts, ir, 2 ** 23, offsets=offs)
    eb = b.advance(ids, ts, ir, 2 ** 23, offsets=offs)
    assert len(ea["t_cross_ns"]) > 7
    for k in ea:
        assert np.array_equal(ea[k], eb[k]), k
    assert a.stats()["overflow_reruns"] >= 1


def test_advance_device_torch_runs():
    """advance_device on runs resident in HBM (torch tensors), then download_events."""
    import torch

    import ballasim_amd as bsx
    W, H = 64, 32
    rng = np.random.default_rng(21)
    p = _params("physical_leak")
    a = bsx.SensorModel(p, W, H, 6, device="cuda")
    ref = O.OracleModel(p.c, W, H, 6, 1, 0)
    irr0 = np.exp(rng.normal(3, 1, (W * H, 1))).astype(np.float32)
    a.init(0, irr0)
    ref.init(0, irr0)
    ids, ts, ir, offs = random_runs(rng, W * H, 0, 2 ** 23)
    dev = torch.device("cuda:0")
    tp = torch.from_numpy(ids.astype(np.int32)).to(dev)
    rs = torch.from_numpy(offs[:-1].copy()).to(dev)
    rc = torch.from_numpy(np.diff(offs).astype(np.int32)).to(dev)
    tt = torch.from_numpy(ts).to(dev)
    ii = torch.from_numpy(ir.reshape(-1)).to(dev)
    torch.cuda.synchronize()
    ptr, count = a.advance_device(tp, rs, rc, tt, ii, 2 ** 23)
    ev = a.download_events(ptr, count)
    want, _ = ref.advance_runs(ids, ts, ir, offs, 2 ** 23)
    assert np.array_equal(events_from_dict(ev, W), want)
