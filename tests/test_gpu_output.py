"""GPU parity of the output stage with its oracle (oracle/output_oracle.py, pinned against the
reference's output_order / readout / label_s1): per window the released events, their order and
every C4 column are identical, and the held count matches; with a crop and a readout with delay,
quantum and clock offset and drift; then flush. Also the C4 file written from the released columns
(t monotone, ms_to_idx)."""
import numpy as np
import pytest

from oracle import oracle as O
from oracle import output_oracle as Q
from tests.helpers import random_runs

pytestmark = pytest.mark.gpu

W = 1 << 23


@pytest.mark.parametrize("crop,readout", [
    (None, {}),
    ((3, 2, 31, 19), dict(readout_delay_s=7e-4, readout_quantum_s=1e-6, clock_offset_ns=-12345, clock_drift_ppm=35.0)),
])
def test_output_stage_matches_oracle(crop, readout, tmp_path):
    import ballasim_amd as bsx
    from ballasim_amd.output import EventFile, OutputStage
    p = bsx.preset("prophesee_gen3")
    p.photoreceptor = "instantaneous"
    p.noise = False
    p.leak_rate_hz = 5.0
    p.refractory_s = 1e-5
    for k, v in readout.items():
        setattr(p, k, v)
    Wd, Hd = 40, 24
    m = bsx.SensorModel(p, Wd, Hd, 3, ["constant"], 0, device="cuda")
    ref = O.OracleModel(p.c, Wd, Hd, 3, 1, 0)
    rng = np.random.default_rng(5)
    irr0 = np.exp(rng.normal(3, 1, (Wd * Hd, 1))).astype(np.float32)
    m.init(0, irr0)
    ref.init(0, irr0)
    info = m._info()
    rd = dict(quantum=round(p.readout_quantum_s * 1e9), delay=round(p.readout_delay_s * 1e9),
              phase=int(info.readout_phase_ns), offset=int(p.clock_offset_ns), drift_ppm=float(p.clock_drift_ppm))
    stage = OutputStage(m, crop=crop, t_offset_us=1000)
    oracle = Q.OutputOracle(Wd, Hd, rd, crop=crop, t_offset_us=1000)
    f = EventFile(tmp_path / "cam.npz", "cam0", Wd, Hd, t_offset_us=1000)
    released = 0
    for w in range(5):
        ids, ts, ir, offs = random_runs(rng, Wd * Hd, w * W, (w + 1) * W, max_n=24, log_sigma=2.0)
        import torch
        dev = torch.device("cuda:0")
        runs = (torch.from_numpy(ids.astype(np.int32)).to(dev), torch.from_numpy(offs[:-1].copy()).to(dev),
                torch.from_numpy(np.diff(offs).astype(np.int32)).to(dev), torch.from_numpy(ts).to(dev),
                torch.from_numpy(np.ascontiguousarray(ir.reshape(-1))).to(dev))
        torch.cuda.synchronize()
        events = m.advance_device(*runs, (w + 1) * W)
        raw, _ = ref.advance_runs(ids, ts, ir, offs, (w + 1) * W)
        assert events[1] == len(raw)
        got = stage.window(events, w * W)
        want = oracle.window(raw, w * W)
        assert len(got["t"]) == len(want["t"]), w
        for k_got, k_want in (("x", "x"), ("y", "y"), ("p", "p"), ("t", "t"), ("t_ns_rem", "t_ns_rem"),
                              ("t_reported_ns", "t_reported"), ("t_cross_ns", "t_cross"), ("label", "label"),
                              ("flags", "flags")):
            assert np.array_equal(got[k_got], want[k_want]), (w, k_got)
        assert stage.held == len(oracle.held["t_cross"])
        f.append({"x": got["x"], "y": got["y"], "p": got["p"], "t": got["t"], "t_ns_rem": got["t_ns_rem"],
                  "label": got["label"], "t_cross_ns": got["t_cross_ns"], "flags": got["flags"]})
        released += len(got["t"])
    got = stage.flush()
    want = oracle.flush()
    assert np.array_equal(got["t_reported_ns"], want["t_reported"]) and np.array_equal(got["x"], want["x"])
    assert stage.held == 0
    f.append({"x": got["x"], "y": got["y"], "p": got["p"], "t": got["t"], "t_ns_rem": got["t_ns_rem"],
              "label": got["label"], "t_cross_ns": got["t_cross_ns"], "flags": got["flags"]})
    released += len(got["t"])
    assert released > 300
    f.close()
    d = np.load(tmp_path / "cam.npz")
    t = d["events/t"]
    assert len(t) == released and np.all(np.diff(t) >= 0)
    ms = d["ms_to_idx"].astype(np.int64)
    k = np.arange(len(ms))
    assert len(ms) == t[-1] // 1000 + 1
    # entry k is the FIRST event with t >= 1000 k
    assert np.all((ms == len(t)) | (t[np.minimum(ms, len(t) - 1)] >= 1000 * k))
    assert np.all((ms == 0) | (t[np.maximum(ms - 1, 0)] < 1000 * k))


def test_output_order_of_millions_of_events():
    """The stage's stable radix passes at scale: about 2.5 M synthetic raw events per window over a
    1280x720 camera (several hundred sort tiles per pass), their crossing times on a coarse grid and a
    1-us readout quantum, so that the reported time, x and y each tie among many events and the order rests
    on every pass being stable. Two windows and the flush, against the oracle's lexsort."""
    import torch

    import ballasim_amd as bsx
    from ballasim_amd.output import OutputStage
    p = bsx.preset("prophesee_gen3")
    p.photoreceptor = "instantaneous"
    p.noise = False
    p.readout_delay_s = 3e-4
    p.readout_quantum_s = 1e-6
    Wd, Hd = 1280, 720
    m = bsx.SensorModel(p, Wd, Hd, 3, ["constant"], 0, device="cuda")
    m.init(0, np.full((Wd * Hd, 1), 50.0, np.float32))
    info = m._info()
    rd = dict(quantum=round(p.readout_quantum_s * 1e9), delay=round(p.readout_delay_s * 1e9),
              phase=int(info.readout_phase_ns), offset=int(p.clock_offset_ns), drift_ppm=float(p.clock_drift_ppm))
    stage = OutputStage(m, t_offset_us=0)
    oracle = Q.OutputOracle(Wd, Hd, rd, t_offset_us=0)
    rng = np.random.default_rng(11)
    for w in range(2):
        n = 2_600_000
        pix = rng.integers(0, Wd * Hd, n).astype(np.uint32)
        tc = (w * W + rng.integers(0, W // 4096, n) * 4096 + rng.integers(0, 3, n)).astype(np.int64)
        key = pix.astype(np.int64) * (1 << 24) + (tc - w * W)
        _, first = np.unique(key, return_index=True)  # one event per (pixel, t_cross): a total order
        first = rng.permutation(first)
        raw = np.zeros(len(first), O.EVENT_DTYPE)
        raw["pixel"] = pix[first]
        raw["t_cross"] = tc[first]
        raw["polarity"] = rng.integers(0, 2, len(first))
        raw["label"] = rng.integers(0, 8, len(first))  # the class is the label's two low bits
        raw["flags"] = rng.integers(0, 4, len(first))
        dev = torch.from_numpy(raw.view(np.uint8).copy()).cuda()
        torch.cuda.synchronize()
        got = stage.window((dev.data_ptr(), len(raw)), w * W)
        want = oracle.window(raw, w * W)
        assert len(got["t"]) == len(want["t"]) > 10 ** 6, w
        for k_got, k_want in (("x", "x"), ("y", "y"), ("p", "p"), ("t_reported_ns", "t_reported"),
                              ("t_cross_ns", "t_cross"), ("label", "label"), ("flags", "flags")):
            assert np.array_equal(got[k_got], want[k_want]), (w, k_got)
        assert stage.held == len(oracle.held["t_cross"]) > 0
        del dev
    got = stage.flush()
    want = oracle.flush()
    for k_got, k_want in (("x", "x"), ("y", "y"), ("t_reported_ns", "t_reported"), ("t_cross_ns", "t_cross")):
        assert np.array_equal(got[k_got], want[k_want]), k_got
