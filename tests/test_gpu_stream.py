"""GPU tests of the event stream (include/bsx/bsx_output.h, bsx_event_stream_*): a moving-rectangle
scene through the plain controller and the instantaneous response with shot and leak noise, each
window's raw events fed to two output stages, one returning its released columns and one streaming
them in double-buffered pinned chunks. The chunks must arrive in order, every one full but the last,
and concatenated equal the returned columns exactly (reported time, x, y, polarity, label, flags),
flush included; as a multiset they equal the model's raw events. Also a slow sink (the producer waits
for a chunk slot) and a sink that raises (the stream stops; the next call raises)."""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

WINDOW = 1 << 23


def _setup(width=320, height=240, n=24, seed=3):
    import ballasim_amd as bsx
    from ballasim_amd import configs
    from ballasim_amd import scene as S
    shapes, bg = S.config2_shapes(seed, n=n, width=width, height=height, v_max=3000.0)
    scene, view = S.Scene(shapes, bg), S.view(width, height)
    p = configs.config2_params("instantaneous")
    p.shot_rate_hz = 0.5
    p.leak_rate_hz = 0.1
    m = bsx.SensorModel(p, width, height, 7, ["constant"], 0, device="cuda")
    m.init(0, scene.render(view, 0).reshape(-1, 1))
    rules = S.rules(float(p.dark_current_fA) / float(p.illuminance_scale), tol=0.01)
    return scene, view, m, rules


def _run(windows, chunk_bytes, sink_delay=0.0, raise_at=None):
    from ballasim_amd.output import EventStream, OutputStage
    scene, view, m, rules = _setup()
    plain, streamed = OutputStage(m), OutputStage(m)
    chunks = []

    def sink(rec, index):
        if raise_at is not None and index == raise_at:
            raise RuntimeError("sink stop")
        if sink_delay:
            time.sleep(sink_delay)
        chunks.append((index, rec.copy()))

    es = EventStream(sink, chunk_bytes=chunk_bytes)
    want, raw = [], []
    for w in range(windows):
        runs, _ = scene.sample_window(view, rules, w * WINDOW, slot=0)
        events = m.advance_device_runs(runs, (w + 1) * WINDOW)
        ev = m.download_events(*events)
        raw.append(ev)
        want.append(plain.window(events, w * WINDOW))
        n = es.window(streamed, events, w * WINDOW)
        assert n == len(want[-1]["t"]) and streamed.held == plain.held, w
    want.append(plain.flush())
    assert es.flush(streamed) == len(want[-1]["t"])
    es.finish()
    return es, chunks, want, raw, m


def _concat(cols, key):
    return np.concatenate([c[key] for c in cols])


def test_streamed_chunks_equal_the_returned_columns():
    chunk_events = 1031
    es, chunks, want, raw, m = _run(4, 16 * chunk_events)
    total = sum(len(c["t"]) for c in want)
    assert total > 20 * chunk_events
    assert [i for i, _ in chunks] == list(range(len(chunks)))
    sizes = [len(r) for _, r in chunks]
    assert all(s == chunk_events for s in sizes[:-1]) and 0 < sizes[-1] <= chunk_events
    got = np.concatenate([r for _, r in chunks])
    assert len(got) == total
    for k_got, k_want in (("t_reported_ns", "t_reported_ns"), ("x", "x"), ("y", "y"), ("p", "p"), ("label", "label"),
                          ("flags", "flags")):
        assert np.array_equal(got[k_got], _concat(want, k_want)), k_got
    st = es.stats()
    assert st["events"] == total and st["chunks"] == len(chunks) and st["bytes"] == 16 * total
    assert st["copy_ms"] > 0
    # as a multiset, the stream holds exactly the model's raw events (pixel, polarity, reported time)
    W = 320
    key_raw = np.sort(np.concatenate([
        ((r["y"].astype(np.int64) * W + r["x"]) << 40) | (r["p"].astype(np.int64) << 39) | r["t_reported_ns"]
        for r in raw]))
    key_got = np.sort(((got["y"].astype(np.int64) * W + got["x"]) << 40) | (got["p"].astype(np.int64) << 39)
                      | got["t_reported_ns"])
    assert np.array_equal(key_raw, key_got)
    es.close()


def test_slow_sink_makes_the_producer_wait():
    es, chunks, want, raw, m = _run(2, 16 * 4096, sink_delay=0.01)
    got = np.concatenate([r for _, r in chunks])
    assert np.array_equal(got["t_reported_ns"], _concat(want, "t_reported_ns"))
    assert es.stats()["wait_ms"] > 5.0 and len(chunks) >= 3
    es.close()


def test_a_sink_that_raises_stops_the_stream():
    with pytest.raises(RuntimeError, match="sink"):
        _run(3, 16 * 1031, raise_at=2)
