"""The C4 event file's index on the CPU: /ms_to_idx entry k must be the index of the FIRST event with
t >= 1000 k (python/ballasim/engine/writer.py:7,102-107), checked against a brute-force scan over
event times with gaps of many milliseconds, runs of equal times and events exactly on millisecond
boundaries; with the run's end given, the entries continue to that end's millisecond, holding the
event count (writer.py:128-130). Also the stream record's layout (16 bytes, the C header's fields) and
the stream's refusal of bad chunk sizes, which needs no device."""
import numpy as np
import pytest

from ballasim_amd.output import STREAM_EVENT, EventFile, EventStream


def _brute(t, n_ms):
    out = []
    for k in range(n_ms):
        i = 0
        while i < len(t) and t[i] < 1000 * k:
            i += 1
        out.append(i)
    return np.array(out, np.uint64)


def _cols(t):
    n = len(t)
    return {"x": np.zeros(n, np.uint16), "y": np.zeros(n, np.uint16), "p": np.zeros(n, np.uint8),
            "t": np.asarray(t, np.int64), "t_ns_rem": np.zeros(n, np.uint16), "label": np.zeros(n, np.uint8),
            "t_cross_ns": np.zeros(n, np.int64), "flags": np.zeros(n, np.uint16)}


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_ms_to_idx_is_the_first_event_at_or_after_each_millisecond(seed, tmp_path):
    rng = np.random.default_rng(seed)
    gaps = rng.choice([0, 0, 1, 7, 999, 1000, 1001, 5000, 23_456], size=400)
    t = np.cumsum(gaps) + int(rng.integers(0, 3000))
    t[::37] = (t[::37] // 1000) * 1000  # some events exactly on a boundary (order kept below)
    t = np.maximum.accumulate(t)
    f = EventFile(tmp_path / "e.npz", "cam", 8, 8, t_offset_us=0)
    f.append(_cols(t[:150]))
    f.append(_cols(t[150:]))
    d = f.datasets()
    n_ms = int(t[-1] // 1000) + 1
    assert np.array_equal(d["ms_to_idx"], _brute(t, n_ms))
    end_ns = (int(t[-1]) + 4321) * 1000
    d2 = f.datasets(end_ns)
    n2 = max(n_ms, (end_ns // 1000) // 1000 + 1)
    assert len(d2["ms_to_idx"]) == n2 and np.array_equal(d2["ms_to_idx"], _brute(t, n2))
    assert np.all(d2["ms_to_idx"][n_ms:] == len(t))


def test_event_file_refuses_events_out_of_order(tmp_path):
    f = EventFile(tmp_path / "e.npz", "cam", 8, 8)
    f.append(_cols([5, 9]))
    with pytest.raises(ValueError):
        f.append(_cols([8]))


def test_stream_record_layout():
    assert STREAM_EVENT.itemsize == 16
    assert [STREAM_EVENT.fields[k][1] for k in ("t_reported_ns", "x", "y", "p", "label", "flags")] == [0, 8, 10, 12, 13,
                                                                                                       14]


@pytest.mark.parametrize("chunk", [0, 15, 24, (1 << 32) + 16])
def test_stream_refuses_bad_chunk_sizes(chunk):
    with pytest.raises(ValueError, match="chunk_bytes"):
        EventStream(lambda rec, i: None, chunk_bytes=chunk)
