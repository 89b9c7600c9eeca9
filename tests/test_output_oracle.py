"""The output stage's CPU oracle against the reference's own output functions (compiled from
src/sensor/output_order.h and readout.h into oracle/_ref): the output order of random events with
ties in every key, reported times and window cuts under quantum, delay, phase, offset and drift, and
the S1 labels. Without the compiled reference the order is checked against a brute-force sort of the
key tuples."""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O
from oracle import output_oracle as Q


def _random_events(rng, n, width, height):
    px = rng.integers(0, width * height, n)
    t = rng.integers(-2000, 5000, n).astype(np.int64) * 997
    # one event per (pixel, t_cross); ties in t_cross across pixels and in t_reported are frequent
    key = np.unique(px * 10 ** 9 + t - t.min())
    px = (key // 10 ** 9).astype(np.int64)
    t = (key % 10 ** 9).astype(np.int64) + t.min()
    perm = rng.permutation(len(px))
    return px[perm], t[perm]


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


@pytest.mark.parametrize("seed", range(3))
def test_output_order_matches_reference(seed):
    rng = np.random.default_rng(seed)
    W, H = 37, 23
    px, tc = _random_events(rng, 5000, W, H)
    x = (px % W).astype(np.uint16)
    y = (px // W).astype(np.uint16)
    tr = Q.reported_time(tc, quantum=5000, delay=1234, phase=777, offset=-40, drift_ppm=55.5)
    got = Q.output_order(tc, x, y, tr)
    keys = sorted(range(len(tc)), key=lambda i: (tr[i], x[i], y[i], tc[i]))
    assert np.array_equal(got, np.array(keys))
    if not O.ref_available():
        pytest.skip("compiled reference not present: checked against the brute-force order only")
    lib = O.ref_lib()
    perm = np.zeros(len(tc), np.uint32)
    tc = np.ascontiguousarray(tc)
    trc = np.ascontiguousarray(tr)
    lib.ref_output_order(_p(tc), _p(x), _p(y), _p(trc), len(tc), H, _p(perm))
    assert np.array_equal(got, perm.astype(np.int64))


def test_readout_and_cut_match_reference():
    if not O.ref_available():
        pytest.skip("compiled reference not present")
    lib = O.ref_lib()
    rng = np.random.default_rng(9)
    t = rng.integers(-10 ** 12, 10 ** 12, 20000).astype(np.int64)
    for q, d, ph, off, dr in ((0, 0, 0, 0, 0.0), (1000, 0, 13, 0, 0.0), (977, 5000, 500, -123456, 0.0),
                              (0, 250, 0, 77, -312.25), (4096, 100000, 4095, 10 ** 9, 1e5)):
        out = np.zeros_like(t)
        cut = C.c_int64()
        lib.ref_readout(q, d, ph, off, dr, _p(t), len(t), _p(out), 12345678901, C.byref(cut))
        assert np.array_equal(Q.reported_time(t, q, d, ph, off, dr), out), (q, d, ph, off, dr)
        assert Q.output_cut(12345678901, off, dr) == cut.value


def test_label_s1_matches_reference():
    c3 = np.arange(8, dtype=np.uint8)
    if not O.ref_available():
        assert list(Q.label_s1(c3)) == [0, 2, 3, 4, 0, 2, 3, 4]
        return
    out = np.zeros(8, np.uint8)
    O.ref_lib().ref_label_s1(_p(c3), 8, _p(out))
    assert np.array_equal(Q.label_s1(c3), out)
