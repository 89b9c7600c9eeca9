"""The advance call that allocates nothing (bsx_sensor_advance_device_noalloc), through the C ABI,
against the existing advance_device on a twin model: two calls in a row, the second with a smaller
input, for the instantaneous response, the cascade and the cascade with its in-path noise. Events,
reported times, saved states and work counters must be identical (the instantaneous path's own
search-evaluation counter aside); a buffer too small for the call is refused and the state kept."""
import ctypes as C

import numpy as np
import pytest

from tests.helpers import random_runs

pytestmark = pytest.mark.gpu

CASES = {
    "instant": ("ideal", dict(reset_mode="physical", refractory_s=1e-4, leak_rate_hz=0.3)),
    "cascade": ("prophesee_gen3", dict(noise=False)),
    "noise": ("prophesee_gen3", dict()),
}


def _tensors(ids, t, irr, offs):
    import torch
    return (torch.from_numpy(ids.astype(np.int32)).cuda(), torch.from_numpy(offs[:-1].copy()).cuda(),
            torch.from_numpy(np.diff(offs).astype(np.int32)).cuda(), torch.from_numpy(t).cuda(),
            torch.from_numpy(np.ascontiguousarray(irr.reshape(-1))).cuda())


@pytest.mark.parametrize("case", sorted(CASES))
def test_noalloc_equals_advance_device(case):
    import torch

    import ballasim_amd as bsx
    from ballasim_amd import _abi
    from ballasim_amd.sensor import _checked_device_runs
    family, over = CASES[case]
    p = bsx.preset(family)
    for k, v in over.items():
        setattr(p, k, v)
    lib = _abi.load_library()
    rng = np.random.default_rng(len(case))
    W, H = 96, 64
    print(f"no-alloc case {case}: {W}x{H}", flush=True)
    a = bsx.SensorModel(p, W, H, 9, ["constant"], 0, device="cuda")
    b = bsx.SensorModel(p, W, H, 9, ["constant"], 0, device="cuda")
    irr0 = np.exp(rng.normal(3, 1, size=(W * H, 1))).astype(np.float32)
    a.init(-5000, irr0)
    b.init(-5000, irr0)
    t = -5000
    cap = 200_000
    buf = torch.zeros(cap * 32, dtype=torch.uint8, device="cuda")
    for call, frac in enumerate((0.9, 0.3)):
        t1 = t + 2 ** 22 + 333 * call
        ids, ts, ir, offs = random_runs(rng, W * H, t, t1, max_n=24, frac=frac)
        dev = _tensors(ids, ts, ir, offs)
        ptr, count = a.advance_device(*dev, t1)
        want = a.download_events(ptr, count)
        runs = _checked_device_runs(*dev, 1, "test")
        n = C.c_uint64(0)
        _abi.check(lib.bsx_sensor_advance_device_noalloc(b._h, C.byref(runs), t1, None, C.c_void_p(buf.data_ptr()),
                                                          cap, C.byref(n)))
        torch.cuda.synchronize()
        assert n.value == count
        got = b.download_events(buf.data_ptr(), n.value)
        for k in want:
            assert np.array_equal(got[k], want[k]), (call, k)
        assert a.save_state() == b.save_state(), call
        sa, sb = a.stats(), b.stats()
        for k in ("substeps", "substep_attempts", "settled_substeps", "cells", "refinements", "fallbacks",
                  "quarantines", "pieces", "events", "input_clamps") + (() if case == "instant" else ("leaf_intervals",)):
            assert sa[k] == sb[k], (call, k, sa[k], sb[k])
        assert count > 10
        t = t1
    # a buffer too small for the call: refused, and the state is the one before the call
    before = b.save_state()
    t1 = t + 2 ** 22
    ids, ts, ir, offs = random_runs(rng, W * H, t, t1, max_n=24)
    runs = _checked_device_runs(*_tensors(ids, ts, ir, offs), 1, "test")
    n = C.c_uint64(0)
    with pytest.raises(ValueError, match="more than the buffer"):
        _abi.check(lib.bsx_sensor_advance_device_noalloc(b._h, C.byref(runs), t1, None, C.c_void_p(buf.data_ptr()), 1,
                                                          C.byref(n)))
    assert b.save_state() == before


def test_noalloc_refuses_oversized_camera():
    import ballasim_amd as bsx
    from ballasim_amd import _abi
    lib = _abi.load_library()
    p = bsx.preset("ideal")
    m = bsx.SensorModel(p, 4096, 1025, 1, ["constant"], 0, device="cuda")  # 2^22 + 4096 pixels
    m.init(0, np.ones((4096 * 1025, 1), np.float32))
    runs = _abi.DeviceRuns()
    n = C.c_uint64(0)
    with pytest.raises(ValueError, match="at most 4194304"):
        _abi.check(lib.bsx_sensor_advance_device_noalloc(m._h, C.byref(runs), 10, None, None, 0, C.byref(n)))
