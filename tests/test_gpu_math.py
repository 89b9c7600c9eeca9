"""The device math the exact searches rely on (GPU): pm::log is monotone on every positive float
and equals the oracle's log bit for bit (an order-independent hash of all 2^31 - 2^23 results
against the one the oracle produced, recorded in the golden manifest)."""
import ctypes as C
import json
import os

import pytest

from ballasim_amd import _abi

pytestmark = pytest.mark.gpu


def test_device_log_monotone_and_equal_to_oracle():
    lib = _abi.load_library()
    v, h = C.c_uint64(0), C.c_uint64(1)
    _abi.check(lib.bsx_math_selfcheck(C.byref(v), C.byref(h)))
    man = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "manifest.json")))["log_check"]
    assert v.value == 0
    assert "0x%016x" % h.value == man["hash"]
