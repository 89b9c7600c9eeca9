"""The oracle's deterministic maths pinned by known answers and by the mathematics (CPU):
Philox4x32-10 by the Random123 known-answer vectors and the reference's pinned stream words; the
constant tables by re-deriving them from their definitions; exp/log/expm1/sincos by ULP bounds
against numpy's fp64 values; the special functions by scipy and closed forms."""
import ctypes as C
import math
import os
import subprocess
import sys

import numpy as np
import pytest

from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def philox(ctr, key):
    lib = O.oracle_lib()
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    out = (C.c_uint32 * 4)()
    lib.orc_philox(c, k, out)
    return list(out)


def test_philox_random123_kats():
    # Random123's kat_vectors for philox4x32_10 (also tests/cpp/test_philox.cpp:11-31 of the reference).
    assert philox([0, 0, 0, 0], [0, 0]) == [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]
    assert philox([0xffffffff] * 4, [0xffffffff] * 2) == [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]
    assert philox([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0]) == [
        0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]


def draw(seed, purpose, camera, pixel, index):
    return philox([(purpose << 8) | (camera & 0xFF), pixel, index & 0xFFFFFFFF, index >> 32],
                  [seed & 0xFFFFFFFF, seed >> 32])


@pytest.mark.parametrize("purpose,origin,word2", [
    (0x0002, 0xcce173ff, 0xa3c44c71),  # kSensorMap
    (0x0010, 0x49a23b22, 0xcaf3d5e0),  # kNoise
    (0x0011, 0xe1acca2d, 0x36d4ec5e),  # kNoiseInit
    (0x0022, 0x1f376e45, 0xe68cd7b4),  # kReadoutPhase
])
def test_draw_layout_pinned_words(purpose, origin, word2):
    """The reference's pinned first words per purpose (tests/cpp/test_philox.cpp:33-70)."""
    assert draw(0, purpose, 0, 0, 0)[0] == origin
    assert draw(0x0123456789abcdef, purpose, 2, 307199, 65536)[2] == word2


def test_constant_tables_rederive():
    """pmath_tables.h and orc_tables.h are exactly what gen_constants.py derives from the definitions."""
    gen = os.path.join(ROOT, "ballasim_amd", "csrc", "core", "gen_constants.py")
    out = subprocess.run([sys.executable, gen], capture_output=True, text=True, check=True).stdout
    assert out == open(os.path.join(ROOT, "ballasim_amd", "csrc", "core", "pmath_tables.h")).read()
    out_c = subprocess.run([sys.executable, gen, "--c"], capture_output=True, text=True, check=True).stdout
    assert out_c == open(os.path.join(ROOT, "oracle", "orc_tables.h")).read()


def test_gauss_legendre_rule_integrates_polynomials():
    """The GL8 nodes and halved weights integrate x^k on [0, 1] exactly to k = 15."""
    sys.path.insert(0, os.path.join(ROOT, "ballasim_amd", "csrc", "core"))
    import gen_constants as g
    rule = g.gauss_legendre8()
    for k in range(16):
        s = sum(w * (((1 - xi) / 2) ** k + ((1 + xi) / 2) ** k) for xi, w in rule)
        assert abs(s - 1.0 / (k + 1)) < 1e-15


def test_cody_waite_split_constants():
    """ln 2 split: hi has 16 (fp32) / 32 (fp64) significant bits and hi + lo = ln 2."""
    from decimal import Decimal, getcontext
    getcontext().prec = 50
    ln2 = sum(Decimal(1) / (k * Decimal(2) ** k) for k in range(1, 180))
    hi32, lo32 = np.float32(float.fromhex("0x1.62e400p-1")), np.float32(float.fromhex("0x1.7f7d1cp-20"))
    assert abs(Decimal(float(hi32)) + Decimal(float(lo32)) - ln2) < Decimal(2) ** -43
    m, _ = math.frexp(float(hi32))
    assert (int(m * 2 ** 16) == m * 2 ** 16)
    hi64 = float.fromhex("0x1.62e42fee00000p-1")
    m, _ = math.frexp(hi64)
    assert int(m * 2 ** 32) == m * 2 ** 32


def ulp_err_f32(got, exact):
    got = np.float32(got)
    exact = float(exact)
    if exact == 0:
        return abs(float(got))
    ulp = np.spacing(np.float32(abs(exact)))
    return abs(float(got) - exact) / float(ulp)


def test_expf_logf_expm1f_ulp():
    lib = O.oracle_lib()
    rng = np.random.default_rng(0)
    xs = np.concatenate([rng.uniform(-87, 88, 3000), rng.uniform(-1, 1, 2000)]).astype(np.float32)
    assert max(ulp_err_f32(lib.orc_expf(float(x)), math.exp(float(x))) for x in xs) <= 3
    ys = np.exp(rng.uniform(-80, 80, 4000)).astype(np.float32)
    ys = ys[ys > 0]
    assert max(ulp_err_f32(lib.orc_logf(float(y)), math.log(float(y))) for y in ys if abs(math.log(float(y))) > 1e-3) <= 3
    zs = rng.uniform(-2, 2, 3000).astype(np.float32)
    assert max(ulp_err_f32(lib.orc_expm1f(float(z)), math.expm1(float(z))) for z in zs) <= 4
    assert lib.orc_expf(-100.0) == 0.0 and lib.orc_logf(1.0) == 0.0


def test_fp64_exp_log_sincos():
    lib = O.oracle_lib()
    rng = np.random.default_rng(1)
    for x in rng.uniform(-700, 700, 500):
        assert abs(lib.orc_exp(x) - math.exp(x)) <= 2 * math.ulp(math.exp(x))
    for x in np.exp(rng.uniform(-700, 700, 500)):
        assert abs(lib.orc_log(x) - math.log(x)) <= 2 * math.ulp(math.log(x)) + 1e-300
    s, c = C.c_double(), C.c_double()
    for x in rng.uniform(-1000, 1000, 500):
        lib.orc_sincos(x, C.byref(s), C.byref(c))
        assert abs(s.value - math.sin(x)) <= 2 * math.ulp(1.0) and abs(c.value - math.cos(x)) <= 2 * math.ulp(1.0)


def test_special_functions():
    scipy_special = pytest.importorskip("scipy.special")
    lib = O.oracle_lib()
    for x in np.concatenate([np.linspace(0, 5, 200), np.geomspace(5, 1e6, 100)]):
        e = float(scipy_special.erfcx(x))
        assert abs(lib.orc_erfcxf(float(x)) - e) <= 3e-6 * e
    for x in np.concatenate([np.linspace(-5, 5, 400), np.geomspace(5, 1e7, 50)]):
        d = float(scipy_special.dawsn(x))
        assert abs(lib.orc_dawsonf(float(x)) - d) <= 3e-6 * abs(d) + 1e-12
    # Residual exponentials: R_j(y) = e^-y - sum_{i<j} (-y)^i/i!, against fp64 series.
    for j in range(5):
        for y in np.concatenate([np.linspace(1e-4, 1, 40), np.linspace(1, 30, 40)]):
            exact = math.exp(-y) - sum((-y) ** i / math.factorial(i) for i in range(j))
            if j >= 1 and y <= 1:
                exact = sum((-y) ** i / math.factorial(i) for i in range(j, j + 25))
            got = lib.orc_resid_expf(j, float(y))
            assert abs(got - exact) <= 2e-6 * abs(exact) + 1e-30, (j, y)


def test_oracle_log_monotone_on_every_float():
    """orc_logf never decreases from one positive float to the next (all 2^31 - 2^23 of them), and its
    bits hash to the value recorded in the golden manifest (the device log is checked against it)."""
    import json
    import os
    man = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "manifest.json")))["log_check"]
    v = np.zeros(1, np.uint64)
    h = np.zeros(1, np.uint64)
    O.oracle_lib().orc_log_check(os.cpu_count() or 1, v.ctypes.data_as(C.c_void_p), h.ctypes.data_as(C.c_void_p))
    assert int(v[0]) == 0 == man["violations"]
    assert "0x%016x" % int(h[0]) == man["hash"]
