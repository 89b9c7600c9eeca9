"""ctypes access to the oracle libraries (TEST INFRASTRUCTURE: only tests/, smoke() and bench.py's
cpu_baseline leg use this module, as the checker or the timed CPU baseline).

- OracleModel: the plain-C restatement (oracle/build/liboracle.so).
- RefModel: the reference's own CPU path compiled from its sources (oracle/_ref/libballa_ref.so),
  present only where the reference was available to build it.
Both take the same inputs as the product's SensorModel and return events as numpy arrays.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_LIB = os.path.join(HERE, "build", "liboracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libballa_ref.so")

EVENT_DTYPE = np.dtype([("pixel", "<u4"), ("polarity", "u1"), ("label", "u1"), ("flags", "<u2"),
                        ("t_cross", "<i8"), ("ds", "<f4"), ("dn", "<f4"), ("dl", "<f4"), ("l_first", "<f4")])
STATE_DTYPE = np.dtype([("z", "<f8"), ("x1", "<f8"), ("s2", "<f8"), ("s3", "<f8"), ("i_scale", "<f8"),
                        ("tv", "<f8"), ("s3_ref", "<f8"), ("n3_ref", "<f8"), ("dr_ref", "<f8"),
                        ("t_ref", "<i8"), ("blind_until", "<i8"), ("t_search", "<i8"), ("flags", "<u2"),
                        ("held", "<f4", (8,)), ("shot_next", "<i8"), ("shot_index", "<u4"),
                        ("nk", "<f8", (3,)), ("nk1", "<f8", (3,)), ("ln_a1", "<f8"), ("cell", "<i8"),
                        ("have_k1", "u1")], align=True)
# The reference's PixelState (200 B; src/sensor/pixel.h:105-121).
REF_STATE_DTYPE = np.dtype([("z", "<f8"), ("x1", "<f8"), ("s2", "<f8"), ("s3", "<f8"), ("i_scale", "<f8"),
                            ("tv", "<f8"), ("nk", "<f8", (3,)), ("nk1", "<f8", (3,)), ("ln_a1", "<f8"),
                            ("cell", "<i8"), ("t_ref", "<i8"), ("s3_ref", "<f8"), ("n3_ref", "<f8"),
                            ("dr_ref", "<f8"), ("blind_until", "<i8"), ("t_search", "<i8"),
                            ("held", "<f4", (8,)), ("flags", "<u2"), ("have_k1", "u1"), ("reserved", "u1"),
                            ("rho_m1", "<f4")])
assert REF_STATE_DTYPE.itemsize == 200


def build():
    """Compile the restatement and, where the reference is present, the reference library."""
    subprocess.run(["make", "-s", "-f", os.path.join(HERE, "Makefile"), "all"], check=True)


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None and a.size else None


STAT_NAMES_ORACLE = ("substeps", "substep_attempts", "settled_substeps", "leaf_intervals", "fallbacks",
                     "input_clamps", "quarantines", "pieces", "cells", "refinements")
STAT_NAMES_REF = ("substeps", "substep_attempts", "settled_substeps", "cells", "refinements", "leaf_intervals",
                  "fallbacks", "input_clamps", "quarantines", "pieces")

_oracle = None
_ref = None


def oracle_lib():
    global _oracle
    if _oracle is None:
        if not os.path.exists(ORACLE_LIB):
            build()
        lib = C.CDLL(ORACLE_LIB)
        P = C.c_void_p
        lib.orc_create.restype = P
        lib.orc_create.argtypes = [P, C.c_uint32, C.c_uint32, C.c_uint64, C.c_uint32, C.c_int32, C.c_int32,
                                   C.c_char_p, C.c_size_t]
        lib.orc_destroy.argtypes = [P]
        lib.orc_draw_maps.argtypes = [P, C.c_uint64, C.c_uint32, C.c_uint32, P, P, P, P, P, P, P]
        lib.orc_init.argtypes = [P, C.c_int64, P]
        lib.orc_advance_runs.argtypes = [P, P, C.c_size_t, P, P, C.c_size_t, P, P, C.c_int64, C.c_char_p, C.c_size_t]
        lib.orc_event_count.restype = C.c_size_t
        lib.orc_event_count.argtypes = [P]
        lib.orc_events.restype = P
        lib.orc_events.argtypes = [P]
        lib.orc_reported.argtypes = [P, P]
        lib.orc_maps.argtypes = [P, P, P, P, P, P, P, P]
        lib.orc_states.argtypes = [P, P]
        lib.orc_get_stats.argtypes = [P, P]
        lib.orc_time.restype = C.c_int64
        lib.orc_time.argtypes = [P]
        lib.orc_philox.argtypes = [P, P, P]
        lib.orc_noise_tables.argtypes = [P, P, P, P, P]
        lib.orc_normals3.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint64, P, P]
        for n in ("orc_expf", "orc_logf", "orc_expm1f", "orc_erfcxf", "orc_dawsonf"):
            getattr(lib, n).restype = C.c_float
            getattr(lib, n).argtypes = [C.c_float]
        lib.orc_resid_expf.restype = C.c_float
        lib.orc_resid_expf.argtypes = [C.c_int, C.c_float]
        for n in ("orc_exp", "orc_log"):
            getattr(lib, n).restype = C.c_double
            getattr(lib, n).argtypes = [C.c_double]
        lib.orc_sincos.argtypes = [C.c_double, P, P]
        lib.orc_sincosf.argtypes = [C.c_float, P, P]
        lib.orc_z_step.restype = C.c_float
        lib.orc_z_step.argtypes = [C.c_float] * 5
        lib.orc_log_check.argtypes = [C.c_int, P, P]
        lib.orc_xc_eval.restype = C.c_float
        lib.orc_xc_eval.argtypes = [P, P, P, C.c_int, C.c_float]
        lib.orc_xc_eval_n.argtypes = [P, P, P, C.c_int, P, C.c_size_t, P]
        lib.orc_xc_critical.argtypes = [P, P, P, C.c_int, C.c_float, C.c_float, P]
        lib.orc_xc_search.argtypes = [P, P, P, C.c_int, C.c_int64, C.c_int64, C.c_int64, C.c_float, C.c_float,
                                      P, P]
        _oracle = lib
    return _oracle


def draw_maps(params, seed, camera, n):
    """The oracle's frozen maps of n pixels."""
    lib = oracle_lib()
    arrs = [np.zeros(n) for _ in range(5)]
    hot, rate = np.zeros(n, np.uint8), np.zeros(n)
    lib.orc_draw_maps(C.byref(params), seed, camera, n, *[_ptr(a) for a in arrs], _ptr(hot), _ptr(rate))
    out = dict(zip(("theta_on", "theta_off", "leak", "dark_fA", "phi0"), arrs))
    out.update(hot=hot, hot_rate_hz=rate)
    return out


def ref_available():
    return os.path.exists(REF_LIB)


def ref_lib():
    global _ref
    if _ref is None:
        lib = C.CDLL(REF_LIB)
        P = C.c_void_p
        lib.ref_create.restype = P
        lib.ref_create.argtypes = [P, C.c_uint32, C.c_uint32, C.c_uint64, C.c_uint32, C.c_int32, C.c_int32,
                                   C.c_int32, C.c_char_p, C.c_size_t]
        lib.ref_destroy.argtypes = [P]
        lib.ref_init.argtypes = [P, C.c_int64, P, C.c_char_p, C.c_size_t]
        lib.ref_advance_runs.argtypes = [P, P, C.c_size_t, P, P, C.c_size_t, P, P, C.c_int64, C.c_char_p,
                                         C.c_size_t]
        lib.ref_event_count.restype = C.c_size_t
        lib.ref_event_count.argtypes = [P]
        lib.ref_events.argtypes = [P, P]
        lib.ref_reported.argtypes = [P, P]
        lib.ref_maps.argtypes = [P, P, P, P, P, P, P, P]
        lib.ref_states.argtypes = [P, P]
        lib.ref_output_order.argtypes = [P, P, P, P, C.c_size_t, C.c_uint32, P]
        lib.ref_readout.argtypes = [C.c_int64, C.c_int64, C.c_int64, C.c_int64, C.c_double, P, C.c_size_t, P,
                                    C.c_int64, P]
        lib.ref_label_s1.argtypes = [P, C.c_size_t, P]
        lib.ref_stats.argtypes = [P, P]
        lib.ref_noise_tables.argtypes = [P, P, P, P, P]
        lib.ref_normals3.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint64, P, P]
        _ref = lib
    return _ref


def noise_tables(params, lib="oracle"):
    """The noise tables of params (frozen diffusion) from the oracle or the compiled reference:
    (fp64 rows, fp32 rows), each (n_nodes, stride)."""
    L = oracle_lib() if lib == "oracle" else ref_lib()
    fn = L.orc_noise_tables if lib == "oracle" else L.ref_noise_tables
    n, s = C.c_int32(), C.c_int32()
    if fn(C.byref(params), None, None, C.byref(n), C.byref(s)) != 0:
        raise ValueError("noise tables: bad leaf level")
    d64 = np.zeros((n.value, s.value), np.float64)
    d32 = np.zeros((n.value, s.value), np.float32)
    fn(C.byref(params), _ptr(d64), _ptr(d32), C.byref(n), C.byref(s))
    return d64, d32


def normals3(seed, purpose, camera, pixel, index, lib="oracle"):
    """normals3 of one key, (fp64 triple, fp32 triple), from the oracle or the compiled reference."""
    L = oracle_lib() if lib == "oracle" else ref_lib()
    fn = L.orc_normals3 if lib == "oracle" else L.ref_normals3
    d = np.zeros(3, np.float64)
    f = np.zeros(3, np.float32)
    fn(seed, purpose, camera, pixel, index, _ptr(d), _ptr(f))
    return d, f


class _Base:
    def advance_runs(self, ids, t, irr, offsets, t_end):
        """Ragged runs: row i of pixel ids[i] is samples offsets[i]:offsets[i + 1] of (t, irr)."""
        ids = np.ascontiguousarray(ids, np.uint32)
        t = np.ascontiguousarray(t, np.int64)
        irr = np.ascontiguousarray(irr, np.float32)
        offsets = np.asarray(offsets, np.int64)
        start = np.ascontiguousarray(offsets[:-1])
        count = np.ascontiguousarray(np.diff(offsets).astype(np.int32))
        return self._advance(ids, t, irr, start, count, int(t_end))

    def advance_padded(self, ids, t, irr, t_end, counts=None):
        """Rectangular runs (n, m) with optional counts, as the reference's advance()."""
        ids = np.ascontiguousarray(ids, np.uint32)
        t = np.asarray(t, np.int64)
        n, m = t.shape
        start = np.arange(n, dtype=np.int64) * m
        count = np.full(n, m, np.int32) if counts is None else np.ascontiguousarray(counts, np.int32)
        return self._advance(ids, np.ascontiguousarray(t.reshape(-1)),
                             np.ascontiguousarray(np.asarray(irr, np.float32).reshape(n * m, self.groups)), start, count,
                             int(t_end))


class OracleModel(_Base):
    def __init__(self, params, width, height, seed, groups=1, camera=0, num_threads=0):
        self.lib = oracle_lib()
        self.width, self.height, self.groups = width, height, groups
        err = C.create_string_buffer(512)
        self.h = self.lib.orc_create(C.byref(params), width, height, seed, camera, groups, num_threads, err, 512)
        if not self.h:
            raise ValueError(err.value.decode())

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.orc_destroy(self.h)
            self.h = None

    def init(self, t0, irr0):
        irr0 = np.ascontiguousarray(irr0, np.float32)
        if self.lib.orc_init(self.h, int(t0), _ptr(irr0)) != 0:
            raise ValueError("init: non-finite irradiance")

    def _advance(self, ids, t, irr, start, count, t_end):
        err = C.create_string_buffer(512)
        r = self.lib.orc_advance_runs(self.h, _ptr(ids), ids.size, _ptr(t), _ptr(irr), t.size, _ptr(start),
                                      _ptr(count), t_end, err, 512)
        if r != 0:
            raise ValueError(err.value.decode())
        n = self.lib.orc_event_count(self.h)
        out = np.zeros(n, EVENT_DTYPE)
        if n:
            C.memmove(out.ctypes.data, self.lib.orc_events(self.h), n * EVENT_DTYPE.itemsize)
        rep = np.zeros(n, np.int64)
        if n:
            self.lib.orc_reported(self.h, _ptr(rep))
        return out, rep

    def maps(self):
        P = self.width * self.height
        arrs = [np.zeros(P) for _ in range(5)]
        hot, rate = np.zeros(P, np.uint8), np.zeros(P)
        self.lib.orc_maps(self.h, *[_ptr(a) for a in arrs], _ptr(hot), _ptr(rate))
        out = dict(zip(("theta_on", "theta_off", "leak", "dark_fA", "phi0"), arrs))
        out.update(hot=hot, hot_rate_hz=rate)
        return out

    def states(self):
        out = np.zeros(self.width * self.height, STATE_DTYPE)
        self.lib.orc_states(self.h, _ptr(out))
        return out

    def stats(self):
        v = np.zeros(len(STAT_NAMES_ORACLE), np.uint64)
        self.lib.orc_get_stats(self.h, _ptr(v))
        return dict(zip(STAT_NAMES_ORACLE, (int(x) for x in v)))


class RefModel(_Base):
    def __init__(self, params, width, height, seed, groups=1, camera=0, num_threads=0, fp32=True):
        self.lib = ref_lib()
        self.width, self.height, self.groups = width, height, groups
        err = C.create_string_buffer(512)
        self.h = self.lib.ref_create(C.byref(params), width, height, seed, camera, groups, num_threads,
                                     1 if fp32 else 0, err, 512)
        if not self.h:
            raise ValueError(err.value.decode())

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.ref_destroy(self.h)
            self.h = None

    def init(self, t0, irr0):
        irr0 = np.ascontiguousarray(irr0, np.float32)
        err = C.create_string_buffer(512)
        if self.lib.ref_init(self.h, int(t0), _ptr(irr0), err, 512) != 0:
            raise ValueError(err.value.decode())

    def _advance(self, ids, t, irr, start, count, t_end):
        err = C.create_string_buffer(512)
        r = self.lib.ref_advance_runs(self.h, _ptr(ids), ids.size, _ptr(t), _ptr(irr), t.size, _ptr(start),
                                      _ptr(count), t_end, err, 512)
        if r != 0:
            raise ValueError(err.value.decode())
        n = self.lib.ref_event_count(self.h)
        out = np.zeros(n, EVENT_DTYPE)
        if n:
            self.lib.ref_events(self.h, _ptr(out))
        rep = np.zeros(n, np.int64)
        if n:
            self.lib.ref_reported(self.h, _ptr(rep))
        return out, rep

    def maps(self):
        P = self.width * self.height
        arrs = [np.zeros(P) for _ in range(5)]
        hot, rate = np.zeros(P, np.uint8), np.zeros(P)
        self.lib.ref_maps(self.h, *[_ptr(a) for a in arrs], _ptr(hot), _ptr(rate))
        out = dict(zip(("theta_on", "theta_off", "leak", "dark_fA", "phi0"), arrs))
        out.update(hot=hot, hot_rate_hz=rate)
        return out

    def states(self):
        out = np.zeros(self.width * self.height, REF_STATE_DTYPE)
        self.lib.ref_states(self.h, _ptr(out))
        return out

    def stats(self):
        v = np.zeros(10, np.uint64)
        self.lib.ref_stats(self.h, _ptr(v))
        return dict(zip(STAT_NAMES_REF, (int(x) for x in v)))
