"""The reference's own host window engine on a bsx scene (TEST INFRASTRUCTURE): oracle/_ref/libwindow_ref.so,
built from src/engine/window.cpp and par_host.cpp with the entry layer window_ref.cpp (oracle/Makefile,
target window) where the reference is present, and shipped with the tree. It renders through the bsx scene
the caller created, so it needs the GPU."""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
WINDOW_LIB = os.path.join(HERE, "_ref", "libwindow_ref.so")

_lib = None


def available():
    return os.path.exists(WINDOW_LIB)


def _load():
    global _lib
    if _lib is None:
        import ballasim_amd  # noqa: F401  (loads libbsx.so first, so both share it)
        lib = C.CDLL(WINDOW_LIB)
        P = C.c_void_p
        lib.wr_run.restype = P
        lib.wr_run.argtypes = [P, P, C.c_int64, C.c_float, C.c_float, C.c_int32, C.c_uint32, C.c_int32, P, P, C.c_size_t,
                               C.c_char_p, C.c_size_t]
        lib.wr_samples.restype = C.c_size_t
        lib.wr_samples.argtypes = [P]
        lib.wr_fetch.argtypes = [P, P, P, P, P]
        lib.wr_free.argtypes = [P]
        _lib = lib
    return _lib


def sample_window(scene, view, t_w, pixels, levels, tol, dark, max_refine_levels=4, min_split=2048, signature=True):
    """The reference engine's runs of one window for the active pixels (in that order) at the given lattice
    levels: (offsets (n + 1,), t (S,), irr (S,), stats dict). Each run starts with the pixel's carried sample
    at t_w, as the reference's runs do."""
    lib = _load()
    px = np.ascontiguousarray(pixels, np.uint32)
    lv = np.ascontiguousarray(levels, np.uint8)
    err = C.create_string_buffer(1024)
    h = lib.wr_run(scene._h, C.byref(view), int(t_w), float(tol), float(dark), int(max_refine_levels), int(min_split),
                   int(bool(signature)), px.ctypes.data_as(C.c_void_p), lv.ctypes.data_as(C.c_void_p), len(px), err,
                   len(err))
    if not h:
        raise RuntimeError("window_ref: " + err.value.decode())
    try:
        s = lib.wr_samples(h)
        offsets = np.empty(len(px) + 1, np.int64)
        t = np.empty(max(s, 1), np.int64)
        irr = np.empty(max(s, 1), np.float32)
        st = np.zeros(12, np.uint64)
        lib.wr_fetch(h, offsets.ctypes.data_as(C.c_void_p), t.ctypes.data_as(C.c_void_p),
                     irr.ctypes.data_as(C.c_void_p), st.ctypes.data_as(C.c_void_p))
    finally:
        lib.wr_free(h)
    stats = {"render_calls": int(st[0]), "flagged_signature": int(st[1]), "flagged_midpoint": int(st[2]),
             "refine_cap_hits": int(st[3]), "samples_by_round": [int(x) for x in st[4:] if x]}
    return offsets, t[:s], irr[:s], stats
