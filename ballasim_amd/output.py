"""The output stage of a camera (include/bsx/bsx_output.h) and the C4 event file.

OutputStage orders a camera's events window by window on the device (t_reported, x, y, t_cross) and
releases those reported before each window's end; EventFile appends the released columns in the C4
layout of the reference's writer (python/ballasim/engine/writer.py:1-40): /events/{x u16, y u16, p u8,
t i64 (us on the sensor clock relative to /t_offset), t_ns_rem u16}, /t_offset, /ms_to_idx, and with
ground truth off /gt/{label, t_cross_ns, flags}. h5py is not installed here, so the same datasets are
written as arrays of one .npz file (names as the HDF5 paths)."""
import ctypes as C
import os

import numpy as np

from . import _abi


class Crop(C.Structure):
    _fields_ = [("x0", C.c_int32), ("y0", C.c_int32), ("x1", C.c_int32), ("y1", C.c_int32)]


class OutputEvents(C.Structure):
    _fields_ = [("count", C.c_size_t), ("x", C.c_void_p), ("y", C.c_void_p), ("p", C.c_void_p), ("t", C.c_void_p),
                ("t_ns_rem", C.c_void_p), ("t_reported_ns", C.c_void_p), ("t_cross_ns", C.c_void_p),
                ("label", C.c_void_p), ("flags", C.c_void_p), ("held", C.c_size_t)]


_COLUMNS = (("x", np.uint16), ("y", np.uint16), ("p", np.uint8), ("t", np.int64), ("t_ns_rem", np.uint16),
            ("t_reported_ns", np.int64), ("t_cross_ns", np.int64), ("label", np.uint8), ("flags", np.uint16))


def _stream(stream):
    return C.c_void_p(stream.cuda_stream if hasattr(stream, "cuda_stream") else int(stream or 0))


class OutputStage:
    def __init__(self, camera, crop=None, t_offset_us=0):
        self._lib = _abi.load_library()
        self._camera = camera
        self._h = C.c_void_p()
        c = Crop(*crop) if crop is not None else None
        _abi.check(self._lib.bsx_output_create(camera._h, C.byref(c) if c is not None else None, int(t_offset_us),
                                               C.byref(self._h)))
        self.held = 0

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self._lib.bsx_output_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _columns(self, out):
        self.held = out.held
        res = {}
        for name, dt in _COLUMNS:
            a = np.empty(out.count, dt)
            if out.count:
                C.memmove(a.ctypes.data, getattr(out, name), a.nbytes)
            res[name] = a
        return res

    def window(self, events, t_w, stream=0):
        """events: (data_ptr, count) of a window's raw device events; returns the released columns."""
        ev = _abi.DeviceEvents(*events)
        out = OutputEvents()
        _abi.check(self._lib.bsx_output_window(self._h, C.byref(ev), int(t_w), _stream(stream), C.byref(out)))
        return self._columns(out)

    def flush(self, stream=0):
        out = OutputEvents()
        _abi.check(self._lib.bsx_output_flush(self._h, _stream(stream), C.byref(out)))
        return self._columns(out)


class EventFile:
    """Appends a camera's released events (in output order) and writes the C4 datasets on close."""

    FORMAT = "ballasim.events/1"

    def __init__(self, path, camera, width, height, t_offset_us=0):
        self.path = os.fspath(path)
        self.camera, self.width, self.height = camera, int(width), int(height)
        self.t_offset_us = int(t_offset_us)
        self.parts = []
        self.last_t = None

    def append(self, cols):
        if len(cols["t"]) == 0:
            return
        if self.last_t is not None and cols["t"][0] < self.last_t:
            raise ValueError("EventFile.append: events must arrive in output order")
        if np.any(np.diff(cols["t"]) < 0):
            raise ValueError("EventFile.append: events must arrive in output order")
        self.last_t = int(cols["t"][-1])
        self.parts.append({k: np.asarray(v).copy() for k, v in cols.items()})

    def datasets(self):
        cat = {k: np.concatenate([p[k] for p in self.parts]) if self.parts else np.zeros(0, dt) for k, dt in
               (("x", np.uint16), ("y", np.uint16), ("p", np.uint8), ("t", np.int64), ("t_ns_rem", np.uint16),
                ("label", np.uint8), ("t_cross_ns", np.int64), ("flags", np.uint16))}
        t = cat["t"]
        n_ms = int(t[-1] // 1000) + 1 if len(t) else 0
        ms_to_idx = np.searchsorted(t, 1000 * np.arange(n_ms, dtype=np.int64), side="left").astype(np.uint64)
        out = {f"events/{k}": cat[k] for k in ("x", "y", "p", "t", "t_ns_rem")}
        out.update({"gt/label": cat["label"], "gt/t_cross_ns": cat["t_cross_ns"], "gt/flags": cat["flags"]})
        out["t_offset"] = np.array(self.t_offset_us, np.int64)
        out["ms_to_idx"] = ms_to_idx
        out["attrs/resolution"] = np.array([self.width, self.height], np.int64)
        out["attrs/camera"] = np.array(self.camera)
        out["attrs/format"] = np.array(self.FORMAT)
        return out

    def close(self):
        tmp = self.path + ".partial.npz"
        np.savez(tmp, **self.datasets())
        os.replace(tmp, self.path)
