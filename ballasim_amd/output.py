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


STREAM_EVENT = np.dtype([("t_reported_ns", np.int64), ("x", np.uint16), ("y", np.uint16), ("p", np.uint8),
                         ("label", np.uint8), ("flags", np.uint16)])
assert STREAM_EVENT.itemsize == 16

_SINK = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_size_t, C.c_uint64, C.c_void_p)


class StreamStats(C.Structure):
    _fields_ = [("chunks", C.c_uint64), ("events", C.c_uint64), ("bytes", C.c_uint64), ("wait_ms", C.c_double),
                ("copy_ms", C.c_double)]


class EventStream:
    """Released events streamed to the host in double-buffered chunks of pinned memory
    (include/bsx/bsx_output.h, bsx_event_stream_*). sink(records, index) is called on the stream's
    delivery thread with chunk `index` as a numpy array of STREAM_EVENT records that views the pinned
    chunk: it is valid only until sink returns, so a sink that keeps it copies it. A sink that raises
    stops the stream (the next call raises RuntimeError)."""

    def __init__(self, sink, chunk_bytes=256 << 20):
        self._lib = _abi.load_library()
        self._user_sink = sink
        self.error = None

        def deliver(ptr, count, index, _user):
            try:
                buf = (C.c_char * (count * 16)).from_address(ptr) if count else b""
                self._user_sink(np.frombuffer(buf, STREAM_EVENT, count), int(index))
                return 0
            except BaseException as e:  # reported by the next call
                self.error = e
                return 1

        self._sink = _SINK(deliver)  # kept alive as long as the stream
        self._h = C.c_void_p()
        _abi.check(self._lib.bsx_event_stream_create(int(chunk_bytes), C.cast(self._sink, C.c_void_p), None,
                                                     C.byref(self._h)))

    def _check(self, rc):
        if rc != 0 and self.error is not None:
            raise RuntimeError(f"EventStream: the sink raised {self.error!r}") from self.error
        _abi.check(rc)

    def window(self, stage, events, t_w, stream=0):
        """stage.window with the released events streamed: returns their count."""
        ev = _abi.DeviceEvents(*events)
        n, held = C.c_size_t(), C.c_size_t()
        self._check(self._lib.bsx_output_window_to_stream(stage._h, C.byref(ev), int(t_w), _stream(stream), self._h,
                                                          C.byref(n), C.byref(held)))
        stage.held = held.value
        return n.value

    def flush(self, stage, stream=0):
        n = C.c_size_t()
        self._check(self._lib.bsx_output_flush_to_stream(stage._h, _stream(stream), self._h, C.byref(n)))
        stage.held = 0
        return n.value

    def finish(self, stream=0):
        """Sends the partial chunk; returns once every chunk has been delivered."""
        self._check(self._lib.bsx_event_stream_finish(self._h, _stream(stream)))

    def stats(self):
        st = StreamStats()
        _abi.check(self._lib.bsx_event_stream_stats(self._h, C.byref(st)))
        return {k: getattr(st, k) for k, _ in StreamStats._fields_}

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self._lib.bsx_event_stream_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


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

    def datasets(self, t_end_reported_ns=None):
        """The file's datasets. /ms_to_idx entry k is the index of the first event with t >= 1000 k; the
        entries run to the last event's millisecond and, given t_end_reported_ns (the run's end on the
        sensor clock), on to that end's, entries past the last event holding the event count
        (python/ballasim/engine/writer.py:102-107,128-130)."""
        cat = {k: np.concatenate([p[k] for p in self.parts]) if self.parts else np.zeros(0, dt) for k, dt in
               (("x", np.uint16), ("y", np.uint16), ("p", np.uint8), ("t", np.int64), ("t_ns_rem", np.uint16),
                ("label", np.uint8), ("t_cross_ns", np.int64), ("flags", np.uint16))}
        t = cat["t"]
        n_ms = int(t[-1] // 1000) + 1 if len(t) else 0
        if t_end_reported_ns is not None:
            n_ms = max(n_ms, (int(t_end_reported_ns) // 1000 - self.t_offset_us) // 1000 + 1)
        ms_to_idx = np.searchsorted(t, 1000 * np.arange(n_ms, dtype=np.int64), side="left").astype(np.uint64)
        out = {f"events/{k}": cat[k] for k in ("x", "y", "p", "t", "t_ns_rem")}
        out.update({"gt/label": cat["label"], "gt/t_cross_ns": cat["t_cross_ns"], "gt/flags": cat["flags"]})
        out["t_offset"] = np.array(self.t_offset_us, np.int64)
        out["ms_to_idx"] = ms_to_idx
        out["attrs/resolution"] = np.array([self.width, self.height], np.int64)
        out["attrs/camera"] = np.array(self.camera)
        out["attrs/format"] = np.array(self.FORMAT)
        return out

    def close(self, t_end_reported_ns=None):
        tmp = self.path + ".partial.npz"
        np.savez(tmp, **self.datasets(t_end_reported_ns))
        os.replace(tmp, self.path)
