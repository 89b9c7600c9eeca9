"""CPU restatement of the output stage (test infrastructure: only tests/ use it).

Follows the reference's engine output stage: crop (src/engine/output.h:36-46), reported times
(src/sensor/readout.h:22-26), S1 labels (src/sensor/output_order.h:22-26), the order of three stable
LSD passes by t_cross, (x, y) and t_reported (output_order.h:27-43, 98-115), which is the lexicographic
order (t_reported, x, y, t_cross), and the window cut (output_order.h:50-56, engine_impl.h:1818-1840).
Pinned against the reference's own functions compiled from output_order.h and readout.h
(oracle/_ref, tests/test_output_oracle.py) and, where the reference is absent, by the lexicographic
identity with numpy's lexsort."""
import numpy as np


def floor_div(a, b):
    return np.floor_divide(a, b)


def reported_time(t_cross, quantum=0, delay=0, phase=0, offset=0, drift_ppm=0.0):
    t = np.asarray(t_cross, np.int64) + delay
    if quantum > 0:
        t = phase + quantum * floor_div(t - phase, quantum)
    # llround of drift 1e-6 t (halves away from zero)
    d = drift_ppm * 1e-6 * t.astype(np.float64)
    r = np.where(d >= 0, np.floor(d + 0.5), np.ceil(d - 0.5)).astype(np.int64)
    return t + offset + r


def output_cut(t_end, offset=0, drift_ppm=0.0):
    return int(reported_time(np.array([t_end]), 0, 0, 0, offset, drift_ppm)[0])


def label_s1(c3):
    return np.array([0, 2, 3, 4], np.uint8)[np.asarray(c3) & 3]


def output_order(t_cross, x, y, t_reported):
    """The permutation into (t_reported, x, y, t_cross) order."""
    return np.lexsort((np.asarray(t_cross), np.asarray(y), np.asarray(x), np.asarray(t_reported)))


class OutputOracle:
    """Window-by-window release of a camera's events (held buffer and cut)."""

    def __init__(self, width, height, readout=None, crop=None, t_offset_us=0):
        self.w, self.h = width, height
        self.r = dict(quantum=0, delay=0, phase=0, offset=0, drift_ppm=0.0)
        self.r.update(readout or {})
        self.crop = crop
        self.t_offset_us = t_offset_us
        self.held = {k: np.zeros(0, dt) for k, dt in self._cols().items()}

    @staticmethod
    def _cols():
        return {"x": np.uint16, "y": np.uint16, "p": np.uint8, "t_reported": np.int64, "t_cross": np.int64,
                "label": np.uint8, "flags": np.uint16}

    def _release(self, ev, cut):
        order = output_order(ev["t_cross"], ev["x"], ev["y"], ev["t_reported"])
        ev = {k: v[order] for k, v in ev.items()}
        m = int(np.searchsorted(ev["t_reported"], cut, side="left"))
        self.held = {k: v[m:] for k, v in ev.items()}
        out = {k: v[:m] for k, v in ev.items()}
        us = floor_div(out["t_reported"], 1000)
        out["t"] = us - self.t_offset_us
        out["t_ns_rem"] = (out["t_reported"] - 1000 * us).astype(np.uint16)
        return out

    def window(self, raw, t_w, window_ns=1 << 23):
        """raw: a structured array of RawEvents (pixel, polarity, label, flags, t_cross)."""
        x = (raw["pixel"] % self.w).astype(np.int64)
        y = (raw["pixel"] // self.w).astype(np.int64)
        keep = np.ones(len(raw), bool)
        if self.crop is not None:
            x0, y0, x1, y1 = self.crop
            keep = (x >= x0) & (x < x1) & (y >= y0) & (y < y1)
        ev = {"x": x[keep].astype(np.uint16), "y": y[keep].astype(np.uint16),
              "p": raw["polarity"][keep].astype(np.uint8),
              "t_reported": reported_time(raw["t_cross"][keep], **self.r),
              "t_cross": raw["t_cross"][keep].astype(np.int64), "label": label_s1(raw["label"][keep]),
              "flags": raw["flags"][keep].astype(np.uint16)}
        ev = {k: np.concatenate([self.held[k], ev[k]]) for k in ev}
        return self._release(ev, output_cut(t_w + window_ns, self.r["offset"], self.r["drift_ppm"]))

    def flush(self):
        return self._release(self.held, np.iinfo(np.int64).max)
