"""The learned sample-interval proposer (include/bsx/bsx_mlp.h): a 12-64-64-1 SiLU perceptron with
BF16 weights on the MFMA units, trained and run by this library's HIP kernels."""
import ctypes as C

import numpy as np

from . import _abi

N_PARAMS = 5057
N_IN = 12


def _addr(x):
    return x.data_ptr() if hasattr(x, "data_ptr") else int(x)


def _stream(stream):
    return C.c_void_p(stream.cuda_stream if hasattr(stream, "cuda_stream") else int(stream or 0))


class Proposer:
    def __init__(self, seed=0):
        self._lib = _abi.load_library()
        self._h = C.c_void_p()
        _abi.check(self._lib.bsx_mlp_create(int(seed), C.byref(self._h)))

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self._lib.bsx_mlp_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def params(self):
        w = np.empty(N_PARAMS, np.float32)
        _abi.check(self._lib.bsx_mlp_get_params(self._h, w.ctypes.data_as(C.c_void_p)))
        return w

    @params.setter
    def params(self, w):
        w = np.ascontiguousarray(w, np.float32)
        if w.size != N_PARAMS:
            raise ValueError(f"Proposer.params: need {N_PARAMS} values")
        _abi.check(self._lib.bsx_mlp_set_params(self._h, w.ctypes.data_as(C.c_void_p)))

    def gradients(self, x, target, stream=0):
        """(gradient [5057], mean squared error) over device rows x [n x 12] (float32) and target [n]."""
        g = np.empty(N_PARAMS, np.float32)
        loss = C.c_double()
        n = x.shape[0] if hasattr(x, "shape") else None
        _abi.check(self._lib.bsx_mlp_gradients(self._h, C.c_void_p(_addr(x)), C.c_void_p(_addr(target)), int(n),
                                               _stream(stream), g.ctypes.data_as(C.c_void_p), C.byref(loss)))
        return g, loss.value

    def train(self, x, target, batch=4096, epochs=1, lr=1e-3, stream=0):
        """Adam over device rows in order; returns the mean squared error of the last pass."""
        loss = C.c_double()
        _abi.check(self._lib.bsx_mlp_train(self._h, C.c_void_p(_addr(x)), C.c_void_p(_addr(target)), int(x.shape[0]),
                                           int(batch), int(epochs), float(lr), _stream(stream), C.byref(loss)))
        return loss.value

    def predict(self, x, y, stream=0):
        """y[:] = the network's output for device rows x (both device tensors)."""
        _abi.check(self._lib.bsx_mlp_predict(self._h, C.c_void_p(_addr(x)), int(x.shape[0]), C.c_void_p(_addr(y)),
                                             _stream(stream)))

    @property
    def handle(self):
        return self._h


def tile_check(a_bits, b_bits):
    """c [32 x 32] = a [32 x 8] . b [8 x 32] through the kernels' MFMA tile (bf16 bit patterns in)."""
    lib = _abi.load_library()
    a = np.ascontiguousarray(a_bits, np.uint16)
    b = np.ascontiguousarray(b_bits, np.uint16)
    c = np.empty((32, 32), np.float32)
    _abi.check(lib.bsx_mlp_tile_check(a.ctypes.data_as(C.c_void_p), b.ctypes.data_as(C.c_void_p),
                                      c.ctypes.data_as(C.c_void_p)))
    return c


class ExamplesSpec(C.Structure):
    _fields_ = [("dark", C.c_float), ("theta_ref", C.c_float), ("theta_band", C.c_float), ("tol_log", C.c_float),
                ("t0_ns", C.c_int64), ("delta_ns", C.c_int64), ("n_dense", C.c_int32), ("max_steps", C.c_int32),
                ("seed", C.c_uint64), ("n", C.c_int64)]


class LearnedRules(C.Structure):
    _fields_ = [("plain", _abi.SamplerRules), ("tol_log", C.c_float), ("theta_band", C.c_float),
                ("theta_ref", C.c_float), ("margin", C.c_float), ("near", C.c_float), ("min_dt_ns", C.c_int64)]


def learned_examples(scene, view, x, target, dark, t0_ns=0, delta_ns=10_000, n_dense=4096, max_steps=838,
                     theta_ref=0.2, theta_band=0.14, tol_log=0.01, seed=0, stream=0):
    """Fills device tensors x [n x 12] and target [n] with training examples from dense traces of the
    scene seen by `view`; returns the number of varying pixels."""
    lib = _abi.load_library()
    spec = ExamplesSpec(dark, theta_ref, theta_band, tol_log, int(t0_ns), int(delta_ns), int(n_dense), int(max_steps),
                        int(seed), int(target.shape[0]))
    n_active = C.c_uint64()
    _abi.check(lib.bsx_scene_learned_examples(scene._h, C.byref(view), C.byref(spec), C.c_void_p(_addr(x)),
                                              C.c_void_p(_addr(target)), _stream(stream), C.byref(n_active)))
    return n_active.value


def learned_rules(plain, tol_log=0.01, theta_band=0.14, theta_ref=0.2, margin=1.0, near=0.0, min_dt_ns=4096):
    return LearnedRules(plain, tol_log, theta_band, theta_ref, margin, near, min_dt_ns)


class LearnedSampler:
    """The learned sample-time controller of one view (include/bsx/bsx_mlp.h)."""

    def __init__(self, scene, view, rules, proposer):
        self._lib = _abi.load_library()
        self._keep = (scene, proposer)
        self.view = view
        self._h = C.c_void_p()
        _abi.check(self._lib.bsx_learned_sampler_create(scene._h, C.byref(view), C.byref(rules), proposer.handle,
                                                        C.byref(self._h)))

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self._lib.bsx_learned_sampler_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def init(self, t0_ns, stream=0):
        _abi.check(self._lib.bsx_learned_sampler_init(self._h, int(t0_ns), _stream(stream)))

    def window(self, t_w, slot=0, stream=0, sensor=None):
        """(DeviceRuns, samples, proposals, fallbacks) of window (t_w, t_w + 2^23]; sensor: the SensorModel
        these runs feed, whose reference state the controller reads at the window start (or None)."""
        runs = _abi.DeviceRuns()
        total, prop, fb = C.c_uint64(), C.c_uint64(), C.c_uint64()
        sh = sensor._h if sensor is not None else None
        _abi.check(self._lib.bsx_learned_sampler_window(self._h, int(t_w), int(slot), sh, _stream(stream),
                                                        C.byref(runs), C.byref(total), C.byref(prop), C.byref(fb)))
        return runs, total.value, prop.value, fb.value


def event_error(test, ref, width):
    """Event error of a run against a reference run of the same sensor (event dicts of SensorModel):
    count mismatch = sum over pixels of |n_test - n_ref| / sum n_ref; timing = the median |t_test - t_ref|
    (ns) over the events of pixels whose polarity sequences agree (their events paired in order), with
    the share of reference events in such pixels."""
    def by_pixel(ev):
        px = ev["y"].astype(np.int64) * width + ev["x"]
        order = np.lexsort((ev["t_cross_ns"], px))
        return px[order], ev["t_cross_ns"][order], ev["p"][order]

    pa, ta, sa = by_pixel(test)
    pb, tb, sb = by_pixel(ref)
    n_pix = int(max(pa.max(initial=-1), pb.max(initial=-1))) + 1
    ca = np.bincount(pa, minlength=n_pix)
    cb = np.bincount(pb, minlength=n_pix)
    mismatch = float(np.abs(ca - cb).sum() / max(cb.sum(), 1))
    same = ca == cb
    # pixels with equal counts: compare the polarity sequences and pair the times in order
    ma = same[pa]
    mb = same[pb]
    pol_ok = np.ones(n_pix, bool)
    if ma.any():
        bad = sa[ma] != sb[mb]
        np.logical_and.at(pol_ok, pa[ma], ~bad)
    keep_a = ma & pol_ok[pa]
    keep_b = mb & pol_ok[pb]
    dt = np.abs(ta[keep_a] - tb[keep_b])
    return {"count_mismatch": mismatch, "n_test": int(ca.sum()), "n_ref": int(cb.sum()),
            "median_abs_dt_ns": float(np.median(dt)) if dt.size else float("nan"),
            "p90_abs_dt_ns": float(np.percentile(dt, 90)) if dt.size else float("nan"),
            "paired_share": float(keep_b.sum() / max(len(tb), 1))}


WINDOW_NS = 1 << 23


def train_schedule(batch):
    """The (learning rate, epochs) schedule of train_proposer: three epochs at 3e-3, 1e-3 and 3e-4 for
    4096-row steps, the rates scaled with the square root of the batch for larger steps."""
    k = (batch / 4096) ** 0.5
    return ((3e-3 * k, 1), (1e-3 * k, 1), (3e-4 * k, 1))


def train_proposer(dark, n_examples=20_000_000, scene_seeds=tuple(range(1000, 1020)), theta_band=0.2, tol_log=0.01,
                   batch=65536, schedule=None):
    """The proposer trained on dense-trace examples of seeded config-1-style scenes (one equal share of
    n_examples per seed, from 30 ms into each trace), the last bias set to the mean target, then Adam
    over the (learning rate, epochs) schedule (default train_schedule(batch)) at `batch` rows per step
    (1,024 groups of 64 rows at the default: every CU busy). Returns (proposer, stats, X, T), the
    examples left on the device for the caller."""
    if schedule is None:
        schedule = train_schedule(batch)
    import time

    import torch

    from . import scene as S
    n_per = n_examples // len(scene_seeds)
    X = torch.empty((n_per * len(scene_seeds), N_IN), dtype=torch.float32, device="cuda")
    T = torch.empty(n_per * len(scene_seeds), dtype=torch.float32, device="cuda")
    t0 = time.perf_counter()
    for k, seed in enumerate(scene_seeds):
        sc, v = S.config1_scene(seed)
        learned_examples(sc, v, X[k * n_per:(k + 1) * n_per], T[k * n_per:(k + 1) * n_per], dark, t0_ns=30_000_000,
                         seed=seed, theta_band=theta_band, tol_log=tol_log)
    torch.cuda.synchronize()
    gen_s = time.perf_counter() - t0
    prop = Proposer(1)
    w = prop.params
    w[-1] = float(T.mean())
    prop.params = w
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    loss = None
    for lr, ep in schedule:
        loss = prop.train(X, T, batch=batch, epochs=ep, lr=lr)
    e1.record()
    torch.cuda.synchronize()
    stats = {"examples": int(X.shape[0]), "generation_s": gen_s, "rows_trained": int(X.shape[0]) * sum(e for _, e in schedule),
             "train_ms": e0.elapsed_time(e1), "final_loss": loss, "batch": batch,
             "learning_rates": [lr for lr, _ in schedule]}
    return prop, stats, X, T


def compare_controllers(scene, view, params, proposer, n_windows, theta_band=0.2, tol_log=0.01, margin=0.5, near=0.05,
                        d_px_curve=(0.125, 0.25, 0.5, 1.0)):
    """The learned controller (linked to the sensor it feeds) against the plain one (tol_log, d_px 0.25)
    on the same scene over n_windows windows from t = 0, both judged against a 10-us dense reference
    through the same sensor (event_error). The plain controller's samples-against-error curve over its
    lattice density d_px gives the plain samples per event at the learned controller's median timing
    error (log-log interpolation between the bracketing points): the gain at equal error."""
    from . import scene as S
    from .sensor import SensorModel
    dark = float(params.dark_current_fA) / float(params.illuminance_scale)
    frame = scene.render(view, 0)

    def run(kind, d_px=0.25):
        m = SensorModel(params, view.width, view.height, 1, ["constant"], 0, device="cuda")
        m.init(0, frame.reshape(-1, 1))
        ls = None
        if kind == "learned":
            ls = LearnedSampler(scene, view, learned_rules(S.rules(dark, tol=tol_log), tol_log=tol_log,
                                                           theta_band=theta_band, margin=margin, near=near), proposer)
            ls.init(0)
        evs, samples, fallbacks, proposals = [], 0, 0, 0
        for i in range(n_windows):
            if kind == "dense":
                r, tot = scene.sample_window(view, S.rules(dark, fixed_period_ns=10_000), i * WINDOW_NS, slot=0)
            elif kind == "plain":
                r, tot = scene.sample_window(view, S.rules(dark, tol=tol_log, d_px=d_px), i * WINDOW_NS, slot=0)
            else:  # the controller reads the sensor's reference state at each window start
                r, tot, pr, fb = ls.window(i * WINDOW_NS, slot=0, sensor=m)
                fallbacks += fb
                proposals += pr
            samples += tot
            ptr, cnt = m.advance_device_runs(r, (i + 1) * WINDOW_NS)
            evs.append(m.download_events(ptr, cnt))
        ev = {k: np.concatenate([e[k] for e in evs]) for k in evs[0]}
        return ev, samples, proposals, fallbacks

    ref, ref_samples, _, _ = run("dense")

    def summary(ev, samples, proposals=0, fallbacks=0):
        err = event_error(ev, ref, view.width)
        return dict(samples=samples, events=err["n_test"], samples_per_event=samples / max(1, err["n_test"]),
                    count_mismatch=err["count_mismatch"], median_abs_dt_ns=err["median_abs_dt_ns"],
                    p90_abs_dt_ns=err["p90_abs_dt_ns"], proposals=proposals, fallbacks=fallbacks)

    res = {"plain": summary(*run("plain")), "learned": summary(*run("learned"))}
    curve = [dict(d_px=d, **summary(*run("plain", d))) for d in d_px_curve]
    target = res["learned"]["median_abs_dt_ns"]
    pts = sorted((c["median_abs_dt_ns"], c["samples_per_event"]) for c in curve)
    equal = None
    for (e0, s0), (e1, s1) in zip(pts, pts[1:]):
        if e0 <= target <= e1 and e1 > e0:
            f = (np.log(target) - np.log(e0)) / (np.log(e1) - np.log(e0))
            equal = float(np.exp(np.log(s0) + f * (np.log(s1) - np.log(s0))))
    res.update(curve=curve, dense={"samples": ref_samples, "events": int(len(ref["t_cross_ns"]))},
               plain_samples_per_event_at_learned_median_error=equal,
               gain_at_equal_median_error=equal / res["learned"]["samples_per_event"] if equal is not None else None,
               gain_vs_plain=res["plain"]["samples_per_event"] / res["learned"]["samples_per_event"])
    return res
