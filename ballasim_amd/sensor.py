"""The sensor path's Python surface, mirroring the reference's _sensor module
(src/sensor/sensor_module.cpp:374-656): SensorParams, preset(), preset_names() and SensorModel with
init / advance (padded or ragged) / advance_device / download_events / save_state / load_state /
stats / sensor_maps. Everything runs in ballasim_amd/lib/libbsx.so on the GPU; this module only
marshals arrays.
"""
import ctypes as C

import numpy as np

from . import _abi

# A walk launch shape's fields (bsx_launch_policy without the event capacity).
TUNING_KEYS = ("block_size", "walk_budget", "search_batch", "queue_chunk", "head_lanes", "head_per_mille")

_ENUMS = {
    "photoreceptor": _abi.PHOTORECEPTOR,
    "reset_mode": _abi.RESET,
    "initial_phase": _abi.PHASE,
}


class SensorParams:
    """The reference's SensorParams fields (src/sensor/params.h:65-172), by name; enum fields take
    the reference's strings ('cascade', 'physical', ...)."""

    def __init__(self, c_params=None):
        object.__setattr__(self, "_p", c_params if c_params is not None else _abi.Params())

    def __getattr__(self, name):
        p = object.__getattribute__(self, "_p")
        if name == "family":
            return p.family.decode()
        if name in _ENUMS:
            inv = {v: k for k, v in _ENUMS[name].items()}
            return inv[getattr(p, name)]
        if name == "noise":
            return bool(p.noise)
        return getattr(p, name)

    def __setattr__(self, name, value):
        p = object.__getattribute__(self, "_p")
        if name == "family":
            p.family = value.encode()
        elif name in _ENUMS:
            setattr(p, name, _ENUMS[name][value] if isinstance(value, str) else int(value))
        elif name == "noise":
            p.noise = 1 if value else 0
        else:
            if not hasattr(p, name):
                raise AttributeError(f"SensorParams has no field {name!r}")
            setattr(p, name, value)

    def validate(self):
        lib = _abi.load_library()
        _abi.check(lib.bsx_params_validate(C.byref(self._p)))

    def copy(self):
        q = _abi.Params()
        C.memmove(C.byref(q), C.byref(self._p), C.sizeof(q))
        return SensorParams(q)

    @property
    def c(self):
        return self._p


def preset(name):
    lib = _abi.load_library()
    p = _abi.Params()
    _abi.check(lib.bsx_params_preset(name.encode(), C.byref(p)))
    return SensorParams(p)


def preset_names():
    lib = _abi.load_library()
    buf = C.create_string_buffer(256)
    _abi.check(lib.bsx_preset_names(buf, 256))
    return buf.value.decode().split(",")


def noise_tables(params):
    """The in-path noise's tables of a parameter set (fp64 and the fp32 copy the kernels read), each
    (n_nodes, stride), built on the host (reference: NoiseTables, src/sensor/noise_tables.cpp:341-411)."""
    lib = _abi.load_library()
    p = params._p if isinstance(params, SensorParams) else params
    n, s = C.c_int32(0), C.c_int32(0)
    _abi.check(lib.bsx_noise_tables(C.byref(p), None, None, 0, C.byref(n), C.byref(s)))
    d64 = np.zeros((n.value, s.value), np.float64)
    d32 = np.zeros((n.value, s.value), np.float32)
    _abi.check(lib.bsx_noise_tables(C.byref(p), _ptr(d64), _ptr(d32), d64.size, C.byref(n), C.byref(s)))
    return d64, d32


def device_count():
    lib = _abi.load_library()
    n = C.c_int32(0)
    _abi.check(lib.bsx_device_count(C.byref(n)))
    return n.value


def _waveform(g):
    w = _abi.Waveform()
    if g is None or g == "constant":
        return w
    if isinstance(g, dict):
        wf = g.get("waveform", g)
        kind = wf.get("type", "constant")
        w.type = _abi.WAVE[kind]
        if kind == "fourier":
            a = list(wf.get("a", []))
            w.n_harmonics = len(a)
            w.f0_hz = float(wf["f0_hz"])
            for i, v in enumerate(a):
                w.a[i] = v
            for i, v in enumerate(wf.get("phi", [0.0] * len(a))):
                w.phi[i] = v
        return w
    raise ValueError(f"light group {g!r}: a waveform dict or 'constant'")


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


class SensorModel:
    """One camera's sensor on the GPU: SensorModel(params, width, height, seed, light_groups, camera,
    num_threads=0, precision=None, overrides=None, device='cuda', roi=None), the reference's signature.
    device must name the GPU ('cuda', 'hip' or 'gpu'); precision is fp32 (None or 'fp32')."""

    def __init__(self, params, width, height, seed, light_groups=None, camera=0, num_threads=0, precision=None,
                 overrides=None, device="cuda", roi=None):
        if device not in ("cuda", "hip", "gpu"):
            raise ValueError(f"device {device!r}: this backend runs the sensor on the GPU only ('cuda')")
        if precision not in (None, "fp32"):
            raise ValueError(f"precision {precision!r}: the GPU kernel runs fp32")
        if roi is not None:
            raise ValueError("roi: region-of-interest models are not supported by this backend")
        p = preset(params) if isinstance(params, str) else params.copy()
        for k, v in (overrides or {}).items():
            setattr(p, k, v)
        groups = light_groups if light_groups is not None else ["constant"]
        if isinstance(groups, int):
            groups = ["constant"] * groups
        arr = (_abi.Waveform * len(groups))(*[_waveform(g) for g in groups])
        self._lib = _abi.load_library()
        self._params = p
        self._h = C.c_void_p()
        _abi.check(self._lib.bsx_sensor_create(C.byref(p.c), width, height, seed, arr, len(groups), camera,
                                               C.byref(self._h)))
        self._w, self._hgt, self._g = width, height, len(groups)

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self._lib.bsx_sensor_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # --- calls -------------------------------------------------------------------------------
    def set_photometry(self, illuminance_scale, gain=None):
        g = None if gain is None else np.ascontiguousarray(gain, np.float32).reshape(-1)
        _abi.check(self._lib.bsx_sensor_set_photometry(self._h, float(illuminance_scale), _ptr(g),
                                                       0 if g is None else g.size))

    def init(self, t0_ns, irradiance0):
        irr = np.ascontiguousarray(irradiance0, np.float32)
        if irr.shape != (self._w * self._hgt, self._g):
            raise ValueError("init: irradiance0 must have shape (height * width, groups)")
        _abi.check(self._lib.bsx_sensor_init(self._h, int(t0_ns), _ptr(irr)))

    def advance(self, pixel_ids, t_ns, irradiance, t_end_ns, counts=None, offsets=None):
        """Advance to t_end_ns; the reference's padded (n, m) or ragged (offsets) forms. Returns the
        events dict sorted by (t_cross_ns, pixel)."""
        ids = np.ascontiguousarray(pixel_ids, np.uint32)
        n = ids.shape[0]
        if offsets is None:
            t = np.ascontiguousarray(t_ns, np.int64)
            irr = np.ascontiguousarray(irradiance, np.float32)
            if t.ndim != 2 or t.shape[0] != n or irr.shape[:2] != t.shape or (irr.ndim == 3 and irr.shape[2] != self._g):
                raise ValueError("advance: t_ns must be (n, m) and irradiance (n, m, groups) for n pixel ids")
            m = t.shape[1]
            cnt = None
            if counts is not None:
                cnt = np.ascontiguousarray(counts, np.int32)
                if cnt.shape[0] != n:
                    raise ValueError("advance: counts must have one entry per pixel id")
            _abi.check(self._lib.bsx_sensor_advance(self._h, _ptr(ids), n, _ptr(t), _ptr(irr), m, _ptr(cnt),
                                                    int(t_end_ns)))
        else:
            if counts is not None:
                raise ValueError("advance: pass either counts or offsets, not both")
            t = np.ascontiguousarray(t_ns, np.int64)
            irr = np.ascontiguousarray(irradiance, np.float32).reshape(t.shape[0], self._g)
            off = np.asarray(offsets, np.int64)
            if off.shape[0] != n + 1 or off[0] != 0 or off[-1] != t.shape[0]:
                raise ValueError("advance: offsets must be (n + 1,) from 0 to the number of samples")
            if np.any(np.diff(off) < 0):
                raise ValueError("advance: offsets must be non-decreasing")
            start = np.ascontiguousarray(off[:-1])
            count = np.ascontiguousarray(np.diff(off).astype(np.int32))
            _abi.check(self._lib.bsx_sensor_advance_runs(self._h, _ptr(ids), n, _ptr(t), _ptr(irr), t.shape[0],
                                                         _ptr(start), _ptr(count), int(t_end_ns)))
        return self._events()

    def advance_device(self, pixel, row_start, row_count, t, irr, t_end_ns, stream=0):
        """advance_device on runs already in device memory, given as contiguous torch CUDA tensors:
        pixel uint32/int32 (n,), row_start int64 (n,), row_count int32 (n,), t int64 (S,), irr
        float32 with S * groups values. Returns (data_ptr, count) of the unordered raw events left in
        device memory. Raw device addresses go through advance_device_runs with a DeviceRuns struct."""
        runs = _checked_device_runs(pixel, row_start, row_count, t, irr, self._g, "SensorModel.advance_device")
        out = _abi.DeviceEvents()
        st = stream.cuda_stream if hasattr(stream, "cuda_stream") else int(stream or 0)
        _abi.check(self._lib.bsx_sensor_advance_device(self._h, C.byref(runs), int(t_end_ns), C.c_void_p(st),
                                                       C.byref(out)))
        return out.data, out.count

    def advance_device_runs(self, runs, t_end_ns, stream=0):
        """advance_device on a bsx_device_runs struct (e.g. Scene.sample_window's); (data_ptr, count)."""
        out = _abi.DeviceEvents()
        st = stream.cuda_stream if hasattr(stream, "cuda_stream") else int(stream or 0)
        _abi.check(self._lib.bsx_sensor_advance_device(self._h, C.byref(runs), int(t_end_ns), C.c_void_p(st),
                                                       C.byref(out)))
        return out.data, out.count

    def held_frame(self, runs=None, stream=0):
        """The input each pixel holds after a call on `runs` (device runs, as advance_device_runs
        takes, or a tuple of torch CUDA tensors): its run's last sample, raw, or else the input it
        holds now. With no runs, the held inputs. (height, width, groups) float32; the model is not
        changed. Reference: SensorModel::held_frame, src/sensor/sensor_model.cpp:748-760, and
        CudaSensorKernel::end_frame, src/sensor/sensor_cuda.cu:989-1010."""
        out = np.empty((self._hgt, self._w, self._g), np.float32)
        r = None if runs is None else _device_runs(runs, self._g, "SensorModel.held_frame")
        st = stream.cuda_stream if hasattr(stream, "cuda_stream") else int(stream or 0)
        _abi.check(self._lib.bsx_sensor_held_frame(self._h, None if r is None else C.byref(r), C.c_void_p(st),
                                                   _ptr(out)))
        return out

    def download_events(self, data_ptr, count):
        ev = _abi.DeviceEvents(data_ptr, count)
        _abi.check(self._lib.bsx_sensor_download_events(self._h, C.byref(ev)))
        return self._events()

    def _events(self):
        e = _abi.Events()
        _abi.check(self._lib.bsx_sensor_events(self._h, C.byref(e)))
        n = e.count
        spec = [("x", np.uint16), ("y", np.uint16), ("p", np.uint8), ("t_cross_ns", np.int64),
                ("t_reported_ns", np.int64), ("label", np.uint8), ("flags", np.uint16), ("ds", np.float32),
                ("dn", np.float32), ("dl", np.float32), ("l_first", np.float32)]
        out = {}
        for name, dt in spec:
            a = np.empty(n, dt)
            if n:
                C.memmove(a.ctypes.data, getattr(e, name), n * a.itemsize)
            out[name] = a
        return out

    # --- state ---------------------------------------------------------------------------------
    def state_bytes(self):
        n = C.c_size_t()
        _abi.check(self._lib.bsx_sensor_state_bytes(self._h, C.byref(n)))
        return n.value

    def save_state(self):
        buf = np.empty(self.state_bytes(), np.uint8)
        _abi.check(self._lib.bsx_sensor_save_state(self._h, _ptr(buf), buf.size))
        return buf.tobytes()

    def load_state(self, blob):
        buf = np.frombuffer(blob, np.uint8).copy()
        _abi.check(self._lib.bsx_sensor_load_state(self._h, _ptr(buf), buf.size))

    # --- introspection -------------------------------------------------------------------------
    def stats(self):
        s = _abi.Stats()
        _abi.check(self._lib.bsx_sensor_stats(self._h, C.byref(s)))
        return {name: getattr(s, name) for name, _ in _abi.Stats._fields_}

    def reset_stats(self):
        _abi.check(self._lib.bsx_sensor_reset_stats(self._h))

    def set_launch_policy(self, event_capacity=0, block_size=0, walk_budget=0, search_batch=0, queue_chunk=0,
                          head_lanes=0, head_per_mille=0):
        pol = _abi.LaunchPolicy(event_capacity, block_size, walk_budget, search_batch, queue_chunk, head_lanes,
                                head_per_mille)
        _abi.check(self._lib.bsx_sensor_set_launch_policy(self._h, C.byref(pol)))

    def _info(self):
        i = _abi.Info()
        _abi.check(self._lib.bsx_sensor_info(self._h, C.byref(i)))
        return i

    def sensor_maps(self):
        P = self._w * self._hgt
        arrs = [np.empty(P) for _ in range(5)]
        hot = np.empty(P, np.uint8)
        _abi.check(self._lib.bsx_sensor_maps(self._h, *[_ptr(a) for a in arrs], _ptr(hot)))
        out = {k: a.reshape(self._hgt, self._w) for k, a in
               zip(("theta_on", "theta_off", "leak", "dark_current_fA", "phi0"), arrs)}
        out["hot"] = hot.reshape(self._hgt, self._w)
        return out

    @property
    def params(self):
        return self._params.copy()

    width = property(lambda self: self._w)
    height = property(lambda self: self._hgt)
    groups = property(lambda self: self._g)
    time_ns = property(lambda self: self._info().time_ns)
    initialized = property(lambda self: bool(self._info().initialized))
    readout_phase_ns = property(lambda self: self._info().readout_phase_ns)
    irr_max = property(lambda self: self._info().irr_max)
    precision = property(lambda self: "fp32")
    device = property(lambda self: "cuda")


def _checked_device_runs(pixel, row_start, row_count, t, irr, groups, where):
    """A bsx_device_runs over torch CUDA tensors, after the checks the C side cannot make: device,
    contiguity, dtypes and lengths. A host tensor, a strided view, a wrong dtype or a short irr would
    otherwise be read as raw device memory."""
    import torch

    named = (("pixel", pixel, (torch.uint32, torch.int32)), ("row_start", row_start, (torch.int64,)),
             ("row_count", row_count, (torch.int32,)), ("t", t, (torch.int64,)), ("irr", irr, (torch.float32,)))
    for name, x, dtypes in named:
        if not isinstance(x, torch.Tensor):
            raise ValueError(f"{where}: {name} must be a torch tensor in device memory")
        if not x.is_cuda:
            raise ValueError(f"{where}: {name} must be on the GPU, not {x.device}")
        if not x.is_contiguous():
            raise ValueError(f"{where}: {name} must be contiguous")
        if x.dtype not in dtypes:
            raise ValueError(f"{where}: {name} must have dtype {' or '.join(str(d) for d in dtypes)}, not {x.dtype}")
    n = pixel.numel()
    if row_start.numel() != n or row_count.numel() != n:
        raise ValueError(f"{where}: pixel, row_start and row_count must have the same length")
    if irr.numel() != t.numel() * groups:
        raise ValueError(f"{where}: irr must hold t.numel() * groups = {t.numel() * groups} values, not {irr.numel()}")
    return _abi.DeviceRuns(pixel.data_ptr(), row_start.data_ptr(), row_count.data_ptr(), n, t.data_ptr(),
                           irr.data_ptr(), t.numel())


def _device_runs(r, groups, where):
    """A bsx_device_runs from a DeviceRuns struct (raw device addresses, taken as given) or a
    (pixel, row_start, row_count, t, irr) tuple of torch CUDA tensors (checked)."""
    if isinstance(r, _abi.DeviceRuns):
        return r
    return _checked_device_runs(*r, groups, where)


class Rig:
    """Several cameras (SensorModel, device 'cuda', one photoreceptor kind) advanced by one launch
    chain: phase A over every camera's segments, one heavy-first order of the joint pixel space and
    one walk launch over all pixels, captured into a hipGraph and replayed. Each camera's events,
    state and statistics equal those of advancing it alone.

    It replaces the reference engine's loop of per-camera advance_device calls on one stream
    (src/engine/engine.h:1-8, src/engine/engine_impl.h:1343-1348). The cameras are borrowed: keep
    them alive while the rig is used."""

    def __init__(self, cameras):
        cameras = list(cameras)
        self._lib = _abi.load_library()
        self._cams = cameras
        hs = (C.c_void_p * len(cameras))(*[c._h.value if c._h else None for c in cameras])
        self._h = C.c_void_p()
        _abi.check(self._lib.bsx_rig_create(hs, len(cameras), C.byref(self._h)))

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self._lib.bsx_rig_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __len__(self):
        return len(self._cams)

    def _runs(self, runs):
        if len(runs) != len(self._cams):
            raise ValueError(f"Rig: need runs for each of the {len(self._cams)} cameras")
        arr = (_abi.DeviceRuns * len(runs))()
        for i, r in enumerate(runs):
            arr[i] = _device_runs(r, self._cams[i].groups, f"Rig.advance_device (camera {i})")
        return arr

    def advance_device(self, runs, t_end_ns, stream=0):
        """Advances camera c on runs[c] to t_end_ns; returns [(data_ptr, count)] of each camera's
        unordered raw events in device memory (download with that camera's download_events)."""
        arr = self._runs(runs)
        out = (_abi.DeviceEvents * len(runs))()
        st = stream.cuda_stream if hasattr(stream, "cuda_stream") else int(stream or 0)
        _abi.check(self._lib.bsx_rig_advance_device(self._h, arr, int(t_end_ns), C.c_void_p(st), out))
        return [(o.data, o.count) for o in out]

    def set_tuning(self, block_size=0, walk_budget=0, search_batch=0, queue_chunk=0, head_lanes=0, head_per_mille=0):
        pol = _abi.LaunchPolicy(0, block_size, walk_budget, search_batch, queue_chunk, head_lanes, head_per_mille)
        _abi.check(self._lib.bsx_rig_set_tuning(self._h, C.byref(pol)))

    def tuning(self):
        pol = _abi.LaunchPolicy()
        _abi.check(self._lib.bsx_rig_get_tuning(self._h, C.byref(pol)))
        return {k: getattr(pol, k) for k in TUNING_KEYS}

    def set_graph(self, on):
        _abi.check(self._lib.bsx_rig_set_graph(self._h, 1 if on else 0))

    def stats(self):
        s = _abi.RigStats()
        _abi.check(self._lib.bsx_rig_stats(self._h, C.byref(s)))
        return {name: getattr(s, name) for name, _ in _abi.RigStats._fields_}

    def autotune(self, runs, t_end_ns, stream=0, reps=3):
        """Times the walk of this call under candidate launch shapes and keeps the fastest (no camera
        changes). Returns the trials as [(tuning dict, walk_ms)]."""
        arr = self._runs(runs)
        cap = 256
        tun = (_abi.LaunchPolicy * cap)()
        ms = (C.c_double * cap)()
        n = C.c_int32()
        st = stream.cuda_stream if hasattr(stream, "cuda_stream") else int(stream or 0)
        _abi.check(self._lib.bsx_rig_autotune(self._h, arr, int(t_end_ns), C.c_void_p(st), int(reps), tun, ms, cap,
                                              C.byref(n)))
        return [({k: getattr(tun[i], k) for k in TUNING_KEYS}, ms[i]) for i in range(min(n.value, cap))]
