"""The procedural scene and sample-time controller on the GPU (include/bsx/bsx_scene.h), and the
seeded scenes of the benchmark workloads.

A workload is a sensor of some size watching a scene of textured rectangles and disks moving
affinely over a uniform background; the controller picks each pixel's sample times per 2^23 ns
window and leaves the runs in HBM, where SensorModel.advance_device reads them.
"""
import ctypes as C

import numpy as np

from . import _abi

WINDOW_NS = 1 << 23
LUX_TO_FA = 100.0  # illuminance_scale of the workloads: fA of photocurrent per lux (DAVIS346-like pixel)


class Scene:
    def __init__(self, shapes, background):
        self._lib = _abi.load_library()
        arr = (_abi.Shape * max(1, len(shapes)))(*shapes)
        self._h = C.c_void_p()
        _abi.check(self._lib.bsx_scene_create(arr, len(shapes), float(background), C.byref(self._h)))
        self.n_shapes = len(shapes)

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self._lib.bsx_scene_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def render(self, view, t_ns):
        out = np.empty(view.width * view.height, np.float32)
        _abi.check(self._lib.bsx_scene_render(self._h, C.byref(view), int(t_ns), out.ctypes.data_as(C.c_void_p)))
        return out

    def sample_window(self, view, rules, t_w, slot=0, stream=0):
        """Runs of every pixel over (t_w, t_w + 2^23] in device buffer set `slot`: (DeviceRuns, samples)."""
        runs = _abi.DeviceRuns()
        total = C.c_uint64()
        st = stream.cuda_stream if hasattr(stream, "cuda_stream") else int(stream or 0)
        _abi.check(self._lib.bsx_scene_sample_window(self._h, C.byref(view), C.byref(rules), int(t_w), int(slot),
                                                     C.c_void_p(st), C.byref(runs), C.byref(total)))
        return runs, total.value

    def set_capture_budget(self, nbytes):
        """Device memory the one-pass sampler may hold (0: count and write in two passes); the runs are the
        same either way."""
        _abi.check(self._lib.bsx_scene_set_capture_budget(self._h, int(nbytes)))

    def capture_budget(self):
        """(budget in bytes, windows whose largest pixel count exceeded the capture slots)."""
        b, o = C.c_uint64(), C.c_uint64()
        _abi.check(self._lib.bsx_scene_capture_budget(self._h, C.byref(b), C.byref(o)))
        return b.value, o.value

    def sample_counts(self, view, counts, t_w, slot=0, stream=0):
        """Runs over (t_w, t_w + 2^23] at imposed per-pixel counts (int32, width*height), samples
        evenly spread in the window: (DeviceRuns, samples)."""
        counts = np.ascontiguousarray(counts, np.int32)
        if counts.size != view.width * view.height:
            raise ValueError("sample_counts: one count per pixel")
        runs = _abi.DeviceRuns()
        total = C.c_uint64()
        st = stream.cuda_stream if hasattr(stream, "cuda_stream") else int(stream or 0)
        _abi.check(self._lib.bsx_scene_sample_counts(self._h, C.byref(view), counts.ctypes.data_as(C.c_void_p),
                                                     int(t_w), int(slot), C.c_void_p(st), C.byref(runs),
                                                     C.byref(total)))
        return runs, total.value

    def render_requests(self, view, t_base_ns, requests, n, out_index, irr, sig, klass=None, groups=1, stream=0):
        """The engine's RenderHook::render on this scene for n requests: device tensors of the 24-byte
        requests (REQUEST_DTYPE records, any dtype holding n * 24 bytes), out_index (n u32 rows, or None for
        row i), irr (rows x groups f32), sig (rows u64) and klass (rows u32, or None). The sizes are checked
        here, before the launch: every output row a request names must exist."""
        n = int(n)
        if n < 0:
            raise ValueError("render_requests: n must be >= 0")

        def nbytes(a):
            return a.numel() * a.element_size()

        rows = n
        if out_index is not None:
            if nbytes(out_index) < 4 * n:
                raise ValueError("render_requests: out_index holds fewer than n entries")
            if n:
                rows = int(out_index[:n].to(dtype=__import__("torch").int64).max()) + 1
                if int(out_index[:n].to(dtype=__import__("torch").int64).min()) < 0:
                    raise ValueError("render_requests: negative output row")
        if nbytes(requests) < 24 * n:
            raise ValueError("render_requests: requests holds fewer than n 24-byte records")
        if nbytes(irr) < 4 * groups * rows or nbytes(sig) < 8 * rows or (klass is not None and nbytes(klass) < 4 * rows):
            raise ValueError("render_requests: an output array is smaller than the rows the requests name")

        def addr(a):
            return None if a is None else C.c_void_p(a.data_ptr())
        st = stream.cuda_stream if hasattr(stream, "cuda_stream") else int(stream or 0)
        _abi.check(self._lib.bsx_scene_render_requests(self._h, C.byref(view), int(t_base_ns), addr(requests), n,
                                                       addr(out_index), int(groups), addr(irr), addr(sig),
                                                       addr(klass), C.c_void_p(st)))


# The engine's SampleRequest (src/backend/api.h:25-33) as a numpy record.
REQUEST_DTYPE = np.dtype([("cam", np.uint16), ("flags", np.uint16), ("pixel", np.uint32), ("k_begin", np.uint32),
                          ("k_count", np.uint32), ("t_off_ns", np.int32), ("pose_slot", np.uint32)])


def runs_to_host(runs, groups=1):
    """A copy of device runs on the host: (ids, t, irr (S, groups), offsets (n + 1))."""
    lib = _abi.load_library()
    n, S = runs.n, runs.n_samples

    def get(ptr, count, dtype):
        a = np.empty(count, dtype)
        if count:
            _abi.check(lib.bsx_copy_to_host(a.ctypes.data_as(C.c_void_p), C.c_void_p(ptr), a.nbytes))
        return a

    ids = get(runs.pixel, n, np.uint32)
    start = get(runs.row_start, n, np.int64)
    count = get(runs.row_count, n, np.int32)
    t = get(runs.t, S, np.int64)
    irr = get(runs.irr, S * groups, np.float32).reshape(S, groups)
    offsets = np.append(start, start[-1] + count[-1] if n else 0).astype(np.int64)
    assert np.array_equal(np.diff(offsets), count)
    return ids, t, irr, offsets


def view(width, height, origin_x=0.0, origin_y=0.0, px_scale=1.0):
    return _abi.View(width, height, origin_x, origin_y, px_scale)


def rules(dark, tol=0.01, d_px=0.25, l_min=1, l_max=13, max_refine_levels=4, min_split_ns=2048, signature=True,
          fixed_period_ns=0):
    """The controller's rules (reference defaults: src/engine/engine.h:41-61; l_min 1 with the midpoint
    test and l_max 13 for min_dt 1024 ns, src/engine/motion.h:301-309)."""
    return _abi.SamplerRules(d_px, tol, dark, l_min, l_max, max_refine_levels, min_split_ns, 1 if signature else 0, 0,
                             fixed_period_ns)


def shape(kind, cx, cy, vx, vy, hx, hy, lux, contrast, seed, angle=0.0, omega=0.0, octaves=6, cell=32.0):
    return _abi.Shape(_abi.SHAPE[kind] if isinstance(kind, str) else kind, seed & 0xFFFFFFFF, cx, cy, vx, vy, angle,
                      omega, hx, hy, lux, contrast, octaves, cell)


_abi.SHAPE = {"rect": 0, "disk": 1}


def config1_scene(seed=0):
    """346x260: a 100 lux background and one disk of radius 40 px, 8-octave value noise of contrast 0.5,
    translating at 800 px/s on a straight line."""
    rng = np.random.default_rng(seed)
    ang = float(rng.uniform(0, 2 * np.pi))
    vx, vy = 800.0 * np.cos(ang), 800.0 * np.sin(ang)
    # Start so that the 100 ms path (80 px) passes the frame centre.
    cx, cy = 173.0 - 0.05 * vx, 130.0 - 0.05 * vy
    disk = shape("disk", cx, cy, vx, vy, 40.0, 40.0, 100.0, 0.5, int(rng.integers(1, 2 ** 31)), octaves=8,
                 cell=64.0)
    return Scene([disk], 100.0), view(346, 260)


def config2_shapes(seed=0, n=64, width=1280, height=720, v_max=2000.0, w_max=3.0, cell_range=(16.0, 64.0),
                   octaves=6, contrast=(0.1, 0.9)):
    rng = np.random.default_rng(seed)
    shapes = []
    for i in range(n):
        hx, hy = rng.uniform(20, 120, 2)
        lux = float(np.exp(rng.uniform(np.log(10), np.log(1000))))
        shapes.append(shape("rect", float(rng.uniform(0, width)), float(rng.uniform(0, height)),
                            float(rng.uniform(-v_max, v_max)), float(rng.uniform(-v_max, v_max)), float(hx), float(hy),
                            lux, float(rng.uniform(*contrast)), int(rng.integers(1, 2 ** 31)),
                            angle=float(rng.uniform(0, 2 * np.pi)), omega=float(rng.uniform(-w_max, w_max)),
                            octaves=octaves, cell=float(rng.uniform(*cell_range))))
    background = float(np.exp(rng.uniform(np.log(10), np.log(1000))))
    return shapes, background


def config2_scene(seed=0):
    """1280x720: 64 textured rectangles (contrast U[0.1, 0.9]), velocity U(+-2000 px/s), rotation
    U(+-3 rad/s), fixed depth order, over a log-uniform 10-1000 lux background."""
    shapes, bg = config2_shapes(seed)
    return Scene(shapes, bg), view(1280, 720)


RIG = [(1280, 720), (640, 480), (346, 260), (1920, 1080)]
# Scene-plane origins of the rig's cameras (1 scene px per camera px): each sees the config-1 disk's
# path (around scene point (173, 130)) at another place of its frame.
RIG_OFFSETS = ((-467.0, -230.0), (-100.0, -50.0), (0.0, 0.0), (-900.0, -400.0))


def rig_views(offsets=RIG_OFFSETS):
    """The 4-camera rig of mixed resolutions viewing one scene, each with its own extrinsic offset."""
    return [view(w, h, ox, oy) for (w, h), (ox, oy) in zip(RIG, offsets)]


def lognormal_counts(rng, n, mu=np.log(20.0), sigma=1.5, lo=1, hi=4000):
    """Per-pixel sample counts of the ragged-batch workload: round(exp(N(mu, sigma))) clipped to [lo, hi]."""
    return np.clip(np.rint(np.exp(rng.normal(mu, sigma, size=n))), lo, hi).astype(np.int32)


def learned_eval_scene(seed=0, width=1280, height=720, n_disks=12):
    """Config-1-style scene at a larger size: n_disks disks like config 1's (radius 40 px, 8-octave value
    noise of contrast 0.5, 800 px/s in a seeded direction) over a 100 lux background, placed so that
    their 500 ms paths stay mostly in the frame."""
    rng = np.random.default_rng(seed)
    shapes = []
    for _ in range(n_disks):
        ang = float(rng.uniform(0, 2 * np.pi))
        vx, vy = 800.0 * np.cos(ang), 800.0 * np.sin(ang)
        cx = float(rng.uniform(0.25, 0.75) * width) - 0.25 * vx
        cy = float(rng.uniform(0.25, 0.75) * height) - 0.25 * vy
        shapes.append(shape("disk", cx, cy, vx, vy, 40.0, 40.0, 100.0, 0.5, int(rng.integers(1, 2 ** 31)), octaves=8,
                            cell=64.0))
    return Scene(shapes, 100.0), view(width, height)
