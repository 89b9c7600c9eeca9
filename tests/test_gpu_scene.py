"""The device scene and sample-time controller (GPU): runs are well formed, the refinement obeys the
controller's rules on the values it rendered (re-derived on the CPU from the run itself), the scene
values agree with a numpy evaluation of the same shapes, and the sensor on these runs equals the
oracle on their host copy bit for bit."""
import math

import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import events_from_dict

pytestmark = pytest.mark.gpu

WINDOW = 1 << 23


def _small_scene(seed=3, n=6, width=96, height=64):
    from ballasim_amd import scene as S
    shapes, bg = S.config2_shapes(seed, n=n, width=width, height=height, v_max=1500.0, w_max=2.0)
    return S.Scene(shapes, bg), S.view(width, height), shapes, bg


def _level_scene(seed=21, n=8, width=80, height=60):
    """Small rectangles and disks (half sizes 5-14 px) at log-uniform speeds of 30-4000 px/s over a
    100 lux background: pixels at many lattice levels, and pixels no shape reaches."""
    from ballasim_amd import scene as S
    rng = np.random.default_rng(seed)
    shapes = []
    for i in range(n):
        speed = float(np.exp(rng.uniform(np.log(30.0), np.log(4000.0))))
        ang = float(rng.uniform(0, 2 * np.pi))
        hx, hy = (float(v) for v in rng.uniform(5, 14, 2))
        shapes.append(S.shape("disk" if i % 3 == 0 else "rect", float(rng.uniform(10, width - 10)),
                              float(rng.uniform(10, height - 10)), speed * np.cos(ang), speed * np.sin(ang), hx, hy,
                              float(np.exp(rng.uniform(np.log(20), np.log(500)))), float(rng.uniform(0.2, 0.8)),
                              int(rng.integers(1, 2 ** 31)), angle=float(rng.uniform(0, 6.28)),
                              omega=float(rng.uniform(-2, 2)), octaves=4, cell=float(rng.uniform(4, 12))))
    return S.Scene(shapes, 100.0), S.view(width, height), shapes, 100.0


def _eta(dark, a, m, b):
    a, m, b = np.float32(a), np.float32(m), np.float32(b)
    dev = np.float32(abs((a + b) * np.float32(-0.5) + m))
    low = min(min(a, m), b)
    return np.float32(np.float32(dev) / np.float32(np.float32(dark) + low))


def _expected_offsets(vals, level, dark, tol, max_levels=4, min_split=2048, sigs=None):
    """The controller's sample offsets for one pixel, re-derived from the rendered values alone: vals
    maps offset -> value (the run plus the carried sample at 0); sigs (None: signature flags off) maps
    offset -> signature, and a flip between an interval's ends flags it, as the reference's
    signature_flag does under its default rules (no jump gate: src/engine/sampling.h:126-134)."""
    n, step = 1 << level, WINDOW >> level
    e = [vals.get(j * step) for j in range(n + 1)]
    if any(v is None for v in e):
        return None
    out = []

    def flip(a, b):
        return sigs is not None and sigs[a] != sigs[b]

    def refine(lo, hi, depth):
        g = (lo + hi) >> 1
        if g not in vals:
            raise KeyError(g)
        bad = _eta(dark, vals[lo], vals[g], vals[hi]) > np.float32(tol)
        if (flip(lo, g) or bad) and g - lo >= min_split and depth < max_levels:
            refine(lo, g, depth + 1)
        out.append(g)
        if (flip(g, hi) or bad) and hi - g >= min_split and depth < max_levels:
            refine(g, hi, depth + 1)

    for j in range(1, n + 1):
        m = (j + 1 <= n and _eta(dark, e[j - 1], e[j], e[j + 1]) > np.float32(tol)) or \
            (j >= 2 and _eta(dark, e[j - 2], e[j - 1], e[j]) > np.float32(tol))
        if (m or flip((j - 1) * step, j * step)) and step >= min_split:
            try:
                refine((j - 1) * step, j * step, 1)
            except KeyError:
                return None
        out.append(j * step)
    return out


def _levels(shapes, view, t_w, d_px=0.25, l_min=1, l_max=13):
    """Each pixel's lattice level under the motion-bound rule l = clamp(ceil(log2(v_hat W / d_px)), l_min,
    l_max) (src/engine/motion.cpp:1330-1340), with v_hat the largest image speed |v| + |omega| r of the
    shapes whose bounding circle (r + 1.5 px), swept over the window, passes the pixel centre; -1: no shape
    reaches the pixel (static, no samples). Restated in numpy from the shapes' parameters."""
    W, H = view.width, view.height
    xs = view.origin_x + (np.arange(W * H) % W + 0.5) * view.px_scale
    ys = view.origin_y + (np.arange(W * H) // W + 0.5) * view.px_scale
    t0, t1 = t_w * 1e-9, t_w * 1e-9 + WINDOW * 1e-9
    v_hat = np.zeros(W * H, np.float32)
    for s in shapes:
        radius = np.float32(s.hx if s.kind == 1 else math.sqrt(s.hx * s.hx + s.hy * s.hy))
        speed = np.float32(math.sqrt(s.vx * s.vx + s.vy * s.vy) + abs(s.omega) * float(radius))
        ax, ay = s.cx + s.vx * t0, s.cy + s.vy * t0
        bx, by = s.cx + s.vx * t1, s.cy + s.vy * t1
        ex, ey = bx - ax, by - ay
        px, py = xs - ax, ys - ay
        ee = ex * ex + ey * ey
        f = np.clip((px * ex + py * ey) / ee, 0.0, 1.0) if ee > 0 else np.zeros_like(px)
        qx, qy = px - f * ex, py - f * ey
        r = float(radius) + 1.5
        v_hat = np.where(qx * qx + qy * qy <= r * r, np.maximum(v_hat, speed), v_hat)
    with np.errstate(divide="ignore"):
        lv = np.ceil(np.log2(v_hat.astype(np.float64) / view.px_scale * (WINDOW * 1e-9) / d_px))
    level = np.clip(lv, l_min, l_max).astype(np.int64)
    return np.where(v_hat > 0, level, -1)


def _signatures(sc, view, t_w, ids, t, offs):
    """The scene's signature (1 + shape index, 0 the background) and value at the carried sample and at
    every sample of every run, rendered on the device through the engine's request interface."""
    import torch

    from ballasim_amd import scene as S
    rows = []
    for i, p in enumerate(ids):
        rows.append((p, 0))
        rows.extend((p, int(x - t_w)) for x in t[offs[i]:offs[i + 1]])
    req = np.zeros(len(rows), S.REQUEST_DTYPE)
    req["pixel"] = [r[0] for r in rows]
    req["k_count"] = 1
    req["t_off_ns"] = [r[1] for r in rows]
    dev = torch.device("cuda:0")
    rq = torch.from_numpy(req.view(np.uint8)).to(dev)
    irr = torch.zeros(len(rows), dtype=torch.float32, device=dev)
    sig = torch.zeros(len(rows), dtype=torch.int64, device=dev)
    sc.render_requests(view, t_w, rq, len(rows), None, irr, sig)
    torch.cuda.synchronize()
    return sig.cpu().numpy(), irr.cpu().numpy(), rows


@pytest.mark.parametrize("signature", [False, True])
def test_runs_well_formed_and_rules_hold(signature):
    """Every run is sorted inside (t_w, t_w + 2^23], on the 1024 ns grid, at the pixel's own lattice level
    (the motion-bound rule, restated in numpy; a pixel no shape reaches takes no samples), and is exactly
    the sample set the controller's rules give on its own values and, with signature flags on, on the
    scene's signatures at those samples (a flip between an interval's ends refines it). The signatures
    come from the engine's request interface, which forms a sample's time as (t_w + dt) 1e-9 s, as the
    controller does, so the values of the two agree bit for bit in every window."""
    from ballasim_amd import scene as S
    sc, view, shapes, _ = _level_scene()
    dark, tol = 0.0021, 0.01
    rules = S.rules(dark, tol=tol, signature=signature)
    checked = refined = flips = static = 0
    levels_seen = set()
    for w in range(3):
        runs, total = sc.sample_window(view, rules, w * WINDOW, slot=0)
        ids, t, irr, offs = S.runs_to_host(runs)
        carried = sc.render(view, w * WINDOW)
        level_of = _levels(shapes, view, w * WINDOW)
        assert np.array_equal(ids, np.arange(view.width * view.height))
        assert total == len(t) == offs[-1]
        sig_at = None
        if signature:
            sig, rendered, rows = _signatures(sc, view, w * WINDOW, ids, t, offs)
            sig_at = {}
            for (p, off), v in zip(rows, sig):
                sig_at.setdefault(p, {})[off] = int(v)
            # The request interface renders every sample's value bit for bit (the carried ones: the frame's).
            want = np.concatenate([np.concatenate([[carried[p]], irr[offs[p]:offs[p + 1], 0]]) for p in ids])
            assert np.array_equal(rendered, want.astype(np.float32))
        for p in range(len(ids)):
            tt = t[offs[p]:offs[p + 1]] - w * WINDOW
            if level_of[p] < 0:
                assert len(tt) == 0, f"window {w} pixel {p}: no shape reaches it, yet it has samples"
                static += 1
                continue
            assert len(tt) > 0
            assert np.all(np.diff(tt) > 0) and tt[0] > 0 and tt[-1] == WINDOW and np.all(tt % 1024 == 0)
            vals = {0: carried[p]}
            vals.update({int(d): v for d, v in zip(tt, irr[offs[p]:offs[p + 1], 0])})
            sigs = sig_at[p] if signature else None
            level = int(level_of[p])
            exp = _expected_offsets(vals, level, dark, tol, sigs=sigs)
            assert exp is not None and exp == [int(x) for x in tt], \
                f"window {w} pixel {p}: the run is not the controller's sample set at level {level}"
            refined += len(tt) > (1 << level)
            levels_seen.add(level)
            if sigs is not None:
                flips += len(set(sigs.values())) > 1
            checked += 1
    assert checked > 100 and refined > 10 and static > 100
    assert len(levels_seen) >= 3, levels_seen
    if signature:
        assert flips > 20


def test_scene_values_match_numpy_evaluation():
    from ballasim_amd import scene as S
    sc, view, shapes, bg = _small_scene(seed=5)
    for t_ns in (0, 3 * WINDOW + 12345):
        img = sc.render(view, t_ns).reshape(view.height, view.width)
        t = t_ns * 1e-9
        ys, xs = np.mgrid[0:view.height, 0:view.width]
        x, y = xs + 0.5, ys + 0.5
        owner = np.full(x.shape, -1)
        for i, s in enumerate(shapes):
            dx = x - (s.cx + s.vx * t)
            dy = y - (s.cy + s.vy * t)
            th = s.angle + s.omega * t
            u = np.cos(th) * dx + np.sin(th) * dy
            v = -np.sin(th) * dx + np.cos(th) * dy
            inside = (np.abs(u) <= s.hx) & (np.abs(v) <= s.hy)
            owner[inside] = i
        bg_mask = owner < 0
        assert np.allclose(img[bg_mask], bg)
        for i, s in enumerate(shapes):
            vals = img[owner == i]
            if len(vals):
                assert np.all(vals >= s.lux * (1 - s.contrast) * 0.999) and np.all(vals <= s.lux * (1 + s.contrast) * 1.001)


def test_sensor_on_scene_runs_matches_oracle():
    import ballasim_amd as bsx
    from ballasim_amd import scene as S
    sc, view, _, _ = _small_scene(seed=7, n=8, width=128, height=80)
    p = bsx.preset("prophesee_gen3")
    p.photoreceptor = "instantaneous"
    p.noise = False
    p.theta_on = p.theta_off = 0.15
    p.refractory_s = 50e-6
    p.leak_rate_hz = 0.05
    p.illuminance_scale = S.LUX_TO_FA
    W, H = view.width, view.height
    gpu = bsx.SensorModel(p, W, H, 9, device="cuda")
    ref = O.OracleModel(p.c, W, H, 9, 1, 0)
    frame = sc.render(view, 0).reshape(-1, 1)
    gpu.init(0, frame)
    ref.init(0, frame)
    rules = S.rules(float(p.dark_current_fA) / S.LUX_TO_FA, tol=0.01)
    total_events = 0
    for w in range(4):
        runs, _ = sc.sample_window(view, rules, w * WINDOW, slot=w % 2)
        ptr, count = gpu.advance_device_runs(runs, (w + 1) * WINDOW)
        ev = gpu.download_events(ptr, count)
        ids, t, irr, offs = S.runs_to_host(runs)
        want, rep = ref.advance_runs(ids, t, irr, offs, (w + 1) * WINDOW)
        got = events_from_dict(ev, W)
        assert np.array_equal(got, want), f"window {w}"
        assert np.array_equal(ev["t_reported_ns"], rep)
        total_events += len(want)
    assert total_events > 100


def test_sample_counts_times_and_values():
    """Imposed per-pixel counts: n samples of pixel p at t_w + floor((k + 1) 2^23 / n), each carrying
    the scene's irradiance at that time (a frame render at the same instant), runs in pixel order."""
    from ballasim_amd import scene as S
    scene, view = S.config1_scene(2)
    P = view.width * view.height
    rng = np.random.default_rng(5)
    counts = S.lognormal_counts(rng, P, mu=np.log(3.0), sigma=1.0, hi=40)
    counts[::97] = 0
    t_w = 3 * S.WINDOW_NS
    runs, total = scene.sample_counts(view, counts, t_w, slot=7)
    ids, t, irr, offs = S.runs_to_host(runs)
    assert total == counts.sum() == len(t)
    assert np.array_equal(ids, np.arange(P)) and np.array_equal(np.diff(offs), counts)
    for p in rng.choice(np.nonzero(counts)[0], size=40, replace=False):
        n = int(counts[p])
        k = np.arange(n, dtype=np.int64)
        assert np.array_equal(t[offs[p]:offs[p + 1]], t_w + (k + 1) * S.WINDOW_NS // n)
    # values: the frame at one of the times, at the pixels sampled then
    tt = int(t_w + S.WINDOW_NS)  # every nonempty run's last sample
    frame = scene.render(view, tt)
    last = offs[1:][counts > 0] - 1
    np.testing.assert_allclose(irr[last, 0], frame[counts > 0], rtol=1e-6)


def test_runs_of_smaller_view_survive_a_larger_view():
    """Runs handed out for one view stay valid when a larger view is sampled into another slot."""
    from ballasim_amd import scene as S
    scene, small = S.config1_scene(3)
    big = S.view(small.width * 2, small.height * 2)
    rng = np.random.default_rng(6)
    r0, _ = scene.sample_counts(small, S.lognormal_counts(rng, small.width * small.height, hi=8), 0, slot=0)
    scene.sample_counts(big, S.lognormal_counts(rng, big.width * big.height, hi=8), 0, slot=1)
    ids, _, _, _ = S.runs_to_host(r0)
    assert np.array_equal(ids, np.arange(small.width * small.height))


def test_render_requests_equal_frames():
    """The RenderHook seam: requests at (pixel, t_base + t_off) in engine order, results in out_index rows,
    equal the frame rendered at each request's time (the same scene evaluation)."""
    import torch

    from ballasim_amd import scene as S
    scene, view = S.config1_scene(4)
    P = view.width * view.height
    rng = np.random.default_rng(8)
    n, G = 4000, 2
    t_base = 5 * S.WINDOW_NS
    offs = np.array([0, 1_000_000, 2_500_000, 8_000_000], np.int64)
    req = np.zeros(n, S.REQUEST_DTYPE)
    req["pixel"] = rng.integers(0, P, n)
    req["t_off_ns"] = offs[rng.integers(0, len(offs), n)]
    req["k_count"] = 64
    out_index = rng.permutation(n).astype(np.uint32)
    dev = torch.device("cuda:0")
    rq = torch.from_numpy(req.view(np.uint8).copy()).to(dev)
    oi = torch.from_numpy(out_index.view(np.int32).copy()).to(dev)
    irr = torch.full((n, G), -1.0, dtype=torch.float32, device=dev)
    sig = torch.zeros(n, dtype=torch.int64, device=dev)
    kl = torch.zeros(n, dtype=torch.int32, device=dev)
    with pytest.raises(ValueError):  # a request count beyond the arrays is refused before any launch
        scene.render_requests(view, t_base, rq, n + 1, oi, irr, sig, kl, groups=G)
    scene.render_requests(view, t_base, rq, n, oi, irr, sig, kl, groups=G)
    torch.cuda.synchronize()
    irr, sig, kl = irr.cpu().numpy(), sig.cpu().numpy(), kl.cpu().numpy()
    frames = {int(o): scene.render(view, t_base + int(o)) for o in offs}
    for i in range(n):
        row = out_index[i]
        assert irr[row, 0] == frames[int(req["t_off_ns"][i])][req["pixel"][i]]
        assert irr[row, 1] == 0.0
        assert (kl[row] != 0) == (sig[row] != 0)
    assert (sig != 0).any() and (sig == 0).any()


RULE_VARIANTS = {
    "default": dict(tol=0.01, signature=True),
    "no_signature": dict(tol=0.01, signature=False),
    "no_midpoint": dict(tol=0.0, signature=True),
    "no_refinement": dict(tol=0.01, signature=True, max_refine_levels=0),
    "deep": dict(tol=0.003, signature=True, max_refine_levels=7, min_split_ns=1024),
}


@pytest.mark.parametrize("rules_name", sorted(RULE_VARIANTS))
@pytest.mark.parametrize("scene", ["levels", "rectangles"])
def test_one_pass_sampler_equals_two_passes(scene, rules_name):
    """The one-pass sampler (each pixel's first k samples captured while it is counted, then copied into
    place; a pixel with more samples sampled again in place) gives the runs of the two-pass sampler (count,
    then write), with slots for every pixel's samples, with 4 slots (nearly every pixel sampled again) and
    with a budget below one slot per pixel (two passes); under the default rules and with signature flags, the
    midpoint test or refinement off, and with seven levels of refinement down to 1024-ns intervals."""
    from ballasim_amd import scene as S
    if scene == "levels":
        sc, view, _, _ = _level_scene()
    else:
        shapes, bg = S.config2_shapes(5, n=16, width=320, height=180)
        sc, view = S.Scene(shapes, bg), S.view(320, 180)
    P = view.width * view.height
    rules = S.rules(0.0021, **RULE_VARIANTS[rules_name])
    budgets = {"two passes": 0, "4 slots": 4 * 8 * P, "default": None}
    for w in range(3):
        got = {}
        for name, b in budgets.items():
            if b is not None:
                sc.set_capture_budget(b)
            else:
                sc.set_capture_budget(16 << 30)
            runs, total = sc.sample_window(view, rules, w * WINDOW, slot=0)
            got[name] = S.runs_to_host(runs) + (total,)
        ref = got["two passes"]
        for name, g in got.items():
            for a, b_ in zip(ref, g):
                assert np.array_equal(np.asarray(a), np.asarray(b_)), f"window {w}: {name} differs from two passes"
    budget, overflow = sc.capture_budget()
    assert budget == 16 << 30
    assert overflow >= 3  # the 4-slot windows
