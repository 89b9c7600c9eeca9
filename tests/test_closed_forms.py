"""The oracle's instantaneous response against what the mathematics fixes, independently of the
reference (CPU): a static pixel never fires; a leak-only pixel fires at t0 + ceil((theta - phi0)/l)
and then every T_r + ceil(theta/l); a linear ramp of photocurrent gives floor(ln(I_b/I_a)/theta)
ideal-reset events at the closed-form times; and a brute-force per-nanosecond scan in fp64 agrees
on small random inputs (the reference's own oracles: tests/sensor/test_pixel.cpp:92-194)."""
import ctypes as C
import math

import numpy as np

from ballasim_amd import _abi
from oracle import oracle as O


def ideal(**kw):
    lib = _abi.load_library()
    p = _abi.Params()
    _abi.check(lib.bsx_params_preset(b"ideal", C.byref(p)))
    p.initial_phase = 2  # zero: the reference starts at the pixel's own level
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def one_pixel(p, irr0, seed=1):
    m = O.OracleModel(p, 1, 1, seed, 1, 0, 1)
    m.init(0, np.array([[irr0]], np.float32))
    return m


def test_static_pixel_never_fires():
    m = one_pixel(ideal(), 100.0)
    t = np.arange(1, 200, dtype=np.int64) * 100003
    ev, _ = m.advance_runs(np.array([0], np.uint32), t, np.full((len(t), 1), 100.0, np.float32),
                           [0, len(t)], int(t[-1]) + 5)
    assert len(ev) == 0


def test_leak_only_events_closed_form():
    theta, leak_hz, T_r = 0.2, 40.0, 300e-6
    p = ideal(reset_mode=0, refractory_s=T_r, leak_rate_hz=leak_hz, theta_on=theta, theta_off=theta,
              initial_phase=0)
    m = one_pixel(p, 50.0, seed=3)
    maps = m.maps()
    l, phi0 = maps["leak"][0], maps["phi0"][0]
    assert abs(l - theta * leak_hz) < 1e-12
    t_end = 200_000_000
    ev, _ = m.advance_runs(np.zeros(0, np.uint32), np.zeros(0, np.int64), np.zeros((0, 1), np.float32), [0], t_end)
    # Delta(t) = phi0 + l (t - t0) until the first crossing; then blind for T_r, re-referenced at release.
    want = [math.ceil((theta - phi0) / l * 1e9)]
    while True:
        nxt = want[-1] + round(T_r * 1e9) + math.ceil(theta / l * 1e9)
        if nxt >= t_end:
            break
        want.append(nxt)
    got = ev["t_cross"]
    assert len(got) == len(want)
    assert np.all(np.abs(got - np.array(want)) <= 1)
    assert np.all(ev["polarity"] == 1) and np.all(ev["label"] == 2)  # leak events


def test_linear_ramp_esim_count_and_times():
    """I(t) linear from I_a to I_b over [0, T]: with the ideal reset and phi0 = 0 the k-th ON event is
    the first ns where I(t) >= I_a e^{k theta} (no dark: dark_current tiny)."""
    theta = 0.2
    p = ideal(theta_on=theta, theta_off=theta, dark_current_fA=1e-6)
    I_a, I_b, T = 10.0, 10.0 * math.exp(5.0), 5_000_000
    m = one_pixel(p, I_a)
    ev, _ = m.advance_runs(np.array([0], np.uint32), np.array([T], np.int64), np.array([[I_b]], np.float32),
                           [0, 1], T + 1)
    n = math.floor(5.0 / theta + 1e-9)
    assert len(ev) in (n, n - 1)  # the last level sits exactly at I_b: within rounding either way
    slope = (np.float32(I_b) - I_a) / T
    for k, t in enumerate(ev["t_cross"], start=1):
        exact = (I_a * math.exp(k * theta) - I_a) / slope
        assert abs(t - math.ceil(exact)) <= 1, (k, t, exact)
    assert np.all(ev["polarity"] == 1)


def _brute_force(seg_t, seg_v, dark, theta_on, theta_off, t_end):
    """Per-ns scan of the ideal-reset comparator on piecewise-linear I, in fp64: the events a
    continuous-valued model gives, with each crossing's margin past the level."""
    out = []
    ref = math.log(dark + seg_v[0])
    for i in range(len(seg_t) - 1):
        t0, t1 = seg_t[i], seg_t[i + 1]
        for t in range(max(t0, 1) if i == 0 else t0, t1):
            I = dark + seg_v[i] + (seg_v[i + 1] - seg_v[i]) * (t - t0) / (t1 - t0)
            d = math.log(I) - ref
            if d >= theta_on:
                out.append((t, 1, d - theta_on))
                ref += theta_on
            elif d <= -theta_off:
                out.append((t, 0, -theta_off - d))
                ref -= theta_off
    return out


def test_brute_force_scan_small_random():
    rng = np.random.default_rng(7)
    checked = 0
    for trial in range(12):
        theta = float(rng.uniform(0.1, 0.3))
        p = ideal(theta_on=theta, theta_off=theta, dark_current_fA=0.5)
        n = 6
        t = np.sort(rng.choice(np.arange(1, 40000), n, replace=False)).astype(np.int64)
        v = np.exp(rng.normal(3, 1.2, n)).astype(np.float32)
        v0 = np.float32(np.exp(rng.normal(3, 1)))
        m = one_pixel(p, float(v0))
        ev, _ = m.advance_runs(np.array([0], np.uint32), t, v.reshape(-1, 1), [0, n], int(t[-1]) + 1)
        dark = float(np.float32(m.maps()["dark_fA"][0]))
        bf = _brute_force([0] + list(t), [float(v0)] + [float(x) for x in v], dark, theta, theta, int(t[-1]))
        # Compare while every crossing is robust (the fp32 and fp64 evaluations cannot disagree).
        for (tb, pb, margin), te, pe in zip(bf, ev["t_cross"], ev["polarity"]):
            if margin < 1e-4:
                break
            assert pb == pe and abs(int(te) - tb) <= 1
            checked += 1
    assert checked > 50
