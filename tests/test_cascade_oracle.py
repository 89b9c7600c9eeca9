"""The oracle's cascade photoreceptor against what the mathematics fixes, independently of the
reference's formulas (CPU):

- stage 1 (z_step) against a stiff ODE solution of dz/dt = k (1 - I(t) z) on linear I, across its
  four branches, to the reference's own fp32 bound of 3e-6 in ln z (tests/sensor/test_stage1.cpp:42-77);
- the comparator input g(u) = cubic + C0 R_j(a0 u) + C1 R_j(a1 u) against fp64, and its critical
  points against a dense fp64 scan of g';
- first firing against a brute-force scan of every integer ns (tests/sensor/test_expcubic.cpp:73-124);
- the whole cascade pixel: the event times of a brightness step against the 3-stage ODE solved
  numerically (ideal reset, no leak: event j where s3(t) + phi0 first reaches j theta_on).
"""
import ctypes as C
import math

import numpy as np
import pytest
from scipy.integrate import solve_ivp

from ballasim_amd import _abi
from oracle import oracle as O

F32 = np.float32


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def z_ode(z0, I0, I1, h, k):
    """z(h) for dz/dt = k (1 - I(t) z), I linear from I0 to I1 over [0, h], in fp64 (Radau)."""
    z0, I0, I1, h, k = map(float, (z0, I0, I1, h, k))
    s = (I1 - I0) / h

    def f(t, y):  # y = z I0 (dimensionless)
        return [k * (I0 - (I0 + s * t) * y[0])]

    def jac(t, y):
        return [[-k * (I0 + s * t)]]

    r = solve_ivp(f, (0.0, h), [z0 * I0], method="Radau", jac=jac, rtol=1e-12, atol=1e-15)
    assert r.success
    return r.y[0, -1] / I0


def test_z_step_against_ode_all_branches():
    lib = O.oracle_lib()
    rng = np.random.default_rng(7)
    k = 10.0  # (fA s)^-1
    seen = set()
    worst = 0.0
    for _ in range(240):
        I0 = float(np.exp(rng.uniform(np.log(1e-2), np.log(1e5))))
        kind = rng.integers(0, 4)
        if kind == 0:
            I1 = I0 * float(np.exp(rng.uniform(-3, 3)))
        elif kind == 1:
            I1 = I0 * (1 + float(rng.uniform(-1e-5, 1e-5)))  # tiny slope (small-slope / ZOH branches)
        elif kind == 2:
            I1 = I0
        else:
            I1 = I0 * float(np.exp(rng.uniform(-0.3, 0.3)))
        h = float(np.exp(rng.uniform(np.log(1e-9), np.log(3e-3))))
        I0f, I1f, hf = F32(I0), F32(I1), F32(h)
        z0f = F32(1.0 / float(I0f) * float(np.exp(rng.uniform(-1, 1))))
        P = k * float(hf) * 0.5 * (float(I0f) + float(I1f))
        seen.add("gl8" if P <= 0.5 else ("hold" if I0f == I1f else ("rise" if I1f > I0f else "fall")))
        got = lib.orc_z_step(z0f, I0f, I1f, hf, F32(k))
        ref = z_ode(z0f, I0f, I1f, hf, k)
        err = abs(math.log(got) - math.log(ref))
        worst = max(worst, err)
        assert err < 3e-6, (float(z0f), float(I0f), float(I1f), float(hf), got, ref)
    assert seen == {"gl8", "hold", "rise", "fall"}
    assert worst > 0.0


def _resid(j, y):
    """R_j(y) = e^-y minus its Taylor polynomial of degree < j, in fp64 (series for small y)."""
    y = np.asarray(y, np.float64)
    if j == 0:
        return np.exp(-y)
    out = np.empty_like(y)
    small = np.abs(y) < 0.5
    ys = y[small]
    term = (-ys) ** j / math.factorial(j)
    acc = term.copy()
    for n in range(j + 1, j + 30):
        term = term * (-ys) / n
        acc += term
    out[small] = acc
    yl = y[~small]
    poly = sum((-yl) ** n / math.factorial(n) for n in range(j))
    out[~small] = np.exp(-yl) - poly
    return out


def _random_g(rng, order):
    w = float(np.exp(rng.uniform(np.log(2e-6), np.log(4e-4))))
    a = np.array([1 / float(np.exp(rng.uniform(np.log(5e-5), np.log(2e-3)))) for _ in range(2)], F32)
    if order == 4:
        # the residual split: an O(1) cubic Taylor part plus large free-mode amplitudes times O((a h)^4)
        C_ = np.array(rng.normal(0, 1, 2) / np.maximum(a * w, 1e-3) ** 2, F32)
    else:
        C_ = np.array(rng.normal(0, 0.2, 2), F32)
    c = np.array([rng.normal(0, 0.1), rng.normal(0, 0.4) / w, rng.normal(0, 0.3) / w ** 2,
                  rng.normal(0, 0.3) / w ** 3], F32)
    return c, C_, a, w


def _g64(c, C_, a, order, u):
    u = np.asarray(u, np.float64)
    cc = c.astype(np.float64)
    out = cc[0] + u * (cc[1] + u * (cc[2] + u * cc[3]))
    for i in range(2):
        if C_[i] != 0:
            out = out + float(C_[i]) * _resid(order, float(a[i]) * u)
    return out


def _eval32(c, C_, a, order, u):
    lib = O.oracle_lib()
    u = np.ascontiguousarray(u, F32)
    out = np.empty_like(u)
    lib.orc_xc_eval_n(_p(c), _p(C_), _p(a), order, _p(u), u.size, _p(out))
    return out


@pytest.mark.parametrize("order", [0, 4])
def test_expcubic_eval_against_fp64(order):
    rng = np.random.default_rng(11 + order)
    for _ in range(60):
        c, C_, a, w = _random_g(rng, order)
        u = np.linspace(0, w, 257).astype(F32)
        got = _eval32(c, C_, a, order, u).astype(np.float64)
        ref = _g64(c, C_, a, order, u.astype(np.float64))
        # fp32 rounding of the terms: a few ulp of the largest term's magnitude
        scale = np.abs(c[0]) + np.abs(c[1]) * w + np.abs(c[2]) * w * w + np.abs(c[3]) * w ** 3 + \
            np.abs(C_).sum() * (1.0 if order == 0 else float(np.max(a)) ** 4 * w ** 4 / 24 + 1e-30)
        assert np.max(np.abs(got - ref)) <= 64 * 2.0 ** -23 * scale + 1e-7, (c, C_, a, w)


@pytest.mark.parametrize("order", [0, 4])
def test_critical_points_split_monotone_pieces(order):
    """Every sign change of g' on a dense fp64 grid lies within one grid step of a returned critical
    point, and no returned point lies where g' keeps its sign far around it."""
    lib = O.oracle_lib()
    rng = np.random.default_rng(21 + order)
    checked = 0
    for _ in range(120):
        c, C_, a, w = _random_g(rng, order)
        # a cubic part with two interior extrema: g' = A (u - r1)(u - r2) plus the free modes
        r1, r2 = np.sort(rng.uniform(0.05, 0.95, 2)) * w
        A = rng.normal(0, 1) * 2.0 / w ** 3
        c = np.array([c[0], A * r1 * r2, -A * (r1 + r2) / 2, A / 3], F32)
        C_ = (C_ * F32(0.05)).astype(F32)
        out = np.zeros(6, F32)
        n = lib.orc_xc_critical(_p(c), _p(C_), _p(a), order, F32(0), F32(w), _p(out))
        crit = np.sort(out[:n].astype(np.float64))
        u = np.linspace(0, w, 20001)
        du = u[1] - u[0]
        d = np.gradient(_g64(c, C_, a, order, u), u)
        sc = np.nonzero(np.sign(d[1:]) * np.sign(d[:-1]) < 0)[0]
        # ignore sign flips of a derivative that is within rounding of zero
        big = np.abs(d).max()
        sc = [i for i in sc if max(abs(d[i]), abs(d[i + 1])) > 1e-4 * big]
        for i in sc:
            x = u[i] + du / 2
            assert n > 0 and np.min(np.abs(crit - x)) <= 3 * du + 1e-3 * w, (x, crit)
            checked += 1
    assert checked > 20


@pytest.mark.parametrize("order", [0, 4])
def test_first_firing_brute_force(order):
    """search(): the first integer ns of [lo, e1) where g - theta_on >= 0 or -g - theta_off >= 0
    (fp32 evaluation of both firing functions), against a scan of every integer; 0 mismatches."""
    lib = O.oracle_lib()
    rng = np.random.default_rng(31 + order)
    hits = 0
    for trial in range(150):
        c, C_, a, w = _random_g(rng, order)
        n_ns = int(w * 1e9)
        e0 = int(rng.integers(-10 ** 6, 10 ** 6))
        e1 = e0 + n_ns + 1
        lo = e0 + int(rng.integers(0, max(1, n_ns // 4)))
        th_on, th_off = F32(rng.uniform(0.05, 0.3)), F32(rng.uniform(0.05, 0.3))
        on = C.c_int(0)
        t = C.c_int64(0)
        found = lib.orc_xc_search(_p(c), _p(C_), _p(a), order, e0, e1, lo, th_on, th_off, C.byref(on), C.byref(t))
        k = np.arange(lo - e0, e1 - e0, dtype=np.int64)
        u = (k.astype(F32) * F32(1e-9)).astype(F32)
        c_on = c.copy()
        c_on[0] = F32(c[0] - th_on)
        c_off = (-c).astype(F32)
        c_off[0] = F32(-c[0] - th_off)
        g_on = _eval32(c_on, C_, a, order, u)
        g_off = _eval32(c_off, (-C_).astype(F32), a, order, u)
        fire = np.nonzero((g_on >= 0) | (g_off >= 0))[0]
        if len(fire) == 0:
            assert found == 0, trial
            continue
        i = fire[0]
        assert found == 1, trial
        assert t.value == e0 + k[i], (trial, t.value, e0 + k[i])
        assert on.value == (1 if g_on[i] >= 0 else 0), trial
        hits += 1
    assert hits > 30


def _cascade_params(**kw):
    lib = _abi.load_library()
    p = _abi.Params()
    _abi.check(lib.bsx_params_preset(b"prophesee_gen3", C.byref(p)))
    p.noise = 0
    p.leak_rate_hz = 0.0
    p.reset_mode = 1  # ideal
    p.refractory_s = 0.0
    p.sigma_theta_on = p.sigma_theta_off = 0.0
    p.dark_sigma_ln = 0.0
    for key, v in kw.items():
        setattr(p, key, v)
    return p


@pytest.mark.parametrize("ratio,ramp_ns", [(math.exp(0.9), 1000), (math.exp(0.55), 200000), (math.exp(-0.8), 5000)])
def test_cascade_step_response_against_ode(ratio, ramp_ns):
    p = _cascade_params()
    assert p.photoreceptor == 0  # cascade
    E0 = 50.0
    E1 = float(F32(E0 * ratio))
    m = O.OracleModel(p, 1, 1, 3, 1, 0, 1)
    m.init(0, np.array([[E0]], np.float32))
    maps = m.maps()
    th_on, th_off, dark, phi0 = (float(F32(maps["theta_on"][0])), float(F32(maps["theta_off"][0])),
                                 float(F32(maps["dark_fA"][0])), float(F32(maps["phi0"][0])))
    scale = float(F32(p.illuminance_scale))
    t_a, t_b, t_end = 10_000, 10_000 + ramp_ns, 40_000_000
    ev, _ = m.advance_runs(np.array([0], np.uint32), np.array([t_a, t_b], np.int64),
                           np.array([[E0], [E1]], np.float32), [0, 2], t_end)
    I0 = float(F32(dark + scale * E0))
    I1 = float(F32(dark + scale * E1))
    k, tau_o, tau_sf = p.k * 1e-15, p.tau_o, p.tau_sf

    def I(t_ns):
        if t_ns <= t_a:
            return I0
        if t_ns >= t_b:
            return I1
        return I0 + (I1 - I0) * (t_ns - t_a) / (t_b - t_a)

    def f(t, y):  # t in s; y = (z I0, s2, s3)
        tn = t * 1e9
        zi = y[0]
        x1 = -math.log(zi)
        return [k * (I0 - I(tn) * zi), (x1 - y[1]) / tau_o, (y[1] - y[2]) / tau_sf]

    sol = solve_ivp(f, (t_a * 1e-9, t_end * 1e-9), [1.0, 0.0, 0.0], method="Radau", rtol=1e-10, atol=1e-13,
                    dense_output=True, max_step=2e-5)
    assert sol.success
    s3 = lambda t_ns: float(sol.sol(t_ns * 1e-9)[2])
    final = math.log(I1 / I0)
    if final > 0:
        expected = int(math.floor((final + phi0) / th_on))
        levels = [j * th_on for j in range(1, expected + 1)]
    else:
        expected = int(math.floor((-final - phi0) / th_off + 1e-12)) if -final > phi0 else 0
        levels = [-(j * th_off) for j in range(1, expected + 1)]
    assert len(ev) == expected, (len(ev), expected)
    assert expected >= 2
    tol = 1e-3  # log units: four times the sub-step tolerance eps_L / 8
    for j, e in enumerate(ev):
        t = int(e["t_cross"])
        assert e["polarity"] == (1 if final > 0 else 0)
        v_at = s3(t) + phi0 - levels[j]
        v_before = s3(t - 1) + phi0 - levels[j]
        if final > 0:
            assert v_at >= -tol and v_before <= tol, (j, t, v_at, v_before)
        else:
            assert v_at <= tol and v_before >= -tol, (j, t, v_at, v_before)
