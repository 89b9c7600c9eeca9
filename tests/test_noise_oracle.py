"""The oracle's in-path noise (oracle/orc_noise.inc) pinned three ways, on the CPU.

1. The tables against mathematics that fixes them: the transition Phi(h) = expm(A h), the
   covariance Q(h) from Van Loan's block exponential, the stationary covariance from the Lyapunov
   equation, the midpoint bridge from Gaussian conditioning and the exclusion margin's whitening
   W C_ee W^T = I with its bridge residual (src/sensor/noise_tables.cpp:1-30 states these laws).
2. The keyed normals against the reference's golden values (tests/sensor/test_noise.cpp:445-454).
3. The tables, the normals and whole noise-on runs (events, fp64 states, work counters) bit for bit
   against the reference compiled from its own sources (oracle/_ref), where it is present.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import first_difference, random_runs

Q_E = 1.602176634e-19
CELL = 1 << 16


def _params(**kw):
    import ballasim_amd as bsx
    p = bsx.preset("prophesee_gen3")
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def _system(p, a1):
    a2, a3 = 1.0 / p.tau_o, 1.0 / p.tau_sf
    A = np.array([[-a1, 0, 0], [a2, -a2, 0], [0, a3, -a3]])
    GG = np.diag([2 * p.k * Q_E * a1, 2 * p.sigma_o ** 2 * a2, 0.0])
    return A, GG


def _transition(A, GG, h):
    from scipy.linalg import expm
    M = np.zeros((6, 6))
    M[:3, :3] = -A
    M[:3, 3:] = GG
    M[3:, 3:] = A.T
    F = expm(M * h)
    phi = F[3:, 3:].T
    return phi, phi @ F[:3, 3:]


def _lower(L):
    return np.array([L[0, 0], L[1, 0], L[1, 1], L[2, 0], L[2, 1], L[2, 2]])


def _full(v):
    return np.array([[v[0], 0, 0], [v[1], v[2], 0], [v[3], v[4], v[5]]])


TABLE_CASES = {
    "prophesee_gen3": dict(),
    "levels3_wide": dict(noise_leaf_level=3, tau_o=40e-6, tau_sf=300e-6, sigma_o=0.15, k=3e15),
}


@pytest.mark.parametrize("case", sorted(TABLE_CASES))
def test_tables_satisfy_their_laws(case):
    """Rows of nodes with a1 h0 in [1e-4, 20] (where the block exponential is accurate): the step,
    its covariance factor, the stationary factor, every level's bridge and the level-0..L-1 margins."""
    from scipy.linalg import solve_continuous_lyapunov
    p = _params(**TABLE_CASES[case])
    d64, d32 = O.noise_tables(p.c)
    L = p.noise_leaf_level
    n_nodes, stride = d64.shape
    assert stride == 18 + 47 * L
    assert np.array_equal(d32, d64.astype(np.float32))
    dln = np.log(10.0) / 64
    h0 = CELL * 1e-9
    checked = 0
    for node in range(0, n_nodes, 37):
        a1 = np.exp(-4.605170185988091 + node * dln)
        if not (1e-4 <= a1 * h0 <= 20):
            continue
        checked += 1
        row = d64[node]
        A, GG = _system(p, a1)
        phi, Q = _transition(A, GG, h0)
        np.testing.assert_allclose(_full(row[0:6]), phi, rtol=1e-9, atol=1e-14)
        C = _full(row[6:12])
        np.testing.assert_allclose(C @ C.T, Q, rtol=1e-7, atol=1e-12 * np.abs(Q).max())
        P = solve_continuous_lyapunov(A, -GG)
        Cp = _full(row[12:18])
        np.testing.assert_allclose(Cp @ Cp.T, P, rtol=1e-9, atol=1e-15)
        for lev in range(1, L + 1):
            ph, qh = _transition(A, GG, h0 / 2 ** lev)
            parent = ph @ qh @ ph.T + qh
            K = qh @ ph.T @ np.linalg.inv(parent)
            Am = ph - K @ ph @ ph
            sig = qh - K @ ph @ qh
            b = row[18 + 24 * (lev - 1): 18 + 24 * lev]
            scale = np.abs(Am).max()
            np.testing.assert_allclose(b[:9].reshape(3, 3), Am, atol=1e-7 * scale)
            np.testing.assert_allclose(b[9:18].reshape(3, 3), K, atol=1e-7 * max(np.abs(K).max(), 1.0))
            Ls = _full(b[18:24])
            np.testing.assert_allclose(Ls @ Ls.T, sig, atol=1e-6 * np.abs(sig).max())
        for lev in range(L):
            mblk = row[18 + 24 * L + 23 * lev: 18 + 24 * L + 23 * (lev + 1)]
            h = h0 / 2 ** lev
            phi_h, _ = _transition(A, GG, h)
            Cee = np.block([[P, P @ phi_h.T], [phi_h @ P, P]])
            W = np.zeros((6, 6))
            W[np.tril_indices(6)] = mblk[2:]
            np.testing.assert_allclose(W @ Cee @ W.T, np.eye(6), atol=1e-6)
            sig_br = 0.0
            for iu in range(1, 20):
                pl, _ = _transition(A, GG, h * iu / 20)
                pr, _ = _transition(A, GG, h * (20 - iu) / 20)
                c = np.concatenate([(pl @ P)[2], (P @ pr.T)[2]])
                sig_br = max(sig_br, np.sqrt(max(P[2, 2] - c @ np.linalg.solve(Cee, c), 0.0)))
            assert mblk[0] == pytest.approx(sig_br, rel=1e-5)
    assert checked >= 5


def test_keyed_normals_goldens():
    """normals3 in fp64 at the reference's golden keys (tests/sensor/test_noise.cpp:449-454); the fp32
    form draws from the same words (23 of their bits)."""
    d, _ = O.normals3(0, 0x10, 0, 0, 0)
    assert d[0] == float.fromhex("0x1.9314987575cccp+0")
    assert d[2] == float.fromhex("0x1.907c8fda1f566p-3")
    d, _ = O.normals3(0x0123456789ABCDEF, 0x10, 1, 307199, 65536)
    assert d[2] == float.fromhex("-0x1.65145c21ebbb9p+1")
    d, f = O.normals3(0x0123456789ABCDEF, 0x11, 1, 307199, 65536)
    assert d[0] == float.fromhex("-0x1.5de1b9dc29e54p-1")
    assert np.all(np.abs(f - d) < 1e-5)
    # standard normal moments over many keys
    rng = np.random.default_rng(0)
    xs = np.array([O.normals3(42, 0x10, 0, int(px), int(t) << 16)[1]
                   for px, t in zip(rng.integers(0, 1000, 4000), rng.integers(0, 10 ** 6, 4000))])
    assert abs(xs.mean()) < 0.04 and abs(xs.std() - 1.0) < 0.03


needs_ref = pytest.mark.skipif(not O.ref_available(), reason="the compiled reference (oracle/_ref) is absent")


@needs_ref
@pytest.mark.parametrize("case", sorted(TABLE_CASES) + ["levels0", "levels10"])
def test_tables_equal_reference(case):
    kw = dict(TABLE_CASES.get(case, {}))
    if case == "levels0":
        kw["noise_leaf_level"] = 0
    if case == "levels10":
        kw["noise_leaf_level"] = 10
    p = _params(**kw)
    a64, a32 = O.noise_tables(p.c)
    b64, b32 = O.noise_tables(p.c, lib="ref")
    assert np.array_equal(a64.view(np.uint64), b64.view(np.uint64))
    assert np.array_equal(a32.view(np.uint32), b32.view(np.uint32))


@needs_ref
def test_normals_equal_reference():
    rng = np.random.default_rng(3)
    for _ in range(3000):
        key = (int(rng.integers(0, 2 ** 63)), int(rng.choice([0x10, 0x11])), int(rng.integers(0, 4)),
               int(rng.integers(0, 2 ** 32)), int(rng.integers(-2 ** 40, 2 ** 40)) & (2 ** 64 - 1))
        a64, a32 = O.normals3(*key)
        b64, b32 = O.normals3(*key, lib="ref")
        assert np.array_equal(a64, b64) and np.array_equal(a32.view(np.uint32), b32.view(np.uint32)), key


STATE_FIELDS = ("z", "x1", "s2", "s3", "tv", "s3_ref", "n3_ref", "dr_ref", "t_ref", "t_search", "blind_until", "nk",
                "nk1", "ln_a1", "cell", "have_k1")

RUN_CASES = {
    "default": dict(),
    "ideal": dict(reset_mode="ideal", refractory_s=1e-5),
    "leak_loud": dict(leak_rate_hz=2.0, refractory_s=5e-5, sigma_o=0.2),
    "levels0": dict(noise_leaf_level=0),
    "levels8_eps": dict(noise_leaf_level=8, eps_ci=0.002, refractory_s=2e-5),
}


def _compare(o, r, where):
    a, _ = o
    b, _ = r
    i, x, y = first_difference(a, b.astype(O.EVENT_DTYPE))
    assert len(a) == len(b) and x is None, f"{where}: events differ at {i}: {x} / {y}"


def _static_runs(rng, P, t0, t1, irr0, frac_static=0.5, n=6):
    """Runs whose first pixels keep their init irradiance (constant, settled: quiet cells) and the rest
    random."""
    ids = np.arange(P, dtype=np.uint32)
    ts, irs, offs = [], [], [0]
    for p in ids:
        tt = np.unique(rng.integers(t0 + 1, t1 + 1, size=n)).astype(np.int64)
        if p < frac_static * P:
            v = np.repeat(irr0[p:p + 1], len(tt), axis=0)
        else:
            v = np.exp(rng.normal(3, 1.0, size=(len(tt), 1))).astype(np.float32)
        ts.append(tt)
        irs.append(v.astype(np.float32))
        offs.append(offs[-1] + len(tt))
    return ids, np.concatenate(ts), np.concatenate(irs), np.array(offs, np.int64)


@needs_ref
@pytest.mark.parametrize("case", sorted(RUN_CASES) + ["quiet_cells"])
def test_noise_runs_equal_reference(case):
    """Noise-on cascade runs, bit for bit: events (all fields), the fp64 states (noise nodes, rate,
    cell included) and the work counters (cells and bridge refinements included), over windows that
    start mid-cell, with init at a time that is not a lattice node."""
    W, H, seed = 20, 14, 5 + len(case)
    p = _params(**RUN_CASES.get(case, {}))
    rng = np.random.default_rng(len(case))
    o = O.OracleModel(p.c, W, H, seed, 1, 0, 2)
    r = O.RefModel(p.c, W, H, seed, 1, 0, 2, True)
    irr0 = np.exp(rng.normal(3, 1, size=(W * H, 1))).astype(np.float32)
    t0 = -70_000
    o.init(t0, irr0)
    r.init(t0, irr0)
    t = t0
    events = 0
    for w in range(3):
        t1 = (w + 1) * (2 ** 21) + (0 if case == "quiet_cells" else 1234 * w)
        if case == "quiet_cells":
            ids, ts, irr, offs = _static_runs(rng, W * H, t, t1, irr0)
        else:
            ids, ts, irr, offs = random_runs(rng, W * H, t, t1, max_n=12)
        a = o.advance_runs(ids, ts, irr, offs, t1)
        b = r.advance_runs(ids, ts, irr, offs, t1)
        _compare(a, b, f"{case} window {w}")
        so, sr = o.states(), r.states()
        for f in STATE_FIELDS:
            assert np.array_equal(so[f], sr[f]), f"{case} window {w}: state {f}"
        assert o.stats() == {k: v for k, v in r.stats().items()}, f"{case} window {w}: counters"
        events += len(a[0])
        t = t1
    st = o.stats()
    assert events > 50 and st["cells"] > 0
    if case == "quiet_cells":
        assert st["settled_substeps"] > 0.3 * st["cells"]
    if case != "levels0":
        assert st["refinements"] > 0
