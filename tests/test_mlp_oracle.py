"""The proposer's CPU oracle on its own: backward pass against central finite differences of the fp64
forward (no rounding), bf16 rounding against known bit patterns, Adam against its closed form on the
first step."""
import numpy as np

from oracle import mlp_oracle as M


def test_bf16_rounding_known_values():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -8 + 2 ** -20, 1.0 + 3 * 2 ** -8, -2.5, 3.0e38], np.float32)
    got = M.bf16(x).view(np.uint32) >> 16
    # 1 + 2^-8 is a tie: rounds to even (1.0); just above the tie rounds up; 1 + 3 * 2^-8 is a tie
    # between 1 + 2^-7 and 1 + 2^-6 ... and rounds to the even mantissa.
    assert list(got) == [0x3F80, 0x3F80, 0x3F81, 0x3F82, 0xC020, 0x7F62]


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(0)
    w = rng.normal(0, 0.3, M.N_PARAMS)
    x = rng.normal(0, 1, (37, M.N_IN))
    t = rng.normal(0, 1, 37)
    g, _ = M.gradients(w, x, t, emulate=False)
    idx = np.concatenate([rng.choice(np.arange(a, b), 12, replace=False) for a, b in
                          ((M.OFF_W1, M.OFF_B1), (M.OFF_B1, M.OFF_W2), (M.OFF_W2, M.OFF_B2), (M.OFF_B2, M.OFF_W3),
                           (M.OFF_W3, M.OFF_B3))] + [[M.OFF_B3]])
    h = 1e-6
    for i in idx:
        wp, wm = w.copy(), w.copy()
        wp[i] += h
        wm[i] -= h
        lp = np.mean((M.forward(wp, x, emulate=False)[0] - t) ** 2)
        lm = np.mean((M.forward(wm, x, emulate=False)[0] - t) ** 2)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * (1 + abs(fd)), (i, fd, g[i])


def test_emulated_forward_is_close_to_exact():
    rng = np.random.default_rng(1)
    w = rng.normal(0, 0.2, M.N_PARAMS)
    x = rng.normal(0, 1, (200, M.N_IN))
    ye, _ = M.forward(w, x, emulate=True)
    yx, _ = M.forward(w, x, emulate=False)
    assert np.max(np.abs(ye - yx)) < 0.05 * (1 + np.max(np.abs(yx)))


def test_adam_first_step_is_signed_lr():
    g = np.array([0.5, -2.0, 1e-3, 0.0])
    w, m, v = M.adam_step(np.zeros(4), g, np.zeros(4), np.zeros(4), 1, 0.01)
    np.testing.assert_allclose(w, -0.01 * g / (np.abs(g) + 1e-8), rtol=1e-12)
