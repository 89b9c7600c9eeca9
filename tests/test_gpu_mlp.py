"""GPU parity of the proposer's kernels with the CPU oracle (oracle/mlp_oracle.py, itself checked
against finite differences): the MFMA tile layout, the forward pass, the gradients of the mean squared
error, the Adam step, and a training run that fits a smooth target.

Tolerances (floating point): the kernels sum bf16 products in fp32 on the MFMA units and use the
hardware exp in silu, the oracle sums in fp64 with the same bf16 rounding points; a bf16 rounding of an
activation can then differ by one unit (2^-8 relative) for inputs near a rounding boundary. Outputs
agree to 1e-2 (absolute and relative) elementwise with a median difference below 1e-4; gradient blocks
to 2e-2 of their norm."""
import numpy as np
import pytest

from oracle import mlp_oracle as M

pytestmark = pytest.mark.gpu


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, np.float32)).to("cuda:0")


def test_mfma_tile_layout():
    from ballasim_amd.learned import tile_check
    rng = np.random.default_rng(0)
    a = M.bf16(rng.normal(0, 1, (32, 8)).astype(np.float32))
    b = M.bf16(rng.normal(0, 1, (8, 32)).astype(np.float32))
    c = tile_check(a.view(np.uint32) >> 16, b.view(np.uint32) >> 16)
    want = a.astype(np.float64) @ b.astype(np.float64)
    np.testing.assert_allclose(c, want, rtol=1e-5, atol=1e-5)


def _random_net(seed, scale=0.3):
    rng = np.random.default_rng(seed)
    w = rng.normal(0, scale, M.N_PARAMS).astype(np.float32)
    return w, rng


def test_forward_matches_oracle():
    import torch

    from ballasim_amd.learned import Proposer
    w, rng = _random_net(1)
    p = Proposer(3)
    p.params = w
    assert np.array_equal(p.params, w)
    x = rng.normal(0, 1, (1000, 12)).astype(np.float32)
    xd = _dev(x)
    yd = torch.empty(1000, dtype=torch.float32, device="cuda:0")
    p.predict(xd, yd)
    y = yd.cpu().numpy()
    want, _ = M.forward(w, x)
    np.testing.assert_allclose(y, want, rtol=1e-2, atol=1e-2)
    assert np.median(np.abs(y - want)) < 1e-4


@pytest.mark.parametrize("n", [64, 1000])
def test_gradients_match_oracle(n):
    from ballasim_amd.learned import Proposer
    w, rng = _random_net(2 + n)
    p = Proposer(0)
    p.params = w
    x = rng.normal(0, 1, (n, 12)).astype(np.float32)
    t = rng.normal(0, 1, n).astype(np.float32)
    g, loss = p.gradients(_dev(x), _dev(t))
    gw, lw = M.gradients(w, x, t)
    assert abs(loss - lw) <= 1e-3 * (1 + lw)
    for a, b in ((M.OFF_W1, M.OFF_B1), (M.OFF_B1, M.OFF_W2), (M.OFF_W2, M.OFF_B2), (M.OFF_B2, M.OFF_W3),
                 (M.OFF_W3, M.OFF_B3), (M.OFF_B3, M.N_PARAMS)):
        d = np.linalg.norm(g[a:b] - gw[a:b])
        assert d <= 2e-2 * (np.linalg.norm(gw[a:b]) + 1e-6), (a, b, d, np.linalg.norm(gw[a:b]))


def test_adam_step_matches_oracle():
    from ballasim_amd.learned import Proposer
    w, rng = _random_net(5)
    p = Proposer(0)
    p.params = w
    x = rng.normal(0, 1, (256, 12)).astype(np.float32)
    t = rng.normal(0, 1, 256).astype(np.float32)
    xd, td = _dev(x), _dev(t)
    g, _ = p.gradients(xd, td)
    p.train(xd, td, batch=256, epochs=1, lr=1e-2)
    w1 = p.params
    want, _, _ = M.adam_step(w.astype(np.float64), g.astype(np.float64), 0.0, 0.0, 1, 1e-2)
    np.testing.assert_allclose(w1, want, rtol=1e-5, atol=1e-6)


def test_training_fits_a_smooth_target():
    from ballasim_amd.learned import Proposer
    rng = np.random.default_rng(7)
    x = rng.uniform(-1, 1, (65536, 12)).astype(np.float32)
    t = (np.sin(2 * x[:, 0]) + 0.5 * x[:, 1] * x[:, 2] - 0.3 * x[:, 3] ** 2).astype(np.float32)
    p = Proposer(11)
    xd, td = _dev(x), _dev(t)
    _, l0 = p.gradients(xd, td)
    loss = p.train(xd, td, batch=1024, epochs=8, lr=3e-3)
    _, l1 = p.gradients(xd, td)
    assert l1 < 0.05 * l0, (l0, loss, l1)
