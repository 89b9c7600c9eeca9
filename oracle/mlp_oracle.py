"""CPU restatement of the learned proposer's network (test infrastructure: only tests/ and bench.py's
checks use it).

The network is this library's own (the reference has no learned controller), so this oracle is pinned
by mathematics, not by the reference: its backward pass is checked against central finite differences
of its own fp64 forward (tests/test_mlp_oracle.py), and its bf16 emulation against bit patterns of
known values. Layout and rounding points are those of include/bsx/bsx_mlp.h:
  z1 = bf16(x) bf16(W1) + b1, h1 = bf16(silu(z1)), z2 = h1 bf16(W2) + b2, h2 = bf16(silu(z2)),
  y = h2 bf16(w3) + b3, loss = mean (y - t)^2,
and the backward pass rounds dz2 and dz1 to bf16 where the kernels feed them to the MFMA units.
With emulate=False every rounding is dropped (the exact fp64 function)."""
import numpy as np

N_IN, N_HID = 12, 64
OFF_W1 = 0
OFF_B1 = OFF_W1 + N_IN * N_HID
OFF_W2 = OFF_B1 + N_HID
OFF_B2 = OFF_W2 + N_HID * N_HID
OFF_W3 = OFF_B2 + N_HID
OFF_B3 = OFF_W3 + N_HID
N_PARAMS = OFF_B3 + 1


def bf16(x):
    """Round float32 values to bfloat16 (nearest even), returned as float32."""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint32) << 16
    return u.view(np.float32)


def unpack(w, dtype=np.float32):
    w = np.asarray(w, dtype)
    return (w[OFF_W1:OFF_B1].reshape(N_IN, N_HID), w[OFF_B1:OFF_W2], w[OFF_W2:OFF_B2].reshape(N_HID, N_HID),
            w[OFF_B2:OFF_W3], w[OFF_W3:OFF_B3], w[OFF_B3])


def pack(W1, b1, W2, b2, w3, b3):
    return np.concatenate([np.ravel(W1), b1, np.ravel(W2), b2, w3, [b3]]).astype(np.float64)


def _silu(z):
    return z / (1.0 + np.exp(-z))


def _silu_grad(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(w, x, emulate=True):
    """Returns y (fp64) and the cache of the backward pass."""
    W1, b1, W2, b2, w3, b3 = unpack(w, np.float32 if emulate else np.float64)
    r = (lambda a: bf16(a).astype(np.float64)) if emulate else (lambda a: np.asarray(a, np.float64))
    xb, W1b, W2b, w3b = r(x), r(W1), r(W2), r(w3)
    z1 = xb @ W1b + b1.astype(np.float64)
    h1 = r(_silu(z1).astype(np.float32)) if emulate else _silu(z1)
    z2 = h1 @ W2b + b2.astype(np.float64)
    h2 = r(_silu(z2).astype(np.float32)) if emulate else _silu(z2)
    y = h2 @ w3b + float(b3)
    return y, (xb, W1b, W2b, w3b, z1, h1, z2, h2, r)


def gradients(w, x, t, emulate=True):
    """(gradient of mean (y - t)^2 as a flat fp64 vector, the loss)."""
    y, (xb, W1b, W2b, w3b, z1, h1, z2, h2, r) = forward(w, x, emulate)
    n = len(y)
    e = y - np.asarray(t, np.float64)
    dy = 2.0 * e / n
    dz2 = dy[:, None] * w3b[None, :] * _silu_grad(z2)
    if emulate:
        dz2 = r(dz2.astype(np.float32))
    dw3 = h2.T @ dy
    db3 = dy.sum()
    db2 = dz2.sum(0)
    dW2 = h1.T @ dz2
    dz1 = (dz2 @ W2b.T) * _silu_grad(z1)
    if emulate:
        dz1 = r(dz1.astype(np.float32))
    db1 = dz1.sum(0)
    dW1 = xb.T @ dz1
    return pack(dW1, db1, dW2, db2, dw3, db3), float(np.mean(e * e))


def adam_step(w, g, m, v, step, lr, beta1=0.9, beta2=0.999, eps=1e-8):
    """One Adam update (fp64): returns (w, m, v)."""
    m = beta1 * m + (1 - beta1) * g
    v = beta2 * v + (1 - beta2) * g * g
    mh = m / (1 - beta1 ** step)
    vh = v / (1 - beta2 ** step)
    return w - lr * mh / (np.sqrt(vh) + eps), m, v
