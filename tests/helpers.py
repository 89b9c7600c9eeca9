"""Shared input generators and comparisons for the tests (seeded, small)."""
import numpy as np

from oracle import oracle as O


def random_runs(rng, P, t0, t1, max_n=12, G=1, frac=0.8, start_at_t0=0.2, log_mu=3.0, log_sigma=1.5):
    """Ragged runs for a random subset of pixels in random order: strictly increasing integer
    sample times in [t0, t1] (some starting exactly at t0), log-normal irradiance."""
    ids = np.nonzero(rng.random(P) < frac)[0].astype(np.uint32)
    rng.shuffle(ids)
    ts, irs, offs = [], [], [0]
    for _ in ids:
        n = int(rng.integers(0, max_n))
        tt = np.unique(rng.integers(t0, t1 + 1, size=n)).astype(np.int64)
        if len(tt) and rng.random() < start_at_t0:
            tt[0] = t0
            tt = np.unique(tt)
        ts.append(tt)
        irs.append(np.exp(rng.normal(log_mu, log_sigma, size=(len(tt), G))).astype(np.float32))
        offs.append(offs[-1] + len(tt))
    t = np.concatenate(ts) if ts else np.zeros(0, np.int64)
    irr = np.concatenate(irs).reshape(-1, G) if irs else np.zeros((0, G), np.float32)
    return ids, t, irr, np.array(offs, np.int64)


def events_from_dict(ev, width):
    """The product's events dict as the oracle's structured array (pixel, polarity, ...)."""
    out = np.zeros(len(ev["t_cross_ns"]), O.EVENT_DTYPE)
    out["pixel"] = ev["y"].astype(np.uint32) * width + ev["x"]
    out["polarity"] = ev["p"]
    out["label"] = ev["label"]
    out["flags"] = ev["flags"]
    out["t_cross"] = ev["t_cross_ns"]
    out["ds"] = ev["ds"]
    out["dn"] = ev["dn"]
    out["dl"] = ev["dl"]
    out["l_first"] = ev["l_first"]
    return out


def first_difference(a, b):
    """The first index where two event arrays of the same dtype differ in any field (value
    comparison per field, as the elements' own comparison), with the two elements; (n, None, None)
    when the common prefix agrees."""
    n = min(len(a), len(b))
    if n:
        ne = np.zeros(n, bool)
        for f in a.dtype.names:
            ne |= a[f][:n] != b[f][:n]
        idx = np.flatnonzero(ne)
        if len(idx):
            i = int(idx[0])
            return i, a[i], b[i]
    return n, None, None


def philox4x32_10(ctr, key):
    """Philox4x32-10 (Salmon et al., SC'11) in plain Python integers: an independent restatement for
    the tests (the library's and the oracle's are pinned by the Random123 known-answer vectors)."""
    M = 0xFFFFFFFF
    c = list(ctr)
    k0, k1 = key
    for _ in range(10):
        pa = 0xD2511F53 * c[0]
        pb = 0xCD9E8D57 * c[2]
        c = [((pb >> 32) ^ c[1] ^ k0) & M, pb & M, ((pa >> 32) ^ c[3] ^ k1) & M, pa & M]
        k0 = (k0 + 0x9E3779B9) & M
        k1 = (k1 + 0xBB67AE85) & M
    return c


def shot_schedule(seed, camera, pixel, t0, rate_hz, t_end):
    """Arrival times and polarities of the shot-noise process of one pixel up to t_end, in fp64 with
    the standard library's log (may differ from the library's by one ns where a ceil is near an
    integer)."""
    import math
    out = []
    t = t0
    k = 0
    while True:
        w = philox4x32_10([(0x101 << 8) | camera, pixel, k & 0xFFFFFFFF, k >> 32],
                          [seed & 0xFFFFFFFF, seed >> 32])
        u = (w[0] + 0.5) * 2.0 ** -32
        gap = max(1, math.ceil(-math.log(u) / rate_hz * 1e9))
        t += gap
        if t >= t_end:
            return out
        out.append((t, w[1] & 1))
        k += 1
