"""Benchmark of the sensor path on one MI355X: pixel-samples/s and events/s.

Workload (the GPU configuration of the metric): a 1280x720 sensor watching 64 textured rectangles
with affine motion (velocity U(+-2000 px/s), rotation U(+-3 rad/s), contrast U[0.1, 0.9]) over a
log-uniform 10-1000 lux background, for 500 ms, sampled by the plain error-tolerance controller
(refine_tol_log 0.01) on the GPU. Thresholds 0.15 +- 0.03, refractory 50 us, leak 0.05 Hz/px, shot
noise 0.1 Hz/px (Philox-keyed Poisson arrivals per pixel).
A step is one advance of every pixel over one 2^23 ns window; all windows' runs are generated on the
device and resident in HBM before the timed region. value = pixel-samples advanced per second.

The JSON line also carries the roofline of the advance kernel (the fixed algorithmic bytes per
launch over its HIP-event time, against 8 TB/s) and the reference's CPU path timed on this host's
cores on a bounded sample of the same runs.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

WINDOW = 1 << 23
HBM_PEAK_GBS = 8000.0
S_STATE, S_MAP, S_EVENT = 48, 16, 16  # the fixed accounting's per-pixel state, map and per-event bytes
# The kernels of one advance call (launched back to back on the call's stream and timed together
# with HIP events around them: the roofline's "launch").
ADVANCE_CHAIN = {
    "instantaneous": "advance chain: chain_init_kernel + chain_prep_kernel + order_count/offsets/scatter + instant_walk_kernel",
    "cascade": "advance chain: chain_init_kernel + chain_cost_kernel + order_count/offsets/scatter + cascade_walk_kernel",
}
# Measured HBM traffic of the instantaneous chain on this workload, bytes per sample: rocprofv3
# FETCH_SIZE x 2 + WRITE_SIZE summed over the chain's kernels, averaged over the timed launches (the
# factor 2 and 1 from a 1 GiB device copy read by the same counters: profiles/advance_chain_1280x720.md).
TRAFFIC_BYTES_PER_SAMPLE = 68.8


def sensor_params(bsx, photoreceptor):
    p = bsx.preset("prophesee_gen3")
    p.photoreceptor = photoreceptor
    p.noise = False
    p.theta_on = p.theta_off = 0.15
    p.sigma_theta_on = p.sigma_theta_off = 0.03
    p.refractory_s = 50e-6
    p.leak_rate_hz = 0.05
    # shot noise 0.1 Hz/px: the library's rate-specified Philox process (the instantaneous response
    # only; the cascade runs noise-free)
    p.shot_rate_hz = 0.1 if photoreceptor == "instantaneous" else 0.0
    from ballasim_amd.scene import LUX_TO_FA
    p.illuminance_scale = LUX_TO_FA
    return p


def algorithmic_bytes(samples, n_listed, n_pixels, events, groups=1):
    return samples * (8 + 4 * groups) + n_listed * 12 + n_pixels * 2 * S_STATE + n_pixels * S_MAP + events * S_EVENT


def cpu_baseline(p, window_runs, frame, width, height, budget_s):
    """The reference's own CPU path (compiled from its sources: oracle/_ref) when present, else the
    oracle's C restatement, on the first windows of the same runs until budget_s of CPU work
    (window_runs yields each window's runs copied to the host)."""
    import numpy as np
    from oracle import oracle as O
    threads = min(16, os.cpu_count() or 1)
    kind = "reference" if O.ref_available() else "port"
    m = (O.RefModel(p.c, width, height, 1, 1, 0, threads, True) if kind == "reference"
         else O.OracleModel(p.c, width, height, 1, 1, 0, threads))
    m.init(0, frame.reshape(-1, 1))
    samples, events, elapsed, windows = 0, 0, 0.0, 0
    for i, (ids, t, irr, offs) in enumerate(window_runs):
        t0 = time.perf_counter()
        ev, _ = m.advance_runs(ids, t, irr, offs, (i + 1) * WINDOW)
        elapsed += time.perf_counter() - t0
        samples += len(t)
        events += len(ev)
        windows += 1
        if elapsed >= budget_s:
            break
    return {"value": samples / elapsed, "unit": "pixel-samples/s", "cores": threads, "kind": kind,
            "sample": f"the first {windows} windows ({windows * WINDOW / 1e6:.1f} ms) of the same runs, "
                      f"{samples} samples, {events} events, {elapsed:.2f} s",
            "events_per_s": events / elapsed}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=58)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--photoreceptor", default="instantaneous", choices=["instantaneous", "cascade"])
    ap.add_argument("--cpu-budget", type=float, default=15.0, help="seconds of CPU-baseline work (0: skip)")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run only a 1 GiB device-to-device copy (the reference kernel for rocprofv3's "
                         "FETCH_SIZE / WRITE_SIZE scale) and exit")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        n = 1 << 28  # 2^28 fp32 = 1 GiB read + 1 GiB written per copy
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": 4 * n, "copy_bytes_written": 4 * n, "copies": 3}), flush=True)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "bench.py runs on one GPU: the sensor path has no multi-GPU form here "
                                   "(pixels shard by tile or camera; see DESIGN.md)", "n_gpus": args.gpus}))
        sys.exit(2)
    import numpy as np

    import ballasim_amd as bsx
    from ballasim_amd import scene as S

    scene, view = S.config2_scene(args.seed)
    W, H = view.width, view.height
    P = W * H
    p = sensor_params(bsx, args.photoreceptor)
    from ballasim_amd.sensor import SensorModel
    sensor = SensorModel(p, W, H, 1, ["constant"], 0, device="cuda")
    frame = scene.render(view, 0)
    sensor.init(0, frame.reshape(-1, 1))
    dark = float(p.dark_current_fA) / float(p.illuminance_scale)
    rules = S.rules(dark, tol=0.01)
    n_win = args.warmup + args.steps
    t_gen = time.perf_counter()
    runs, totals = [], []
    for i in range(n_win):
        r, total = scene.sample_window(view, rules, i * WINDOW, slot=i)
        runs.append(r)
        totals.append(total)
    gen_s = time.perf_counter() - t_gen
    print(f"# generated {n_win} windows, {sum(totals)} samples in {gen_s:.1f} s", file=sys.stderr, flush=True)
    for i in range(args.warmup):
        sensor.advance_device_runs(runs[i], (i + 1) * WINDOW)
    sensor.reset_stats()
    ev_counts = []
    t0 = time.perf_counter()
    for i in range(args.warmup, n_win):
        _, count = sensor.advance_device_runs(runs[i], (i + 1) * WINDOW)
        ev_counts.append(count)
    t1 = time.perf_counter()  # advance_device returns after its launch and counter read-back (synchronized)
    st = sensor.stats()
    elapsed = t1 - t0
    samples = sum(totals[args.warmup:])
    events = sum(ev_counts)
    kernel_s = st["kernel_ms"] * 1e-3
    bytes_total = sum(algorithmic_bytes(totals[i], P, P, ev_counts[i - args.warmup])
                      for i in range(args.warmup, n_win))
    achieved = bytes_total / kernel_s / 1e9
    line = {
        "metric": "pixel-samples/s + events/s on 1 MI355X; irradiance samples/event at fixed error",
        "value": samples / elapsed,
        "unit": "pixel-samples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (procedural scene, seed %d)" % args.seed,
        "config": {"workload": "1280x720 sensor, 64 textured affine rectangles, %d windows of 2^23 ns "
                               "(%.1f ms), plain controller tol 0.01, %s response, leak 0.05 Hz/px, %s"
                               % (args.steps, args.steps * WINDOW / 1e6, args.photoreceptor,
                                  "shot noise 0.1 Hz/px" if p.shot_rate_hz > 0 else "noise off"),
                   "pixels": P, "samples": samples, "events": events},
        "events_per_s": events / elapsed,
        "samples_per_event": samples / max(events, 1),
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": (TRAFFIC_BYTES_PER_SAMPLE * samples / args.steps
                                 if args.photoreceptor == "instantaneous" and TRAFFIC_BYTES_PER_SAMPLE else None),
                     "kernel": ADVANCE_CHAIN[args.photoreceptor],
                     "kernel_ms_per_launch": st["kernel_ms"] / args.steps,
                     "bytes_per_launch": bytes_total / args.steps},
    }
    if args.cpu_budget > 0:
        window_runs = (S.runs_to_host(runs[i]) for i in range(n_win))
        line["cpu_baseline"] = cpu_baseline(p, window_runs, frame, W, H, args.cpu_budget)
    else:
        line["cpu_baseline"] = None
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
