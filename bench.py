"""Benchmark of the sensor path on one MI355X: pixel-samples/s and events/s.

Workload (the GPU configuration of the metric): a 1280x720 sensor watching 64 textured rectangles
with affine motion (velocity U(+-2000 px/s), rotation U(+-3 rad/s), contrast U[0.1, 0.9]) over a
log-uniform 10-1000 lux background, for 500 ms, sampled by the plain error-tolerance controller
(refine_tol_log 0.01) on the GPU. Thresholds 0.15 +- 0.03, refractory 50 us, leak 0.05 Hz/px, shot
noise 0.1 Hz/px (Philox-keyed Poisson arrivals per pixel).
A step is three advances of every pixel over consecutive 2^23 ns windows (--windows-per-step), so the
default 20 steps cover the whole 500-ms run from t = 0 (60 windows, 503 ms); warm-up and the
autotuner run on a separate sensor first. All windows' runs are generated on the device and resident
in HBM before the timed region. value = pixel-samples advanced per second. The same line carries the
cascade photoreceptor's figures on the first windows of the same runs, without noise and with the
reference's in-path noise ("cascade", "cascade_noise").

The JSON line also carries the roofline of the advance kernel (the fixed algorithmic bytes per
launch over its HIP-event time, against 8 TB/s; the measured HBM traffic comes from separate
rocprofv3 counter runs, so "traffic" is null here) and the reference's CPU path timed on every core
this process may use, on a bounded sample of the same runs.

--workload rig: the ragged batch. A 4-camera rig (1280x720, 640x480, 346x260, 1920x1080, each with
its own offset in the scene plane) watching the 346x260 scene of one textured disk (100 lux
background, radius 40 px, 8 octaves, contrast 0.5, 800 px/s); per-pixel sample counts log-normal
(mu ln 20, sigma 1.5, clipped to [1, 4000]), evenly spread over each window; thresholds 0.2 +- 0.02,
refractory 100 us, noise off. A step advances all four cameras over one window by one graph-captured
launch chain, after the autotuner has picked the walk's launch shape on the warm-up window; the
same windows advanced camera by camera (four calls per window, the reference engine's loop) are
timed beside it.

--workload stress: a 2048x1536 sensor for 1 s (120 windows) watching 256 textured rectangles
(velocity U(+-5000 px/s), rotation U(+-3 rad/s), value noise of 7 octaves from a 32-64 px cell down
to 0.25-0.5 px), the instantaneous response with shot noise 0.5 Hz/px and leak 0.1 Hz/px, sampled by
the plain controller window by window on the device (the runs of a whole second do not fit at once),
and every released event streamed to the host in output order through the double-buffered 256 MiB
pinned chunks of the event stream. A step is one window: sampling, the advance, the output stage and
the stream. The sink on the host checks that the chunks arrive in order and their events in output
order.
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
    "instantaneous": "advance chain: chain_init_kernel + chain_screen_kernel + order_count/offsets/scatter + instant_walk_kernel",
    "cascade": "advance chain: chain_init_kernel + chain_cost_kernel + order_count/offsets/scatter + cascade_walk_kernel",
}


def sensor_params(bsx, photoreceptor, noise=False):
    from ballasim_amd.configs import config2_params
    return config2_params(photoreceptor, noise)


def algorithmic_bytes(samples, n_listed, n_pixels, events, groups=1):
    return samples * (8 + 4 * groups) + n_listed * 12 + n_pixels * 2 * S_STATE + n_pixels * S_MAP + events * S_EVENT


def cpu_baseline(p, window_runs, frame, width, height, budget_s):
    """The reference's own CPU path (compiled from its sources: oracle/_ref) when present, else the
    oracle's C restatement, on the first windows of the same runs until budget_s of CPU work
    (window_runs yields each window's runs copied to the host)."""
    import numpy as np
    from oracle import oracle as O
    threads = len(os.sched_getaffinity(0))
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
            "host_cpus": os.cpu_count(),
            "sample": f"the first {windows} windows ({windows * WINDOW / 1e6:.1f} ms) of the same runs, "
                      f"{samples} samples, {events} events, {elapsed:.2f} s",
            "events_per_s": events / elapsed}


def rig_params(bsx, photoreceptor):
    from ballasim_amd.configs import config3_params
    return config3_params(photoreceptor)


def bench_rig(args):
    import numpy as np

    import ballasim_amd as bsx
    from ballasim_amd import scene as S
    from ballasim_amd.sensor import Rig, SensorModel

    scene, _ = S.config1_scene(args.seed)
    views = S.rig_views()
    p = rig_params(bsx, args.photoreceptor)
    rng = np.random.default_rng(args.seed + 1)
    counts = [S.lognormal_counts(rng, v.width * v.height) for v in views]
    allc = np.sort(np.concatenate(counts))[::-1]
    top1 = float(allc[: max(1, len(allc) // 100)].sum() / allc.sum())
    n_win = args.warmup + args.steps
    runs, totals = [], []
    for i in range(n_win):
        rw, tw = [], 0
        for c, v in enumerate(views):
            r, tot = scene.sample_counts(v, counts[c], i * WINDOW, slot=i * len(views) + c)
            rw.append(r)
            tw += tot
        runs.append(rw)
        totals.append(tw)
    print(f"# generated {n_win} windows of the rig, {sum(totals)} samples", file=sys.stderr, flush=True)

    def cameras():
        cams = []
        for c, v in enumerate(views):
            m = SensorModel(p, v.width, v.height, 1 + c, ["constant"], c, device="cuda")
            m.init(0, scene.render(v, 0).reshape(-1, 1))
            cams.append(m)
        return cams

    cams = cameras()
    rig = Rig(cams)
    for i in range(args.warmup):
        rig.advance_device(runs[i], (i + 1) * WINDOW)
    trials = rig.autotune(runs[args.warmup], (args.warmup + 1) * WINDOW, reps=2)
    tuning = rig.tuning()
    for m in cams:
        m.reset_stats()
    k0 = rig.stats()["kernel_ms"]
    ev_counts = []
    t0 = time.perf_counter()
    for i in range(args.warmup, n_win):
        out = rig.advance_device(runs[i], (i + 1) * WINDOW)
        ev_counts.append(sum(n for _, n in out))
    elapsed = time.perf_counter() - t0
    kernel_ms = rig.stats()["kernel_ms"] - k0
    # The same windows camera by camera (a fresh set of cameras with the same tuning).
    solo = cameras()
    for m in solo:
        m.set_launch_policy(**{k: (v if v != 0 or k not in ("queue_chunk", "head_per_mille") else -1)
                               for k, v in tuning.items()})
    for i in range(args.warmup):
        for c, m in enumerate(solo):
            m.advance_device_runs(runs[i][c], (i + 1) * WINDOW)
    t0 = time.perf_counter()
    solo_events = 0
    for i in range(args.warmup, n_win):
        for c, m in enumerate(solo):
            solo_events += m.advance_device_runs(runs[i][c], (i + 1) * WINDOW)[1]
    solo_elapsed = time.perf_counter() - t0
    samples = sum(totals[args.warmup:])
    events = sum(ev_counts)
    assert events == solo_events, (events, solo_events)
    P = sum(v.width * v.height for v in views)
    bytes_total = sum(algorithmic_bytes(totals[i], P, P, ev_counts[i - args.warmup]) for i in range(args.warmup, n_win))
    achieved = bytes_total / (kernel_ms * 1e-3) / 1e9
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
        "data": "synthetic (procedural scene and log-normal sample counts, seed %d)" % args.seed,
        "config": {"workload": "4-camera rig 1280x720 + 640x480 + 346x260 + 1920x1080 (%d px) with scene-plane "
                               "offsets, one textured disk at 800 px/s over 100 lux, per-pixel sample counts "
                               "log-normal(ln 20, 1.5) clipped to [1, 4000], %d windows of 2^23 ns, %s response, "
                               "noise off, one graph-captured launch chain per window"
                               % (P, args.steps, args.photoreceptor),
                   "pixels": P, "samples": samples, "events": events,
                   "top_1pct_pixels_sample_share": top1},
        "events_per_s": events / elapsed,
        "per_camera_calls_value": samples / solo_elapsed,
        "rig_speedup_vs_per_camera_calls": solo_elapsed / elapsed,
        "tuning": tuning,
        "autotune_trials": len(trials),
        "rig_stats": rig.stats(),
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": ADVANCE_CHAIN[args.photoreceptor] + " (4 cameras, one graph)",
                     "kernel_ms_per_launch": kernel_ms / args.steps,
                     "bytes_per_launch": bytes_total / args.steps},
    }
    if args.cpu_budget > 0:
        line["cpu_baseline"] = cpu_baseline_rig(p, runs, views, scene, args.cpu_budget)
    else:
        line["cpu_baseline"] = None
    print(json.dumps(line), flush=True)


def cpu_baseline_rig(p, runs, views, scene, budget_s):
    """The CPU path on the rig's cameras in turn (window by window) until budget_s of CPU work."""
    from ballasim_amd import scene as S
    from oracle import oracle as O
    threads = len(os.sched_getaffinity(0))
    kind = "reference" if O.ref_available() else "port"
    models = []
    for c, v in enumerate(views):
        m = (O.RefModel(p.c, v.width, v.height, 1 + c, 1, c, threads, True) if kind == "reference"
             else O.OracleModel(p.c, v.width, v.height, 1 + c, 1, c, threads))
        m.init(0, scene.render(v, 0).reshape(-1, 1))
        models.append(m)
    samples, events, elapsed, done = 0, 0, 0.0, []
    for i in range(len(runs)):
        for c, m in enumerate(models):
            ids, t, irr, offs = S.runs_to_host(runs[i][c])
            t0 = time.perf_counter()
            ev, _ = m.advance_runs(ids, t, irr, offs, (i + 1) * WINDOW)
            elapsed += time.perf_counter() - t0
            samples += len(t)
            events += len(ev)
            done.append((i, c))
            if elapsed >= budget_s:
                break
        if elapsed >= budget_s:
            break
    return {"value": samples / elapsed, "unit": "pixel-samples/s", "cores": threads, "kind": kind,
            "sample": f"{len(done)} camera-windows of the same runs (window, camera) {done[0]}..{done[-1]}, "
                      f"{samples} samples, {events} events, {elapsed:.2f} s",
            "events_per_s": events / elapsed}


MLP_FWD_FLOP = 2 * (12 * 64 + 64 * 64 + 64)  # 9,856 FLOP per forward row (multiply-adds of the three layers)
MLP_TRAIN_FLOP = 3 * MLP_FWD_FLOP  # forward + backward (input and weight gradients) per training row
BF16_PEAK_TFLOPS = 2500.0


def bench_learned(args):
    """Trains the proposer on dense-trace examples of seeded config-1-style scenes, then compares the
    learned controller with the plain one on a held-out 1280x720 scene against a dense-sampled reference."""
    import torch

    import ballasim_amd as bsx
    from ballasim_amd import scene as S
    from ballasim_amd.learned import compare_controllers, train_proposer

    p = rig_params(bsx, "instantaneous")
    dark = float(p.dark_current_fA) / float(p.illuminance_scale)
    band, tol, margin, near = 0.2, 0.01, 0.5, 0.05
    prop, training, X, T = train_proposer(dark, n_examples=args.train_examples, theta_band=band, tol_log=tol,
                                          batch=args.train_batch)
    rows = training["rows_trained"]
    train_ms = training["train_ms"]
    training["rows_per_s"] = rows / (train_ms * 1e-3)
    train_tflops = rows * MLP_TRAIN_FLOP / (train_ms * 1e-3) / 1e12
    # inference throughput of the batched forward kernel
    Y = torch.empty(X.shape[0], dtype=torch.float32, device="cuda")
    prop.predict(X, Y)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(3):
        prop.predict(X, Y)
    e1.record()
    torch.cuda.synchronize()
    infer_ms = e0.elapsed_time(e1) / 3
    n_rows = X.shape[0]
    infer_tflops = n_rows * MLP_FWD_FLOP / (infer_ms * 1e-3) / 1e12
    del X, T, Y
    torch.cuda.empty_cache()

    sc, v = S.learned_eval_scene(args.seed + 7)
    n_win = max(1, int(round(args.eval_ms * 1e6 / WINDOW)))
    res = compare_controllers(sc, v, p, prop, n_win, theta_band=band, tol_log=tol, margin=margin, near=near)
    line = {
        "metric": "pixel-samples/s + events/s on 1 MI355X; irradiance samples/event at fixed error",
        "value": res["learned"]["samples_per_event"],
        "unit": "irradiance samples/event",
        "n_gpus": 1, "steps": n_win, "warmup": 0, "ms_per_step": None,
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "bf16 (MFMA, fp32 accumulation)",
        "data": "synthetic (config-1-style procedural scenes: training seeds 1000-1019, evaluation seed %d)"
                % (args.seed + 7),
        "config": {"workload": "learned proposer 12-64-64-1 SiLU trained on %d dense-trace examples (dt 10 us) of "
                               "20 seeded 346x260 disk scenes; evaluated at 1280x720 (12 textured disks, 800 px/s, "
                               "100 lux) over %d windows (%.0f ms) against a 10-us dense reference; thresholds "
                               "0.2 +- 0.02, refractory 100 us, instantaneous response, noise off"
                               % (training["examples"], n_win, n_win * WINDOW / 1e6),
                   "band": band, "tol_log": tol, "margin": margin, "near": near,
                   "sensor_linked": True},
        "plain": res["plain"], "learned": res["learned"], "plain_density_curve": res["curve"],
        "dense_reference": res["dense"],
        "samples_per_event_gain_vs_plain": res["gain_vs_plain"],
        "plain_samples_per_event_at_learned_median_error": res["plain_samples_per_event_at_learned_median_error"],
        "gain_at_equal_median_error": res["gain_at_equal_median_error"],
        "training": training,
        "roofline": {"bound": "mfma", "achieved": train_tflops, "peak": BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": train_tflops / BF16_PEAK_TFLOPS, "traffic": None,
                     "kernel": "gradients_kernel + adam_kernel (training step)",
                     "inference_tflops": infer_tflops, "inference_rows_per_s": X_rows_per_s(infer_ms, n_rows)},
        "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)

def X_rows_per_s(ms, rows):
    return rows / (ms * 1e-3)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None, help="timed steps (default 20; rig: 8 windows)")
    ap.add_argument("--windows-per-step", type=int, default=3,
                    help="plain: 2^23-ns windows per step (default 3: 20 steps cover the 500-ms run)")
    ap.add_argument("--cascade-windows", type=int, default=3,
                    help="plain: windows timed with the cascade photoreceptor, without and with its noise (0: skip)")
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--photoreceptor", default="instantaneous", choices=["instantaneous", "cascade"])
    ap.add_argument("--noise", action="store_true",
                    help="plain, cascade: the reference's in-path noise on (the instantaneous response carries the "
                         "shot process instead)")
    ap.add_argument("--workload", default="plain", choices=["plain", "rig", "learned", "stress"],
                    help="plain: the 1280x720 sensor (default); rig: the 4-camera ragged batch; learned: the "
                         "proposer's training and the learned controller against the plain one; stress: the "
                         "2048x1536 sensor for 1 s with the event stream")
    ap.add_argument("--chunk-mib", type=int, default=256, help="stress: event stream chunk size (MiB)")
    ap.add_argument("--serial-sampling", dest="overlap_sampling", action="store_false",
                    help="stress: sample on the advance's stream (no overlap of sampling with the advance)")
    ap.add_argument("--no-autotune", dest="autotune", action="store_false",
                    help="keep the default launch shape (no autotuner trials before the timed windows)")
    ap.add_argument("--train-examples", type=int, default=20_000_000, help="learned: training examples")
    ap.add_argument("--train-batch", type=int, default=65536,
                    help="learned: rows per optimizer step (a multiple of 64; 64 rows per group of four waves; the "
                         "learning rates scale with its square root)")
    ap.add_argument("--eval-ms", type=float, default=500.0, help="learned: evaluation span (ms)")
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
                                   "(pixels shard by tile or camera)", "n_gpus": args.gpus}))
        sys.exit(2)
    if args.steps is None:
        args.steps = {"rig": 8, "stress": 120}.get(args.workload, 20)
    if args.workload == "rig":
        return bench_rig(args)
    if args.workload == "learned":
        return bench_learned(args)
    if args.workload == "stress":
        return bench_stress(args)
    return bench_plain(args)


def time_windows(sensor, runs, totals, first, count, tuning=None):
    """Advances a fresh-initialised sensor over windows [first, first + count) (a rig of one camera: one
    graph-captured launch chain per window); (elapsed s, samples, events per window, stats with kernel_ms)."""
    from ballasim_amd.sensor import Rig
    rig = Rig([sensor])
    if tuning is not None:
        rig.set_tuning(**tuning)
    sensor.reset_stats()
    k0 = rig.stats()["kernel_ms"]
    ev_counts = []
    t0 = time.perf_counter()
    for i in range(first, first + count):
        (_, n_ev), = rig.advance_device([runs[i]], (i + 1) * WINDOW)
        ev_counts.append(n_ev)
    elapsed = time.perf_counter() - t0  # a rig call returns after its counters are read back (synchronized)
    st = dict(sensor.stats())
    st["kernel_ms"] = rig.stats()["kernel_ms"] - k0
    rig.close()
    return elapsed, sum(totals[first:first + count]), ev_counts, st


def bench_stress(args):
    import threading

    import numpy as np

    import ballasim_amd as bsx
    from ballasim_amd import configs
    from ballasim_amd import scene as S
    from ballasim_amd.output import EventStream, OutputStage
    from ballasim_amd.sensor import Rig, SensorModel

    W, H = 2048, 1536
    P = W * H
    shapes, bg = S.config2_shapes(args.seed, n=256, width=W, height=H, v_max=5000.0, cell_range=(32.0, 64.0),
                                  octaves=7)
    scene, view = S.Scene(shapes, bg), S.view(W, H)
    p = configs.config2_params("instantaneous")
    p.shot_rate_hz = 0.5
    p.leak_rate_hz = 0.1
    rules = S.rules(float(p.dark_current_fA) / float(p.illuminance_scale), tol=0.01)
    frame = scene.render(view, 0).reshape(-1, 1)
    lock = threading.Lock()
    seen = {"chunks": 0, "events": 0, "last_t": None, "ordered": True}

    def sink(rec, index):  # on the stream's delivery thread
        t = rec["t_reported_ns"]
        with lock:
            ok = index == seen["chunks"] and bool(np.all(t[1:] >= t[:-1]))
            if len(t) and seen["last_t"] is not None:
                ok = ok and int(t[0]) >= seen["last_t"]
            seen["ordered"] &= ok
            seen["chunks"] += 1
            seen["events"] += len(t)
            if len(t):
                seen["last_t"] = int(t[-1])

    def run(n_windows, stream_chunks):
        import queue

        import torch
        m = SensorModel(p, W, H, 1, ["constant"], 0, device="cuda")
        m.init(0, frame)
        rig = Rig([m])
        stage = OutputStage(m)
        es = EventStream(sink, chunk_bytes=args.chunk_mib << 20) if stream_chunks else None
        samples = events = released = 0
        busy = [0.0]
        k0 = rig.stats()["kernel_ms"]
        # The sampler runs one window ahead on a stream of its own, in a thread of its own (the library's
        # calls release the interpreter), so that a window's sampling overlaps the previous window's advance
        # and output stage; three slots hold the window being advanced, the next one and the one after.
        sampler_stream = torch.cuda.Stream()
        ready = queue.Queue(maxsize=1)

        def produce():
            try:
                for w in range(n_windows):
                    ts = time.perf_counter()
                    item = scene.sample_window(view, rules, w * WINDOW, slot=w % 3,
                                               stream=sampler_stream if args.overlap_sampling else 0)
                    busy[0] += time.perf_counter() - ts
                    ready.put(item)
            except BaseException as e:  # handed to the consumer, which raises it
                ready.put(e)

        producer = threading.Thread(target=produce, daemon=True)
        t0 = time.perf_counter()
        producer.start()
        for w in range(n_windows):
            item = ready.get()
            if isinstance(item, BaseException):
                raise item
            runs, n_s = item
            (ptr, n_ev), = rig.advance_device([runs], (w + 1) * WINDOW)
            samples += n_s
            events += n_ev
            if es is not None:
                released += es.window(stage, (ptr, n_ev), w * WINDOW)
            if (w + 1) % 10 == 0:
                print(f"# stress window {w + 1}/{n_windows}: {samples} samples, {events} events", file=sys.stderr,
                      flush=True)
        producer.join()
        if es is not None:
            released += es.flush(stage)
            es.finish()
        elapsed = time.perf_counter() - t0
        t_sample = busy[0]
        kernel_ms = rig.stats()["kernel_ms"] - k0
        st = es.stats() if es is not None else None
        if es is not None:
            es.close()
        rig.close()
        return elapsed, t_sample, samples, events, released, kernel_ms, st

    run(max(1, args.warmup), False)  # warm-up: kernels loaded, the sampler's buffers sized
    seen.update(chunks=0, events=0, last_t=None, ordered=True)
    n = args.steps
    elapsed, t_sample, samples, events, released, kernel_ms, st = run(n, True)
    assert released == events == seen["events"], (released, events, seen["events"])
    bytes_total = algorithmic_bytes(samples, P * n, P * n, events)  # every pixel listed in every window
    achieved = bytes_total / (kernel_ms * 1e-3) / 1e9
    line = {
        "metric": "pixel-samples/s + events/s on 1 MI355X; irradiance samples/event at fixed error",
        "value": samples / elapsed,
        "unit": "pixel-samples/s",
        "n_gpus": 1, "steps": n, "warmup": args.warmup, "ms_per_step": elapsed / n * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
        "data": "synthetic (procedural scene, seed %d)" % args.seed,
        "config": {"workload": "stress: 2048x1536 sensor, 256 textured affine rectangles (+-5000 px/s, 7 octaves "
                               "down to 0.25-0.5 px), %d windows of 2^23 ns (%.1f ms), plain controller tol 0.01, "
                               "instantaneous response, shot 0.5 Hz/px, leak 0.1 Hz/px, events streamed in %d MiB "
                               "double-buffered pinned chunks; a step = sampling + advance + output + stream of one "
                               "window%s" % (n, n * WINDOW / 1e6, args.chunk_mib,
                                             ", the next window sampled on its own stream meanwhile"
                                             if args.overlap_sampling else ""),
                   "pixels": P, "samples": samples, "events": events, "windows": n},
        "events_per_s": events / elapsed,
        "samples_per_event": samples / max(events, 1),
        "sampling_s": t_sample,
        "stream": {"chunks": st["chunks"], "bytes": st["bytes"], "copy_ms": st["copy_ms"],
                   "copy_GBps": st["bytes"] / max(st["copy_ms"], 1e-9) / 1e6, "producer_wait_ms": st["wait_ms"],
                   "in_order": seen["ordered"]},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None, "kernel": ADVANCE_CHAIN["instantaneous"],
                     "kernel_ms_per_launch": kernel_ms / n, "bytes_per_launch": bytes_total / n},
        "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)


def bench_plain(args):
    import numpy as np

    import ballasim_amd as bsx
    from ballasim_amd import scene as S
    from ballasim_amd.sensor import Rig, SensorModel

    scene, view = S.config2_scene(args.seed)
    W, H = view.width, view.height
    P = W * H
    p = sensor_params(bsx, args.photoreceptor, args.noise)
    frame = scene.render(view, 0)
    dark = float(p.dark_current_fA) / float(p.illuminance_scale)
    rules = S.rules(dark, tol=0.01)
    wps = args.windows_per_step
    n_timed = args.steps * wps
    n_win = max(n_timed, args.warmup * wps, 1)
    t_gen = time.perf_counter()
    runs, totals = [], []
    for i in range(n_win):
        r, total = scene.sample_window(view, rules, i * WINDOW, slot=i)
        runs.append(r)
        totals.append(total)
    gen_s = time.perf_counter() - t_gen
    print(f"# generated {n_win} windows, {sum(totals)} samples in {gen_s:.1f} s", file=sys.stderr, flush=True)

    def fresh(params):
        m = SensorModel(params, W, H, 1, ["constant"], 0, device="cuda")
        m.init(0, frame.reshape(-1, 1))
        return m

    # Warm-up and the autotuner on a separate sensor over the first windows; the timed sensor then runs
    # windows 0 .. n_timed - 1 from its own init, so the timed steps cover the run from its start.
    warm = fresh(p)
    rig = Rig([warm])
    for i in range(args.warmup * wps):
        rig.advance_device([runs[i]], (i + 1) * WINDOW)
    k = args.warmup * wps
    trials = rig.autotune([runs[k]], (k + 1) * WINDOW, reps=2) if args.autotune and k < n_win else []
    tuning = rig.tuning()
    rig.close()
    del warm
    sensor = fresh(p)
    elapsed, samples, ev_counts, st = time_windows(sensor, runs, totals, 0, n_timed, tuning)
    print(f"# timed {n_timed} windows in {elapsed:.3f} s", file=sys.stderr, flush=True)
    events = sum(ev_counts)
    kernel_s = st["kernel_ms"] * 1e-3
    bytes_total = sum(algorithmic_bytes(totals[i], P, P, ev_counts[i]) for i in range(n_timed))
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
        "config": {"workload": "1280x720 sensor, 64 textured affine rectangles, %d windows of 2^23 ns from t = 0 "
                               "(%.1f ms; %d windows per step), plain controller tol 0.01, %s response, leak "
                               "0.05 Hz/px, %s"
                               % (n_timed, n_timed * WINDOW / 1e6, wps, args.photoreceptor,
                                  "shot noise 0.1 Hz/px" if p.shot_rate_hz > 0 else
                                  "the in-path noise" if p.noise else "noise off"),
                   "pixels": P, "samples": samples, "events": events, "windows": n_timed,
                   "windows_per_step": wps},
        "events_per_s": events / elapsed,
        "samples_per_event": samples / max(events, 1),
        "tuning": tuning,
        "autotune_trials": len(trials),
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": (ADVANCE_CHAIN[args.photoreceptor].replace("cascade_walk_kernel", "noise_walk_kernel")
                                if p.noise else ADVANCE_CHAIN[args.photoreceptor]),
                     "kernel_ms_per_launch": st["kernel_ms"] / n_timed,
                     "bytes_per_launch": bytes_total / n_timed},
    }
    # The cascade photoreceptor on the first windows of the same runs: without noise, and with the
    # reference's in-path noise (its default sensor model) in place of the shot process.
    if args.cascade_windows > 0:
        extra = {}
        for name, noise in (("cascade", False), ("cascade_noise", True)):
            m = fresh(sensor_params(bsx, "cascade", noise))
            nw = min(args.cascade_windows, n_win)
            el, sm, evc, stc = time_windows(m, runs, totals, 0, nw, tuning)
            extra[name] = {"value": sm / el, "unit": "pixel-samples/s", "events_per_s": sum(evc) / el,
                           "windows": nw, "ms_per_window": el / nw * 1e3,
                           "kernel_ms_per_launch": stc["kernel_ms"] / nw, "samples": sm, "events": sum(evc),
                           "noise": "in-path LNA noise (prophesee_gen3 model)" if noise else "off",
                           "cells": stc["cells"], "refinements": stc["refinements"],
                           "substeps": stc["substeps"], "substep_attempts": stc["substep_attempts"],
                           "settled_substeps": stc["settled_substeps"], "pieces": stc["pieces"],
                           "leaf_intervals": stc["leaf_intervals"]}
            print(f"# {name}: {nw} windows in {el:.3f} s", file=sys.stderr, flush=True)
            del m
        line["cascade"] = extra["cascade"]
        line["cascade_noise"] = extra["cascade_noise"]
    if args.cpu_budget > 0:
        window_runs = (S.runs_to_host(runs[i]) for i in range(n_win))
        line["cpu_baseline"] = cpu_baseline(p, window_runs, frame, W, H, args.cpu_budget)
    else:
        line["cpu_baseline"] = None
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
