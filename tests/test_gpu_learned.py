"""The learned controller end to end on the GPU: training examples from dense traces of seeded
config-1-style scenes, training by the library's kernels, and the learned sampler on a held-out seed
against the plain error-tolerance controller and against the same sampler with the untrained network.
Event error is measured against a dense-sampled reference (10 us): count mismatch and median timing error
(ballasim_amd.learned.event_error). The learned sampler reads the sensor's reference state at each window
start (its band of no event start) and samples as the plain controller near the band edges.

What the test pins: the examples are well formed (finite features, targets within the step range);
training lowers the loss; every learned run is a valid sample stream the sensor accepts; the learned
controller needs fewer irradiance samples per event than the plain one, and the trained network fewer
than the untrained one, with the count mismatch bounded. The timing error is reported, not bounded
(printed with -s)."""
import json

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

W = 1 << 23


def _sensor_params(bsx):
    from ballasim_amd import scene as S
    p = bsx.preset("prophesee_gen3")
    p.photoreceptor = "instantaneous"
    p.noise = False
    p.theta_on = p.theta_off = 0.2
    p.sigma_theta_on = p.sigma_theta_off = 0.02
    p.refractory_s = 100e-6
    p.leak_rate_hz = 0.0
    p.illuminance_scale = S.LUX_TO_FA
    return p


def test_learned_controller_gain():
    import torch

    import ballasim_amd as bsx
    from ballasim_amd import scene as S
    from ballasim_amd.learned import LearnedSampler, Proposer, event_error, learned_examples, learned_rules
    p = _sensor_params(bsx)
    dark = float(p.dark_current_fA) / float(p.illuminance_scale)
    band, tol, margin, near = 0.2, 0.01, 0.5, 0.05
    n_per, seeds = 250_000, [101, 102]
    X = torch.empty((n_per * len(seeds), 12), dtype=torch.float32, device="cuda:0")
    T = torch.empty(n_per * len(seeds), dtype=torch.float32, device="cuda:0")
    for i, s in enumerate(seeds):
        sc, v = S.config1_scene(s)
        n_active = learned_examples(sc, v, X[i * n_per:(i + 1) * n_per], T[i * n_per:(i + 1) * n_per], dark,
                                    t0_ns=30_000_000, seed=s, theta_band=band, tol_log=tol)
        assert n_active > 1000
    Xn, Tn = X.cpu().numpy(), T.cpu().numpy()
    assert np.isfinite(Xn).all() and np.isfinite(Tn).all()
    lo, hi = (np.log(0.5 * 10_000) - 11.512925464970229) / 2, (np.log(838 * 10_000) - 11.512925464970229) / 2
    assert Tn.min() >= lo - 1e-4 and Tn.max() <= hi + 1e-4
    trained = Proposer(1)
    w = trained.params
    w[-1] = Tn.mean()
    trained.params = w
    untrained = Proposer(1)
    untrained.params = w
    _, l0 = trained.gradients(X, T)
    for lr in (3e-3, 1e-3):
        trained.train(X, T, batch=4096, epochs=2, lr=lr)
    _, l1 = trained.gradients(X, T)
    assert l1 < l0

    sc, v = S.config1_scene(7)
    frame = sc.render(v, 0)
    n_win = 12

    def run(kind, prop=None):
        m = bsx.SensorModel(p, v.width, v.height, 1, ["constant"], 0, device="cuda")
        m.init(0, frame.reshape(-1, 1))
        ls = None
        if prop is not None:
            ls = LearnedSampler(sc, v, learned_rules(S.rules(dark, tol=tol), tol_log=tol, theta_band=band,
                                                     margin=margin, near=near), prop)
            ls.init(0)
        evs, samples, fallbacks = [], 0, 0
        for i in range(n_win):
            if kind == "dense":
                r, tot = sc.sample_window(v, S.rules(dark, fixed_period_ns=10_000), i * W, slot=0)
            elif kind == "plain":
                r, tot = sc.sample_window(v, S.rules(dark, tol=0.01), i * W, slot=0)
            else:
                r, tot, _, fb = ls.window(i * W, slot=0, sensor=m)
                fallbacks += fb
                # every sample inside the window, strictly increasing per pixel (the device path does not
                # check sample times against the model time)
                _, ts, _, offs = S.runs_to_host(r)
                assert len(ts) == tot and (tot == 0 or (ts.min() > i * W and ts.max() <= (i + 1) * W))
                d = np.diff(ts)
                starts = np.zeros(len(ts), bool)
                starts[offs[:-1][offs[:-1] < len(ts)]] = True
                assert np.all(d[~starts[1:]] > 0)
            samples += tot
            ptr, cnt = m.advance_device_runs(r, (i + 1) * W)
            evs.append(m.download_events(ptr, cnt))
        ev = {k: np.concatenate([e[k] for e in evs]) for k in evs[0]}
        return ev, samples, fallbacks

    ref, _, _ = run("dense")
    assert len(ref["t_cross_ns"]) > 500
    out = {}
    for name, kind, prop in (("plain", "plain", None), ("learned", "learned", trained),
                             ("untrained", "learned", untrained)):
        ev, samples, fb = run(kind, prop)
        err = event_error(ev, ref, v.width)
        out[name] = dict(samples_per_event=samples / max(1, err["n_test"]), fallbacks=fb, **err)
    print(json.dumps(out, indent=1))
    assert out["learned"]["samples_per_event"] < out["plain"]["samples_per_event"]
    assert out["learned"]["samples_per_event"] < out["untrained"]["samples_per_event"] or \
        out["learned"]["fallbacks"] < out["untrained"]["fallbacks"]
    assert out["learned"]["count_mismatch"] < 0.01


def test_learned_gain_at_equal_error_config4():
    """Config 4 at its own workload: the proposer trained on 20 M dense-trace examples of 20 seeded
    scenes, then the learned controller against the plain one at 1280x720 over 503 ms (60 windows)
    against a 10-us dense reference. The plain controller's samples-against-error curve over its lattice
    density gives its samples per event at the learned controller's median timing error; the learned
    controller must need at least 2x fewer there, with a count mismatch of at most 1 %. The north star's
    1-us median timing error is not reached by either controller (printed, not asserted)."""
    from ballasim_amd import scene as S
    from ballasim_amd.configs import config3_params
    from ballasim_amd.learned import compare_controllers, train_proposer
    p = config3_params("instantaneous")
    dark = float(p.dark_current_fA) / float(p.illuminance_scale)
    prop, training, X, T = train_proposer(dark, n_examples=20_000_000)
    del X, T
    sc, v = S.learned_eval_scene(7)
    res = compare_controllers(sc, v, p, prop, 60)
    print(json.dumps({k: res[k] for k in ("plain", "learned", "gain_at_equal_median_error", "gain_vs_plain")},
                     indent=1))
    print(f"median timing error: learned {res['learned']['median_abs_dt_ns'] / 1e3:.1f} us, plain "
          f"{res['plain']['median_abs_dt_ns'] / 1e3:.1f} us (target 1 us: not met)")
    assert res["gain_at_equal_median_error"] is not None and res["gain_at_equal_median_error"] >= 2.0
    assert res["learned"]["count_mismatch"] <= 0.01 and res["plain"]["count_mismatch"] <= 0.01
    assert res["gain_vs_plain"] >= 2.0
