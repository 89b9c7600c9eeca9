"""The plain controller against the reference's own window engine: WindowEngine<kHost> compiled from
src/engine/window.cpp and par_host.cpp (oracle/_ref/libwindow_ref.so), rendering through the engine hook
of ballasim_amd/seam/scene_hook_bsx.h on the same bsx scene, is given each active pixel, its lattice level
(the motion-bound rule l = clamp(ceil(log2(v_hat W / d_px)), 1, 13), restated in numpy from the shapes'
parameters, src/engine/motion.cpp:1330-1340) and its carried sample at t_w. Its lattice slots, round-0
flags, refinement rounds and run assembly (src/engine/sampling.h:355-839, window_impl.h:462-716) decide
the samples. Every run of bsx_scene_sample_window must equal the reference engine's bit for bit: the
same times and values in the same order, without the carried sample the reference's run starts with,
and the pixels no shape reaches (not active) take no samples. Signature flags on and off, the eta_L
test on, several windows."""
import numpy as np
import pytest

from oracle import window_ref
from tests.test_gpu_scene import WINDOW, _level_scene, _levels

pytestmark = pytest.mark.gpu

needs_ref = pytest.mark.skipif(not window_ref.available(),
                               reason="oracle/_ref/libwindow_ref.so was not built (the reference is absent)")


def _compare(sc, view, shapes, windows, tol, dark, signature, min_refined, max_levels=4, min_split=2048):
    from ballasim_amd import scene as S
    rules = S.rules(dark, tol=tol, signature=signature, max_refine_levels=max_levels, min_split_ns=min_split)
    total_refined = total_pixels = 0
    for w in windows:
        t_w = w * WINDOW
        print(f"window {w} signature {signature}", flush=True)
        level = _levels(shapes, view, t_w)
        active = np.nonzero(level >= 0)[0].astype(np.uint32)
        offsets, t, irr, stats = window_ref.sample_window(sc, view, t_w, active, level[active], tol, dark,
                                                          max_refine_levels=max_levels, min_split=min_split,
                                                          signature=signature)
        runs, total = sc.sample_window(view, rules, t_w, slot=0)
        ids, tt, ii, offs = S.runs_to_host(runs)
        carried = sc.render(view, t_w)
        # The reference's runs hold the carried sample first: one more sample per active pixel.
        assert len(t) == total + len(active), (len(t), total, len(active))
        is_active = np.zeros(len(ids), bool)
        is_active[active] = True
        assert np.all(np.diff(offs)[~is_active] == 0), "a pixel no shape reaches has samples"
        for a, p in enumerate(active):
            r0, r1 = offsets[a], offsets[a + 1]
            assert t[r0] == t_w and irr[r0] == carried[p], f"window {w} pixel {p}: carried sample"
            mine_t, mine_v = tt[offs[p]:offs[p + 1]], ii[offs[p]:offs[p + 1], 0]
            ref_t, ref_v = t[r0 + 1:r1], irr[r0 + 1:r1]
            assert np.array_equal(mine_t, ref_t), f"window {w} pixel {p} level {level[p]}: times {mine_t} {ref_t}"
            assert np.array_equal(mine_v, ref_v), f"window {w} pixel {p}: values"
            total_refined += (r1 - r0 - 1) > (1 << int(level[p]))
        total_pixels += len(active)
        print(f"  {len(active)} active pixels, {total} samples, reference stats {stats}", flush=True)
    assert total_pixels > 0 and total_refined >= min_refined, (total_pixels, total_refined)


@needs_ref
@pytest.mark.parametrize("signature", [False, True])
def test_runs_equal_reference_window_engine_level_scene(signature):
    sc, view, shapes, _ = _level_scene()
    _compare(sc, view, shapes, range(3), 0.01, 0.0021, signature, min_refined=50)


@needs_ref
def test_runs_equal_reference_window_engine_textured_rectangles():
    """Rectangles of the benchmark's kind (6-octave texture, up to 2000 px/s and 3 rad/s) at 320x180."""
    from ballasim_amd import scene as S
    shapes, bg = S.config2_shapes(5, n=16, width=320, height=180)
    sc, view = S.Scene(shapes, bg), S.view(320, 180)
    _compare(sc, view, shapes, range(2), 0.01, 0.0021, True, min_refined=200)


@needs_ref
def test_runs_equal_reference_window_engine_deep_refinement():
    """Seven refinement levels down to 1024-ns intervals at a tighter tolerance, and the midpoint test off."""
    sc, view, shapes, _ = _level_scene()
    _compare(sc, view, shapes, range(2), 0.003, 0.0021, True, min_refined=50, max_levels=7, min_split=1024)
    _compare(sc, view, shapes, range(1), 0.0, 0.0021, True, min_refined=0)
