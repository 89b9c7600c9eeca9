"""The C-ABI library on the CPU (no device): it loads, exports every entry point include/bsx/*.h
declares, fills presets, refuses invalid parameters by name, and reports a missing device as a
runtime error rather than running anything else."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import ballasim_amd as bsx
from ballasim_amd import _abi
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    inc = os.path.join(ROOT, "include", "bsx")
    for f in os.listdir(inc):
        if f.endswith(".h"):
            names |= set(re.findall(r"^\s*int\s+(bsx_\w+)\s*\(", open(os.path.join(inc, f)).read(), re.M))
    return names


def test_library_exports_every_declared_symbol():
    lib = _abi.load_library()
    names = declared_symbols()
    assert len(names) >= 20
    for n in sorted(names):
        assert hasattr(lib, n), n


def test_abi_signature_table_covers_header():
    assert declared_symbols() <= set(_abi.SIGNATURES) | {"bsx_draw_maps"} | {
        n for n in declared_symbols() if n.startswith(("bsx_scene", "bsx_sampler", "bsx_mlp", "bsx_batch"))}


def test_struct_sizes_match_header_layout():
    assert C.sizeof(_abi.RawEvent) == 32
    assert C.sizeof(_abi.Waveform) == 8 + 8 + 64 + 64 + 48


def test_presets_and_params_surface():
    assert bsx.preset_names() == ["prophesee_gen3", "samsung_gen3", "ideal"]
    p = bsx.preset("prophesee_gen3")
    assert p.photoreceptor == "cascade" and p.noise and p.theta_on == 0.25 and p.family == "prophesee_gen3"
    q = bsx.preset("ideal")
    assert q.photoreceptor == "instantaneous" and q.reset_mode == "ideal" and q.refractory_s == 0.0
    q.reset_mode = "physical"
    q.refractory_s = 1e-4
    q.validate()
    with pytest.raises(ValueError):
        bsx.preset("nope")


@pytest.mark.parametrize("field,value,msg", [
    ("theta_on", 0.005, "theta_on must exceed theta_floor"),
    ("tau_sf", 1.05e-6, "differ by at least 10"),
    ("eps_ci", 0.01, "eps_ci"),
    ("refractory_s", -1.0, "refractory_s"),
    ("veiling_glare", 0.1, "not supported"),
    ("tau_hp", 0.01, "not supported"),
    ("hot_rate_law", 2, "hot_rate_law"),
    ("hot_leak_fraction", 1.5, "hot_leak_fraction"),
    ("family", "foo", "family"),
])
def test_validate_refuses_by_name(field, value, msg):
    p = bsx.preset("ideal")
    setattr(p, field, value)
    with pytest.raises(ValueError, match=msg):
        p.validate()


def test_noise_on_cascade_is_accepted_and_its_unbuilt_variants_refused():
    """The reference's default sensor (the cascade with its in-path noise) validates unmodified; the
    noise variants this backend does not run are refused by name, never approximated."""
    bsx.preset("prophesee_gen3").validate()
    p = bsx.preset("prophesee_gen3")
    p.noise_diffusion = 1
    with pytest.raises(ValueError, match="transient"):
        p.validate()
    p = bsx.preset("prophesee_gen3")
    p.hot_fraction = 0.01
    p.hot_leak_fraction = 0.5
    with pytest.raises(ValueError, match="threshold-type hot pixels"):
        p.validate()
    p = bsx.preset("prophesee_gen3")
    p.shot_rate_hz = 0.1
    with pytest.raises(ValueError, match="instantaneous"):
        p.validate()


def test_noise_tables_equal_oracle():
    """The product's host tables (bsx_noise_tables) bit for bit against the oracle's, which the
    reference pins (test_noise_oracle.py), at the default and the extreme tree depths."""
    for levels in (5, 0, 10):
        p = bsx.preset("prophesee_gen3")
        p.noise_leaf_level = levels
        a64, a32 = bsx.noise_tables(p)
        b64, b32 = O.noise_tables(p.c)
        assert a64.shape == (705, 18 + 47 * levels)
        assert np.array_equal(a64.view(np.uint64), b64.view(np.uint64))
        assert np.array_equal(a32.view(np.uint32), b32.view(np.uint32))


def test_no_device_is_a_runtime_error():
    if bsx.device_count() > 0:
        pytest.skip("a device is visible")
    with pytest.raises(RuntimeError, match="no HIP device"):
        bsx.SensorModel("ideal", 8, 8, 1, device="cuda")
    with pytest.raises(ValueError, match="GPU only"):
        bsx.SensorModel("ideal", 8, 8, 1, device="cpu")


def test_draw_maps_host_entry_point():
    lib = _abi.load_library()
    p = bsx.preset("prophesee_gen3")
    n = 100
    a = [np.zeros(n) for _ in range(5)]
    _abi.check(lib.bsx_draw_maps(C.byref(p.c), 1, 0, n, *[x.ctypes.data_as(C.c_void_p) for x in a]))
    assert np.all(a[0] > 0.01) and np.all(a[3] > 0)
    bad = bsx.preset("prophesee_gen3")
    bad.theta_on = -1.0
    with pytest.raises(ValueError):
        _abi.check(lib.bsx_draw_maps(C.byref(bad.c), 1, 0, n, None, None, None, None, None))


@pytest.mark.parametrize("law,reset", [(0, "physical"), (1, "ideal"), (1, "physical")])
def test_draw_maps_hot_pixels_equal_oracle(law, reset):
    """Hot pixels in the product's host maps (bsx_draw_maps): the raised leaks of leak-type hot pixels
    equal the oracle's (itself pinned to the compiled reference by test_oracle_vs_ref.py)."""
    from oracle import oracle as O
    lib = _abi.load_library()
    p = bsx.preset("ideal")
    p.hot_fraction = 0.25
    p.hot_rate_law = law
    p.reset_mode = reset
    p.refractory_s = 1e-4
    p.hot_rate_min_hz, p.hot_rate_max_hz = 5.0, 3000.0
    p.hot_rate_median_hz, p.hot_rate_sigma_ln = 80.0, 1.2
    p.leak_rate_hz = 0.1
    p.validate()
    n = 4000
    a = [np.zeros(n) for _ in range(5)]
    _abi.check(lib.bsx_draw_maps(C.byref(p.c), 77, 3, n, *[x.ctypes.data_as(C.c_void_p) for x in a]))
    om = O.draw_maps(p.c, 77, 3, n)
    for key, x in zip(("theta_on", "theta_off", "leak", "dark_fA", "phi0"), a):
        assert np.array_equal(x, om[key]), key
    hot = om["hot"] == 1
    assert 0.2 < hot.mean() < 0.3
    # a leak-type hot pixel's leak is theta_on R_L with R_L = R / (1 - R T_r) (physical) or R (ideal)
    R = om["hot_rate_hz"][hot]
    RL = R / (1.0 - R * 1e-4) if reset == "physical" else R
    assert np.allclose(a[2][hot], a[0][hot] * RL, rtol=1e-12)
