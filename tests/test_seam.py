"""The reference's own SensorModel built against this library (CPU side): the reference's sources
compiled with BALLASIM_SENSOR_CUDA=1 and ballasim_amd/seam/sensor_bsx.cpp in place of its CUDA backend
(src/sensor/sensor_cuda.cu), linked to libbsx.so (oracle/Makefile, target seam). Here, without a
device, SensorModel(device = kCuda) must reach this library's device count and raise the reference's
own "no CUDA device" error; tests/test_gpu_seam.py runs the same binary's parity cases on the GPU."""
import os
import subprocess

import pytest

from oracle import oracle as O

SEAM = os.path.join(os.path.dirname(O.REF_LIB), "seam_check")


@pytest.mark.skipif(not os.path.isdir("/root/reference/src/sensor"), reason="the reference's sources are absent")
def test_reference_sensor_model_builds_against_bsx():
    r = subprocess.run(["make", "-s", "-f", os.path.join(os.path.dirname(O.__file__), "Makefile"), "seam"],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert os.path.exists(SEAM)
    out = subprocess.run([SEAM, "build"], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    import ballasim_amd as bsx
    if bsx.device_count() == 0:
        assert "no device: SensorModel: no CUDA device available" in out.stdout, out.stdout
    else:
        assert "constructed on the device" in out.stdout, out.stdout


@pytest.mark.skipif(not os.path.isdir("/root/reference/src/engine"), reason="the reference's sources are absent")
def test_reference_window_engine_builds_against_bsx():
    """The reference's host WindowEngine (src/engine/window.cpp, par_host.cpp) with the engine hook of
    ballasim_amd/seam/scene_hook_bsx.h, linked to libbsx.so (oracle/Makefile, target window): it builds,
    loads next to libbsx.so and exports its entry layer. tests/test_gpu_window_engine.py runs it."""
    import ctypes

    from oracle import window_ref
    r = subprocess.run(["make", "-s", "-f", os.path.join(os.path.dirname(O.__file__), "Makefile"), "window"],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert window_ref.available()
    lib = window_ref._load()
    for name in ("wr_run", "wr_samples", "wr_fetch", "wr_free"):
        assert isinstance(getattr(lib, name), ctypes._CFuncPtr)
