"""The reference's own SensorModel(precision = kFp32, device = kCuda) running on this library through
its backend seam (oracle/_ref/seam_check: the reference's sources with BALLASIM_SENSOR_CUDA=1 and
ballasim_amd/seam/sensor_bsx.cpp, built where the reference is present and shipped with the tree),
against the same model on the reference's CPU path: events in every field (reported times included),
saved states and work counters identical, for the default prophesee_gen3 sensor (the cascade with its
in-path noise), the noise-free cascade, the noise with the ideal reset and leak, and the instantaneous
response; host runs, device runs (advance_device and download_events) and a GPU state resumed on the
CPU."""
import os
import subprocess

import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

SEAM = os.path.join(os.path.dirname(O.REF_LIB), "seam_check")


@pytest.mark.skipif(not os.path.exists(SEAM), reason="oracle/_ref/seam_check was not built (the reference is absent)")
def test_reference_sensor_model_on_bsx_equals_its_cpu_path():
    print("seam_check run", flush=True)
    r = subprocess.run([SEAM, "run"], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = [l for l in r.stdout.splitlines() if l.endswith("identical")]
    assert len(lines) == 4, r.stdout
