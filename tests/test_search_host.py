"""The instantaneous lane's exact searches compiled for the host (g++, the same fp32 policy as the
device build: -ffp-contract=off, no fast-math): on falling segments under leak where Delta runs
level with the OFF threshold, the monotone transition search must equal the generic time-ordered
bisection (the reference's instant_search, src/sensor/pixel.h:1884-1916) in every field, and need
far fewer evaluations (the reason it exists)."""
import pathlib
import re
import shutil
import subprocess

import pytest

HERE = pathlib.Path(__file__).resolve().parent
SENSOR = HERE.parent / "ballasim_amd" / "csrc" / "sensor"


@pytest.fixture(scope="module")
def search_check(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ not found")
    exe = tmp_path_factory.mktemp("search") / "search_check"
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", f"-I{SENSOR}", "-o", str(exe),
                    str(HERE / "host" / "search_check.cpp")], check=True)
    return exe


@pytest.mark.parametrize("seed", [12345, 777])
def test_transition_search_equals_bisection(search_check, seed):
    r = subprocess.run([str(search_check), "20000", str(seed)], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout
    m = re.search(r"trials (\d+) found (\d+) generic (\d+) mismatches (\d+) evaluations transition (\d+) "
                  r"bisection (\d+)", r.stdout)
    assert m, r.stdout
    trials, found, generic, bad, ev_t, ev_b = map(int, m.groups())
    assert bad == 0
    assert generic == 0                  # every case here is decided by the transition search
    assert 0.1 * trials < found < 0.9 * trials   # both outcomes are exercised
    assert ev_t * 10 < ev_b
