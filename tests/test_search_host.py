"""The instantaneous lane's searches compiled for the host (g++, the same fp32 policy as the device
build: -ffp-contract=off, no fast-math), against the generic time-ordered bisection (the
reference's instant_search, src/sensor/pixel.h:1884-1916):
  - on falling segments under leak where Delta runs level with the OFF threshold, or with the ON
    threshold that the leak may reach, the monotone transition searches must equal it in every
    field, with far fewer evaluations (the reason they exist);
  - the photocurrent screen (phase A's and the walk's) must agree with it wherever it decides;
  - the walk from phase A's first loud segment, screening each piece, and also passing phase A's
    16-segment screen blocks at once, must give exactly the events and final state of the
    segment-by-segment walk on exact searches alone."""
import pathlib
import re
import shutil
import subprocess

import pytest

HERE = pathlib.Path(__file__).resolve().parent
SENSOR = HERE.parent / "ballasim_amd" / "csrc" / "sensor"


def _build(tmp_path_factory, name):
    if shutil.which("g++") is None:
        pytest.skip("g++ not found")
    exe = tmp_path_factory.mktemp(name) / name
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", f"-I{SENSOR}", "-o", str(exe),
                    str(HERE / "host" / f"{name}.cpp")], check=True)
    return exe


@pytest.fixture(scope="module")
def search_check(tmp_path_factory):
    return _build(tmp_path_factory, "search_check")


@pytest.fixture(scope="module")
def walk_check(tmp_path_factory):
    return _build(tmp_path_factory, "walk_check")


@pytest.mark.parametrize("seed", [12345, 777])
def test_transition_search_equals_bisection(search_check, seed):
    r = subprocess.run([str(search_check), "20000", str(seed)], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout
    m = re.search(r"trials (\d+) found (\d+) on (\d+) generic (\d+) mismatches (\d+) evaluations transition (\d+) "
                  r"bisection (\d+)", r.stdout)
    assert m, r.stdout
    trials, found, on, generic, bad, ev_t, ev_b = map(int, m.groups())
    assert bad == 0
    assert generic == 0                  # every case here is decided by the transition search
    assert 0.1 * trials < found < 0.9 * trials   # both outcomes are exercised
    assert 0.02 * trials < on < found    # ON crossings of the two-sided bisection too
    assert ev_t * 10 < ev_b


@pytest.mark.parametrize("seed", [1, 2])
def test_photocurrent_screen_agrees_with_bisection(search_check, seed):
    """The log-free photocurrent screen (phase A's and the walk's) on rising, falling and mixed segments,
    with and without leak, over eight decades of photocurrent, Delta placed a few margins from either
    level: wherever it passes a segment, the generic bisection finds no crossing there."""
    r = subprocess.run([str(search_check), "60000", str(seed), "screen"], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout
    m = re.search(r"screen trials (\d+) quiet (\d+) near_quiet (\d+) loud (\d+) mismatches (\d+)", r.stdout)
    assert m, r.stdout
    trials, quiet, near_quiet, loud, bad = map(int, m.groups())
    assert bad == 0
    assert near_quiet > 0.1 * trials and loud > 0.1 * trials  # passes right at its margin are exercised


@pytest.mark.parametrize("seed", [7, 8])
def test_screened_walk_equals_exact_walk(walk_check, seed):
    r = subprocess.run([str(walk_check), "3000", str(seed)], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout
    m = re.search(r"walk trials (\d+) events (\d+) screened (\d+) block_skips (\d+) differences (\d+)", r.stdout)
    assert m, r.stdout
    trials, events, screened, blocks, bad = map(int, m.groups())
    assert bad == 0
    assert events > 5 * trials       # runs that fire (a third of them in steady light, which fires less)
    assert screened > trials         # and quiet prefixes that phase A's screen decides
    assert blocks > trials           # and screen blocks the walk passes at once
