"""The noise lane program (ballasim_amd/csrc/sensor/noise.h) built for the host as a small C library
(tests/host/noise_lane.cpp), with the arrays of one camera in the kernels' layout: the CPU run of the
device program that the tests compare with the oracle before the GPU runs it."""
import ctypes as C
import os
import pathlib
import shutil
import subprocess

import numpy as np

from oracle import oracle as O

HERE = pathlib.Path(__file__).resolve().parent
SENSOR = HERE.parent / "ballasim_amd" / "csrc" / "sensor"

EVENT = O.EVENT_DTYPE


class HostConst(C.Structure):
    _fields_ = [("n_pixels", C.c_int64), ("groups", C.c_int64), ("refractory_ns", C.c_int64),
                ("ideal_reset", C.c_int64), ("seed", C.c_uint64), ("camera", C.c_int64)] + \
               [(n, C.c_float) for n in ("k", "a2", "a3", "tau_o", "tau_sf", "eps_accept", "eps_drop", "eps_ci",
                                         "label_cut", "irr_max", "ln_min", "inv_dln", "margin_z")] + \
               [("levels", C.c_int32), ("n_nodes", C.c_int32), ("stride", C.c_int32), ("pad", C.c_int32)]


class HostMaps(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in ("theta_on", "theta_off", "leak", "dark", "scale", "phi0", "hot")]


STATE_F32 = ("z", "x1", "s2", "s3", "i_scale", "tv", "s3_ref", "n3_ref", "dr_ref")
STATE_I64 = ("t_ref", "blind_until", "t_search")


class HostState(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in STATE_F32 + STATE_I64 + ("flags", "held", "nk", "nk1", "ln_a1", "cell",
                                                                   "have_k1")]


COUNTERS = ("cells", "refinements", "substeps", "substep_attempts", "settled_substeps", "leaf_intervals",
            "fallbacks", "quarantines", "pieces")

_lib = None


def lib(tmp_dir):
    global _lib
    if _lib is None:
        if shutil.which("g++") is None:
            return None
        out = os.path.join(tmp_dir, "libnoise_lane.so")
        subprocess.run(["g++", "-O2", "-std=c++20", "-ffp-contract=off", "-fno-fast-math", "-shared", "-fPIC",
                        f"-I{SENSOR}", "-o", out, str(HERE / "host" / "noise_lane.cpp")], check=True)
        L = C.CDLL(out)
        P = C.c_void_p
        L.noise_lane_init.argtypes = [C.POINTER(HostConst), P, HostMaps, P, C.c_int64, HostState, P]
        L.noise_lane_advance.restype = C.c_int64
        L.noise_lane_advance.argtypes = [C.POINTER(HostConst), P, HostMaps, P, P, P, P, P, C.c_int64, C.c_int64,
                                         HostState, P, C.c_int64, P]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


class HostCamera:
    """One camera of the lane program on the host: the product's host maps and tables, the fp32 call
    constants formed as the model forms them, and the fp32 structure-of-arrays state."""

    def __init__(self, L, params, width, height, seed, camera=0, groups=1):
        import ballasim_amd as bsx
        from ballasim_amd import _abi
        self.L, self.P, self.G = L, width * height, groups
        p = params
        P = self.P
        lib = _abi.load_library()
        maps = [np.zeros(P) for _ in range(5)]
        _abi.check(lib.bsx_draw_maps(C.byref(p.c), seed, camera, P, *[_p(a) for a in maps]))
        self.maps = {k: v.astype(np.float32) for k, v in zip(("theta_on", "theta_off", "leak", "dark", "phi0"), maps)}
        self.maps["scale"] = np.full(P, np.float32(p.illuminance_scale), np.float32)
        self.maps["hot"] = np.zeros(P, np.uint8)
        _, self.tab = bsx.noise_tables(p)
        self.tab = np.ascontiguousarray(self.tab)
        eps_h = p.eps_L / 8.0
        log10 = O.oracle_lib().orc_log(10.0)
        self.c = HostConst(n_pixels=P, groups=groups, refractory_ns=int(round(p.refractory_s * 1e9)),
                           ideal_reset=1 if p.reset_mode == "ideal" else 0, seed=seed, camera=camera,
                           k=np.float32(p.k * 1e-15), a2=np.float32(1.0 / p.tau_o), a3=np.float32(1.0 / p.tau_sf),
                           tau_o=np.float32(p.tau_o), tau_sf=np.float32(p.tau_sf),
                           eps_accept=np.float32(eps_h / 1.5),
                           eps_drop=np.float32(max(eps_h / 1024.0, 32.0 * 2.0 ** -23)),
                           eps_ci=np.float32(p.eps_ci), label_cut=np.float32(p.label_share_cut),
                           irr_max=np.float32(1e9 / p.illuminance_scale),
                           ln_min=np.float32(-4.605170185988091), inv_dln=np.float32(1.0 / (log10 / 64)),
                           margin_z=np.float32(p.noise_margin_z), levels=p.noise_leaf_level,
                           n_nodes=self.tab.shape[0], stride=self.tab.shape[1])
        self.m = HostMaps(*[_p(self.maps[k]) for k in ("theta_on", "theta_off", "leak", "dark", "scale", "phi0", "hot")])
        self.st = {k: np.zeros(P, np.float32) for k in STATE_F32}
        self.st.update({k: np.zeros(P, np.int64) for k in STATE_I64})
        self.st["flags"] = np.zeros(P, np.uint16)
        self.st["held"] = np.zeros(P * groups, np.float32)
        self.st["nk"] = np.zeros(3 * P, np.float32)
        self.st["nk1"] = np.zeros(3 * P, np.float32)
        self.st["ln_a1"] = np.zeros(P, np.float32)
        self.st["cell"] = np.zeros(P, np.int64)
        self.st["have_k1"] = np.zeros(P, np.uint8)
        self.s = HostState(*[_p(self.st[k]) for k in STATE_F32 + STATE_I64 + ("flags", "held", "nk", "nk1", "ln_a1",
                                                                              "cell", "have_k1")])
        self.counters = np.zeros(len(COUNTERS), np.uint64)
        self.t_state = None

    def init(self, t0, irr0):
        irr0 = np.ascontiguousarray(irr0, np.float32)
        self.L.noise_lane_init(C.byref(self.c), _p(self.tab), self.m, _p(irr0), int(t0), self.s, _p(self.counters))
        self.t_state = int(t0)

    def advance_runs(self, ids, t, irr, offsets, t_end):
        row_of = np.full(self.P, -1, np.int32)
        row_of[np.asarray(ids, np.int64)] = np.arange(len(ids), dtype=np.int32)
        start = np.ascontiguousarray(np.asarray(offsets[:-1], np.int64))
        count = np.ascontiguousarray(np.diff(offsets).astype(np.int32))
        t = np.ascontiguousarray(t, np.int64)
        irr = np.ascontiguousarray(irr, np.float32)
        cap = 1 << 20
        ev = np.zeros(cap, EVENT)
        n = self.L.noise_lane_advance(C.byref(self.c), _p(self.tab), self.m, _p(row_of), _p(start), _p(count), _p(t),
                                      _p(irr), self.t_state, int(t_end), self.s, _p(ev), cap, _p(self.counters))
        assert n <= cap
        self.t_state = int(t_end)
        ev = ev[:n]
        return ev[np.lexsort((ev["pixel"], ev["t_cross"]))]

    def states(self):
        P = self.P
        out = dict(self.st)
        out["nk"] = self.st["nk"].reshape(3, P).T
        out["nk1"] = self.st["nk1"].reshape(3, P).T
        return out

    def stats(self):
        return dict(zip(COUNTERS, (int(x) for x in self.counters)))


def compare_states(host, oracle_states, where):
    """Every field the noise lane keeps, fp32 against the oracle's fp64 storage (exact widening)."""
    st = host.states()
    for f in STATE_F32 + ("ln_a1",):
        a = st[f]
        b = oracle_states[f].astype(np.float32)
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), f"{where}: state {f}"
    for f in STATE_I64 + ("cell",):
        assert np.array_equal(st[f], oracle_states[f]), f"{where}: state {f}"
    assert np.array_equal(st["flags"], oracle_states["flags"]), f"{where}: state flags"
    assert np.array_equal(st["have_k1"], oracle_states["have_k1"]), f"{where}: state have_k1"
    for f in ("nk", "nk1"):
        assert np.array_equal(st[f].view(np.uint32), oracle_states[f].astype(np.float32).view(np.uint32)), \
            f"{where}: state {f}"
