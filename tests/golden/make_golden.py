"""Writes the golden fixtures of tests/golden from the reference's own CPU sensor path.

The reference is compiled from its sources by oracle/Makefile into oracle/_ref/libballa_ref.so (only
where /root/reference is present); this script drives it through oracle/oracle.py's RefModel on
small seeded inputs and stores inputs and outputs as .npz data, with a manifest of the parameters.
Run it from the repository root: python tests/golden/make_golden.py
"""
import ctypes as C
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from ballasim_amd import _abi  # noqa: E402
from oracle import oracle as O  # noqa: E402
from tests.golden.cases import (CASES, MAP_CASES, NOISE_STATE_FIELDS, STATE_FIELDS, case_inputs,  # noqa: E402
                                case_params, noise_case)


def params_dict(p):
    out = {}
    for f, _ in _abi.Params._fields_:
        v = getattr(p, f)
        out[f] = v.decode() if isinstance(v, bytes) else v
    return out


def main():
    if not O.ref_available():
        raise SystemExit("oracle/_ref/libballa_ref.so is missing: build it with make -f oracle/Makefile ref")
    manifest = {"source": "reference CPU sensor path, fp32 (oracle/_ref)", "cases": {}, "maps": {}}
    for name, spec in MAP_CASES.items():
        p = case_params(spec)
        m = O.RefModel(p, spec["n"], 1, spec["seed"], 1, spec["camera"], 1, True)
        maps = m.maps()
        np.savez_compressed(os.path.join(HERE, f"maps_{name}.npz"), **maps)
        manifest["maps"][name] = {"params": params_dict(p), "n": spec["n"], "seed": spec["seed"],
                                  "camera": spec["camera"]}
    for name, spec in CASES.items():
        p = case_params(spec)
        W, H, G = spec["width"], spec["height"], spec["groups"]
        m = O.RefModel(p, W, H, spec["seed"], G, spec["camera"], 1, True)
        irr0, calls = case_inputs(spec)
        m.init(spec["t0"], irr0)
        arrays = {"irr0": irr0}
        for k, (ids, t, irr, offs, t_end) in enumerate(calls):
            arrays[f"ids_{k}"] = ids
            arrays[f"t_{k}"] = t
            arrays[f"irr_{k}"] = irr
            arrays[f"offsets_{k}"] = offs
            arrays[f"t_end_{k}"] = np.int64(t_end)
            ev, rep = m.advance_runs(ids, t, irr, offs, t_end)
            arrays[f"events_{k}"] = ev
            arrays[f"reported_{k}"] = rep
        st = m.states()
        for f in STATE_FIELDS + (NOISE_STATE_FIELDS if noise_case(spec) else ()):
            arrays[f"state_{f}"] = st[f]
        arrays["stats"] = np.array([m.stats()[k] for k in sorted(m.stats())], np.uint64)
        arrays["stat_names"] = np.array(sorted(m.stats()))
        np.savez_compressed(os.path.join(HERE, f"case_{name}.npz"), **arrays)
        manifest["cases"][name] = {"params": params_dict(p), **{k: v for k, v in spec.items() if k != "overrides"},
                                   "overrides": spec.get("overrides", {}),
                                   "events": int(sum(len(arrays[f"events_{k}"]) for k in range(len(calls))))}
    # The oracle's log over every positive float: monotone, and the hash the device log must match.
    v = np.zeros(1, np.uint64)
    h = np.zeros(1, np.uint64)
    O.oracle_lib().orc_log_check(os.cpu_count() or 1, v.ctypes.data_as(C.c_void_p), h.ctypes.data_as(C.c_void_p))
    manifest["log_check"] = {"violations": int(v[0]), "hash": "0x%016x" % int(h[0]),
                             "source": "oracle orc_logf over bit patterns 1 .. 0x7f7fffff"}
    with open(os.path.join(HERE, "manifest.json"), "w") as f:
        json.dump(manifest, f, indent=1, sort_keys=True)
    print("wrote", len(manifest["cases"]), "cases and", len(manifest["maps"]), "map sets")


if __name__ == "__main__":
    main()
