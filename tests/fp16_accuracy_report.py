"""Errors of the storage forms against an FP64 fixture of a jet.

    python tests/fp16_accuracy_report.py --fixture F.npz \
        --workload jet_d3q27_128x128x512_fp16 [--forms fp32,fp16,fp16_ne,mom]

F.npz holds the FP64 oracle's interior u_x, u_y, u_z and phi at the steps
its "steps" array lists, with the FP64 sums of phi (the format of
tests/golden/fp64_jet_d3q27_40x40x128.npz; `tests/fp64_snapshots.py --pack`
writes one from a snapshot directory).  For every storage form the library
runs the workload from t = 0 on cuda:0 and the errors of
tests/test_fp16_accuracy.py (rel_l2 and rel_linf of u and phi, phase-mass
drift) at every fixture step are printed as a table and one JSON line.

Test infrastructure: the 128x128x512 fixture takes about 5 h of CPU to
write and some 250 MB, so this comparison is a script, not a test; the
committed 40x40x128 fixture is checked the same way by the GPU tests.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, HERE)

import phasefieldlbm_amd as pf  # noqa: E402
from phasefieldlbm_amd import presets  # noqa: E402
from test_fp16_accuracy import _reduced_metrics  # noqa: E402

FORMS = {"fp32": pf.STORE_FP32, "fp16": pf.STORE_FP16, "fp16_ne": pf.STORE_FP16_NE, "mom": pf.STORE_MOM}


def run(fixture: str, workload: str, forms) -> dict:
    fix = np.load(fixture)
    steps = [int(s) for s in fix["steps"]]
    make, _ = presets.WORKLOADS[workload]
    rep = {}
    for name in forms:
        cfg = make()
        cfg.storage = FORMS[name]
        t0 = time.time()
        with pf.Simulation(cfg) as sim:
            sim.init()
            done = 0
            for s in steps:
                sim.step(s - done, save=True)
                done = s
                got = {k: sim.field(k) for k in ("ux", "uy", "uz", "phi")}
                m = _reduced_metrics(got, fix, s)
                rep.setdefault(s, {})[name] = m
                print(f"[{name}] step {s}: rel_l2 u {m['rel_l2_u']:.3e} phi {m['rel_l2_phi']:.3e} "
                      f"({time.time() - t0:.1f} s)", flush=True)
    return rep


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--fixture", required=True)
    ap.add_argument("--workload", required=True)
    ap.add_argument("--forms", default="fp32,fp16,fp16_ne,mom")
    args = ap.parse_args()
    forms = [f for f in args.forms.split(",") if f]
    rep = run(args.fixture, args.workload, forms)
    print("| step | form | rel_l2 u | rel_linf u | rel_l2 phi | rel_linf phi | mass drift |")
    print("|---|---|---|---|---|---|---|")
    for s in sorted(rep):
        for f in forms:
            m = rep[s][f]
            print(f"| {s} | {f} | {m['rel_l2_u']:.2e} | {m['rel_linf_u']:.2e} | {m['rel_l2_phi']:.2e} | "
                  f"{m['rel_linf_phi']:.2e} | {m['phase_mass_drift']:.1e} |")
    print(json.dumps({"accuracy_vs_fp64": {"workload": args.workload, "fixture": os.path.basename(args.fixture),
                                           "errors": rep}}), flush=True)


if __name__ == "__main__":
    main()
