"""Run the CPU oracle on a named workload and save field snapshots.

    python tests/fp64_snapshots.py --workload jet_d3q27_128x128x512_fp16 \
        --precision 64 --steps 100,500,1000,2000,3000,5000 --out <dir>

Writes <dir>/<field>_<step>.npy (float32 copies of rho, ux, uy, uz, phi)
and <dir>/meta_<step>.json (FP64 sums).  Test infrastructure: these
snapshots are the accuracy reference of the binary16-storage form, read by
tests/test_fp16_accuracy.py.  The FP64 oracle needs about 3.8 s per step of
the 128x128x512 D3Q27 jet on 7 cores, so the 5000-step set takes ~5 h.
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

import oracle  # noqa: E402
from phasefieldlbm_amd import presets  # noqa: E402

FIELDS = ("rho", "ux", "uy", "uz", "phi")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", required=True)
    ap.add_argument("--precision", type=int, default=64)
    ap.add_argument("--steps", default="500,1000,2000,5000")
    ap.add_argument("--out", required=True)
    ap.add_argument("--threads", type=int, default=0)
    args = ap.parse_args()
    if args.threads:
        oracle.set_num_threads(args.threads)
    os.makedirs(args.out, exist_ok=True)
    make, _ = presets.WORKLOADS[args.workload]
    cfg = make()
    stops = sorted(int(s) for s in args.steps.split(","))
    o = oracle.Oracle(cfg, args.precision).init()
    t0 = time.time()
    done = 0
    for s in stops:
        o.step(s - done)
        done = s
        meta = {"step": s, "precision": args.precision, "workload": args.workload,
                "elapsed_s": time.time() - t0, "threads": oracle.num_threads()}
        I = (slice(1, -1),) * 3
        for k in FIELDS:
            a = o.field(k)
            meta["sum_" + k] = float(a[I].sum())
            np.save(os.path.join(args.out, f"{k}_{s}.npy"), a.astype(np.float32))
        with open(os.path.join(args.out, f"meta_{s}.json"), "w") as fh:
            json.dump(meta, fh)
        print(json.dumps(meta), flush=True)


if __name__ == "__main__":
    main()
