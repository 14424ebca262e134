"""Run the CPU oracle on a named workload and save field snapshots.

    python tests/fp64_snapshots.py --workload jet_d3q27_128x128x512_fp16 \
        --precision 64 --steps 100,500,1000,2000,3000,5000 --out <dir>

Writes <dir>/<field>_<step>.npy (float32 copies of rho, ux, uy, uz, phi)
and <dir>/meta_<step>.json (FP64 sums).

    python tests/fp64_snapshots.py --pack <dir> --out F.npz

packs a snapshot directory into the fixture format that
tests/fp16_accuracy_report.py reads (interior u_x, u_y, u_z, phi rounded
to 16 mantissa bits, the FP64 sums, the step list).  Test infrastructure:
these snapshots are the accuracy reference of the binary16 storage forms.
The FP64 oracle needs about 3.8 s per step of the 128x128x512 D3Q27 jet on
8 cores, so the 5000-step set takes ~5 h.
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

import oracle  # noqa: E402
from phasefieldlbm_amd import presets  # noqa: E402

FIELDS = ("rho", "ux", "uy", "uz", "phi")


def pack(src: str, out: str) -> None:
    from golden.make_fp64_jet import FIELDS as PACKED, MANTISSA_BITS, round_mantissa
    steps = sorted(int(os.path.basename(p)[5:-5]) for p in os.listdir(src) if p.startswith("meta_"))
    res = {"steps": np.array(steps, dtype=np.int32)}
    I = (slice(1, -1),) * 3
    for s in steps:
        with open(os.path.join(src, f"meta_{s}.json")) as fh:
            meta = json.load(fh)
        res[f"sum_phi_{s}"] = np.float64(meta["sum_phi"])
        res[f"sum_rho_{s}"] = np.float64(meta["sum_rho"])
        for k in PACKED:
            a = np.load(os.path.join(src, f"{k}_{s}.npy"))[I]
            res[f"{k}_{s}"] = round_mantissa(np.ascontiguousarray(a, dtype=np.float32), MANTISSA_BITS)
    np.savez_compressed(out, **res)
    print(out, steps, os.path.getsize(out), "bytes", flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--pack", default="")
    ap.add_argument("--workload", default="")
    ap.add_argument("--precision", type=int, default=64)
    ap.add_argument("--steps", default="500,1000,2000,5000")
    ap.add_argument("--out", required=True)
    ap.add_argument("--threads", type=int, default=0)
    args = ap.parse_args()
    if args.pack:
        pack(args.pack, args.out)
        return
    if not args.workload:
        ap.error("--workload is required")
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
