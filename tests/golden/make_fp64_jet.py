"""Write the FP64 oracle fixture of the reduced precision-form jet.

    python tests/golden/make_fp64_jet.py [--threads N]

Workload: D3Q27 round jet on 40x40x128 cells (ghost shell included), nozzle
D = 8, u = 0.05, Re = 5000, We = 500, uniform +-5 % u_z inflow noise with the
default seed: the 128x128x512 / D = 16 accuracy case of the binary16 form at
a quarter of the resolution, small enough for the FP64 oracle to run in a few
minutes.  Saved: u_x, u_y, u_z, phi and the density deviation rho - 1 of
the interior cells after 500, 1000, 2000 and 5000 steps as float32 rounded
to 16 mantissa bits (relative rounding <= 2^-17 = 7.6e-6 per value, which
keeps the file at a few MB; the errors the fixture measures are 1e-5 and
above), and the FP64 sums of phi and rho.  Test infrastructure: the accuracy reference of
tests/test_fp16_accuracy.py.
"""
from __future__ import annotations

import argparse
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle  # noqa: E402
from phasefieldlbm_amd import D3Q27, presets  # noqa: E402

OUT = os.path.join(HERE, "fp64_jet_d3q27_40x40x128.npz")
STEPS = (500, 1000, 2000, 5000)
FIELDS = ("ux", "uy", "uz", "phi")
MANTISSA_BITS = 16


def round_mantissa(a, bits):
    """Round float32 values to `bits` mantissa bits (to nearest, ties up)."""
    drop = 23 - bits
    b = a.view(np.uint32).astype(np.uint64)
    b = ((b + np.uint64(1 << (drop - 1))) >> np.uint64(drop)) << np.uint64(drop)
    return b.astype(np.uint32).view(np.float32)


def config():
    return presets.round_jet(40, 40, 128, 8, D3Q27)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--threads", type=int, default=0)
    args = ap.parse_args()
    if args.threads:
        oracle.set_num_threads(args.threads)
    o = oracle.Oracle(config(), 64).init()
    I = (slice(1, -1),) * 3
    out = {"steps": np.array(STEPS, dtype=np.int32)}
    done = 0
    t0 = time.time()
    for s in STEPS:
        o.step(s - done)
        done = s
        out[f"sum_rho_{s}"] = np.float64(o.field("rho")[I].sum())
        for k in FIELDS:
            a = o.field(k)[I]
            if k == "phi":
                out[f"sum_phi_{s}"] = np.float64(a.sum())
            out[f"{k}_{s}"] = round_mantissa(a.astype(np.float32), MANTISSA_BITS)
        # rho - 1 in FP64 before the rounding: rho itself is 1 to ~1e-3
        out[f"drho_{s}"] = round_mantissa((o.field("rho")[I] - 1.0).astype(np.float32), MANTISSA_BITS)
        print(f"step {s}: {time.time() - t0:.0f} s", flush=True)
    np.savez_compressed(OUT, **out)
    print(OUT, os.path.getsize(OUT), "bytes")


if __name__ == "__main__":
    main()
