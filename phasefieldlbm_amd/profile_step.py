"""Run a named workload for a few timesteps, for profilers.

    python -m phasefieldlbm_amd.profile_step --workload jet_d3q27_256x256x1024 --steps 4
    python -m phasefieldlbm_amd.profile_step --workload jet_d3q27_256x256x1024 --closure --warmup 3000 --steps 20
    python -m phasefieldlbm_amd.profile_step --copy-bytes 1073741824

The first form initialises the workload and runs --steps timesteps (an even
number runs both in-place step kinds equally often); the second runs one
streaming copy of a known size, whose memory-traffic counters calibrate the
counters of the timestep kernels.
"""
from __future__ import annotations

import argparse
import os
import sys

if __package__ in (None, ""):  # run as a script: make the package importable
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default=None)
    ap.add_argument("--steps", type=int, default=4)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--zchunk", type=int, default=0)
    ap.add_argument("--slot-pad", type=int, default=-1)
    ap.add_argument("--copy-bytes", type=int, default=0)
    ap.add_argument("--closure", action="store_true", help="learned curvature closure (committed weights)")
    args = ap.parse_args(argv)

    import phasefieldlbm_amd as pf
    from phasefieldlbm_amd import presets

    if args.copy_bytes:
        gbs = pf.copy_bandwidth(0, args.copy_bytes, 1)
        print(f"copy {args.copy_bytes} bytes: {gbs:.1f} GB/s", flush=True)
    if args.workload:
        make, _ = presets.WORKLOADS[args.workload]
        cfg = make()
        cfg.zchunk, cfg.slot_pad = args.zchunk, args.slot_pad
        if args.closure:
            cfg.closure = pf.CLOSURE_MLP
        with pf.Simulation(cfg) as sim:
            if args.closure:
                from phasefieldlbm_amd import closure
                sim.set_closure_params(closure.load())
            sim.init().step(args.warmup)
            sim.sync()
            print(f"{args.workload}: warmup {args.warmup} done, timing {args.steps} steps", flush=True)
            sim.step(args.steps)
            sim.sync()
        print("done", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
