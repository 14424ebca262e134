"""Regenerate tests/golden/oracle_trajectories.npz.

The reference ships no golden vectors, known-answer tests or fixtures for
its timestep, and it cannot be built here (CUDA).  These fixtures are the
FP32 oracle's own trajectories on three tiny cases, committed so that a
change to the oracle shows up as a test failure and so that the GPU parity
tests have a fixed target that does not depend on rebuilding the oracle:

  droplet_q19   static droplet, D3Q19, 16^3, R = 4, 12 steps
  droplet_q27   static droplet, D3Q27, 16^3, R = 4, 12 steps
  jet_q27       round jet, D3Q27, 12x12x24, D = 4, uniform +-5 % u_z noise,
                13 steps (odd, so the fixture ends after an A-A odd step)

Each case stores rho, ux, uy, uz and phi (float32, z-major).  Run from the
repository root:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

FIELDS = ("rho", "ux", "uy", "uz", "phi")


def cases():
    import phasefieldlbm_amd as pf
    from phasefieldlbm_amd import presets
    return {
        "droplet_q19": (presets.static_droplet(n=16, radius=4.0, velocity_set=pf.D3Q19), 12),
        "droplet_q27": (presets.static_droplet(n=16, radius=4.0, velocity_set=pf.D3Q27), 12),
        "jet_q27": (presets.round_jet(12, 12, 24, 4, pf.D3Q27, noise=pf.NOISE_UNIFORM_Z), 13),
    }


def main():
    import oracle
    out = {}
    for name, (cfg, steps) in cases().items():
        o = oracle.Oracle(cfg, 32).init().step(steps)
        for k in FIELDS:
            out[f"{name}/{k}"] = o.field(k).astype(np.float32)
        out[f"{name}/steps"] = np.array(steps)
        o.close()
    path = os.path.join(ROOT, "tests", "golden", "oracle_trajectories.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, sum(v.nbytes for v in out.values()), "bytes")


if __name__ == "__main__":
    main()
