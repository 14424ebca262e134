"""Accuracy of binary16 population storage against the FP64 oracle.

Workload: 128x128x512 D3Q27 round jet (D = 16, u = 0.05, Re = 5000,
We = 500, uniform +-5 % u_z inflow noise).  The FP64 reference fields come
from tests/fp64_snapshots.py (about 5 h of CPU for 5000 steps), so the tests
run only when PF_FP64_SNAPSHOTS names a directory holding them; otherwise they
are skipped.  For every snapshot step the library runs with FP16 and with
FP32 storage, and the error over interior cells is

    rel_l2(u)   = ||u - u64||_2 / ||u64||_2       (vector field)
    rel_l2(phi) = ||phi - phi64||_2 / ||phi64||_2
    rel_linf    = max |a - a64| / max |a64|
    rel_l2(rho - 1), phase-mass drift (sum phi - sum phi64) / sum phi64

FP32 storage is the floor set by round-off growth in a jet at Re = 5000.
The bound is 1e-3 on u and phi, or 3x that floor where the floor itself is
larger.  Up to 1000 steps binary16 storage holds it.  Later, the jet's
shear-layer instability amplifies the larger binary16 rounding (about 40x
the FP32 error at 1000 steps) until the flow decorrelates from the FP64
trajectory: at 5000 steps the error is some 7e-2 while FP32 stays near
2.5e-4.  That case is an expected failure and fails loudly if it passes.
"""
import glob
import json
import os

import numpy as np
import pytest

SNAP = os.environ.get("PF_FP64_SNAPSHOTS", "")


def _steps(d):
    return sorted(int(os.path.basename(p)[5:-5]) for p in glob.glob(os.path.join(d, "meta_*.json")))


def _metrics(got, ref):
    I = (slice(1, -1),) * 3
    u = np.stack([got[k][I] for k in ("ux", "uy", "uz")]).astype(np.float64)
    u64 = np.stack([ref[k][I] for k in ("ux", "uy", "uz")]).astype(np.float64)
    ph, ph64 = got["phi"][I].astype(np.float64), ref["phi"][I].astype(np.float64)
    r, r64 = got["rho"][I].astype(np.float64) - 1.0, ref["rho"][I].astype(np.float64) - 1.0
    return {
        "rel_l2_u": float(np.linalg.norm(u - u64) / np.linalg.norm(u64)),
        "rel_linf_u": float(np.abs(u - u64).max() / np.abs(u64).max()),
        "rel_l2_phi": float(np.linalg.norm(ph - ph64) / np.linalg.norm(ph64)),
        "rel_linf_phi": float(np.abs(ph - ph64).max() / np.abs(ph64).max()),
        "rel_l2_rho_minus_1": float(np.linalg.norm(r - r64) / max(np.linalg.norm(r64), 1e-30)),
        "phase_mass_drift": float((ph.sum() - ph64.sum()) / ph64.sum()),
    }


def test_metrics_definition():
    """The error measures on hand-made fields (no GPU, no snapshots)."""
    rng = np.random.default_rng(0)
    shape = (6, 5, 4)
    ref = {k: rng.normal(size=shape).astype(np.float32) for k in ("ux", "uy", "uz")}
    ref["phi"] = rng.uniform(size=shape).astype(np.float32)
    ref["rho"] = (1 + 0.01 * rng.normal(size=shape)).astype(np.float32)
    m0 = _metrics(ref, ref)
    assert all(v == 0.0 for v in m0.values())
    got = dict(ref)
    got["phi"] = ref["phi"] * np.float32(1.001)
    m = _metrics(got, ref)
    assert abs(m["rel_l2_phi"] - 1e-3) < 1e-6 and abs(m["phase_mass_drift"] - 1e-3) < 1e-6
    assert m["rel_l2_u"] == 0.0


needs_snapshots = pytest.mark.skipif(not SNAP or not os.path.isdir(SNAP),
                                     reason="PF_FP64_SNAPSHOTS not set (FP64 snapshots take ~5 h)")


@pytest.fixture(scope="module")
def report():
    """Errors of FP16 and FP32 storage against every FP64 snapshot."""
    import phasefieldlbm_amd as pf
    from phasefieldlbm_amd import presets
    steps = _steps(SNAP)
    assert steps, f"no snapshots in {SNAP}"
    rep = {}
    for name, storage in (("fp16", pf.STORE_FP16), ("fp32", pf.STORE_FP32)):
        cfg = presets.jet_d3q27_128x128x512(storage)
        with pf.Simulation(cfg) as sim:
            sim.init()
            done = 0
            for s in steps:
                sim.step(s - done, save=True)
                done = s
                got = {k: sim.field(k) for k in ("rho", "ux", "uy", "uz", "phi")}
                ref = {k: np.load(os.path.join(SNAP, f"{k}_{s}.npy")) for k in got}
                rep.setdefault(s, {})[name] = _metrics(got, ref)
    print(json.dumps({"fp16_accuracy_vs_fp64": rep}), flush=True)
    return rep


def _check(r, s):
    for k in ("rel_l2_u", "rel_l2_phi"):
        assert np.isfinite(r["fp16"][k])
        assert r["fp16"][k] <= max(1e-3, 3.0 * r["fp32"][k]), (s, k, r)


@pytest.mark.gpu
@needs_snapshots
def test_fp16_storage_against_fp64_to_1000_steps(report):
    early = [s for s in report if s <= 1000]
    assert early, "no snapshot at or before step 1000"
    for s in early:
        _check(report[s], s)


@pytest.mark.gpu
@needs_snapshots
@pytest.mark.xfail(strict=True, reason="binary16 storage decorrelates from the FP64 jet between 1000 and 3000 "
                                       "steps (shear-layer instability); 5000-step error ~7e-2")
def test_fp16_storage_against_fp64_at_5000_steps(report):
    if 5000 not in report:
        pytest.skip("no snapshot at step 5000")
    _check(report[5000], 5000)
