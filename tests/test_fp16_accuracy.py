"""Accuracy of binary16 population storage against the FP64 oracle.

Two workloads, both D3Q27 round jets (u = 0.05, Re = 5000, We = 500,
uniform +-5 % u_z inflow noise):
  * 40x40x128, D = 8: the committed FP64 fixture
    tests/golden/fp64_jet_d3q27_40x40x128.npz (written by
    tests/golden/make_fp64_jet.py, a few minutes of CPU) at 500, 1000 and
    2000 steps; these tests always run on the GPU;
  * 128x128x512, D = 16, the benchmark's accuracy case: FP64 snapshots from
    tests/fp64_snapshots.py (about 5 h of CPU for 5000 steps), run only when
    PF_FP64_SNAPSHOTS names a directory holding them.
For every snapshot step the library runs with FP16 and with FP32 storage,
and the error over interior cells is

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
FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "fp64_jet_d3q27_40x40x128.npz")


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


# ---------------------------------------------------------------------------
# The reduced jet of the committed fixture
# ---------------------------------------------------------------------------

def _reduced_config(storage=0):
    from phasefieldlbm_amd import D3Q27, presets
    return presets.round_jet(40, 40, 128, 8, D3Q27, storage=storage)


def _reduced_metrics(got, fix, s):
    """Errors of interior fields `got` (ux, uy, uz, phi; ghost shell included
    in the arrays) against the fixture at step s."""
    I = (slice(1, -1),) * 3
    u = np.stack([got[k][I] for k in ("ux", "uy", "uz")]).astype(np.float64)
    u64 = np.stack([fix[f"{k}_{s}"] for k in ("ux", "uy", "uz")]).astype(np.float64)
    ph, ph64 = got["phi"][I].astype(np.float64), fix[f"phi_{s}"].astype(np.float64)
    return {
        "rel_l2_u": float(np.linalg.norm(u - u64) / np.linalg.norm(u64)),
        "rel_linf_u": float(np.abs(u - u64).max() / np.abs(u64).max()),
        "rel_l2_phi": float(np.linalg.norm(ph - ph64) / np.linalg.norm(ph64)),
        "rel_linf_phi": float(np.abs(ph - ph64).max() / np.abs(ph64).max()),
        "phase_mass_drift": float((ph.sum() - float(fix[f"sum_phi_{s}"])) / float(fix[f"sum_phi_{s}"])),
    }


@pytest.mark.slow
def test_reduced_fixture_matches_oracle():
    """The fixture is the oracle's: the FP32 oracle follows it to FP32 round-off
    growth after 500 steps (no GPU)."""
    import oracle
    fix = np.load(FIXTURE)
    assert list(fix["steps"]) == [500, 1000, 2000]
    o = oracle.Oracle(_reduced_config(), 32).init().step(500)
    got = {k: o.field(k) for k in ("ux", "uy", "uz", "phi")}
    m = _reduced_metrics(got, fix, 500)
    print(m)
    assert m["rel_l2_u"] < 1e-4 and m["rel_l2_phi"] < 1e-4
    # the jet has developed: phi = 1 fluid well inside the domain
    assert fix["phi_2000"][40:].max() > 0.9 and np.abs(fix["uz_2000"]).max() > 0.04


@pytest.fixture(scope="module")
def reduced_report():
    import phasefieldlbm_amd as pf
    fix = np.load(FIXTURE)
    steps = [int(s) for s in fix["steps"]]
    rep = {}
    for name, storage in (("fp16", pf.STORE_FP16), ("fp16_ne", pf.STORE_FP16_NE), ("fp32", pf.STORE_FP32)):
        with pf.Simulation(_reduced_config(storage)) as sim:
            sim.init()
            done = 0
            for s in steps:
                sim.step(s - done, save=True)
                done = s
                got = {k: sim.field(k) for k in ("ux", "uy", "uz", "phi")}
                rep.setdefault(s, {})[name] = _reduced_metrics(got, fix, s)
    print(json.dumps({"fp16_accuracy_vs_fp64_40x40x128": rep}), flush=True)
    return rep


@pytest.mark.gpu
def test_fp16_reduced_jet_against_fp64(reduced_report):
    """Binary16 storage on the reduced jet against the committed FP64 fixture:
    within max(1e-3, 3 x the FP32 floor) on u and phi up to 1000 steps, and
    FP32 storage itself within 1e-3."""
    for s in (500, 1000):
        r = reduced_report[s]
        assert r["fp32"]["rel_l2_u"] < 1e-3 and r["fp32"]["rel_l2_phi"] < 1e-3, (s, r)
        _check(r, s)


@pytest.mark.gpu
def test_fp16_ne_reduced_jet_against_fp64(reduced_report):
    """The binary16 non-equilibrium form on the reduced jet: within 3e-4 of
    FP64 on u and phi up to 1000 steps (measured 5e-5 / 2e-5 at 500,
    7e-5 / 3e-5 at 1000), and at least 3 times closer to FP64 than the
    plain binary16 form on u at every snapshot."""
    for s in (500, 1000):
        r = reduced_report[s]["fp16_ne"]
        assert r["rel_l2_u"] < 3e-4 and r["rel_l2_phi"] < 3e-4, (s, r)
    for s, r in reduced_report.items():
        assert r["fp16_ne"]["rel_l2_u"] * 3 < r["fp16"]["rel_l2_u"], (s, r)
