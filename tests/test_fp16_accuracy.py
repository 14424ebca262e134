"""Accuracy of the binary16 storage forms against the FP64 oracle.

Workload: D3Q27 round jet on 40x40x128 cells, D = 8, u = 0.05, Re = 5000,
We = 500, uniform +-5 % u_z inflow noise: the 128x128x512 / D = 16 accuracy
case at a quarter of the resolution.  The committed FP64 fixture
tests/golden/fp64_jet_d3q27_40x40x128.npz (written by
tests/golden/make_fp64_jet.py, 6 minutes of CPU) holds the FP64 oracle's
fields at 500, 1000, 2000 and 5000 steps; the library runs the jet with
FP32, binary16 and binary16 non-equilibrium storage and the error over
interior cells is

    rel_l2(u)   = ||u - u64||_2 / ||u64||_2       (vector field)
    rel_l2(phi) = ||phi - phi64||_2 / ||phi64||_2
    rel_l2(rho) = ||rho - rho64||_2 / ||rho64||_2, and the same of the
                  deviation rho - 1 (rho itself is 1 to ~1e-3)
    rel_linf    = max |a - a64| / max |a64|
    phase-mass drift (sum phi - sum phi64) / sum phi64

FP32 storage is the floor set by round-off growth in a jet at Re = 5000
(measured rel_l2 u / phi: 6e-6 / 3e-6 at 500 steps, 3.0e-4 / 2.8e-4 at
5000).  The plain binary16 form holds max(1e-3, 3x that floor) up to 1000
steps (5.5e-4 / 3.1e-4); later the jet's shear-layer instability amplifies
its rounding until the flow decorrelates from the FP64 trajectory (3.1e-2 /
2.7e-2 at 5000 steps).  The non-equilibrium form stores only what the
writer's second-order equilibrium leaves, and stays 3-13x closer: 8.1e-5 /
4.7e-5 at 1000 steps, 2.2e-4 / 5.5e-4 at 2000.  Past 2000 steps the
comparison measures decorrelation more than precision: two builds of the
non-equilibrium form that differ only in the float rounding of its shift
measured 1.6e-3 / 2.4e-3 and 6.8e-3 / 5.8e-3 at 5000 steps (and 6.9e-5 /
8.1e-5 on u at 1000), so the 5000-step bounds are loose.

The 128x128x512 case itself is compared by tests/fp16_accuracy_report.py
against FP64 snapshots from tests/fp64_snapshots.py (about 5 h of CPU and
250 MB, so not a committed fixture).
"""
import json
import os

import numpy as np
import pytest

FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "fp64_jet_d3q27_40x40x128.npz")


def test_metrics_definition():
    """The error measures on hand-made fields (no GPU): interior cells only,
    phi scaled by 1.001 gives 1e-3 in rel_l2, rel_linf and the mass drift,
    and a change in the ghost shell changes nothing."""
    rng = np.random.default_rng(0)
    shape = (6, 5, 4)
    I = (slice(1, -1),) * 3
    got = {k: rng.normal(size=shape).astype(np.float32) for k in ("ux", "uy", "uz")}
    got["phi"] = rng.uniform(0.1, 1.0, size=shape).astype(np.float32)
    fix = {f"{k}_7": v[I].copy() for k, v in got.items()}
    fix["sum_phi_7"] = np.float64(got["phi"][I].astype(np.float64).sum())
    drho = (1e-3 * rng.normal(size=shape)).astype(np.float64)
    got["rho"] = 1.0 + drho
    fix["drho_7"] = drho[I].copy()
    m0 = _reduced_metrics(got, fix, 7)
    assert all(abs(v) < 1e-12 for v in m0.values()), m0
    got["phi"] = got["phi"] * np.float32(1.001)
    got["ux"][0] += 5.0  # ghost shell
    got["rho"] = 1.0 + 1.001 * drho
    m = _reduced_metrics(got, fix, 7)
    for k in ("rel_l2_phi", "rel_linf_phi", "phase_mass_drift", "rel_l2_drho"):
        assert abs(m[k] - 1e-3) < 1e-6, (k, m)
    assert m["rel_l2_u"] == 0.0
    assert m["rel_l2_rho"] < 2e-6  # rho itself is 1 to ~1e-3


def _check(r, s):
    for k in ("rel_l2_u", "rel_l2_phi"):
        assert np.isfinite(r["fp16"][k])
        assert r["fp16"][k] <= max(1e-3, 3.0 * r["fp32"][k]), (s, k, r)


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
    m = {
        "rel_l2_u": float(np.linalg.norm(u - u64) / np.linalg.norm(u64)),
        "rel_linf_u": float(np.abs(u - u64).max() / np.abs(u64).max()),
        "rel_l2_phi": float(np.linalg.norm(ph - ph64) / np.linalg.norm(ph64)),
        "rel_linf_phi": float(np.abs(ph - ph64).max() / np.abs(ph64).max()),
        "phase_mass_drift": float((ph.sum() - float(fix[f"sum_phi_{s}"])) / float(fix[f"sum_phi_{s}"])),
    }
    if "rho" in got and f"drho_{s}" in fix:
        d, d64 = got["rho"][I].astype(np.float64) - 1.0, fix[f"drho_{s}"].astype(np.float64)
        m["rel_l2_rho"] = float(np.linalg.norm(d - d64) / np.linalg.norm(d64 + 1.0))
        m["rel_linf_rho"] = float(np.abs(d - d64).max() / np.abs(d64 + 1.0).max())
        m["rel_l2_drho"] = float(np.linalg.norm(d - d64) / np.linalg.norm(d64))
    return m


@pytest.mark.slow
def test_reduced_fixture_matches_oracle():
    """The fixture is the oracle's: the FP32 oracle follows it to FP32 round-off
    growth after 500 steps (no GPU)."""
    import oracle
    fix = np.load(FIXTURE)
    assert list(fix["steps"]) == [500, 1000, 2000, 5000]
    o = oracle.Oracle(_reduced_config(), 32).init().step(500)
    got = {k: o.field(k) for k in ("ux", "uy", "uz", "phi")}
    m = _reduced_metrics(got, fix, 500)
    print(m)
    assert m["rel_l2_u"] < 1e-4 and m["rel_l2_phi"] < 1e-4
    # the jet has developed: phi = 1 fluid well inside the domain
    for s in (2000, 5000):
        assert fix[f"phi_{s}"][40:].max() > 0.9 and np.abs(fix[f"uz_{s}"]).max() > 0.04
    # the phase mass grows by the inflow: sum phi increases from snapshot to snapshot
    sums = [float(fix[f"sum_phi_{s}"]) for s in fix["steps"]]
    assert all(b > a for a, b in zip(sums, sums[1:])), sums


@pytest.fixture(scope="module")
def reduced_report():
    import phasefieldlbm_amd as pf
    fix = np.load(FIXTURE)
    steps = [int(s) for s in fix["steps"]]
    rep = {}
    for name, storage in (("fp16", pf.STORE_FP16), ("fp16_ne", pf.STORE_FP16_NE), ("fp32", pf.STORE_FP32),
                          ("mom", pf.STORE_MOM)):
        with pf.Simulation(_reduced_config(storage)) as sim:
            sim.init()
            done = 0
            for s in steps:
                sim.step(s - done, save=True)
                done = s
                got = {k: sim.field(k) for k in ("ux", "uy", "uz", "phi", "rho")}
                rep.setdefault(s, {})[name] = _reduced_metrics(got, fix, s)
    print(json.dumps({"fp16_accuracy_vs_fp64_40x40x128": rep}), flush=True)
    return rep


@pytest.mark.gpu
def test_fp16_reduced_jet_against_fp64(reduced_report):
    """FP32 storage within 1e-3 of FP64 on u and phi at every snapshot (to
    5000 steps), and binary16 storage within max(1e-3, 3 x that floor) up
    to 1000 steps; every form keeps the phase mass within 1e-4 of FP64 up
    to 2000 steps."""
    for s, r in reduced_report.items():
        assert r["fp32"]["rel_l2_u"] < 1e-3 and r["fp32"]["rel_l2_phi"] < 1e-3, (s, r)
        if s <= 1000:
            _check(r, s)
        if s <= 2000:
            assert all(abs(v["phase_mass_drift"]) < 1e-4 for v in r.values()), (s, r)


@pytest.mark.gpu
def test_fp16_ne_reduced_jet_against_fp64(reduced_report):
    """The binary16 non-equilibrium form on the reduced jet: within 3e-4 of
    FP64 on u and phi up to 1000 steps (measured 5.2e-5 / 2.2e-5 at 500,
    8.1e-5 / 4.7e-5 at 1000), at least 3 times closer to FP64 than the
    plain binary16 form on u and phi up to 2000 steps (measured 3.8-13x),
    and at 5000 steps, where both have partly decorrelated, within 2e-2 and
    closer than the plain form (measured 6.8e-3 / 5.8e-3 against 3.1e-2 /
    2.7e-2)."""
    for s in (500, 1000):
        r = reduced_report[s]["fp16_ne"]
        assert r["rel_l2_u"] < 3e-4 and r["rel_l2_phi"] < 3e-4, (s, r)
    for s, r in reduced_report.items():
        for k in ("rel_l2_u", "rel_l2_phi"):
            if s <= 2000:
                assert r["fp16_ne"][k] * 3 < r["fp16"][k], (s, k, r)
            else:
                assert r["fp16_ne"][k] < min(2e-2, r["fp16"][k]), (s, k, r)


@pytest.mark.gpu
def test_mom_reduced_jet_against_fp64(reduced_report):
    """The FP32 moment storage form on the reduced jet: within the FP32 bound
    (1e-3 of FP64 on u and phi) at every snapshot up to 5000 steps, and the
    phase mass within 1e-4 of FP64 up to 2000 steps, as FP32 population
    storage."""
    for s, r in reduced_report.items():
        m = r["mom"]
        assert m["rel_l2_u"] < 1e-3 and m["rel_l2_phi"] < 1e-3, (s, r)
        if s <= 2000:
            assert abs(m["phase_mass_drift"]) < 1e-4, (s, r)
