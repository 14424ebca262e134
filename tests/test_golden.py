"""The committed fixtures of tests/golden/ (FP32 oracle trajectories; see
tests/golden/make_golden.py for the cases and why the reference has none).

CPU: the oracle still reproduces them.  GPU: the library reproduces them to
the parity bar, rel L-inf <= 1e-5 per field (velocity scaled by the inflow
speed or the capillary velocity, as in the parity tests)."""
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import make_golden  # noqa: E402

GOLDEN = np.load(os.path.join(HERE, "golden", "oracle_trajectories.npz"))


def _scale(cfg, k, ref):
    import phasefieldlbm_amd as pf
    m = float(np.abs(ref).max())
    if k in ("ux", "uy", "uz"):
        u = cfg.u_inf if cfg.flow_case == pf.JET else cfg.sigma / ((cfg.tau - 0.5) / 3.0)
        return max(m, u)
    return max(m, 1e-30)


@pytest.mark.parametrize("name", sorted(make_golden.cases()))
def test_oracle_reproduces_golden(name):
    import oracle
    cfg, steps = make_golden.cases()[name]
    assert int(GOLDEN[f"{name}/steps"]) == steps
    o = oracle.Oracle(cfg, 32).init().step(steps)
    for k in make_golden.FIELDS:
        want = GOLDEN[f"{name}/{k}"]
        got = o.field(k)
        assert got.shape == want.shape
        assert np.abs(got - want).max() <= 1e-6 * _scale(cfg, k, want), (name, k)
    o.close()


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(make_golden.cases()))
def test_library_reproduces_golden(name):
    import phasefieldlbm_amd as pf
    cfg, steps = make_golden.cases()[name]
    print(f"\n[golden] {name} steps={steps}", flush=True)
    with pf.Simulation(cfg) as sim:
        sim.init().step(steps, save=True)
        got = {k: sim.field(k) for k in make_golden.FIELDS}
    I = (slice(1, -1),) * 3
    for k in make_golden.FIELDS:
        want = GOLDEN[f"{name}/{k}"][I]
        err = np.abs(got[k][I].astype(np.float64) - want).max() / _scale(cfg, k, want)
        assert err <= 1e-5, (name, k, err)
