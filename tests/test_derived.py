"""Derived fields: running means, Reynolds moments and vorticity
(src/derivedFields/operators/timeAverage.cuh:34-88, reynoldsMoments.cuh:34-104,
vorticityFields.cuh:34-100).

The device results are compared with the definitions evaluated in numpy on
the same run's per-step fields: a_t = a_{t-1} + (v_t - a_{t-1}) / t in FP32
over every cell (ghosts included) for the means, central differences at
interior cells for the vorticity.
"""
import numpy as np
import pytest

import phasefieldlbm_amd as pf
from phasefieldlbm_amd import presets


def _vorticity(ux, uy, uz):
    h = np.float32(0.5)
    w = np.zeros((3,) + ux.shape, np.float32)
    I = (slice(1, -1),) * 3

    def d(a, axis):  # central difference along numpy axis (z=0, y=1, x=2)
        sl_p = [slice(1, -1)] * 3
        sl_m = [slice(1, -1)] * 3
        sl_p[axis] = slice(2, None)
        sl_m[axis] = slice(None, -2)
        return h * (a[tuple(sl_p)] - a[tuple(sl_m)])

    w[0][I] = d(uz, 1) - d(uy, 0)
    w[1][I] = d(ux, 0) - d(uz, 2)
    w[2][I] = d(uy, 2) - d(ux, 1)
    mag = np.sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2])
    return w, mag


def test_vorticity_reference_formula():
    """The numpy vorticity used below is the curl: exact on a linear field
    u = A x (vorticity = the axial vector of A - A^T)."""
    n = 6
    z, y, x = np.meshgrid(*(np.arange(n, dtype=np.float32),) * 3, indexing="ij")
    A = np.array([[0.1, 0.3, -0.2], [0.05, 0.0, 0.4], [-0.3, 0.2, 0.1]], np.float32)
    ux = A[0, 0] * x + A[0, 1] * y + A[0, 2] * z
    uy = A[1, 0] * x + A[1, 1] * y + A[1, 2] * z
    uz = A[2, 0] * x + A[2, 1] * y + A[2, 2] * z
    w, _ = _vorticity(ux, uy, uz)
    want = [A[2, 1] - A[1, 2], A[0, 2] - A[2, 0], A[1, 0] - A[0, 1]]
    for k in range(3):
        np.testing.assert_allclose(w[k][1:-1, 1:-1, 1:-1], want[k], atol=1e-5)
        assert np.all(w[k][0] == 0) and np.all(w[k][:, :, -1] == 0)


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["jet", "droplet", "jet_moments"])
def test_running_means_and_vorticity(case):
    if case == "jet":
        cfg = presets.round_jet(24, 24, 48, 6, pf.D3Q19)
    elif case == "jet_moments":  # the moment storage form materialises every field each step too
        cfg = presets.round_jet(24, 24, 48, 6, pf.D3Q27, storage=pf.STORE_MOM)
    else:
        cfg = presets.static_droplet(n=20, radius=5.0)
    steps = 12
    ref = {}
    with pf.Simulation(cfg) as sim:  # per-step fields, derived fields off
        sim.init()
        acc = {k: np.zeros(cfg.shape, np.float32) for k in pf.DERIVED_FIELDS[:10]}
        for t in range(steps):
            sim.step(1, save=True, t0=t)
            ux, uy, uz, phi = (sim.field(k) for k in ("ux", "uy", "uz", "phi"))
            vals = [phi, ux, uy, uz, ux * ux, uy * uy, uz * uz, ux * uy, ux * uz, uy * uz]
            tt = np.float32(t + 1)
            for k, v in zip(pf.DERIVED_FIELDS[:10], vals):
                acc[k] = acc[k] + (v - acc[k]) / tt
        ref.update(acc)
        w, mag = _vorticity(ux, uy, uz)
    cfg.derived = pf.DERIVED_AVERAGE | pf.DERIVED_REYNOLDS | pf.DERIVED_VORTICITY
    with pf.Simulation(cfg) as sim:
        sim.init().step(steps, t0=0)
        for k in pf.DERIVED_FIELDS[:10]:
            got = sim.field(k)
            scale = max(np.abs(ref[k]).max(), 1e-30)
            assert np.abs(got - ref[k]).max() <= 1e-6 * scale, k
        for k, want in zip(("vort_x", "vort_y", "vort_z"), w):
            np.testing.assert_allclose(sim.field(k), want, rtol=0, atol=1e-6 * max(np.abs(want).max(), 1e-30))
        np.testing.assert_allclose(sim.field("vort_mag"), mag, rtol=1e-6, atol=1e-9)


@pytest.mark.gpu
def test_derived_not_enabled():
    with pf.Simulation(presets.static_droplet(n=12)) as sim:
        sim.init().step(1, save=True)
        with pytest.raises(pf.PfError) as e:
            sim.field("avg_ux")
        assert e.value.status == -1
    cfg = presets.static_droplet(n=12)
    cfg.derived = 8
    with pytest.raises(pf.PfError):
        pf.Simulation(cfg)
