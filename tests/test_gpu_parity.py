"""GPU parity: the HIP timestep through the C-ABI against the CPU oracle.

Tolerances: FP32 plain path within 1e-5 relative L-infinity on rho, u and
phi (normalised by the field's max magnitude); the
binary16-storage form within 2e-3, since one differently-rounded product
can move a stored population by one binary16 ulp.
"""
import numpy as np
import pytest

import phasefieldlbm_amd as pf
from phasefieldlbm_amd import D3Q19, D3Q27, NOISE_GAUSSIAN_XY, NOISE_NONE, NOISE_UNIFORM_Z, STORE_FP16, presets
import oracle

pytestmark = pytest.mark.gpu

I = (slice(1, -1),) * 3
CORE = ("rho", "ux", "uy", "uz", "phi")


def rel_linf(a, b):
    scale = max(float(np.abs(b).max()), 1e-30)
    return float(np.abs(a.astype(np.float64) - b).max() / scale)


def compare(cfg, steps, prec=32, fields=CORE, tol=1e-5, region=I):
    print(f"\n[parity] Q={cfg.velocity_set} case={cfg.flow_case} grid={cfg.nx}x{cfg.ny}x{cfg.nz} "
          f"storage={cfg.storage} noise={cfg.noise} steps={steps}", flush=True)
    with pf.Simulation(cfg) as sim:
        sim.init()
        sim.step(steps, save=True)
        got = {k: sim.field(k) for k in fields}
    o = oracle.Oracle(cfg, prec).init().step(steps)
    errs = {k: rel_linf(got[k][region], o.field(k)[region]) for k in fields}
    print("  errors", errs, flush=True)
    for k, e in errs.items():
        assert e <= tol, (k, e)
    return errs


@pytest.mark.parametrize("Q", [D3Q19, D3Q27])
def test_droplet_fp32(Q):
    compare(presets.static_droplet(n=32, radius=8.0, velocity_set=Q), 200)


@pytest.mark.parametrize("steps", [1, 2, 7])
def test_droplet_all_fields_short(steps):
    """Every reference field, including normals and forces, after 1/2/7 steps
    (both A-A step kinds and an odd count)."""
    compare(presets.static_droplet(n=20, radius=5.0, velocity_set=D3Q27), steps, fields=pf.FIELDS)


@pytest.mark.parametrize("Q", [D3Q19, D3Q27])
@pytest.mark.parametrize("noise", [NOISE_NONE, NOISE_UNIFORM_Z])
def test_jet_fp32(Q, noise):
    compare(presets.round_jet(24, 20, 48, 6, Q, noise=noise), 120)


@pytest.mark.parametrize("Q", [D3Q19, D3Q27])
def test_jet_all_fields_with_ghosts(Q):
    """Whole arrays including the ghost shell written by the boundary
    conditions (nozzle plane, zero-gradient outflow plane)."""
    cfg = presets.round_jet(20, 24, 40, 8, Q, noise=NOISE_UNIFORM_Z)
    compare(cfg, 33, fields=pf.FIELDS, region=(slice(None),) * 3)


def test_jet_gaussian_noise():
    """Reference Box-Muller inflow noise: device logf/cosf are not the C
    library's, so this path is held to 1e-4."""
    compare(presets.round_jet(24, 24, 48, 8, D3Q19, noise=NOISE_GAUSSIAN_XY), 60, tol=1e-4)


def test_reference_jet_geometry():
    """The reference's own JET build (64x64x128, D=10, Gaussian noise)."""
    compare(presets.reference_jet(D3Q27), 30, tol=1e-4)


def test_reference_droplet_geometry():
    """The reference's own DROPLET build (128^3, R=35 with x halved, sigma=0.03)."""
    compare(presets.reference_droplet(D3Q19), 20)


@pytest.mark.parametrize("Q", [D3Q19, D3Q27])
def test_fp16_storage_against_fp16_oracle(Q):
    compare(presets.static_droplet(n=24, radius=6.0, velocity_set=Q, storage=STORE_FP16), 100, prec=16,
            tol=2e-3)
    compare(presets.round_jet(20, 20, 40, 6, Q, storage=STORE_FP16, noise=NOISE_UNIFORM_Z), 60, prec=16,
            tol=2e-3)


@pytest.mark.parametrize("steps", [4, 5])
@pytest.mark.parametrize("case", ["droplet", "jet"])
def test_populations(steps, case):
    """Post-streaming populations in the reference layout after an even and
    an odd number of in-place steps."""
    cfg = (presets.static_droplet(n=16, radius=4.0, velocity_set=D3Q27) if case == "droplet"
           else presets.round_jet(16, 16, 32, 6, D3Q27, noise=NOISE_UNIFORM_Z))
    with pf.Simulation(cfg) as sim:
        sim.init().step(steps)
        f, g = sim.populations("f"), sim.populations("g")
    o = oracle.Oracle(cfg, 32).init().step(steps)
    J = (slice(None),) + I
    assert rel_linf(f[J], o.populations("f")[J]) < 1e-5
    assert rel_linf(g[J], o.populations("g")[J]) < 1e-5


def test_phase_mass_and_diagnostics():
    cfg = presets.static_droplet(n=32, radius=8.0)
    with pf.Simulation(cfg) as sim:
        sim.init()
        m0 = sim.phase_mass()
        sim.step(300, save=True)
        m1 = sim.phase_mass()
        d = sim.diagnostics()
        rho, phi = sim.field("rho")[I], sim.field("phi")[I]
        u = np.stack([sim.field(k)[I] for k in ("ux", "uy", "uz")]).astype(np.float64)
    o = oracle.Oracle(cfg, 64).init()
    want0 = o.field("phi")[I].sum()
    assert abs(m0 - want0) / want0 < 1e-6
    assert abs(m1 - m0) / m0 < 1e-6  # FP32 phase conservation
    assert abs(d["mass"] - rho.astype(np.float64).sum()) < 1e-6 * rho.size
    assert abs(d["phi_mass"] - phi.astype(np.float64).sum()) < 1e-9 * phi.size
    np.testing.assert_allclose(d["momentum"], (rho * u).sum(axis=(1, 2, 3)), atol=1e-9 * rho.size)
    assert abs(d["umax"] - np.sqrt((u ** 2).sum(0)).max()) < 1e-9


def test_step_chunking_is_invisible():
    """pf_step(t0, n) equals n calls of pf_step(t, 1) (noise uses t)."""
    cfg = presets.round_jet(16, 16, 32, 6, D3Q19, noise=NOISE_UNIFORM_Z)
    with pf.Simulation(cfg) as a, pf.Simulation(cfg) as b:
        a.init().step(9, save=True)
        b.init()
        for _ in range(8):
            b.step(1)
        b.step(1, save=True)
        for k in CORE:
            np.testing.assert_array_equal(a.field(k), b.field(k))


@pytest.mark.parametrize("zchunk", [1, 3, 64])
def test_zchunk_invariance(zchunk):
    """The z-marching tile length changes the schedule, not the result."""
    cfg = presets.round_jet(24, 24, 40, 6, D3Q27, noise=NOISE_UNIFORM_Z)
    ref = presets.round_jet(24, 24, 40, 6, D3Q27, noise=NOISE_UNIFORM_Z)
    cfg.zchunk = zchunk
    ref.zchunk = 0
    with pf.Simulation(cfg) as a, pf.Simulation(ref) as b:
        a.init().step(10, save=True)
        b.init().step(10, save=True)
        for k in CORE:
            np.testing.assert_array_equal(a.field(k), b.field(k))


def test_state_errors():
    cfg = presets.static_droplet(n=16)
    with pf.Simulation(cfg) as sim:
        with pytest.raises(pf.PfError) as e:
            sim.step(1)
        assert e.value.status == -7
        sim.init()
        with pytest.raises(pf.PfError) as e:
            sim._check(sim._lib.pf_get_field(sim._h, 99, None, 0), "bad id")
        assert e.value.status == -1
    mcfg = presets.static_droplet(n=16)
    mcfg.closure = pf.CLOSURE_MLP
    with pf.Simulation(mcfg) as sim:
        sim.init()
        with pytest.raises(pf.PfError) as e:
            sim.step(1)
        assert e.value.status == -7
