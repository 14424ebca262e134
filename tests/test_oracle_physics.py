"""Invariants of the full oracle timestep (no reference outputs exist).

Each test fixes a property the mathematics of the reference scheme
guarantees: phase and mass conservation, the quiescent fixed point, mirror
symmetry of the centred jet, the jet boundary planes, and the Laplace law.
"""
import numpy as np
import pytest

from phasefieldlbm_amd import NOISE_NONE, NOISE_UNIFORM_Z, D3Q19, D3Q27, presets
import oracle

I = (slice(1, -1),) * 3


@pytest.mark.parametrize("Q", [19, 27])
def test_droplet_conservation_fp64(Q):
    cfg = presets.static_droplet(n=24, radius=6.0, velocity_set=Q)
    o = oracle.Oracle(cfg, 64).init()
    phi0 = o.field("phi")[I].sum()
    o.step(60)
    # sum phi after step n is computed from the populations of step n
    phi1 = o.field("phi")[I].sum()
    assert abs(phi1 - phi0) / phi0 < 1e-12
    f = o.populations("f")[(slice(None),) + I]
    mass = (f.sum(0) + 1.0).sum()
    assert abs(mass / f[0].size - 1.0) < 1e-12
    mom = [(o.field(k)[I] * o.field("rho")[I]).sum() for k in ("ux", "uy", "uz")]
    assert np.abs(mom).max() < 1e-12


def test_quiescent_fixed_point():
    """Uniform phi (ghost shell included: the gradient stencil reads the
    ghost values unwrapped), no flow: normals, force and f - w stay zero."""
    cfg = presets.static_droplet(n=12, radius=0.0)
    o = oracle.Oracle(cfg, 32)
    o.set_field("rho", np.ones(cfg.cells))
    o.set_field("phi", np.full(cfg.shape, 0.5))
    o.init_distros()
    o.step(5)
    f = o.populations("f")[(slice(None),) + I]
    assert np.all(f == 0.0)
    assert np.all(o.field("rho")[I] == 1.0)
    g = o.populations("g")[(slice(None),) + I]
    assert np.allclose(g.sum(0), 0.5, atol=0, rtol=1e-7)


@pytest.mark.parametrize("Q", [19, 27])
def test_fp32_tracks_fp64(Q):
    cfg = presets.static_droplet(n=24, radius=6.0, velocity_set=Q)
    a = oracle.Oracle(cfg, 32).init().step(100)
    b = oracle.Oracle(cfg, 64).init().step(100)
    for k in ("rho", "phi"):
        d = np.abs(a.field(k) - b.field(k)).max() / np.abs(b.field(k)).max()
        assert d < 1e-5, k
    u = np.stack([a.field(k) for k in ("ux", "uy", "uz")])
    v = np.stack([b.field(k) for k in ("ux", "uy", "uz")])
    assert np.abs(u - v).max() / np.abs(v).max() < 1e-3


@pytest.mark.parametrize("Q", [19, 27])
def test_jet_mirror_symmetry(Q):
    """Centred nozzle, no noise: the flow is symmetric under x -> nx-1-x."""
    cfg = presets.round_jet(20, 20, 40, 6, Q, noise=NOISE_NONE)
    o = oracle.Oracle(cfg, 64).init().step(40)
    uz = o.field("uz")[I]
    ux = o.field("ux")[I]
    assert np.abs(uz).max() > 1e-3
    assert np.abs(uz - uz[:, :, ::-1]).max() < 1e-12
    assert np.abs(ux + ux[:, :, ::-1]).max() < 1e-12
    assert np.abs(uz - uz[:, ::-1, :]).max() < 1e-12


def test_jet_boundary_planes():
    cfg = presets.round_jet(20, 20, 40, 6, D3Q19, noise=NOISE_UNIFORM_Z, noise_amp=0.05)
    o = oracle.Oracle(cfg, 32).init().step(30)
    uz, phi, rho = o.field("uz"), o.field("phi"), o.field("rho")
    # nozzle cells at z = 0: phi = 1, rho = 1, |u_z/u - 1| <= 5 %
    yy, xx = np.mgrid[0:20, 0:20]
    noz = (xx - 9.5) ** 2 + (yy - 9.5) ** 2 <= 9
    assert np.all(phi[0][noz] == 1.0) and np.all(rho[0][noz] == 1.0)
    r = uz[0][noz] / 0.05 - 1
    assert np.abs(r).max() <= 0.05 + 1e-6 and r.std() > 0.01
    assert np.all(phi[0][~noz] == 0.0)
    # zero-gradient outflow ghost plane
    for k in ("rho", "phi", "uz"):
        f = o.field(k)
        np.testing.assert_array_equal(f[-1, 1:-1, 1:-1], f[-2, 1:-1, 1:-1])


@pytest.mark.slow
def test_laplace_law_fp64():
    """Delta p = cs^2 (rho_in - rho_out) approaches 2 sigma / R."""
    R, sigma = 8.0, 0.01
    cfg = presets.static_droplet(n=32, radius=R, sigma=sigma)
    o = oracle.Oracle(cfg, 64).init().step(1500)
    rho = o.field("rho")
    z, y, x = np.mgrid[0:32, 0:32, 0:32]
    r = np.sqrt((x - 15.5) ** 2 + (y - 15.5) ** 2 + (z - 15.5) ** 2)
    inside = rho[(r < R - 4) & (r > 0)].mean()
    outside = rho[(r > R + 4) & (x > 0) & (x < 31) & (y > 0) & (y < 31) & (z > 0) & (z < 31)].mean()
    dp = (inside - outside) / 3.0
    err = abs(dp - 2 * sigma / R) / (2 * sigma / R)
    assert err < 0.25, (dp, 2 * sigma / R)
