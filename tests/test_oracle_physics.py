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


@pytest.mark.parametrize("case", ["droplet", "jet"])
def test_phase_flux_identity_fp64(case):
    """The D3Q7 populations carry no non-equilibrium part
    (src/velocitySet/D3Q7.cuh:139-143), so the g that reaches x from its
    neighbour s = x - c_q is phi(s)/8 + c_q.h(s), with
    h = (4 phi u + gamma phi (1 - phi) n) / 8 of that neighbour, and
    g_0 = phi(x)/4.  The library stores h instead of the seven g and rebuilds
    them this way; checked here on the FP64 oracle's own populations, which
    store and stream g as the reference does."""
    if case == "droplet":
        cfg = presets.static_droplet(n=20, radius=5.0, velocity_set=D3Q19)
        steps = 7
    else:
        cfg = presets.round_jet(16, 16, 32, 6, D3Q19, noise=NOISE_UNIFORM_Z)
        steps = 9
    o = oracle.Oracle(cfg, 64).init().step(steps)
    phi = o.field("phi")
    u = [o.field(k) for k in ("ux", "uy", "uz")]
    n = [o.field(k) for k in ("normx", "normy", "normz")]
    # gamma = 4 (tau_g - 1/2) / W, from the float32 values of pf_config
    tau_g, width = float(np.float32(cfg.tau_g)), float(np.float32(cfg.interface_width))
    gamma = (4.0 * (0.25 * (tau_g - 0.5)) / width) / 0.25
    h = [(4.0 * phi * u[a] + gamma * phi * (1.0 - phi) * n[a]) / 8.0 for a in range(3)]
    g = o.populations("g")
    o.close()
    nz, ny, nx = phi.shape
    z0, z1 = (2, nz - 2) if case == "jet" else (1, nz - 1)  # jet: z = 1, nz-2 take boundary g
    z, y, x = np.meshgrid(np.arange(z0, z1), np.arange(1, ny - 1), np.arange(1, nx - 1), indexing="ij")

    def wrap(i, m):  # ghost-shell periodicity over interior cells
        return np.where(i < 1, m - 2, np.where(i > m - 2, 1, i))

    assert np.abs(g[0][z, y, x] - phi[z, y, x] / 4.0).max() < 1e-15
    dirs = [(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)]  # D3Q7 order q = 1..6
    for q, c in enumerate(dirs, start=1):
        sx, sy, sz = wrap(x - c[0], nx), wrap(y - c[1], ny), wrap(z - c[2], nz)
        a = int(np.flatnonzero(c)[0])
        want = phi[sz, sy, sx] / 8.0 + c[a] * h[a][sz, sy, sx]
        assert np.abs(g[q][z, y, x] - want).max() < 1e-15, q
    # and a planted error is seen: the flux of the cell itself instead of the source
    bad = phi[z, y, x] / 8.0 + h[0][z, y, x]
    assert np.abs(g[1][z, y, x] - bad).max() > 1e-6


@pytest.mark.parametrize("case", ["droplet", "jet"])
def test_binary16_flux_mode(case):
    """Precision 17 (binary16 f and phase flux, the library's binary16 form):
    every streamed g equals phi(s)/8 +- h16_k(s) of its source cell s, where
    h16 is h = (4 phi u + gamma phi (1 - phi) n) / 8 formed in FP32 and
    rounded to binary16 (numpy float32 arithmetic, numpy float16 rounding),
    and g_0 = phi/4.  Total phi stays conserved (droplet)."""
    if case == "droplet":
        cfg = presets.static_droplet(n=20, radius=5.0, velocity_set=D3Q27)
        steps = 7
    else:
        cfg = presets.round_jet(16, 16, 32, 6, D3Q27, noise=NOISE_UNIFORM_Z)
        steps = 9
    o = oracle.Oracle(cfg, 17).init()
    m0 = o.field("phi")[I].sum()
    o.step(steps)
    f32 = np.float32
    phi = o.field("phi").astype(f32)
    u = [o.field(k).astype(f32) for k in ("ux", "uy", "uz")]
    n = [o.field(k).astype(f32) for k in ("normx", "normy", "normz")]
    gamma = f32(o.consts()["gamma"])
    sharp = gamma * phi * (f32(1) - phi)
    h16 = [(f32(0.125) * (f32(4) * phi * u[a] + sharp * n[a])).astype(np.float16).astype(np.float64)
           for a in range(3)]
    g = o.populations("g")
    nz, ny, nx = phi.shape
    z0, z1 = (2, nz - 2) if case == "jet" else (1, nz - 1)
    z, y, x = np.meshgrid(np.arange(z0, z1), np.arange(1, ny - 1), np.arange(1, nx - 1), indexing="ij")

    def wrap(i, m):
        return np.where(i < 1, m - 2, np.where(i > m - 2, 1, i))

    phi64 = phi.astype(np.float64)
    assert np.array_equal(g[0][z, y, x], (f32(0.25) * phi[z, y, x]).astype(np.float64))
    dirs = [(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)]
    for q, c in enumerate(dirs, start=1):
        sx, sy, sz = wrap(x - c[0], nx), wrap(y - c[1], ny), wrap(z - c[2], nz)
        a = int(np.flatnonzero(c)[0])
        base = f32(0.125) * phi[sz, sy, sx]
        hv = h16[a][sz, sy, sx].astype(f32)
        want = (base + hv if c[a] > 0 else base - hv).astype(np.float64)
        assert np.array_equal(g[q][z, y, x], want), q
        # the FP32 flux (precision 16's g) differs from it
        h32 = (f32(0.125) * (f32(4) * phi * u[a] + sharp * n[a]))[sz, sy, sx]
        assert np.abs((base + c[a] * h32) - want).max() > 0 or np.abs(h32).max() == 0
    if case == "droplet":
        assert abs(o.field("phi")[I].sum() - m0) / m0 < 1e-6
    # it stays within binary16 round-off of the reference's FP16 mode
    ref = oracle.Oracle(cfg, 16).init().step(steps)
    assert np.abs(ref.field("phi") - phi64).max() < 2e-3


def ne_shift_np(cfg, fields, q, c, w):
    """Shift S_q of the binary16 non-equilibrium form, per interior cell, from
    the moments of the cell that wrote direction q, x - c_q (ghost-shell wrap
    in x, y; in z for the droplet): M = (A, B) = (rho (1 - 1.5|u|^2) - 1,
    3 rho u) rounded to binary16, S_q = w_q fma(b, fma(0.5, b, 1), A),
    b = c_q.B, in float32 in the library's operation order (each fma formed
    in float64, where the float32 product is exact, and rounded once)."""
    f32 = np.float32
    nz, ny, nx = fields["rho"].shape
    z, y, x = np.meshgrid(np.arange(1, nz - 1), np.arange(1, ny - 1), np.arange(1, nx - 1), indexing="ij")

    def wrap(i, m):
        return np.where(i < 1, m - 2, np.where(i > m - 2, 1, i))

    sx, sy = wrap(x - c[q][0], nx), wrap(y - c[q][1], ny)
    sz = z - c[q][2] if cfg.flow_case == 0 else wrap(z - c[q][2], nz)
    rho = fields["rho"][sz, sy, sx].astype(f32)
    u = [fields[k][sz, sy, sx].astype(f32) for k in ("ux", "uy", "uz")]
    v = f32(1) - f32(1.5) * (u[0] * u[0] + u[1] * u[1] + u[2] * u[2])
    r3 = f32(3) * rho
    h = lambda a: a.astype(np.float16).astype(f32)
    A = h(rho * v - f32(1))
    B = [h(r3 * u[a]) for a in range(3)]
    if q == 0:
        return f32(w[q]) * A
    b = None
    for a in range(3):
        if c[q][a] == 0:
            continue
        t = B[a] if c[q][a] > 0 else -B[a]
        b = t if b is None else b + t
    return f32(w[q]) * fma32(b, fma32(f32(0.5), b, f32(1)), A)


def fma32(a, b, c):
    """float32 fma: a*b exact in float64, one rounding of the sum (a double
    rounding differs from the single one only at float64 ties, which the
    callers' tolerances cover)."""
    return (np.float64(a) * np.float64(b) + np.float64(c)).astype(np.float32)


@pytest.mark.parametrize("case", ["droplet", "jet"])
def test_ne_storage_mode(case, oracle_lib):
    """Precision 18 (the library's binary16 non-equilibrium form): every
    population held after a step is S_q(M of its writer) + v / 2^12 with v a
    binary16 value; taking the reader's own moments instead of the writer's
    leaves values off that grid.  It stays within binary16 round-off of the
    FP32 oracle and conserves the droplet's phase."""
    if case == "droplet":
        cfg = presets.static_droplet(n=20, radius=5.0, velocity_set=D3Q27)
        steps = 7
    else:
        cfg = presets.round_jet(16, 16, 32, 6, D3Q27, noise=NOISE_UNIFORM_Z)
        steps = 9
    o = oracle.Oracle(cfg, 18).init()
    m0 = o.field("phi")[I].sum()
    o.step(steps)
    fields = {k: o.field(k) for k in ("rho", "ux", "uy", "uz")}
    f = o.populations("f")
    Q = cfg.velocity_set
    cx = np.zeros(Q, np.int32); cy = np.zeros(Q, np.int32); cz = np.zeros(Q, np.int32); w = np.zeros(Q)
    assert oracle_lib.pfo_lattice(Q, cx.ctypes.data, cy.ctypes.data, cz.ctypes.data, w.ctypes.data) == 0
    c = np.stack([cx, cy, cz], 1)
    nz = cfg.nz
    zlo, zhi = (1, nz - 3) if case == "jet" else (0, nz - 2)  # jet: planes 1 and nz-2 take boundary values
    K = np.float32(4096.0)
    off_own = 0.0
    for q in range(Q):
        S = ne_shift_np(cfg, fields, q, c, w)[zlo:zhi]
        fq = f[q][I][zlo:zhi].astype(np.float32)
        v = (fq - S) * K
        on_grid = v.astype(np.float16).astype(np.float32)
        spacing = np.maximum(np.abs(on_grid) * 2.0 ** -10, 2.0 ** -24)
        assert np.all(np.abs(v - on_grid) <= 1e-3 * spacing + 4.0 * K * np.spacing(np.abs(S) + np.abs(fq))), q
    # a planted slip: the shift of direction +z from the reader's own moments
    q = 5
    fq = f[q][I][zlo:zhi].astype(np.float32)
    f32 = np.float32
    rho = fields["rho"][I][zlo:zhi].astype(f32)
    ux, uy, uz = (fields[k][I][zlo:zhi].astype(f32) for k in ("ux", "uy", "uz"))
    vv = f32(1) - f32(1.5) * (ux * ux + uy * uy + uz * uz)
    A = (rho * vv - f32(1)).astype(np.float16).astype(f32)
    b = (f32(3) * rho * uz).astype(np.float16).astype(f32)
    S_bad = f32(w[q]) * fma32(b, fma32(f32(0.5), b, f32(1)), A)
    v_bad = (fq - S_bad) * K
    off = np.abs(v_bad - v_bad.astype(np.float16).astype(f32)) / np.maximum(np.abs(v_bad) * 2.0 ** -10, 2.0 ** -24)
    assert np.mean(off > 1e-2) > 0.2
    o32 = oracle.Oracle(cfg, 32).init().step(steps)
    assert np.abs(o32.field("rho")[I] - o.field("rho")[I]).max() < 1e-5
    if case == "droplet":
        assert abs(o.field("phi")[I].sum() - m0) / m0 < 1e-6
