"""Closed-form pins of the oracle's constants, relaxation, sponge, initial
condition and inflow noise.

Each quantity is recomputed here in numpy FP64 from its defining formula in
the reference and compared with what the oracle uses; each check also shows
that a plausible slip (a linear ramp, (nz-1)/12 sponge cells, tanh(arg),
the cell-centre at n/2, sin for cos, a gain of 10) would fail it.

Reference formulas:
  nu = u D / Re, tau = 1/2 + 3 nu, omega_from_nu = 1 / (1/2 + 3 nu)
      (src/functions/globalFunctions.cuh:298-393)
  sigma = u^2 D / We; gamma = 4 cs2_g (tau_g - 1/2) / W / cs2_g
      (src/constants.cuh:92-109)
  sponge: nz/12 cells (integer division), z_start = (nz-1-cells)/(nz-1),
      cubic smoothstep s^2 (3 - 2 s), tau_zmax from nu (1 + 100)
      (src/functions/globalFunctions.cuh:260-292,
       src/functions/deviceFunctions.cuh:94-101)
  jet omega: tau_phi = (1-phi) tau_w + phi tau_o,
      tau_eff = tau_phi + r (tau_zmax(phi) - tau_phi), omega = 1/tau_eff
      (src/lbm.cu:158-170)
  droplet: phi = 1/2 + 1/2 tanh(2 (R - L) / W), L the distance from the
      centre (n-1)/2 with x divided by 2 (src/initialConditions.cu:58-83,
      src/functions/globalFunctions.cuh:111-124)
  inflow noise: sqrt(-2 ln max(u1, 1e-12)) cos(2 pi_f u2), sigma_u = 0.08 u
      (src/BoundaryConditions.cuh:60-69,183-213)
"""
import numpy as np
import pytest

import oracle
from phasefieldlbm_amd import D3Q19, D3Q27, NOISE_GAUSSIAN_XY, presets

I = (slice(1, -1),) * 3


def f32(v):
    """The pf_config field as the oracle receives it (a float32)."""
    return float(np.float32(v))


def jet_consts(nz, diam, nx=8, ny=8, re_oil=5000.0, prec=64):
    """Oracle constants of a round jet; nx, ny only set the nozzle centre,
    so the full-size grids are represented by narrow ones."""
    cfg = presets.round_jet(nx, ny, nz, diam, D3Q19)
    cfg.reynolds_oil = re_oil
    o = oracle.Oracle(cfg, prec)
    return cfg, o, o.consts()


# The 128x128x512 (D = 16) and 256x256x1024 (D = 32) jets at Re = 5000, We = 500, the reference's
# own 64x64x128 jet, and nz = 120, where nz/12 and (nz-1)/12 differ.
CASES = [(512, 16, 1.6e-4, 0.50048, 8e-5, 42), (1024, 32, 3.2e-4, 0.50096, 1.6e-4, 85),
         (128, 10, 1.0e-4, 0.5003, 5e-5, 10), (120, 12, 1.2e-4, 0.50036, 6e-5, 10)]


@pytest.mark.parametrize("nz,diam,nu,tau,sigma,cells", CASES)
def test_jet_constants(nz, diam, nu, tau, sigma, cells):
    cfg, o, k = jet_consts(nz, diam)
    u = f32(cfg.u_inf)
    # u_inf enters as a float32 (0.05 -> 0.0500000007), hence 1e-7
    assert k["visc_water"] == pytest.approx(nu, rel=1e-7)
    assert k["visc_oil"] == pytest.approx(nu, rel=1e-7)
    assert k["tau_water"] == pytest.approx(tau, rel=1e-9)
    assert k["sigma"] == pytest.approx(sigma, rel=1e-7)
    assert k["visc_water"] == pytest.approx(u * diam / 5000.0, rel=1e-15)
    assert k["sigma"] == pytest.approx(u * u * diam / 500.0, rel=1e-15)
    assert k["tau_oil"] == pytest.approx(0.5 + 3.0 * k["visc_oil"], rel=1e-15)
    assert k["omco_zmin"] == pytest.approx(1.0 - 1.0 / (0.5 + 3.0 * k["visc_oil"]), rel=1e-13)
    # sponge geometry: nz/12 cells
    got_cells = (1.0 - k["sp_z_start"]) * (nz - 1)
    assert abs(got_cells - cells) < 1e-9 and cells == nz // 12
    assert k["sp_inv_nz_m1"] == pytest.approx(1.0 / (nz - 1), rel=1e-15)
    assert k["sp_inv_sponge"] == pytest.approx((nz - 1) / cells, rel=1e-15)
    # nozzle: centre (n-1)/2 and the integer radius D/2
    assert k["cx"] == (cfg.nx - 1) / 2 and k["cy"] == (cfg.ny - 1) / 2
    assert k["R2"] == (diam // 2) ** 2
    # gamma = 4 (tau_g - 1/2) / W = 0.5 at tau_g = 1, W = 4
    assert k["gamma"] == pytest.approx(0.5, rel=1e-15)
    # planted slips are visible at this precision
    assert abs(k["tau_water"] - (0.5 + k["visc_water"])) > 1e-5        # nu without 3 (= 1/cs^2)
    assert abs(k["sigma"] - u * diam / 500.0) > 1e-4                  # u instead of u^2
    if nz % 12 == 0:
        assert (nz - 1) // 12 != cells                                 # (nz-1)/12 would differ here


@pytest.mark.parametrize("tau_g,width,gamma", [(1.0, 4.0, 0.5), (0.8, 4.0, 0.3), (0.8, 5.0, 0.24)])
def test_sharpening_gamma(tau_g, width, gamma):
    cfg = presets.static_droplet(n=8, radius=2.0, tau_g=tau_g)
    cfg.interface_width = width
    k = oracle.Oracle(cfg, 64).consts()
    assert k["gamma"] == pytest.approx(4.0 * (f32(tau_g) - 0.5) / f32(width), rel=1e-15)
    assert k["gamma"] == pytest.approx(gamma, rel=1e-6)
    assert abs(k["gamma"] - 4.0 * (f32(tau_g) - 0.5) / f32(width) * 0.25) > 0.05  # cs^2 dropped


def test_droplet_relaxation():
    """Droplet: omega = omega_from_nu((tau - 1/2)/3) = 1/tau everywhere."""
    cfg = presets.static_droplet(n=8, radius=2.0, tau=0.55)
    o = oracle.Oracle(cfg, 64)
    k = o.consts()
    assert k["omega_ref"] == pytest.approx(1.0 / f32(0.55), rel=1e-12)
    assert k["sigma"] == pytest.approx(f32(0.01), rel=1e-15)
    for z in (1, 3, 6):
        assert o.omega(0.3, z) == k["omega_ref"]


def smoothstep(s):
    s = np.clip(s, 0.0, 1.0)
    return s * s * (3.0 - 2.0 * s)


@pytest.mark.parametrize("nz", [120, 512, 1024])
def test_sponge_ramp(nz):
    _, o, k = jet_consts(nz, 16 if nz >= 512 else 12)
    cells = nz // 12
    start = nz - 1 - cells
    for z in (1, start // 2, start - 1, start):
        assert o.sponge_ramp(z) == 0.0, z
    assert o.sponge_ramp(nz - 1) == pytest.approx(1.0, abs=1e-15)
    for j in (cells // 4, cells // 2, (3 * cells) // 4, cells - 1):
        s = j / cells
        got = o.sponge_ramp(start + j)
        assert got == pytest.approx(smoothstep(s), rel=1e-12, abs=1e-15), (j, got)
        # a linear ramp (or a quintic smoothstep) would be seen
        assert abs(got - s) > 1e-3 or abs(s - 0.5) < 1e-12
        assert abs(got - s ** 3 * (10 - 15 * s + 6 * s * s)) > 1e-4 or abs(s - 0.5) < 1e-12


@pytest.mark.parametrize("prec", [64, 32])
def test_jet_omega(prec):
    """Per-cell collision frequency of the jet against the closed form, with
    unequal water/oil viscosities so the phase blend of tau is visible."""
    nz, diam = 120, 12
    cfg, o, k = jet_consts(nz, diam, re_oil=2000.0, prec=prec)
    u = f32(cfg.u_inf)
    nu_w, nu_o = u * diam / 5000.0, u * diam / 2000.0
    tau_w, tau_o = 0.5 + 3 * nu_w, 0.5 + 3 * nu_o
    cells = nz // 12
    start = nz - 1 - cells
    tol = 1e-12 if prec == 64 else 2e-6
    for phi in (0.0, 0.3, 1.0):
        tau_phi = (1 - phi) * tau_w + phi * tau_o
        nu_phi = (1 - phi) * nu_w + phi * nu_o
        tau_zmax = 0.5 + 3.0 * nu_phi * (1.0 + 100.0)
        assert o.omco_zmax(phi) == pytest.approx(1.0 - 1.0 / tau_zmax, rel=tol)
        for z in (1, start, start + cells // 3, start + (2 * cells) // 3, nz - 2):
            r = smoothstep((z - start) / cells)
            want = 1.0 / (tau_phi + r * (tau_zmax - tau_phi))
            got = o.omega(phi, z)
            assert got == pytest.approx(want, rel=tol), (phi, z)
            if 0 < r:
                # a sponge gain of 10 or a blend of omega instead of tau is seen
                bad_k = 1.0 / (tau_phi + r * (0.5 + 3.0 * nu_phi * 11.0 - tau_phi))
                assert abs(got - bad_k) > 1e-3 * got
            if 0 < phi < 1 and r == 0:
                bad_blend = (1 - phi) / tau_w + phi / tau_o
                assert abs(got - bad_blend) > 1e-9


@pytest.mark.parametrize("axes", [(2.0, 1.0, 1.0), (1.0, 1.0, 1.0), (1.0, 1.5, 0.8)])
@pytest.mark.parametrize("prec", [64, 32])
def test_droplet_initial_condition(axes, prec):
    """phi = 1/2 + 1/2 tanh(2 (R - L)/W) on interior cells, 0 on the ghost
    shell; L with the coordinate divisors (the reference's x/2 by default)."""
    n, R = 26, 7.0
    cfg = presets.static_droplet(n=n, radius=R)
    cfg.axis_x, cfg.axis_y, cfg.axis_z = axes
    cfg.interface_width = 4.0
    phi = oracle.Oracle(cfg, prec).init().field("phi")
    z, y, x = np.mgrid[0:n, 0:n, 0:n].astype(np.float64)
    c = (n - 1) / 2
    L = np.sqrt(((x - c) / f32(axes[0])) ** 2 + ((y - c) / f32(axes[1])) ** 2 + ((z - c) / f32(axes[2])) ** 2)
    want = 0.5 + 0.5 * np.tanh(2.0 * (R - L) / 4.0)
    tol = 1e-14 if prec == 64 else 2e-6
    assert np.abs(phi[I] - want[I]).max() < tol
    ghost = np.ones(phi.shape, bool)
    ghost[I] = False
    assert np.all(phi[ghost] == 0.0)
    assert phi[I].max() > 0.99 and phi[I].min() < 0.01
    # tanh(arg) for tanh(2 arg), or the centre at n/2, would be seen
    assert np.abs(phi[I] - (0.5 + 0.5 * np.tanh((R - L) / 4.0))[I]).max() > 0.05
    Lb = np.sqrt(((x - n / 2) / f32(axes[0])) ** 2 + ((y - n / 2) / f32(axes[1])) ** 2 + ((z - n / 2) / f32(axes[2])) ** 2)
    assert np.abs(phi[I] - (0.5 + 0.5 * np.tanh(2.0 * (R - Lb) / 4.0))[I]).max() > 0.05


def _lowbias32(x):
    x = np.asarray(x, dtype=np.uint64)
    m = np.uint64(0xFFFFFFFF)
    x = x ^ (x >> np.uint64(16))
    x = (x * np.uint64(0x7FEB352D)) & m
    x = x ^ (x >> np.uint64(15))
    x = (x * np.uint64(0x846CA68B)) & m
    x = x ^ (x >> np.uint64(16))
    return x


def _box_muller(seed, salt, x, y, t):
    m = np.uint64(0xFFFFFFFF)
    base = (np.uint64(seed ^ salt) ^ x.astype(np.uint64) ^ ((y.astype(np.uint64) * np.uint64(0x85EBCA6B)) & m)
            ^ np.uint64((t * 0xC2B2AE35) & 0xFFFFFFFF))
    u1 = (_lowbias32(base).astype(np.float64) + 0.5) * 2.3283064365386963e-10
    u2 = (_lowbias32(base ^ np.uint64(0x68BC21EB)).astype(np.float64) + 0.5) * 2.3283064365386963e-10
    two_pi = 2.0 * float(np.float32(np.pi))  # 2 * CUDART_PI_F
    return np.sqrt(-2.0 * np.log(np.maximum(u1, 1e-12))), two_pi * u2


@pytest.mark.parametrize("t", [0, 7, 123456])
def test_box_muller_inflow(t):
    """The reference's Gaussian inflow noise at the nozzle plane after one
    step at time index t: u_x, u_y = 0.08 u N(x, y, t) with the two salts,
    u_z = u, phi = rho = 1; zero off the nozzle."""
    cfg = presets.round_jet(24, 24, 16, 10, D3Q19, noise=NOISE_GAUSSIAN_XY)
    o64 = oracle.Oracle(cfg, 64).init().step(1, t0=t)
    o32 = oracle.Oracle(cfg, 32).init().step(1, t0=t)
    y, x = np.mgrid[0:cfg.ny, 0:cfg.nx]
    c = (cfg.nx - 1) / 2
    noz = (x - c) ** 2 + (y - c) ** 2 <= (cfg.diam // 2) ** 2
    u = f32(cfg.u_inf)
    sig = 0.08 * u
    comps = {}
    for name, salt in (("ux", 0xA341316C), ("uy", 0xC8013EA4)):
        r, theta = _box_muller(cfg.seed, salt, x[noz], y[noz], t)
        want = sig * r * np.cos(theta)
        got = o64.field(name)[0][noz]
        # libm log/cos: a few ulp
        assert np.abs(got - want).max() < 1e-14 * sig * 8, name
        # FP32 oracle (float log/cos) agrees to float round-off
        assert np.abs(o32.field(name)[0][noz] - want).max() < 2e-6 * sig * 8, name
        # sin for cos, or a missing sqrt, would be seen
        assert np.abs(got - sig * r * np.sin(theta)).max() > 0.1 * sig
        assert np.abs(got - sig * r * r * np.cos(theta)).max() > 0.1 * sig
        comps[name] = got
        assert np.all(o64.field(name)[0][~noz] == 0.0)
    # the two components are different draws
    assert np.abs(comps["ux"] - comps["uy"]).max() > 0.1 * sig
    assert np.all(o64.field("uz")[0][noz] == u)
    assert np.all(o64.field("phi")[0][noz] == 1.0) and np.all(o64.field("rho")[0][noz] == 1.0)
    assert np.all(o64.field("phi")[0][~noz] == 0.0)
