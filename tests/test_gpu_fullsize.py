"""The benchmark jets at their full sizes, through size-independent
properties (the CPU oracle cannot run them in test time).

Workloads (the reference's JET case, src/constants.cuh:82-109,
src/flowCase/Jet.cuh:48-64, at the benchmark grids):
  D3Q19 128x128x512, D = 16, FP32
  D3Q27 256x256x1024, D = 32, FP32
  D3Q27 256x256x1024, D = 32, binary16 storage
  D3Q27 256x256x1024, D = 32, binary16 non-equilibrium storage
  D3Q27 256x256x1024, D = 32, FP32 moment storage
u = 0.05, Re = 5000, We = 500, uniform +-5 % u_z inflow noise, 200 steps
(the jet front is ~10 cells from the nozzle).

Checked:
  * every field finite; phi in [-1e-3, 1 + 1e-2], rho within 1 +- 1e-2
  * phase-mass balance in FP64: from one step to the next the phase mass
    (pf_phase_mass, the FP64 sum of the populations the next phase sum adds)
    changes by exactly what crosses the two boundary planes:
        + sum over nozzle cells of g_5 = (1 + 4 u_z,in)/8 entering z = 1
        - sum over z = 1 of g_6 = phi/8 - h_z leaving towards z = 0
        + sum over z = nz-2 of g_6 = phi_b (1 - 4 u_z,b)/8 entering
        - sum over z = nz-2 of g_5 = phi/8 + h_z leaving towards z = nz-1
    with h = (4 phi u + gamma phi (1 - phi) n)/8
    (src/BoundaryConditions.cuh:104,168, src/lbm.cu:209-238); lateral
    faces are periodic.  Tolerance 1e-6 of the mass (FP32 round-off of the
    populations).
  * noise off: the centred nozzle makes the flow symmetric under x -> nx-1-x,
    y -> ny-1-y and x <-> y; the FP32 sums run in a fixed direction order,
    so mirrored cells round differently and the bar is 1e-5 u (FP32).  In
    the binary16 form those FP32 differences now and then flip the rounding
    of a stored population by one binary16 ulp, and the bar is 2e-3 u
    (measured 9.5e-4 after 200 steps).
  * z-chunk invariance at the chunk the automatic choice makes for each grid
    (2 for the FP32 256x256x1024 jet, 16 for its binary16 and moment
    forms, 4 for 128x128x512) against 4 or 8-plane chunks, bit for bit.
"""
import numpy as np
import pytest

import phasefieldlbm_amd as pf
from phasefieldlbm_amd import D3Q19, D3Q27, NOISE_NONE, STORE_FP16, STORE_FP16_NE, STORE_FP32, STORE_MOM, presets

pytestmark = pytest.mark.gpu

I = (slice(1, -1),) * 3

CASES = {
    "d3q19_128x128x512": lambda **kw: presets.round_jet(128, 128, 512, 16, D3Q19, **kw),
    "d3q27_256x256x1024": lambda **kw: presets.round_jet(256, 256, 1024, 32, D3Q27, **kw),
    "d3q27_256x256x1024_fp16": lambda **kw: presets.round_jet(256, 256, 1024, 32, D3Q27, storage=STORE_FP16, **kw),
    "d3q27_256x256x1024_fp16_ne": lambda **kw: presets.round_jet(256, 256, 1024, 32, D3Q27, storage=STORE_FP16_NE,
                                                                 **kw),
    "d3q27_256x256x1024_mom": lambda **kw: presets.round_jet(256, 256, 1024, 32, D3Q27, storage=STORE_MOM, **kw),
}


def _gamma(cfg):
    # gamma = 4 (tau_g - 1/2) / W with the float32 values of pf_config
    tau_g, w = np.float32(cfg.tau_g), np.float32(cfg.interface_width)
    return float((np.float32(4) * (np.float32(0.25) * (tau_g - np.float32(0.5))) / w) / np.float32(0.25))


def _boundary_flux(sim, cfg):
    """Net phase crossing the z = 1 and z = nz-2 planes in the step just run
    (saved fields), in FP64."""
    gam = np.float32(_gamma(cfg))
    phi32, uz32, nz32 = sim.field("phi"), sim.field("uz"), sim.field("normz")
    phi, uz = phi32.astype(np.float64), uz32.astype(np.float64)
    J = (slice(1, -1), slice(1, -1))

    def hz(k):
        # the stored flux: formed in FP32, rounded to binary16 in that form
        p = phi32[k][J]
        h = np.float32(0.125) * (np.float32(4) * p * uz32[k][J] + gam * p * (np.float32(1) - p) * nz32[k][J])
        if cfg.storage in (STORE_FP16, STORE_FP16_NE):
            h = h.astype(np.float16)
        return h.astype(np.float64)

    ny, nx = cfg.ny, cfg.nx
    yy, xx = np.mgrid[0:ny, 0:nx]
    c = (nx - 1) / 2
    noz = ((xx - c) ** 2 + (yy - (ny - 1) / 2) ** 2 <= (cfg.diam // 2) ** 2)
    noz[0, :] = noz[-1, :] = False
    noz[:, 0] = noz[:, -1] = False
    g5_in = ((1.0 + 4.0 * uz[0][noz]) / 8.0).sum()  # phi = 1 at the nozzle
    g6_out = (phi[1][J] / 8.0 - hz(1)).sum()
    k = cfg.nz - 2
    phib = phi[k][J]
    g6_in = (phib * (1.0 - 4.0 * uz[k][J]) / 8.0).sum()
    g5_out = (phib / 8.0 + hz(k)).sum()
    return g5_in - g6_out + g6_in - g5_out, g5_in


@pytest.mark.parametrize("case", list(CASES))
def test_fullsize_jet_properties(case):
    cfg = CASES[case]()
    print(f"\n[fullsize] {case} grid={cfg.nx}x{cfg.ny}x{cfg.nz} storage={cfg.storage}", flush=True)
    with pf.Simulation(cfg) as sim:
        sim.init()
        checked = 0
        for stop in (100, 200):
            sim.step(stop - 1 - sim.t)
            m0 = sim.phase_mass()
            sim.step(1, save=True)
            m1 = sim.phase_mass()
            flux, inflow = _boundary_flux(sim, cfg)
            err = abs((m1 - m0) - flux)
            print(f"  step {sim.t}: mass {m1:.6f} dM {m1 - m0:.6f} boundary flux {flux:.6f} "
                  f"(inflow {inflow:.3f}) |diff| {err:.3e}", flush=True)
            assert inflow > 0.1 * np.pi * (cfg.diam / 2) ** 2
            assert err <= 1e-6 * m1 + 1e-5, (sim.t, m1 - m0, flux)
            # a planted slip in the balance (inflow without the 4 u_z term) is far outside it
            assert abs((m1 - m0) - (flux - inflow + inflow / (1 + 4 * 0.05))) > 100 * (1e-6 * m1 + 1e-5)
            checked += 1
        d = sim.diagnostics()
        assert np.isfinite(d["mass"]) and np.isfinite(d["umax"]) and d["umax"] < 0.2
        for k in pf.FIELDS:
            a = sim.field(k)
            assert np.isfinite(a).all(), k
        phi, rho = sim.field("phi")[I], sim.field("rho")[I]
        print(f"  phi [{phi.min():.4g}, {phi.max():.4g}]  rho [{rho.min():.6f}, {rho.max():.6f}]  "
              f"umax {d['umax']:.4g}", flush=True)
        assert phi.min() > -1e-3 and phi.max() < 1.01
        assert np.abs(rho - 1.0).max() < 1e-2
        assert phi.max() > 0.99  # the jet has entered the domain
    assert checked == 2


@pytest.mark.parametrize("case", list(CASES))
def test_fullsize_mirror_symmetry(case):
    cfg = CASES[case](noise=NOISE_NONE)
    with pf.Simulation(cfg) as sim:
        sim.init().step(200, save=True)
        uz, ux, uy, phi = (sim.field(k)[I].astype(np.float64) for k in ("uz", "ux", "uy", "phi"))
    u = cfg.u_inf
    errs = {
        "uz x-mirror": np.abs(uz - uz[:, :, ::-1]).max(),
        "ux x-antisym": np.abs(ux + ux[:, :, ::-1]).max(),
        "uz y-mirror": np.abs(uz - uz[:, ::-1, :]).max(),
        "uy y-antisym": np.abs(uy + uy[:, ::-1, :]).max(),
        "uz transpose": np.abs(uz - uz.transpose(0, 2, 1)).max(),
        "ux-uy transpose": np.abs(ux - uy.transpose(0, 2, 1)).max(),
    }
    print(f"\n[mirror] {case}: " + ", ".join(f"{k} {v / u:.2e}" for k, v in errs.items()), flush=True)
    assert np.abs(uz).max() > 0.5 * u
    bar = 2e-3 if cfg.storage in (STORE_FP16, STORE_FP16_NE) else 1e-5
    for k, v in errs.items():
        assert v <= bar * u, (k, v)
    dphi = np.abs(phi - phi[:, :, ::-1]).max()
    print(f"  phi x-mirror {dphi:.2e}", flush=True)
    assert dphi <= (5e-3 if cfg.storage in (STORE_FP16, STORE_FP16_NE) else 1e-5)


@pytest.mark.parametrize("case,chunk", [("d3q27_256x256x1024", 4), ("d3q27_256x256x1024_fp16", 8),
                                        ("d3q27_256x256x1024_fp16_ne", 8), ("d3q27_256x256x1024_mom", 8),
                                        ("d3q19_128x128x512", 8)])
def test_fullsize_zchunk_invariance(case, chunk):
    ref, alt = CASES[case](), CASES[case]()
    alt.zchunk = chunk
    with pf.Simulation(ref) as a:
        a.init().step(11, save=True)
        want = {k: a.field(k) for k in ("rho", "ux", "uz", "phi", "pxz", "fsz")}
    with pf.Simulation(alt) as b:
        b.init().step(11, save=True)
        for k, v in want.items():
            assert np.array_equal(b.field(k), v), k
