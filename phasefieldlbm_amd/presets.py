"""Run configurations of the benchmark and validation workloads.

The reference hard-codes one jet (64x64x128, D=10) and one droplet
(128^3, R=35) as constexpr values (src/constants.cuh:82-141); these presets
are the same physics at the grids used for validation and measurement.
Relations kept from the reference: nu = u*D/Re, sigma = u^2*D/We
(src/constants.cuh:96-103, src/functions/globalFunctions.cuh:298-305).
"""
from __future__ import annotations

from . import (CLOSURE_FD, D3Q19, D3Q27, DROPLET, JET, NOISE_GAUSSIAN_XY, NOISE_UNIFORM_Z, STORE_FP16,
               STORE_FP16_NE, STORE_FP32, STORE_MOM, Config)


def reference_jet(velocity_set: int = D3Q19) -> Config:
    """The reference's JET build: 64x64x128, D=10, Gaussian inflow noise."""
    return Config(velocity_set=velocity_set, flow_case=JET, nx=64, ny=64, nz=128, diam=10, u_inf=0.05,
                  reynolds_water=5000.0, reynolds_oil=5000.0, weber=500.0, noise=NOISE_GAUSSIAN_XY)


def reference_droplet(velocity_set: int = D3Q19) -> Config:
    """The reference's DROPLET build: 128^3, R=35, x-stretched, sigma=0.03."""
    return Config(velocity_set=velocity_set, flow_case=DROPLET, nx=128, ny=128, nz=128, radius=35.0,
                  axis_x=2.0, axis_y=1.0, axis_z=1.0, sigma=0.03, tau=0.55)


def static_droplet(n: int = 64, radius: float = 12.0, velocity_set: int = D3Q19, sigma: float = 0.01,
                   tau: float = 0.55, tau_g: float = 0.8, storage: int = STORE_FP32) -> Config:
    """Static spherical droplet in a periodic box (64^3, R=12, W=4, tau=0.55,
    sigma=0.01, phase relaxation 0.8 -- which, as in the reference, sets the
    sharpening strength gamma = 4(tau_g - 1/2)/W)."""
    return Config(velocity_set=velocity_set, flow_case=DROPLET, storage=storage, nx=n, ny=n, nz=n,
                  radius=radius, axis_x=1.0, axis_y=1.0, axis_z=1.0, sigma=sigma, tau=tau, tau_g=tau_g,
                  interface_width=4.0)


def round_jet(nx: int, ny: int, nz: int, diam: int, velocity_set: int, storage: int = STORE_FP32,
              noise: int = NOISE_UNIFORM_Z, noise_amp: float = 0.05) -> Config:
    """Round jet along z: nozzle of diameter D at z=0 with u_z=0.05 and phi=1,
    walls elsewhere on z=0, zero-gradient outflow, periodic laterals;
    Re=5000, We=500; per-cell uniform +-5% perturbation on u_z each step."""
    return Config(velocity_set=velocity_set, flow_case=JET, storage=storage, nx=nx, ny=ny, nz=nz, diam=diam,
                  u_inf=0.05, reynolds_water=5000.0, reynolds_oil=5000.0, weber=500.0, noise=noise,
                  noise_amp=noise_amp, closure=CLOSURE_FD)


def jet_d3q19_128x128x512() -> Config:
    return round_jet(128, 128, 512, 16, D3Q19)


def jet_d3q27_256x256x1024(storage: int = STORE_FP32) -> Config:
    return round_jet(256, 256, 1024, 32, D3Q27, storage=storage)


def jet_d3q27_128x128x512(storage: int = STORE_FP32) -> Config:
    """Accuracy geometry of the precision form (128x128x512, D=16, D3Q27)."""
    return round_jet(128, 128, 512, 16, D3Q27, storage=storage)


# Named workloads (steps are the full run lengths).
WORKLOADS = {
    "droplet_d3q19_64": (lambda: static_droplet(), 2000),
    "jet_d3q19_128x128x512": (jet_d3q19_128x128x512, 10000),
    "jet_d3q19_128x128x512_mom": (lambda: round_jet(128, 128, 512, 16, D3Q19, storage=STORE_MOM), 10000),
    "jet_d3q27_256x256x1024": (jet_d3q27_256x256x1024, 20000),
    "jet_d3q27_256x256x1024_fp16": (lambda: jet_d3q27_256x256x1024(STORE_FP16), 20000),
    "jet_d3q27_128x128x512_fp16": (lambda: jet_d3q27_128x128x512(STORE_FP16), 5000),
    "jet_d3q27_256x256x1024_fp16_ne": (lambda: jet_d3q27_256x256x1024(STORE_FP16_NE), 20000),
    "jet_d3q27_128x128x512_fp16_ne": (lambda: jet_d3q27_128x128x512(STORE_FP16_NE), 5000),
    "jet_d3q27_256x256x1024_mom": (lambda: jet_d3q27_256x256x1024(STORE_MOM), 20000),
    "jet_d3q27_128x128x512_mom": (lambda: jet_d3q27_128x128x512(STORE_MOM), 5000),
    # the 128x128x512 accuracy case at a quarter of the resolution (D = 8)
    "jet_d3q27_40x40x128": (lambda: round_jet(40, 40, 128, 8, D3Q27), 5000),
}


def bytes_per_update(cfg: Config) -> int:
    """Algorithmic HBM bytes of one cell update in the reference's scheme:
    read and write every f and g population once, write and read phi once
    (f and g in the storage type)."""
    s = 2 if cfg.storage in (STORE_FP16, STORE_FP16_NE) else 4
    return 2 * cfg.velocity_set * s + 2 * 7 * s + 8


def bytes_moved_per_update(cfg: Config) -> int:
    """Algorithmic HBM bytes of one cell update as this library stores the
    phase field: read and write every f population once; the timestep reads
    phi and writes the 3-component phase flux, the phase sum reads phi and
    the flux and writes phi (2 Q s + 3 s + 4 + 3 s + 8; f and the flux in
    the storage type, phi in FP32); the non-equilibrium form also writes and
    reads the 8-byte binary16 moments of every cell; the moment form reads
    and writes 10 (D3Q19) or 17 (D3Q27) FP32 coefficients per cell instead of
    the Q populations."""
    if cfg.storage == STORE_MOM:
        nc = 17 if cfg.velocity_set == 27 else 10
        return 2 * nc * 4 + 6 * 4 + 12
    s = 2 if cfg.storage != STORE_FP32 else 4
    extra = 16 if cfg.storage == STORE_FP16_NE else 0  # binary16 moments written and read
    return 2 * cfg.velocity_set * s + 6 * s + 12 + extra


__all__ = ["reference_jet", "reference_droplet", "static_droplet", "round_jet", "jet_d3q19_128x128x512",
           "jet_d3q27_256x256x1024", "jet_d3q27_128x128x512", "WORKLOADS", "bytes_per_update", "bytes_moved_per_update",
           "STORE_FP32", "STORE_FP16", "STORE_FP16_NE", "STORE_MOM"]
