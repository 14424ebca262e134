"""MI355X-native phase-field lattice Boltzmann timestep (phaseFieldLBM drop-in).

The host side mirrors the reference's driver interface (src/main.cu:35-215):
a ``Simulation`` owns the fields of one run, ``init()`` applies the initial
conditions, ``step()`` advances the timestep (phase field, moments,
collide-stream, flow-case boundary conditions) and ``field()`` reads the
macroscopic arrays named as in host::FieldID (src/fileIO/fields.cuh:33-68).
All work runs in the HIP library ``libphasefieldlbm_amd.so`` through its
C-ABI; nothing here computes physics.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, fields as dc_fields
from typing import Optional

import numpy as np

from . import _abi
from ._abi import (CLOSURE_FD, CLOSURE_MLP, D3Q19, D3Q27, DERIVED_AVERAGE, DERIVED_FIELDS, DERIVED_REYNOLDS,
                   DERIVED_VORTICITY, DROPLET, FIELDS, JET, KERNEL_LITERAL, KERNEL_PAIR, NOISE_GAUSSIAN_XY,
                   NOISE_NONE, NOISE_UNIFORM_Z, STORE_FP16, STORE_FP16_NE, STORE_FP32, STORE_MOM)

__all__ = ["Config", "Simulation", "PfError", "presets", "FIELDS", "D3Q19", "D3Q27", "JET", "DROPLET",
           "STORE_FP32", "STORE_FP16", "STORE_FP16_NE", "STORE_MOM", "NOISE_GAUSSIAN_XY", "NOISE_UNIFORM_Z", "NOISE_NONE",
           "CLOSURE_FD", "CLOSURE_MLP", "KERNEL_PAIR", "KERNEL_LITERAL", "library_path", "copy_bandwidth",
           "stream_bandwidth", "MLP",
           "split_params", "mfma_check", "vtk", "DERIVED_FIELDS", "DERIVED_AVERAGE", "DERIVED_REYNOLDS",
           "DERIVED_VORTICITY"]


def library_path() -> str:
    return _abi.LIB_PATH


def copy_bandwidth(device: int = 0, nbytes: int = 2 << 30, iters: int = 20) -> float:
    """Measured HBM bandwidth (GB/s, read + write) of a streaming copy."""
    g = ctypes.c_double()
    rc = _abi.load().pf_copy_bandwidth(device, nbytes, iters, ctypes.byref(g))
    if rc:
        raise PfError(rc, "pf_copy_bandwidth")
    return g.value


def stream_bandwidth(device: int = 0, nbytes: int = 4 << 30, iters: int = 10) -> dict:
    """Measured HBM bandwidth (GB/s) of read-only, write-only and copy
    streaming kernels over `nbytes`-byte arrays."""
    out = np.zeros(3, dtype=np.float64)
    rc = _abi.load().pf_stream_bandwidth(device, nbytes, iters, out.ctypes.data)
    if rc:
        raise PfError(rc, "pf_stream_bandwidth")
    return {"read_gbs": float(out[0]), "write_gbs": float(out[1]), "copy_gbs": float(out[2])}


class PfError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"phasefieldlbm status {status} ({_abi.STATUS.get(status, '?')}): {msg}")
        self.status = status


@dataclass
class Config:
    """Run configuration; the fields of ``pf_config`` (include/phasefieldlbm.h)."""
    velocity_set: int = D3Q19
    flow_case: int = JET
    storage: int = STORE_FP32
    closure: int = CLOSURE_FD
    nx: int = 64
    ny: int = 64
    nz: int = 128
    diam: int = 10
    u_inf: float = 0.05
    reynolds_water: float = 5000.0
    reynolds_oil: float = 5000.0
    weber: float = 500.0
    noise: int = NOISE_GAUSSIAN_XY
    noise_amp: float = 0.05
    seed: int = 0x9E3779B9
    radius: float = 35.0
    axis_x: float = 2.0
    axis_y: float = 1.0
    axis_z: float = 1.0
    tau: float = 0.55
    sigma: float = 0.03
    interface_width: float = 4.0
    tau_g: float = 1.0
    reserved0: float = 0.0  # reserved: must be 0
    device: int = -1
    zchunk: int = 0
    mlp_lo: float = 0.001
    mlp_hi: float = 0.999
    kernel: int = KERNEL_PAIR
    slot_pad: int = -1
    mlp_kmin: float = -2.0
    mlp_kmax: float = -0.03
    derived: int = 0  # DERIVED_AVERAGE | DERIVED_REYNOLDS | DERIVED_VORTICITY
    tile_x: int = 0  # timestep tile width: 0 = automatic, 64 or 128

    @classmethod
    def reference(cls, flow_case: int, velocity_set: int) -> "Config":
        """The reference's compile-time defaults (src/constants.cuh:82-141)."""
        lib = _abi.load()
        c = _abi.PfConfig()
        rc = lib.pf_config_default(ctypes.byref(c), flow_case, velocity_set)
        if rc:
            raise PfError(rc, "pf_config_default")
        return cls(**{f.name: getattr(c, f.name) for f in dc_fields(cls)})

    def to_c(self) -> _abi.PfConfig:
        c = _abi.PfConfig()
        for f in dc_fields(self):
            setattr(c, f.name, getattr(self, f.name))
        return c

    @property
    def shape(self):
        """numpy shape of a field: (nz, ny, nx), x fastest."""
        return (self.nz, self.ny, self.nx)

    @property
    def cells(self) -> int:
        return self.nx * self.ny * self.nz

    @property
    def q(self) -> int:
        return self.velocity_set


class Simulation:
    """One simulation on one GPU (the reference's LBMFields + stream)."""

    def __init__(self, cfg: Config):
        self.cfg = cfg
        self._lib = _abi.load()
        self._c = cfg.to_c()
        h = ctypes.c_void_p()
        rc = self._lib.pf_create(ctypes.byref(self._c), ctypes.byref(h))
        if rc:
            raise PfError(rc, "pf_create")
        self._h = h
        self.t = 0

    # -- lifecycle ---------------------------------------------------------
    def close(self):
        if getattr(self, "_h", None):
            self._lib.pf_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, rc: int, what: str):
        if rc:
            msg = self._lib.pf_last_error(self._h)
            raise PfError(rc, f"{what}: {msg.decode() if msg else ''}")

    # -- reference driver surface -------------------------------------------
    def init(self):
        """setInitialDensity + setJet|setDroplet + setDistros."""
        self._check(self._lib.pf_init(self._h), "pf_init")
        self.t = 0
        return self

    def set_field(self, name: str, values: np.ndarray):
        a = np.ascontiguousarray(values, dtype=np.float32).reshape(-1)
        self._check(self._lib.pf_set_field(self._h, _abi.FIELD_ID[name], a.ctypes.data, a.size),
                    "pf_set_field")

    def init_distros(self):
        self._check(self._lib.pf_init_distros(self._h), "pf_init_distros")
        self.t = 0

    def step(self, n: int = 1, save: bool = False, t0: Optional[int] = None, graph: bool = False):
        """Advance n timesteps; the inflow noise uses time index t0+s.

        graph=True replays the kernels before the boundary conditions as a
        captured hipGraph per A-A parity (bit-identical to direct launches).
        """
        t0 = self.t if t0 is None else t0
        flags = (_abi.STEP_SAVE_LAST if save else 0) | (_abi.STEP_GRAPH if graph else 0)
        self._check(self._lib.pf_step(self._h, t0, n, flags), "pf_step")
        self.t = t0 + n
        return self

    def sync(self):
        self._check(self._lib.pf_sync(self._h), "pf_sync")

    def field(self, name: str) -> np.ndarray:
        out = np.empty(self.cfg.cells, dtype=np.float32)
        self._check(self._lib.pf_get_field(self._h, _abi.FIELD_ID[name], out.ctypes.data, out.size),
                    "pf_get_field")
        return out.reshape(self.cfg.shape)

    def populations(self, which: str = "f") -> np.ndarray:
        nq = self.cfg.q if which == "f" else 7
        out = np.empty(nq * self.cfg.cells, dtype=np.float32)
        self._check(self._lib.pf_get_populations(self._h, 0 if which == "f" else 1, out.ctypes.data,
                                                 out.size), "pf_get_populations")
        return out.reshape((nq,) + self.cfg.shape)

    def diagnostics(self) -> dict:
        out = np.zeros(6, dtype=np.float64)
        self._check(self._lib.pf_diagnostics(self._h, out.ctypes.data), "pf_diagnostics")
        return {"mass": out[0], "momentum": out[1:4].copy(), "phi_mass": out[4], "umax": out[5]}

    def phase_mass(self) -> float:
        m = ctypes.c_double()
        self._check(self._lib.pf_phase_mass(self._h, ctypes.byref(m)), "pf_phase_mass")
        return m.value

    def set_closure_params(self, params):
        """Closure weights from one 6017-value parameter vector."""
        self.set_closure(*split_params(params))

    def set_closure(self, W1, b1, W2, b2, W3, b3):
        arrs = [np.ascontiguousarray(a, dtype=np.float32).reshape(-1) for a in (W1, b1, W2, b2, W3, b3)]
        sizes = [64 * 27, 64, 64 * 64, 64, 64, 1]
        for a, n in zip(arrs, sizes):
            if a.size != n:
                raise ValueError(f"closure array of size {a.size}, expected {n}")
        self._check(self._lib.pf_set_closure(self._h, *[a.ctypes.data for a in arrs]), "pf_set_closure")

    def time_steps(self, n: int) -> dict:
        """HIP-event timing of n steps on the simulation's stream (ms)."""
        out = np.zeros(4, dtype=np.float32)
        self._check(self._lib.pf_time_steps(self._h, self.t, n, out.ctypes.data), "pf_time_steps")
        self.t += 2 * n
        return {"total_ms": float(out[0]), "phase_sum_ms": float(out[1]),
                "step_ms": float(out[2]), "bc_ms": float(out[3])}

    def device_bytes(self) -> int:
        return int(self._lib.pf_device_bytes(self._h))

    def write(self, directory: str, step: int, fields=("rho", "uz"), formats=("bin", "vtr")):
        """Write fields as <name><step6>.bin, step_<step6>.vti and/or
        step_<step6>.vtr in an existing directory (the reference's output
        step, src/main.cu:178-185 and src/postProcess/PostProcess.cuh)."""
        ids = (ctypes.c_int32 * len(fields))(*[_abi.FIELD_ID[f] for f in fields])
        mask = 0
        for f in formats:
            mask |= {"bin": _abi.OUT_BIN, "vti": _abi.OUT_VTI, "vtr": _abi.OUT_VTR}[f]
        self._check(self._lib.pf_write_fields(self._h, str(directory).encode(), int(step), ids, len(fields), mask),
                    "pf_write_fields")


class MLP:
    """The learned curvature closure's network on the device (pf_mlp).

    Parameters are one FP32 vector of 6017 values, W1[64][27] | b1 | W2[64][64]
    | b2 | w3[64] | b3; samples are the 27 normalised stencil features of
    ``closure.stencils`` (feature (dz+1)*9 + (dy+1)*3 + (dx+1)) and a target
    4*curvature.
    """

    NPARAMS = _abi.MLP_PARAMS
    PREDICT_MAX = _abi.MLP_PREDICT_MAX

    def __init__(self, device: int = 0, params: Optional[np.ndarray] = None):
        self._lib = _abi.load()
        h = ctypes.c_void_p()
        rc = self._lib.pf_mlp_create(device, ctypes.byref(h))
        if rc:
            raise PfError(rc, "pf_mlp_create")
        self._h = h
        if params is not None:
            self.set_params(params)

    def close(self):
        if getattr(self, "_h", None):
            self._lib.pf_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, what):
        if rc:
            raise PfError(rc, what)

    def set_params(self, params):
        a = np.ascontiguousarray(params, dtype=np.float32).reshape(-1)
        assert a.size == self.NPARAMS
        self._check(self._lib.pf_mlp_set_params(self._h, a.ctypes.data), "pf_mlp_set_params")

    def params(self) -> np.ndarray:
        out = np.empty(self.NPARAMS, dtype=np.float32)
        self._check(self._lib.pf_mlp_get_params(self._h, out.ctypes.data), "pf_mlp_get_params")
        return out

    def set_data(self, x, t):
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, 27)
        t = np.ascontiguousarray(t, dtype=np.float32).reshape(-1)
        assert x.shape[0] == t.size
        self._keep = (x, t)
        self._check(self._lib.pf_mlp_set_data(self._h, x.ctypes.data, t.ctypes.data, t.size), "pf_mlp_set_data")

    def train(self, epochs: int, batch: int = 512, lr: float = 1e-3, seed: int = 1, augment: bool = True):
        out = np.zeros(epochs, dtype=np.float32)
        self._check(self._lib.pf_mlp_train(self._h, epochs, batch, lr, seed, 1 if augment else 0,
                                           out.ctypes.data), "pf_mlp_train")
        return out

    def gradients(self, n: int):
        g = np.empty(self.NPARAMS, dtype=np.float32)
        mse = ctypes.c_double()
        self._check(self._lib.pf_mlp_gradients(self._h, n, g.ctypes.data, ctypes.byref(mse)), "pf_mlp_gradients")
        return g, mse.value

    def predict(self, x) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, 27)
        y = np.empty(x.shape[0], dtype=np.float32)
        self._check(self._lib.pf_mlp_predict(self._h, x.ctypes.data, x.shape[0], y.ctypes.data), "pf_mlp_predict")
        return y


    def predict_noalloc(self, x) -> np.ndarray:
        """As predict(), staged in the library's fixed device buffer: no
        allocation by the library (at most PREDICT_MAX samples per call)."""
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, 27)
        y = np.empty(x.shape[0], dtype=np.float32)
        self._check(self._lib.pf_mlp_predict_noalloc(self._h, x.ctypes.data, x.shape[0], y.ctypes.data),
                    "pf_mlp_predict_noalloc")
        return y


def split_params(params):
    """(W1[64,27], b1, W2[64,64], b2, w3[64], b3[1]) views of a parameter vector."""
    p = np.asarray(params, dtype=np.float32).reshape(-1)
    o = [0, 64 * 27, 64 * 27 + 64, 64 * 27 + 64 + 4096, 64 * 27 + 128 + 4096, 64 * 27 + 192 + 4096, 6017]
    return (p[o[0]:o[1]].reshape(64, 27), p[o[1]:o[2]], p[o[2]:o[3]].reshape(64, 64), p[o[3]:o[4]],
            p[o[4]:o[5]], p[o[5]:o[6]])


def mfma_check(a, b, device: int = 0) -> np.ndarray:
    """One BF16 MFMA tile C = A B (A 16x32, B 32x16) computed as the kernels do."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    b = np.ascontiguousarray(b, dtype=np.float32)
    c = np.empty((16, 16), dtype=np.float32)
    rc = _abi.load().pf_mfma_check(device, a.ctypes.data, b.ctypes.data, c.ctypes.data)
    if rc:
        raise PfError(rc, "pf_mfma_check")
    return c


from . import presets, vtk  # noqa: E402
