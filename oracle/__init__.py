"""CPU oracle -- TEST INFRASTRUCTURE ONLY.

ctypes wrapper of ``libpforacle.so`` (built from ``pf_oracle.cpp`` by the
Makefile in this directory): a plain C++/OpenMP restatement of the reference
timestep in FP32, FP32 with binary16 f (the reference's FP16 mode), FP32
with binary16 f and phase flux (the library's binary16 form) and FP64.  Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg may use it, as the
checker; the product never calls it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpforacle.so")
_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    lib = ctypes.CDLL(LIB_PATH)
    vp, i32, u32, u64, dp = ctypes.c_void_p, ctypes.c_int, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_void_p
    sig = {
        "pfo_create": (vp, [vp, i32]),
        "pfo_destroy": (None, [vp]),
        "pfo_init": (i32, [vp]),
        "pfo_init_distros": (i32, [vp]),
        "pfo_step": (i32, [vp, u64, u32]),
        "pfo_get_field": (i32, [vp, i32, dp]),
        "pfo_set_field": (i32, [vp, i32, dp]),
        "pfo_get_pop": (i32, [vp, i32, dp]),
        "pfo_set_pop": (i32, [vp, i32, dp]),
        "pfo_num_threads": (i32, []),
        "pfo_set_num_threads": (None, [i32]),
        "pfo_lattice": (i32, [i32, vp, vp, vp, vp]),
        "pfo_cell_feq": (i32, [i32, ctypes.c_double, vp, vp]),
        "pfo_cell_fneq": (i32, [i32, vp, vp, vp]),
        "pfo_cell_force": (i32, [i32, vp, vp, vp]),
        "pfo_cell_geq": (i32, [ctypes.c_double, vp, vp, ctypes.c_double, vp, vp]),
        "pfo_iso_grad": (i32, [i32, vp, vp]),
        "pfo_f2h": (ctypes.c_uint16, [ctypes.c_float]),
        "pfo_h2f": (ctypes.c_float, [ctypes.c_uint16]),
        "pfo_hash32": (u32, [u32]),
        "pfo_consts": (i32, [vp, dp]),
        "pfo_omega": (ctypes.c_double, [vp, ctypes.c_double, i32]),
        "pfo_omco_zmax": (ctypes.c_double, [vp, ctypes.c_double]),
        "pfo_sponge_ramp": (ctypes.c_double, [vp, i32]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


FIELDS = ("rho", "ux", "uy", "uz", "pxx", "pyy", "pzz", "pxy", "pxz", "pyz",
          "phi", "normx", "normy", "normz", "ind", "fsx", "fsy", "fsz")


CONSTS = ("sigma", "gamma", "u_inf", "visc_water", "visc_oil", "tau_water", "tau_oil", "omco_zmin",
          "omega_ref", "sp_z_start", "sp_inv_nz_m1", "sp_inv_sponge", "cx", "cy", "cz", "R2")


class Oracle:
    """One CPU simulation: precision 32 (FP32), 16 (binary16 f, the
    reference's FP16 mode), 17 (binary16 f and phase flux, the library's
    binary16 form) or 64."""

    def __init__(self, cfg, precision: int = 32):
        self.lib = load()
        self.cfg = cfg
        self._c = cfg.to_c()
        self.h = self.lib.pfo_create(ctypes.byref(self._c), precision)
        if not self.h:
            raise ValueError("oracle rejected the configuration")
        self.t = 0

    def close(self):
        if self.h:
            self.lib.pfo_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def init(self):
        assert self.lib.pfo_init(self.h) == 0
        self.t = 0
        return self

    def init_distros(self):
        assert self.lib.pfo_init_distros(self.h) == 0
        self.t = 0

    def set_field(self, name, values):
        a = np.ascontiguousarray(values, dtype=np.float64).reshape(-1)
        assert a.size == self.cfg.cells
        assert self.lib.pfo_set_field(self.h, FIELDS.index(name), a.ctypes.data) == 0

    def step(self, n=1, t0=None):
        t0 = self.t if t0 is None else t0
        assert self.lib.pfo_step(self.h, t0, n) == 0
        self.t = t0 + n
        return self

    def consts(self) -> dict:
        """Derived constants of the configuration (pfo_consts)."""
        out = np.zeros(len(CONSTS), dtype=np.float64)
        assert self.lib.pfo_consts(self.h, out.ctypes.data) == 0
        return dict(zip(CONSTS, out.tolist()))

    def omega(self, phi: float, z: int) -> float:
        """Collision frequency of a cell with phase phi in plane z."""
        return float(self.lib.pfo_omega(self.h, float(phi), int(z)))

    def omco_zmax(self, phi: float) -> float:
        return float(self.lib.pfo_omco_zmax(self.h, float(phi)))

    def sponge_ramp(self, z: int) -> float:
        return float(self.lib.pfo_sponge_ramp(self.h, int(z)))

    def field(self, name) -> np.ndarray:
        out = np.empty(self.cfg.cells, dtype=np.float64)
        assert self.lib.pfo_get_field(self.h, FIELDS.index(name), out.ctypes.data) == 0
        return out.reshape(self.cfg.shape)

    def set_populations(self, which, values):
        """Replace the post-streaming populations f (Q x shape, shifted by -w)
        or g (7 x shape), e.g. with a GPU state."""
        nq = self.cfg.velocity_set if which == "f" else 7
        a = np.ascontiguousarray(values, dtype=np.float64).reshape(-1)
        assert a.size == nq * self.cfg.cells
        assert self.lib.pfo_set_pop(self.h, 0 if which == "f" else 1, a.ctypes.data) == 0

    def populations(self, which="f") -> np.ndarray:
        nq = self.cfg.velocity_set if which == "f" else 7
        out = np.empty(nq * self.cfg.cells, dtype=np.float64)
        assert self.lib.pfo_get_pop(self.h, 0 if which == "f" else 1, out.ctypes.data) == 0
        return out.reshape((nq,) + self.cfg.shape)


def num_threads() -> int:
    return int(load().pfo_num_threads())


def set_num_threads(n: int) -> None:
    load().pfo_set_num_threads(int(n))
