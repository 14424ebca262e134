"""ctypes declarations of the C-ABI in include/phasefieldlbm.h.

The library is the in-tree ``libphasefieldlbm_amd.so`` built by
``phasefieldlbm_amd/csrc/Makefile``.  Loading fails loudly when it is missing:
there is no fallback implementation.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libphasefieldlbm_amd.so")

# enums (include/phasefieldlbm.h)
D3Q19, D3Q27 = 19, 27
JET, DROPLET = 0, 1
STORE_FP32, STORE_FP16, STORE_FP16_NE, STORE_MOM = 0, 1, 2, 3
NOISE_GAUSSIAN_XY, NOISE_UNIFORM_Z, NOISE_NONE = 0, 1, 2
CLOSURE_FD, CLOSURE_MLP = 0, 1
KERNEL_PAIR, KERNEL_LITERAL = 0, 1
STEP_SAVE_LAST = 1
STEP_GRAPH = 2
OUT_BIN, OUT_VTI, OUT_VTR = 1, 2, 4
MLP_PARAMS = 6017
MLP_PREDICT_MAX = 8192  # PF_MLP_PREDICT_MAX

# host::FieldID order (src/fileIO/fields.cuh:33-68): the timestep's fields,
# then the derived fields (enabled through Config.derived)
FIELDS = ("rho", "ux", "uy", "uz", "pxx", "pyy", "pzz", "pxy", "pxz", "pyz",
          "phi", "normx", "normy", "normz", "ind", "fsx", "fsy", "fsz")
DERIVED_FIELDS = ("avg_phi", "avg_ux", "avg_uy", "avg_uz", "avg_uxux", "avg_uyuy", "avg_uzuz", "avg_uxuy",
                  "avg_uxuz", "avg_uyuz", "vort_x", "vort_y", "vort_z", "vort_mag")
FIELD_ID = {n: i for i, n in enumerate(FIELDS + DERIVED_FIELDS)}
DERIVED_AVERAGE, DERIVED_REYNOLDS, DERIVED_VORTICITY = 1, 2, 4

STATUS = {0: "ok", -1: "invalid configuration or argument", -2: "device unavailable",
          -3: "device allocation failed", -4: "kernel launch or runtime error",
          -5: "file could not be written", -6: "non-finite value", -7: "call out of order"}


class PfConfig(ctypes.Structure):
    """Mirror of ``pf_config``; field order must match the C header."""
    _fields_ = [
        ("velocity_set", ctypes.c_int32),
        ("flow_case", ctypes.c_int32),
        ("storage", ctypes.c_int32),
        ("closure", ctypes.c_int32),
        ("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
        ("diam", ctypes.c_int32),
        ("u_inf", ctypes.c_float),
        ("reynolds_water", ctypes.c_float),
        ("reynolds_oil", ctypes.c_float),
        ("weber", ctypes.c_float),
        ("noise", ctypes.c_int32),
        ("noise_amp", ctypes.c_float),
        ("seed", ctypes.c_uint32),
        ("radius", ctypes.c_float),
        ("axis_x", ctypes.c_float), ("axis_y", ctypes.c_float), ("axis_z", ctypes.c_float),
        ("tau", ctypes.c_float),
        ("sigma", ctypes.c_float),
        ("interface_width", ctypes.c_float),
        ("tau_g", ctypes.c_float),
        ("reserved0", ctypes.c_float),
        ("device", ctypes.c_int32),
        ("zchunk", ctypes.c_int32),
        ("mlp_lo", ctypes.c_float), ("mlp_hi", ctypes.c_float),
        ("kernel", ctypes.c_int32),
        ("slot_pad", ctypes.c_int32),
        ("mlp_kmin", ctypes.c_float), ("mlp_kmax", ctypes.c_float),
        ("derived", ctypes.c_int32),
        ("tile_x", ctypes.c_int32),
    ]


# Every exported symbol with its ctypes signature.
SIGNATURES = {
    "pf_config_default": (ctypes.c_int, [ctypes.POINTER(PfConfig), ctypes.c_int32, ctypes.c_int32]),
    "pf_create": (ctypes.c_int, [ctypes.POINTER(PfConfig), ctypes.POINTER(ctypes.c_void_p)]),
    "pf_destroy": (None, [ctypes.c_void_p]),
    "pf_init": (ctypes.c_int, [ctypes.c_void_p]),
    "pf_set_field": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p, ctypes.c_size_t]),
    "pf_init_distros": (ctypes.c_int, [ctypes.c_void_p]),
    "pf_step": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32]),
    "pf_sync": (ctypes.c_int, [ctypes.c_void_p]),
    "pf_get_field": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p, ctypes.c_size_t]),
    "pf_get_populations": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p, ctypes.c_size_t]),
    "pf_diagnostics": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "pf_phase_mass": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)]),
    "pf_set_closure": (ctypes.c_int, [ctypes.c_void_p] + [ctypes.c_void_p] * 6),
    "pf_time_steps": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_void_p]),
    "pf_device_bytes": (ctypes.c_size_t, [ctypes.c_void_p]),
    "pf_copy_bandwidth": (ctypes.c_int, [ctypes.c_int32, ctypes.c_size_t, ctypes.c_uint32,
                                         ctypes.POINTER(ctypes.c_double)]),
    "pf_stream_bandwidth": (ctypes.c_int, [ctypes.c_int32, ctypes.c_size_t, ctypes.c_uint32, ctypes.c_void_p]),
    "pf_mlp_create": (ctypes.c_int, [ctypes.c_int32, ctypes.POINTER(ctypes.c_void_p)]),
    "pf_mlp_destroy": (None, [ctypes.c_void_p]),
    "pf_mlp_set_params": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "pf_mlp_get_params": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "pf_mlp_set_data": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t]),
    "pf_mlp_train": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_float,
                                    ctypes.c_uint32, ctypes.c_int32, ctypes.c_void_p]),
    "pf_mlp_gradients": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p,
                                        ctypes.POINTER(ctypes.c_double)]),
    "pf_mlp_predict": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p]),
    "pf_mlp_predict_noalloc": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p]),
    "pf_mfma_check": (ctypes.c_int, [ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "pf_write_fields": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_uint64, ctypes.c_void_p,
                                       ctypes.c_int32, ctypes.c_int32]),
    "pf_write_vtk": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                    ctypes.POINTER(ctypes.c_char_p), ctypes.POINTER(ctypes.c_void_p), ctypes.c_int32]),
    "pf_field_name": (ctypes.c_char_p, [ctypes.c_int32]),
    "pf_get_config": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(PfConfig)]),
    "pf_device_name": (ctypes.c_int, [ctypes.c_int32, ctypes.c_char_p, ctypes.c_size_t]),
    "pf_last_error": (ctypes.c_char_p, [ctypes.c_void_p]),
    "pf_status_string": (ctypes.c_char_p, [ctypes.c_int]),
}

_lib = None


def load() -> ctypes.CDLL:
    """Load the HIP library (once).  Raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `make -C phasefieldlbm_amd/csrc` "
            "(or __graft_entry__.build()); there is no CPU fallback")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib
