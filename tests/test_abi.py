"""The C-ABI library loads and exports every declared symbol (no GPU work)."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import phasefieldlbm_amd as pf
from phasefieldlbm_amd import _abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "phasefieldlbm.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pf_[a-z_0-9]+)\s*\(", text)))


def test_header_symbols_exported():
    lib = _abi.load()
    syms = declared_symbols()
    assert len(syms) >= 15
    for s in syms:
        assert hasattr(lib, s), s
    assert set(syms) == set(_abi.SIGNATURES), set(syms) ^ set(_abi.SIGNATURES)


def test_config_struct_layout():
    # the ctypes mirror has the C size: four-byte members only, in header order
    assert ctypes.sizeof(_abi.PfConfig) == 4 * len(_abi.PfConfig._fields_)
    text = open(os.path.join(ROOT, "include", "phasefieldlbm.h")).read()
    body = text[text.index("typedef struct pf_config {"):text.index("} pf_config;")]
    names = re.findall(r"\b(?:int32_t|uint32_t|float)\s+([a-z_0-9, ]+);", body)
    flat = [n.strip() for grp in names for n in grp.split(",")]
    assert flat == [f[0] for f in _abi.PfConfig._fields_]


def test_reference_defaults():
    jet = pf.Config.reference(pf.JET, pf.D3Q19)
    assert (jet.nx, jet.ny, jet.nz, jet.diam) == (64, 64, 128, 10)
    assert jet.reynolds_water == 5000 and jet.weber == 500 and abs(jet.u_inf - 0.05) < 1e-9
    dr = pf.Config.reference(pf.DROPLET, pf.D3Q27)
    assert (dr.nx, dr.radius, dr.axis_x) == (128, 35.0, 2.0)
    assert abs(dr.sigma - 0.03) < 1e-9 and abs(dr.tau - 0.55) < 1e-7


@pytest.mark.parametrize("field,value", [("velocity_set", 15), ("flow_case", 7), ("storage", 9), ("nx", 2),
                                         ("noise", 5), ("interface_width", 0.0), ("tile_x", 32),
                                         ("reserved0", 1.0)])
def test_bad_config_rejected(field, value):
    cfg = pf.presets.static_droplet(n=16)
    setattr(cfg, field, value)
    with pytest.raises(pf.PfError) as e:
        pf.Simulation(cfg)
    assert e.value.status == -1


def test_ne_storage_constraints():
    """The binary16 non-equilibrium form takes the pair collision only, and a
    jet nozzle inside the lateral ghost shell."""
    cfg = pf.presets.round_jet(24, 24, 40, 6, pf.D3Q27, storage=pf.STORE_FP16_NE)
    cfg.kernel = pf.KERNEL_LITERAL
    with pytest.raises(pf.PfError) as e:
        pf.Simulation(cfg)
    assert e.value.status == -1
    cfg = pf.presets.round_jet(12, 12, 40, 10, pf.D3Q27, storage=pf.STORE_FP16_NE)
    with pytest.raises(pf.PfError) as e:
        pf.Simulation(cfg)
    assert e.value.status == -1


def test_moment_storage_constraints():
    """The moment storage form takes the pair collision and the FD curvature
    only, and planes whose coefficients fit a 32-bit buffer range
    (17 x 4 B x nx*ny < 2^32); checked before any device is touched."""
    for field, value in (("kernel", pf.KERNEL_LITERAL), ("closure", pf.CLOSURE_MLP)):
        cfg = pf.presets.round_jet(24, 24, 40, 6, pf.D3Q27, storage=pf.STORE_MOM)
        setattr(cfg, field, value)
        with pytest.raises(pf.PfError) as e:
            pf.Simulation(cfg)
        assert e.value.status == -1, field
    cfg = pf.presets.static_droplet(n=8, radius=2.0, velocity_set=pf.D3Q27, storage=pf.STORE_MOM)
    cfg.nx, cfg.ny, cfg.nz = 8000, 8000, 4
    with pytest.raises(pf.PfError) as e:
        pf.Simulation(cfg)
    assert e.value.status == -1


def test_status_strings():
    lib = _abi.load()
    for code in range(0, -8, -1):
        assert lib.pf_status_string(code).decode() == _abi.STATUS[code]


def test_no_device_reports_device_error():
    from tests.conftest import HAVE_GPU
    if HAVE_GPU:
        pytest.skip("a GPU is present")
    with pytest.raises(pf.PfError) as e:
        pf.Simulation(pf.presets.static_droplet(n=16))
    assert e.value.status == -2


MINIMAL = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "phasefieldlbm_amd", "minimal")


def test_c_example_builds_and_fails_cleanly_without_device(tmp_path):
    """The plain-C host program (driver/minimal.c) compiles as C99 against the
    header alone and links against the library; without a GPU it stops at
    pf_create with the device error."""
    from tests.conftest import HAVE_GPU
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    exe = tmp_path / "minimal"
    subprocess.run(["cc", "-std=c99", "-Wall", "-Wextra", "-Werror", "-I", os.path.join(root, "include"),
                    os.path.join(root, "phasefieldlbm_amd", "driver", "minimal.c"), "-o", str(exe),
                    "-L", os.path.join(root, "phasefieldlbm_amd"), "-lphasefieldlbm_amd",
                    "-Wl,-rpath," + os.path.join(root, "phasefieldlbm_amd")], check=True)
    if HAVE_GPU:
        pytest.skip("a GPU is present")
    r = subprocess.run([str(exe), str(tmp_path), "10"], capture_output=True, text=True)
    assert r.returncode == 1
    assert "pf_create: device unavailable" in r.stderr


@pytest.mark.gpu
def test_c_example_runs(tmp_path):
    """driver/minimal.c on the GPU: 120 steps of the reference's D3Q27 jet,
    .vtr output every 50 steps and at the end, finite diagnostics."""
    from phasefieldlbm_amd import vtk
    r = subprocess.run([MINIMAL, str(tmp_path), "120"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "max |u|" in r.stdout
    names = sorted(p.name for p in tmp_path.glob("*.vtr"))
    assert names == ["step_000049.vtr", "step_000099.vtr", "step_000119.vtr"], names
    meta, arrays = vtk.read_vtk(str(tmp_path / "step_000119.vtr"))
    assert arrays["phi"].shape == (128, 64, 64) and np.isfinite(arrays["uz"]).all()
    assert arrays["phi"].max() > 0.5  # the nozzle plane and the jet carry phi = 1
