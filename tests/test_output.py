"""Output files and the host driver.

The reference writes <name><step6>.bin raw FP32 dumps and step_<step6>.vti
VTK ImageData with raw appended data (src/postProcess/PostProcess.cuh:43-261)
into bin/<CASE>/<VS>/<ID>/ with an <ID>_info.txt (src/functions/
hostFunctions.cuh:33-85), from the executable `<case> <velocity set> <ID>`
(src/main.cu:35-215).  The byte-level checks here follow the VTK XML file
format itself (appended raw blocks, each a UInt32 or UInt64 byte count and
the data), not this package's reader.
"""
import os
import re
import struct
import subprocess
import xml.etree.ElementTree as ET

import numpy as np
import pytest

from phasefieldlbm_amd import _abi, vtk

PFSIM = os.path.join(os.path.dirname(_abi.LIB_PATH), "pfsim")


def _parse_appended(raw):
    """XML root and the raw appended blocks in file order, by the format's
    own rules: '_' starts the data; each block is a UInt32 count + bytes."""
    start = raw.index(b"<AppendedData")
    mark = raw.index(b"_", raw.index(b">", start)) + 1
    root = ET.fromstring(raw[:start].decode() + "</VTKFile>")
    blocks, pos = [], mark
    end = raw.rindex(b"</AppendedData>")
    while pos < end - 8:
        n = struct.unpack_from("<I", raw, pos)[0]
        blocks.append((pos - mark, raw[pos + 4:pos + 4 + n]))
        pos += 4 + n
    return root, blocks


@pytest.mark.parametrize("fmt", ["vti", "vtr"])
def test_vtk_layout(tmp_path, fmt):
    rng = np.random.default_rng(1)
    nz, ny, nx = 6, 5, 7
    arrays = {"rho": rng.normal(size=(nz, ny, nx)).astype(np.float32),
              "uz": rng.normal(size=(nz, ny, nx)).astype(np.float32),
              "phi": rng.uniform(size=(nz, ny, nx)).astype(np.float32)}
    path = str(tmp_path / f"step_000010.{fmt}")
    vtk.write_vtk(path, arrays, fmt)
    raw = open(path, "rb").read()
    root, blocks = _parse_appended(raw)
    kind = "ImageData" if fmt == "vti" else "RectilinearGrid"
    assert root.tag == "VTKFile" and root.get("type") == kind
    assert root.get("byte_order") == "LittleEndian" and root.get("header_type") == "UInt32"
    grid = root.find(kind)
    assert grid.get("WholeExtent") == f"0 {nx - 1} 0 {ny - 1} 0 {nz - 1}"
    assert grid.find("Piece").get("Extent") == grid.get("WholeExtent")
    if fmt == "vti":
        assert grid.get("Origin") == "0 0 0" and grid.get("Spacing") == "1 1 1"
    pdata = grid.findall("Piece/PointData/DataArray")
    assert [d.get("Name") for d in pdata] == list(arrays)
    assert grid.find("Piece/PointData").get("Scalars") == "rho"
    coords = grid.findall("Piece/Coordinates/DataArray")
    assert len(coords) == (3 if fmt == "vtr" else 0)
    assert len(blocks) == len(pdata) + len(coords)
    for (off, data), da, name in zip(blocks, pdata, arrays):
        assert int(da.get("offset")) == off and da.get("type") == "Float32"
        np.testing.assert_array_equal(np.frombuffer(data, "<f4").reshape(nz, ny, nx), arrays[name])
    for (off, data), da, n in zip(blocks[len(pdata):], coords, (nx, ny, nz)):
        assert int(da.get("offset")) == off
        np.testing.assert_array_equal(np.frombuffer(data, "<f4"), np.arange(n, dtype=np.float32))
    # the package's reader agrees
    meta, back = vtk.read_vtk(path)
    assert meta["shape"] == (nz, ny, nx)
    for k, v in arrays.items():
        np.testing.assert_array_equal(back[k], v)


def test_vtk_rejects_bad_arguments(tmp_path):
    lib = _abi.load()
    import ctypes
    a = np.zeros(8, np.float32)
    names = (ctypes.c_char_p * 1)(b"rho")
    ptrs = (ctypes.c_void_p * 1)(a.ctypes.data)
    p = str(tmp_path / "x.vtr").encode()
    assert lib.pf_write_vtk(p, 3, 2, 2, 2, names, ptrs, 1) == -1        # not a single VTK format
    assert lib.pf_write_vtk(p, _abi.OUT_VTR, 0, 2, 2, names, ptrs, 1) == -1
    bad = (ctypes.c_char_p * 1)(b'r"ho')
    assert lib.pf_write_vtk(p, _abi.OUT_VTR, 2, 2, 2, bad, ptrs, 1) == -1
    missing = str(tmp_path / "no_such_dir" / "x.vtr").encode()
    assert lib.pf_write_vtk(missing, _abi.OUT_VTR, 2, 2, 2, names, ptrs, 1) == -5


def test_field_names():
    """Names in host::FieldID order (src/fileIO/fields.cuh:33-68), derived
    fields after the timestep's; the reference's passive scalar is not offered."""
    lib = _abi.load()
    names = list(_abi.FIELDS + _abi.DERIVED_FIELDS)
    assert [lib.pf_field_name(i).decode() for i in range(len(names))] == names
    assert names[18] == "avg_phi" and names[28] == "vort_x" and names[31] == "vort_mag"
    assert lib.pf_field_name(-1) is None and lib.pf_field_name(len(names)) is None


def _run(args, cwd, timeout=300):
    return subprocess.run([PFSIM] + args, cwd=cwd, capture_output=True, text=True, timeout=timeout)


def test_driver_argument_errors(tmp_path):
    assert os.path.exists(PFSIM), "build the driver (make -C phasefieldlbm_amd/csrc)"
    r = _run([], tmp_path)
    assert r.returncode == 1 and "Usage" in r.stderr
    r = _run(["JET", "D3Q15", "000"], tmp_path)
    assert r.returncode == 1 and "velocity set" in r.stderr
    r = _run(["SPRAY", "D3Q19", "000"], tmp_path)
    assert r.returncode == 1 and "flow case" in r.stderr
    r = _run(["JET", "D3Q19", "000", "--preset", "droplet_d3q19_64"], tmp_path)
    assert r.returncode == 1 and "preset" in r.stderr
    r = _run(["DROPLET", "D3Q19", "000", "--fields", "rho,vorticity"], tmp_path)
    assert r.returncode == 1 and "unknown field" in r.stderr
    r = _run(["DROPLET", "D3Q19", "000", "--format", "vtk"], tmp_path)
    assert r.returncode == 1 and "unknown format" in r.stderr


@pytest.mark.gpu
def test_simulation_write_matches_fields(tmp_path):
    import phasefieldlbm_amd as pf
    cfg = pf.presets.static_droplet(n=24, radius=6.0)
    with pf.Simulation(cfg) as sim:
        sim.init().step(5, save=True)
        want = {k: sim.field(k) for k in ("rho", "uz", "phi")}
        sim.write(str(tmp_path), 5, fields=("rho", "uz", "phi"), formats=("bin", "vti", "vtr"))
    for k, v in want.items():
        np.testing.assert_array_equal(vtk.read_bin(str(tmp_path / f"{k}000005.bin"), v.shape), v)
    for fmt in ("vti", "vtr"):
        meta, arrays = vtk.read_vtk(str(tmp_path / f"step_000005.{fmt}"))
        for k, v in want.items():
            np.testing.assert_array_equal(arrays[k], v)


@pytest.mark.gpu
@pytest.mark.parametrize("graph", ["off", "on"])
def test_driver_droplet_run(tmp_path, graph):
    """pfsim DROPLET D3Q19 <ID>: directory layout, info file, output cadence
    (steps 0, 10, 20 for NSTEPS=20, MACRO_SAVE=10) and the final fields equal
    those of the same 21 steps run through the Python binding, with the step
    launched directly or replayed as a captured graph (--graph)."""
    import phasefieldlbm_amd as pf
    r = _run(["DROPLET", "D3Q19", "007", "--preset", "droplet_d3q19_64", "--steps", "20", "--save", "10",
              "--fields", "rho,uz,phi", "--format", "bin,vti,vtr", "--graph", graph], tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    d = tmp_path / "bin" / "DROPLET" / "D3Q19" / "007"
    info = (d / "007_info.txt").read_text()
    assert re.search(r"NX:\s+64", info) and re.search(r"Timesteps:\s+20", info)
    assert re.search(r"Output interval:\s+10", info)
    for s in ("000000", "000010", "000020"):
        for k in ("rho", "uz", "phi"):
            assert (d / f"{k}{s}.bin").exists()
        assert (d / f"step_{s}.vti").exists() and (d / f"step_{s}.vtr").exists()
    assert not (d / "rho000005.bin").exists()
    assert "MLUPS" in r.stdout and "Using GPU" in r.stdout
    cfg = pf.presets.static_droplet()
    with pf.Simulation(cfg) as sim:
        sim.init().step(21, save=True, t0=0)
        for k in ("rho", "uz", "phi"):
            np.testing.assert_array_equal(vtk.read_bin(str(d / f"{k}000020.bin"), cfg.shape), sim.field(k))


@pytest.mark.gpu
def test_driver_jet_benchmark(tmp_path):
    r = _run(["JET", "D3Q27", "001", "--mode", "benchmark", "--steps", "50"], tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    m = re.search(r"Performance\s+:\s+([0-9.eE+]+) MLUPS", r.stdout)
    assert m and float(m.group(1)) > 0
    d = tmp_path / "bin" / "JET" / "D3Q27" / "001"
    assert sorted(os.listdir(d)) == ["001_info.txt"]


@pytest.mark.gpu
def test_driver_fp16_jet(tmp_path):
    """pfsim JET D3Q27 <ID> --fp16: binary16 storage through the driver; the
    metadata names it and the final fields equal the same steps run through
    the Python binding in the binary16 form."""
    import phasefieldlbm_amd as pf
    r = _run(["JET", "D3Q27", "002", "--preset", "jet_d3q27_128x128x512", "--fp16", "--steps", "30", "--save", "30",
              "--fields", "uz,phi", "--format", "bin"], tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "FP16" in r.stdout
    d = tmp_path / "bin" / "JET" / "D3Q27" / "002"
    cfg = pf.presets.jet_d3q27_128x128x512(pf.STORE_FP16)
    with pf.Simulation(cfg) as sim:
        sim.init().step(31, save=True, t0=0)
        for k in ("uz", "phi"):
            got = vtk.read_bin(str(d / f"{k}000030.bin"), cfg.shape)
            assert np.isfinite(got).all()
            np.testing.assert_array_equal(got, sim.field(k))


@pytest.mark.gpu
def test_driver_moments_jet(tmp_path):
    """pfsim JET D3Q27 <ID> --moments: the FP32 moment storage form through
    the driver; the metadata names it and the final fields equal the same
    steps run through the Python binding in that form."""
    import phasefieldlbm_amd as pf
    r = _run(["JET", "D3Q27", "005", "--preset", "jet_d3q27_128x128x512", "--moments", "--steps", "30", "--save",
              "30", "--fields", "uz,phi", "--format", "bin"], tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "moments" in r.stdout
    d = tmp_path / "bin" / "JET" / "D3Q27" / "005"
    cfg = pf.presets.jet_d3q27_128x128x512(pf.STORE_MOM)
    with pf.Simulation(cfg) as sim:
        sim.init().step(31, save=True, t0=0)
        for k in ("uz", "phi"):
            got = vtk.read_bin(str(d / f"{k}000030.bin"), cfg.shape)
            assert np.isfinite(got).all()
            np.testing.assert_array_equal(got, sim.field(k))


@pytest.mark.gpu
def test_driver_learned_closure(tmp_path):
    from phasefieldlbm_amd import closure
    r = _run(["DROPLET", "D3Q19", "002", "--preset", "droplet_d3q19_64", "--steps", "30", "--save", "30",
              "--closure", closure.WEIGHTS_RAW], tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "learned" in r.stdout
    d = tmp_path / "bin" / "DROPLET" / "D3Q19" / "002"
    rho = vtk.read_bin(str(d / "rho000030.bin"), (64, 64, 64))
    assert np.all(np.isfinite(rho))


@pytest.mark.gpu
def test_driver_derived_fields(tmp_path):
    """--derived appends the enabled derived fields to the output, as the
    reference's main appends DerivedFields::makeOutputFields (src/main.cu:130-146)."""
    r = _run(["DROPLET", "D3Q19", "003", "--preset", "droplet_d3q19_64", "--steps", "4", "--save", "4",
              "--derived", "average,vorticity", "--format", "bin"], tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    d = tmp_path / "bin" / "DROPLET" / "D3Q19" / "003"
    for k in ("rho", "uz", "avg_phi", "avg_ux", "avg_uy", "avg_uz", "vort_x", "vort_y", "vort_z", "vort_mag"):
        assert (d / f"{k}000004.bin").exists(), k
    assert not (d / "avg_uxux000004.bin").exists()
    r = _run(["DROPLET", "D3Q19", "004", "--derived", "scalar"], tmp_path)
    assert r.returncode == 1 and "derived" in r.stderr
