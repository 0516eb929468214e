"""The C ABI library: it loads, exports what the headers declare, and refuses
to run without a GPU (no CPU fallback)."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    text = open(os.path.join(ROOT, header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(f3d_\w+)\s*\(", text)))


@pytest.fixture(scope="module")
def lib():
    import fargo3d_amd
    path = fargo3d_amd.library_path()
    if not os.path.exists(path):
        fargo3d_amd.build()
    return fargo3d_amd.lib()


def test_exports_every_declared_symbol(lib):
    names = declared("include/fargo3d_hip.h") + declared("fargo3d_amd/csrc/host/f3d_driver.h")
    assert len(names) > 60
    out = subprocess.run(["nm", "-D", "--defined-only", lib._name], stdout=subprocess.PIPE, text=True).stdout
    exported = set(line.split()[-1] for line in out.splitlines() if line.strip())
    missing = [n for n in names if n not in exported]
    assert not missing, missing


def test_no_silent_cpu_path(lib):
    import fargo3d_amd
    if lib.f3d_hip_device_count() > 0:
        pytest.skip("a GPU is visible")
    with pytest.raises(fargo3d_amd.HipError):
        fargo3d_amd.Simulation("fargo3d_amd/setups/c1_fargo2d.par")


def test_routine_entry_points_cover_the_reference_table():
    """Every function-pointer group of std/func_arch.cfg on the hydro path has an entry point."""
    names = set(declared("include/fargo3d_hip.h"))
    for r in ["SubStep1_x", "SubStep1_y", "SubStep1_z", "SubStep2_a", "SubStep2_b", "SubStep3",
              "ComputePressureFieldIso", "ComputePressureFieldAd", "DivideByRho",
              "VanLeerX_a", "VanLeerX_b", "VanLeerY_a", "VanLeerY_b", "VanLeerZ_a", "VanLeerZ_b",
              "VanLeerX_PPA_a", "VanLeerX_PPA_b", "VanLeerX_PPA_c", "VanLeerX_PPA_d",
              "momenta_x", "momenta_y", "momenta_z", "UpdateX", "UpdateY", "UpdateZ",
              "UpdateDensityX", "UpdateDensityY", "UpdateDensityZ", "NewVelocity_x", "NewVelocity_y",
              "NewVelocity_z", "ComputeResidual", "AdvectSHIFT", "ComputeVweight", "ComputeVmed",
              "reduction_SUM", "reduction_MIN", "reduction_full_SUM", "reduction_full_MIN", "cfl",
              "Potential", "CorrectVtheta", "ComputeForce", "visctensor_cyl", "visctensor_sph",
              "addviscosity_cyl", "addviscosity_sph", "StockholmBoundary", "init_stockholm",
              "copy_velocities", "boundaries", "ComputeTotalDensity", "Reset_field"]:
        assert "f3d_" + r in names, r


def test_par_reader_and_overrides():
    import oracle
    o = oracle.Oracle("fargo3d_amd/setups/c1_fargo2d.par", "nx=32 NY=16 cfl=0.3")
    o.set_shape(32, 16, 1, 0)
    assert o.field("Density").shape == (1, 22, 32)
    with pytest.raises(RuntimeError):
        oracle.Oracle("fargo3d_amd/setups/c1_fargo2d.par", "NoSuchParameter=1")
    with pytest.raises(RuntimeError):
        oracle.Oracle("fargo3d_amd/setups/does_not_exist.par")


def test_nsnap_is_refused(lib):
    """NSNAP > 0 (a SPECIFIC output every NSNAP steps with TimeStep renumbered,
    src/main.c:347-353) is not implemented by the main loop: the set-up is
    refused with a message naming it, before any device is touched."""
    err = ctypes.create_string_buffer(512)
    h = lib.f3d_sim_create(b"fargo3d_amd/setups/c1_fargo2d.par", b"Nx=32 Ny=20 Nsnap=5", err, 512)
    assert not h
    assert b"NSNAP" in err.value


def test_binding_compiles_against_the_pointer_types():
    """tests/bind/hip_bind.c assigns every *_hip wrapper to a pointer of the
    reference's table (restated from src/global.h:303-416); with
    -Werror=incompatible-pointer-types a wrapper whose type differs from its
    pointer's fails to compile."""
    d = os.path.join(ROOT, "tests", "bind")
    r = subprocess.run(["gcc", "-fsyntax-only", "-std=gnu11", "-Wall", "-Werror=incompatible-pointer-types",
                        "-Werror=implicit-function-declaration", "-I" + os.path.join(ROOT, "include"),
                        "-I" + os.path.join(ROOT, "fargo3d_amd", "csrc", "host"),
                        os.path.join(d, "hip_bind.c"), os.path.join(d, "fargo_host.c")],
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert r.returncode == 0, r.stdout
