"""The CFL time step with no allocation (f3d_fused_cfl_fluids_noalloc)
against f3d_fused_cfl_fluids, through the C ABI: the same step bit for bit
and the same ring means (VxMed), on two calls in a row, the second on a
smaller mesh, and a mesh larger than its fixed scratch refused."""
import ctypes
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

C1 = "fargo3d_amd/setups/c1_fargo2d.par"
C2 = "fargo3d_amd/setups/c2_p3d.par"
C3 = "fargo3d_amd/setups/c3_adiabatic2d.par"
C4 = "fargo3d_amd/setups/c4_multifluid2d.par"


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _tables(s):
    P = ctypes.c_void_p * 8
    cs, rho, vx, vy, vz = P(), P(), P(), P(), P()
    for k in range(s.nfluids):
        s.select_fluid(k)
        cs[k], rho[k] = s.device_ptr("Energy"), s.device_ptr("Density")
        vx[k], vy[k] = s.device_ptr("Vx"), s.device_ptr("Vy")
        vz[k] = s.device_ptr("Vz") if s.dimz else s.device_ptr("Vy")
    vxm = s.device_ptr("Vx")  # the last fluid's (src/main.c:366 then :381)
    s.select_fluid(0)
    return cs, rho, vx, vy, vz, vxm


def _both(fa, s):
    L = fa.lib()
    for f in (L.f3d_fused_cfl_fluids, L.f3d_fused_cfl_fluids_noalloc):
        f.restype = ctypes.c_int
        f.argtypes = [ctypes.c_void_p, ctypes.c_int] + [ctypes.c_void_p] * 6 + [ctypes.POINTER(ctypes.c_double)]
    cs, rho, vx, vy, vz, vxm = _tables(s)
    a, b = ctypes.c_double(), ctypes.c_double()
    assert L.f3d_fused_cfl_fluids(s.ctx, s.nfluids, cs, rho, vx, vy, vz, vxm, ctypes.byref(a)) == 0
    vm_a = s.field("VxMed").copy()
    s.set_field("VxMed", np.zeros_like(vm_a))
    rc = L.f3d_fused_cfl_fluids_noalloc(s.ctx, s.nfluids, cs, rho, vx, vy, vz, vxm, ctypes.byref(b))
    assert rc == 0, L.f3d_hip_last_error(s.ctx)
    return a.value, b.value, vm_a, s.field("VxMed")


@pytest.fixture(scope="module")
def fa():
    import fargo3d_amd
    return fargo3d_amd


@pytest.mark.parametrize("cases", [[(C3, "Nx=1024 Ny=200"), (C1, "Nx=96 Ny=48")],
                                   [(C2, "Nx=64 Ny=40 Nz=16"), (C4, "Nx=128 Ny=64")]],
                         ids=["long_rings_then_short", "3d_then_five_fluids"])
def test_noalloc_cfl_matches_the_allocating_form(fa, cases):
    for par, ov in cases:
        s = fa.Simulation(par, ov)
        for n in range(3):
            say("noalloc cfl", par, ov, "step", n)
            a, b, vm_a, vm_b = _both(fa, s)
            assert a == b and a > 0, (par, ov, n, a, b)
            assert np.array_equal(vm_a, vm_b)
            s.step()
        s.close()


def test_noalloc_cfl_refuses_meshes_beyond_its_scratch(fa):
    """8 fluids x 34,816 rings exceed the 262,144 ring minima the fixed array
    holds: -2 and a message, before any launch."""
    s = fa.Simulation(C2, "Nx=4 Ny=272 Nz=128")
    L = fa.lib()
    L.f3d_fused_cfl_fluids_noalloc.restype = ctypes.c_int
    cs, rho, vx, vy, vz, vxm = _tables(s)
    for k in range(1, 8):
        cs[k], rho[k], vx[k], vy[k], vz[k] = cs[0], rho[0], vx[0], vy[0], vz[0]
    out = ctypes.c_double()
    rc = L.f3d_fused_cfl_fluids_noalloc(s.ctx, 8, cs, rho, vx, vy, vz, vxm, ctypes.byref(out))
    assert rc == -2
    assert b"exceed the fixed scratch" in L.f3d_hip_last_error(s.ctx)
    s.close()
