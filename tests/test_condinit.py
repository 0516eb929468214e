"""CondInit on the device (f3d_CondInit / f3d32_CondInit) against the host's.

The set-ups' initial states are per-ring profiles times seeded noise drawn
from glibc's drand48 in the set-up's call order.  The device reproduces the
stream by jump-ahead of the 48-bit generator and applies the same per-cell
arithmetic:

* CPU: the jump-ahead (f3d_drand48_skip, a host routine of the library) lands
  on glibc's own drand48 values after any number of draws, from srand48's
  state and from glibc's unseeded state;
* GPU: the fields of a run initialised on the device equal those of the same
  run initialised on the host, bit for bit, for the Keplerian disk (one draw
  per cell), the FARGO disk (three draws), the 3-D spherical disk and the
  gas + dust disk (no noise), in FP64 and in the precision form; the
  log-normal disk evaluates exp/log/cos with the device's own functions and
  agrees to a few ulp;
* GPU: a batch built with device initial states equals one built with the
  host's, bit for bit, after steps.
"""
import ctypes
import sys

import numpy as np
import pytest

C1 = "fargo3d_amd/setups/c1_fargo2d.par"
C2 = "fargo3d_amd/setups/c2_p3d.par"
C3 = "fargo3d_amd/setups/c3_adiabatic2d.par"
C4 = "fargo3d_amd/setups/c4_multifluid2d.par"


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _lib():
    import fargo3d_amd as fa
    L = fa.lib()
    L.f3d_drand48_skip.restype = ctypes.c_ulonglong
    L.f3d_drand48_skip.argtypes = [ctypes.c_ulonglong, ctypes.c_ulonglong]
    return L


@pytest.mark.parametrize("seed", [20240601, 7, 0xFFFFFFFF + 12345])
def test_jump_ahead_lands_on_glibc_drand48(seed):
    L = _lib()
    libc = ctypes.CDLL("libc.so.6")
    libc.drand48.restype = ctypes.c_double
    libc.srand48.argtypes = [ctypes.c_long]
    libc.srand48(seed)
    x0 = (((seed & 0xFFFFFFFF) << 16) | 0x330E)
    draws = [libc.drand48() for _ in range(5000)]
    for n in (1, 2, 3, 17, 1000, 4999, 5000):
        assert L.f3d_drand48_skip(x0, n) * 2.0 ** -48 == draws[n - 1], (seed, n)
    # and far ahead: the composition of jumps equals stepping
    a = L.f3d_drand48_skip(x0, 123456789)
    b = L.f3d_drand48_skip(L.f3d_drand48_skip(x0, 123456000), 789)
    assert a == b


def _fields(s):
    names = ["Density", "Energy", "Vx", "Vy"] + (["Vz"] if s.dimz else [])
    out = {}
    for k in range(s.nfluids):
        s.select_fluid(k)
        for n in names:
            out[(k, n)] = s.field(n)
    s.select_fluid(0)
    return out


CASES = {
    "keplerian_2d": (C1, "Nx=192 Ny=80"),
    "fargo_noise_2d": (C1, "Nx=192 Ny=80 Setup=fargo"),
    "spherical_3d": (C2, "Nx=64 Ny=24 Nz=12"),
    "gas_4dust_2d": (C4, "Nx=128 Ny=64"),
    "keplerian_2d_fp32": (C1, "Nx=192 Ny=80 Precision=32"),
    "spherical_3d_fp32": (C2, "Nx=64 Ny=24 Nz=12 Precision=32"),
}


@pytest.mark.gpu
@pytest.mark.parametrize("case", list(CASES))
def test_device_condinit_equals_host(case):
    import fargo3d_amd as fa
    par, ov = CASES[case]
    say("condinit", case)
    host = _fields(fa.Simulation(par, ov))
    dev = _fields(fa.Simulation(par, ov + " CondInitOnDevice=yes"))
    for key, a in host.items():
        b = dev[key]
        nb = int(np.count_nonzero(a.view(np.int64) != b.view(np.int64)))
        assert nb == 0, (case, key, nb)
    say("  %d fields bit-identical" % len(host))


@pytest.mark.gpu
def test_device_condinit_lognormal():
    import fargo3d_amd as fa
    ov = "Nx=512 Ny=128"
    host = _fields(fa.Simulation(C3, ov))
    dev = _fields(fa.Simulation(C3, ov + " CondInitOnDevice=yes"))
    for key, a in host.items():
        b = dev[key]
        rel = float(np.max(np.abs(a - b) / np.maximum(np.abs(a), 1e-300)))
        nb = int(np.count_nonzero(a != b))
        say("  %s: %d cells differ, max rel %.2e" % (key[1], nb, rel))
        # the device's exp/log/cos against glibc's: a few ulp
        assert rel <= 1e-15, (key, rel)
    # the velocities carry no log-normal noise: equal
    for n in ("Vx", "Vy"):
        assert np.array_equal(host[(0, n)], dev[(0, n)])


@pytest.mark.gpu
def test_batch_with_device_condinit():
    import fargo3d_amd as fa
    specs = fa.sweep_specs(6)
    a = fa.Batch(specs)
    b = fa.Batch(specs, condinit_on_device=False)
    a.step(10)
    b.step(10)
    a.sync()
    b.sync()
    for k in range(len(specs)):
        for n in ("Density", "Vx", "Vy"):
            x, y = a.instances[k].field(n), b.instances[k].field(n)
            assert np.array_equal(x.view(np.int64), y.view(np.int64)), (k, n)
    say("6 disks initialised on the device equal the host-initialised ones after 10 steps")
    a.close()
    b.close()
