"""FARGO3D's function-pointer table bound to the library, driven like the
reference's main loop.

tests/bind/ holds the binding INTEGRATION.md describes (hip_bind.c),
compiled against FARGO3D's pointer types (fargo3d_api.h restates
src/global.h:303-416), and a miniature FARGO3D host (fargo_host.c) whose DT
loop body calls every routine through the table.  Its run of the 2-D
drop-in disk and of the 3-D disk must equal the CPU restatement bit for bit.
"""
import ctypes
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "tests", "bind", "lib", "libfargo_bind.so")
C1 = "fargo3d_amd/setups/c1_fargo2d.par"
C2 = "fargo3d_amd/setups/c2_p3d.par"


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def bind():
    if not os.path.exists(LIB):
        pytest.fail("tests/bind/lib/libfargo_bind.so is not built (make -C tests/bind)")
    return LIB


def run_bound(path, par, ov):
    """The host's table and globals are process state: one run at a time,
    released with bind_teardown."""
    L = ctypes.CDLL(path)
    L.bind_setup.restype = ctypes.c_int
    L.bind_setup.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int]
    L.bind_step.restype = ctypes.c_double
    L.bind_error.restype = ctypes.c_char_p
    L.bind_field.argtypes = [ctypes.c_char_p, ctypes.c_void_p]
    L.bind_planets.argtypes = [ctypes.c_void_p]
    err = ctypes.create_string_buffer(512)
    assert L.bind_setup(par.encode(), ov.encode(), err, 512) == 0, err.value
    return L


@pytest.mark.parametrize("par,ov,nsteps", [(C1, "Nx=192 Ny=64", 8), (C2, "Nx=64 Ny=24 Nz=12", 4)],
                         ids=["c1_2d", "c2_3d"])
def test_pointer_table_run_matches_oracle(bind, par, ov, nsteps):
    import oracle
    L = run_bound(bind, par, ov)
    o = oracle.Oracle(par, ov)
    names = ["Density", "Vx", "Vy"] + (["Vz"] if o.dimz else [])
    for n in range(1, nsteps + 1):
        dts, dto = L.bind_step(), o.step()
        say("bound step %d dt %r" % (n, dts))
        assert dts == dto, (n, dts, dto, L.bind_error())
    for f in names:
        ref = o.field(f)
        got = np.empty(ref.size)
        assert L.bind_field(f.encode(), got.ctypes.data) == 0
        got = got.reshape(ref.shape)
        nb = int(np.count_nonzero(got.view(np.int64) != ref.view(np.int64)))
        say("  %s: %d bits differ" % (f, nb))
        assert nb == 0, f
    pl = np.zeros(7 * 16)
    L.bind_planets(pl.ctypes.data)
    L.bind_teardown()
    assert np.array_equal(pl[:7], o.planets()[0])
