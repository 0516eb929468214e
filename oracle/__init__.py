"""CPU restatement of FARGO3D's hydro time step -- TEST INFRASTRUCTURE ONLY.

This package is the checker for the HIP product path.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
load it; the product package ``fargo3d_amd`` never imports it.

The C sources (``oracle_kernels.c``, ``oracle_step.c``) restate the
reference's sequential CPU routines (each function cites its reference
file:line).  ``make -C oracle`` builds ``oracle/lib/liboracle.so``.

Parity status (how each part is pinned; see tests/test_oracle.py):
  * mesh / metric arrays          -- closed forms (linear spacing, exact areas)
  * van Leer / PPA reconstructions -- closed forms, constants, monotonicity
  * Update*/UpdateDensity*        -- flux-form conservation (exact telescoping)
  * AdvectSHIFT                   -- brute-force permutation, bit-exact
  * ComputeResidual / ComputeVmed -- brute-force sequential sums, |n~-n| <= 1/2
  * cfl                           -- closed form per cell
  * SubStep1_x/y/z                -- closed forms of the pressure, potential,
                                     centrifugal and frame terms
  * SubStep2_a/b                  -- one compressed cell: q, opposite face kicks,
                                     heating
  * SubStep3                      -- closed form e(1-t)/(1+t)
  * viscosity                     -- solid-body rotation gives tau_rphi = 0
  * Collisions / dust diffusion   -- two-fluid closed form, direction-wise drag;
                                     uniform dust-to-gas ratio kept, mass conserved
  * boundaries                    -- closed forms of the ghost formulas
  * planets (host code)           -- Kepler orbit closed form
  * whole step                    -- invariants (mass conservation without
                                     damping, axisymmetric steady states);
                                     the reference binary itself is not
                                     buildable here (its build needs generated
                                     code), so no bit-level reference output pins
                                     the whole step ("parity unpinned" at that
                                     level beyond the invariants above).
  * parity unpinned individually: the Y/Z parts of SubStep2_b (pinned only through the whole-step invariants and the
    bit-exact agreement of two independent implementations, the oracle and
    the HIP kernels, written from the reference routines).
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "lib", "liboracle.so")
_lib = None

NAMES3D = ["Density", "Energy", "Vx", "Vy", "Vz", "Vx_temp", "Vy_temp", "Vz_temp",
           "Mpx", "Mmx", "Mpy", "Mmy", "Mpz", "Mmz", "Pot", "Slope", "DivRho",
           "DensStar", "Qs", "Pressure", "Total_Density", "QL", "QR"]
NAMES2D = ["Density0", "Energy0", "Vx0", "Vy0", "Vz0", "Reduction2D", "VxMed"]


def build():
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        vp, d, i = ctypes.c_void_p, ctypes.c_double, ctypes.c_int
        L.orc_create.restype = vp
        L.orc_create.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, i]
        L.orc_destroy.argtypes = [vp]
        L.orc_step.restype = d
        L.orc_step.argtypes = [vp]
        L.orc_field.restype = ctypes.POINTER(d)
        L.orc_field.argtypes = [vp, ctypes.c_char_p]
        L.orc_nshift.restype = ctypes.POINTER(i)
        L.orc_nshift.argtypes = [vp]
        L.orc_get_scalars.argtypes = [vp, ctypes.POINTER(d)]
        L.orc_set_scalars.argtypes = [vp, ctypes.POINTER(d)]
        L.orc_get_planets.argtypes = [vp, ctypes.POINTER(d), i]
        _lib = L
    return _lib


class Oracle:
    """One CPU simulation (a FARGO3D run restated)."""

    def __init__(self, parfile, overrides=""):
        L = lib()
        err = ctypes.create_string_buffer(512)
        self._s = L.orc_create(parfile.encode(), overrides.encode(), err, 512)
        if not self._s:
            raise RuntimeError("oracle: " + err.value.decode())
        self.L = L
        # grid shape from the state struct: read via Density length is not
        # available, so mirror the host layer's rules
        self.shape = None

    def close(self):
        if self._s:
            self.L.orc_destroy(self._s)
            self._s = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_shape(self, nx, ny, nz, dimz):
        sy = ny + 6
        sz = nz + 6 if dimz else 1
        self.shape = (sz, sy, nx)
        self.shape2d = (sz, sy)

    def select_fluid(self, n):
        """SelectFluid (src/LowTasks.c:565-593): field() then reads fluid n."""
        self.L.orc_select_fluid.argtypes = [ctypes.c_void_p, ctypes.c_int]
        self.L.orc_select_fluid(self._s, int(n))

    def field(self, name):
        p = self.L.orc_field(self._s, name.encode())
        if not p:
            raise KeyError(name)
        shp = self.shape2d if name in NAMES2D else self.shape
        n = int(np.prod(shp))
        return np.ctypeslib.as_array(p, shape=(n,)).reshape(shp)

    def nshift(self):
        p = self.L.orc_nshift(self._s)
        n = int(np.prod(self.shape2d))
        return np.ctypeslib.as_array(p, shape=(n,)).reshape(self.shape2d)

    def step(self):
        dt = self.L.orc_step(self._s)
        if dt < 0:
            raise RuntimeError("oracle: null time step")
        return dt

    def scalars(self):
        out = (ctypes.c_double * 8)()
        self.L.orc_get_scalars(self._s, out)
        keys = ["PhysicalTime", "OMEGAFRAME", "OMEGAFRAME0", "Domega", "dtemp",
                "StepTime", "XAxisRotationAngle", "Timestepcount"]
        return dict(zip(keys, list(out)))

    def set_scalars(self, d):
        cur = self.scalars()
        cur.update(d)
        keys = ["PhysicalTime", "OMEGAFRAME", "OMEGAFRAME0", "Domega", "dtemp",
                "StepTime", "XAxisRotationAngle", "Timestepcount"]
        arr = (ctypes.c_double * 8)(*[cur[k] for k in keys])
        self.L.orc_set_scalars(self._s, arr)

    def planets(self, maxn=16):
        out = (ctypes.c_double * (8 * maxn))()
        self.L.orc_get_planets(self._s, out, maxn)
        a = np.array(out).reshape(maxn, 8)
        nb = int(a[0, 7]) if a[0, 7] > 0 else 0
        return a[:nb, :7]

    def call(self, fname, *args):
        """Call one restated routine by its reference name (orc_<name>)."""
        f = getattr(self.L, "orc_" + fname)
        f.restype = None
        cargs = [ctypes.c_void_p(self._s)]
        for a in args:
            if isinstance(a, float):
                cargs.append(ctypes.c_double(a))
            elif isinstance(a, int):
                cargs.append(ctypes.c_int(a))
            elif isinstance(a, str):
                cargs.append(self.L.orc_field(self._s, a.encode()))
            else:
                cargs.append(a)
        return f(*cargs)
