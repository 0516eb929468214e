"""fargo3d_amd -- FARGO3D's hydro time step on AMD MI355X (gfx950).

The product is the shared library ``fargo3d_amd/lib/libfargo3d_amd.so``:
hand-written HIP kernels behind the C ABI of ``include/fargo3d_hip.h`` and a C
step driver (``csrc/host/f3d_driver.h``) that keeps FARGO3D's host structure
(.par set-up, DT loop, per-routine function table).  This module is a thin
ctypes binding over that library; it loads nothing else and has no fallback:
without the built library, or without a HIP device, it raises.
"""
import ctypes
import os

import numpy as np

__all__ = ["Simulation", "Batch", "sweep_specs", "lib", "library_path", "build", "HipError", "SCALAR_KEYS"]

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "lib", "libfargo3d_amd.so")
_lib = None

SCALAR_KEYS = ["PhysicalTime", "OMEGAFRAME", "OMEGAFRAME0", "Domega", "dtemp",
               "StepTime", "XAxisRotationAngle", "Timestepcount"]
FIELDS_2D = {"Density0", "Energy0", "Vx0", "Vy0", "Vz0", "Reduction2D", "VxMed"}


class HipError(RuntimeError):
    pass


def library_path():
    return _LIB


def build(force=True):
    from . import buildlib as _b
    return _b.build(force=force)


def lib():
    """The loaded product library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            raise HipError("fargo3d_amd: %s is missing; run fargo3d_amd.build() first" % _LIB)
        L = ctypes.CDLL(_LIB, mode=ctypes.RTLD_GLOBAL)
        vp, d, i, cp = ctypes.c_void_p, ctypes.c_double, ctypes.c_int, ctypes.c_char_p
        L.f3d_sim_create.restype = vp
        L.f3d_sim_create.argtypes = [cp, cp, cp, i]
        L.f3d_sim_destroy.argtypes = [vp]
        L.f3d_sim_error.restype = cp
        L.f3d_sim_error.argtypes = [vp]
        L.f3d_sim_step.argtypes = [vp, ctypes.POINTER(d)]
        L.f3d_sim_run.argtypes = [vp, i]
        L.f3d_sim_set_arch.argtypes = [vp, cp, cp]
        L.f3d_sim_read_arch.argtypes = [vp, cp]
        L.f3d_sim_field_size.restype = ctypes.c_long
        L.f3d_sim_field_size.argtypes = [vp, cp]
        L.f3d_sim_get_field.argtypes = [vp, cp, vp]
        L.f3d_sim_set_field.argtypes = [vp, cp, vp]
        L.f3d_sim_get_nshift.argtypes = [vp, vp]
        L.f3d_sim_set_nshift.argtypes = [vp, vp]
        L.f3d_sim_set_planets.argtypes = [vp, vp, i]
        L.f3d_sim_device_field.restype = vp
        L.f3d_sim_device_field.argtypes = [vp, cp]
        L.f3d_sim_get_scalars.argtypes = [vp, vp]
        L.f3d_sim_set_scalars.argtypes = [vp, vp]
        L.f3d_sim_get_planets.argtypes = [vp, vp, i]
        L.f3d_sim_shape.argtypes = [vp, vp]
        L.f3d_sim_eos.argtypes = [vp]
        L.f3d_sim_force_dt.argtypes = [vp, d]
        L.f3d_sim_nfluids.argtypes = [vp]
        L.f3d_sim_select_fluid.argtypes = [vp, i]
        L.f3d_sim_ctx.restype = vp
        L.f3d_sim_ctx.argtypes = [vp]
        L.f3d_sim_push_phys.argtypes = [vp]
        L.f3d_sim_write_field.argtypes = [vp, cp, cp]
        L.f3d_sim_write_outputs.argtypes = [vp, i, cp]
        L.f3d_sim_write_init.argtypes = [vp, cp]
        L.f3d_sim_restart.argtypes = [vp, i, cp]
        L.f3d_sim_advance.argtypes = [vp]
        L.f3d_sim_main_loop.argtypes = [vp, i, cp]
        L.f3d_sim_monitor.argtypes = [vp, i, i, i, i, cp]
        for fn in (L.f3d_write_active, L.f3d_read_active):
            fn.argtypes = [vp, i, i, i, i, i, ctypes.c_long, ctypes.c_long, cp]
        L.f3d_hip_last_error.restype = cp
        L.f3d_hip_last_error.argtypes = [vp]
        L.f3d_hip_sync.argtypes = [vp]
        L.f3d_hip_device_count.restype = i
        _lib = L
    return _lib


class Simulation:
    """One device-resident run of the hydro step (a FARGO3D run on MI355X)."""

    def __init__(self, parfile, overrides="", arch=None, _handle=None):
        L = lib()
        if L.f3d_hip_device_count() < 1:
            raise HipError("fargo3d_amd: no HIP device is visible")
        self._owned = _handle is None
        if _handle is None:
            err = ctypes.create_string_buffer(512)
            self._s = L.f3d_sim_create(parfile.encode(), overrides.encode(), err, 512)
            if not self._s:
                raise HipError("fargo3d_amd: " + err.value.decode())
        else:
            self._s = _handle
        self.L = L
        shp = (ctypes.c_int * 4)()
        L.f3d_sim_shape(self._s, shp)
        self.nx, self.ny, self.nz, self.dimz = list(shp)
        self.adiabatic = (L.f3d_sim_eos(self._s) == 2)
        self.nfluids = L.f3d_sim_nfluids(self._s)
        sy = self.ny + 6
        sz = self.nz + 6 if self.dimz else 1
        self.shape = (sz, sy, self.nx)
        self.shape2d = (sz, sy)
        for g, m in (arch or {}).items():
            self.set_arch(g, m)

    def _check(self, rc):
        if rc != 0:
            raise HipError(self.L.f3d_sim_error(self._s).decode())

    def close(self):
        if getattr(self, "_s", None):
            if self._owned:
                self.L.f3d_sim_destroy(self._s)
            self._s = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def select_fluid(self, n):
        """SelectFluid (src/LowTasks.c:565): field access then refers to fluid n (0 = gas)."""
        self._check(self.L.f3d_sim_select_fluid(self._s, int(n)))

    @property
    def ctx(self):
        return self.L.f3d_sim_ctx(self._s)

    def set_arch(self, group, mode):
        self._check(self.L.f3d_sim_set_arch(self._s, group.encode(), mode.encode()))

    def read_arch(self, path):
        self._check(self.L.f3d_sim_read_arch(self._s, path.encode()))

    def step(self):
        dt = ctypes.c_double()
        self._check(self.L.f3d_sim_step(self._s, ctypes.byref(dt)))
        return dt.value

    def force_dt(self, dt):
        """Use dt as the next step's time step before the output-time clamp
        (paired-run measurements: pass another run's step_time(), so that
        both clamp to the output times identically)."""
        self._check(self.L.f3d_sim_force_dt(self._s, float(dt)))

    def step_time(self):
        """The last step's time step before the output-time clamp (StepTime);
        step() returns the clamped one."""
        return self.scalars()["StepTime"]

    def run(self, nsteps):
        self._check(self.L.f3d_sim_run(self._s, int(nsteps)))

    def sync(self):
        if self.L.f3d_hip_sync(self.ctx) != 0:
            raise HipError(self.L.f3d_hip_last_error(self.ctx).decode())

    def field(self, name):
        shp = self.shape2d if name in FIELDS_2D else self.shape
        out = np.empty(shp, dtype=np.float64)
        self._check(self.L.f3d_sim_get_field(self._s, name.encode(), out.ctypes.data))
        return out

    def set_field(self, name, arr):
        shp = self.shape2d if name in FIELDS_2D else self.shape
        a = np.ascontiguousarray(arr, dtype=np.float64).reshape(shp)
        self._check(self.L.f3d_sim_set_field(self._s, name.encode(), a.ctypes.data))

    def device_ptr(self, name):
        p = self.L.f3d_sim_device_field(self._s, name.encode())
        if not p:
            raise KeyError(name)
        return p

    def nshift(self):
        out = np.empty(self.shape2d, dtype=np.int32)
        self._check(self.L.f3d_sim_get_nshift(self._s, out.ctypes.data))
        return out

    def set_nshift(self, arr):
        a = np.ascontiguousarray(arr, dtype=np.int32).reshape(self.shape2d)
        self._check(self.L.f3d_sim_set_nshift(self._s, a.ctypes.data))

    def set_planets(self, arr):
        a = np.zeros((len(arr), 8))
        a[:, :7] = arr
        self._check(self.L.f3d_sim_set_planets(self._s, a.ctypes.data, len(arr)))

    def scalars(self):
        out = (ctypes.c_double * 8)()
        self.L.f3d_sim_get_scalars(self._s, out)
        return dict(zip(SCALAR_KEYS, list(out)))

    def set_scalars(self, d):
        cur = self.scalars()
        cur.update(d)
        arr = (ctypes.c_double * 8)(*[cur[k] for k in SCALAR_KEYS])
        self.L.f3d_sim_set_scalars(self._s, arr)
        self._check(self.L.f3d_sim_push_phys(self._s))

    def push_phys(self):
        self._check(self.L.f3d_sim_push_phys(self._s))

    def planets(self, maxn=16):
        out = (ctypes.c_double * (8 * maxn))()
        self.L.f3d_sim_get_planets(self._s, out, maxn)
        a = np.array(out).reshape(maxn, 8)
        nb = int(a[0, 7]) if a[0, 7] > 0 else 0
        return a[:nb, :7]

    def write_field(self, name, path):
        self._check(self.L.f3d_sim_write_field(self._s, name.encode(), path.encode()))

    def write_outputs(self, n, outdir):
        """Planet files and WriteOutputs(ALL) of every fluid for output n
        (src/main.c:328-334): <fluid>dens<n>.dat, ... in merged raw format."""
        self._check(self.L.f3d_sim_write_outputs(self._s, int(n), str(outdir).encode()))

    def write_init(self, outdir):
        """Empty planet files and the damping references' 0_2d files, as a
        fresh run writes them (src/main.c:262-266, src/init_stockholm.c)."""
        self._check(self.L.f3d_sim_write_init(self._s, str(outdir).encode()))

    def restart(self, n, outdir):
        """Restart from output n of outdir (src/main.c:249-260); returns
        begin_i = n * NINTERM."""
        rc = self.L.f3d_sim_restart(self._s, int(n), str(outdir).encode())
        self._check(-1 if rc < 0 else 0)
        return rc

    def advance(self):
        """One pass of the DT loop (src/main.c:361-448); returns its step count."""
        rc = self.L.f3d_sim_advance(self._s)
        self._check(-1 if rc < 0 else 0)
        return rc

    def monitor(self, m2d, y_raw, scalar, timestep, outdir):
        """MonitorGlobal of every fluid (src/monitor.c:236-254) for the
        quantity bits (MASS 1, MOM_X 64, MOM_Y 128, MOM_Z 256, TORQ 512) of
        each kind, into outdir/monitor/<fluid>/."""
        self._check(self.L.f3d_sim_monitor(self._s, int(m2d), int(y_raw), int(scalar), int(timestep),
                                           str(outdir).encode()))

    def main_loop(self, begin_i, outdir):
        """The main loop from begin_i to NTOT with its outputs (src/main.c:324-356)."""
        self._check(self.L.f3d_sim_main_loop(self._s, int(begin_i), str(outdir).encode()))

    def call(self, routine, *args):
        """Call one C-ABI entry point f3d_<routine>(ctx, ...) by its reference
        name.  str arguments name a bound field (its device pointer)."""
        f = getattr(self.L, "f3d_" + routine)
        f.restype = ctypes.c_int
        cargs = [ctypes.c_void_p(self.ctx)]
        for a in args:
            if isinstance(a, float):
                cargs.append(ctypes.c_double(a))
            elif isinstance(a, bool) or isinstance(a, int):
                cargs.append(ctypes.c_int(int(a)))
            elif isinstance(a, str):
                cargs.append(ctypes.c_void_p(self.device_ptr(a)))
            else:
                cargs.append(a)
        rc = f(*cargs)
        if rc != 0:
            raise HipError("%s: %s" % (routine, self.L.f3d_hip_last_error(self.ctx).decode()))
        return rc


class Batch:
    """Many independent 2-D isothermal disks advanced together in the FP32
    form: each stage of the step is one launch over every instance, and the
    step (time steps included) is one hipGraph replay, planets included.
    The initial states are computed on the device (CondInitOnDevice) unless
    condinit_on_device is False."""

    def __init__(self, specs, condinit_on_device=True):
        L = lib()
        vp, cp, i = ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int
        L.f3d_batch_create.restype = vp
        L.f3d_batch_create.argtypes = [i, ctypes.POINTER(cp), ctypes.POINTER(cp), cp, i]
        L.f3d_batch_destroy.argtypes = [vp]
        L.f3d_batch_error.restype = cp
        L.f3d_batch_error.argtypes = [vp]
        L.f3d_batch_instance.restype = vp
        L.f3d_batch_instance.argtypes = [vp, i]
        L.f3d_batch_step.argtypes = [vp, i]
        L.f3d_batch_sync.argtypes = [vp]
        L.f3d_batch_clock.argtypes = [vp, i, vp]
        L.f3d_batch_set_tiling.argtypes = [vp, i, i]
        L.f3d_batch_tune.argtypes = [vp, i, vp, vp]
        L.f3d_batch_planets.argtypes = [vp, i, vp, i, vp]
        self.L = L
        n = len(specs)
        pars = (cp * n)(*[p.encode() for p, _ in specs])
        extra = " CondInitOnDevice=%s" % ("yes" if condinit_on_device else "no")
        ovs = (cp * n)(*[(o + extra).encode() for _, o in specs])
        err = ctypes.create_string_buffer(512)
        self._b = L.f3d_batch_create(n, pars, ovs, err, 512)
        if not self._b:
            raise HipError("fargo3d_amd: " + err.value.decode())
        self.n = n
        self.instances = [Simulation(None, _handle=L.f3d_batch_instance(self._b, k)) for k in range(n)]
        self.cells = [s.nx * s.ny for s in self.instances]

    def _check(self, rc):
        if rc != 0:
            raise HipError(self.L.f3d_batch_error(self._b).decode())

    def step(self, nsteps=1):
        self._check(self.L.f3d_batch_step(self._b, int(nsteps)))

    def sync(self):
        self._check(self.L.f3d_batch_sync(self._b))

    def clock(self, k):
        out = (ctypes.c_double * 3)()
        self._check(self.L.f3d_batch_clock(self._b, k, out))
        return {"dt": out[0], "dtemp": out[1], "PhysicalTime": out[2]}

    def planets(self, k, maxn=8):
        """Planets of instance k (x, y, z, vx, vy, vz, mass per row) and the
        frame rate, from the device."""
        out = (ctypes.c_double * (8 * maxn))()
        of = ctypes.c_double()
        nb = self.L.f3d_batch_planets(self._b, k, out, maxn, ctypes.byref(of))
        if nb < 0:
            raise HipError(self.L.f3d_batch_error(self._b).decode())
        return np.array(out).reshape(maxn, 8)[:nb, :7], of.value

    def set_tiling(self, rows_y, tile):
        self._check(self.L.f3d_batch_set_tiling(self._b, int(rows_y), int(tile)))

    def tune(self, reps=5):
        r, t = ctypes.c_int(), ctypes.c_int()
        self._check(self.L.f3d_batch_tune(self._b, int(reps), ctypes.byref(r), ctypes.byref(t)))
        return r.value, t.value

    def close(self):
        if getattr(self, "_b", None):
            for s in self.instances:
                s._s = None
            self.L.f3d_batch_destroy(self._b)
            self._b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass



C1_PAR = os.path.join("fargo3d_amd", "setups", "c1_fargo2d.par")


def sweep_specs(n, seed=20240605, parfile=C1_PAR):
    """A parameter sweep of n 2-D isothermal planet-disk set-ups for the
    batched form: grids Nx in {256, 384, 768, 1536} x Ny in {128, 192, 384},
    planet mass log-uniform in [1e-5, 3e-3] M*, viscosity log-uniform in
    [1e-6, 1e-4] and aspect ratio uniform in [0.03, 0.08], drawn from a seeded
    generator.  Returns (parfile, overrides) pairs."""
    rng = np.random.default_rng(seed)
    specs = []
    for _ in range(n):
        nx = int(rng.choice([256, 384, 768, 1536]))
        ny = int(rng.choice([128, 192, 384]))
        mp = 10 ** rng.uniform(np.log10(1e-5), np.log10(3e-3))
        nu = 10 ** rng.uniform(-6, -4)
        h = rng.uniform(0.03, 0.08)
        specs.append((parfile, "Nx=%d Ny=%d PlanetMass=%.6e Nu=%.6e AspectRatio=%.6f" % (nx, ny, mp, nu, h)))
    return specs
