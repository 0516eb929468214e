"""The reduced-precision form (FP32 fields and stencils; FP64 ring sums,
residual/shift split and CFL time step) against the FP64 path.

The FP64 fused path is bit-identical to the CPU restatement of the reference
(test_gpu_parity.py), so it stands in for it here.  Two measurements, each
printing the relative L1 error sum|a-b|/sum|b| and the relative L-infinity
error max|a-b|/max|b| of every field over the active cells:

* paired: the FP32 run takes the FP64 run's time step each step, so the
  difference is the arithmetic alone;
* free: each run picks its own time step and both stop at the same physical
  time (the output-time clamp of src/main.c:394-395).
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

C1 = "fargo3d_amd/setups/c1_fargo2d.par"
C2 = "fargo3d_amd/setups/c2_p3d.par"
C3 = "fargo3d_amd/setups/c3_adiabatic2d.par"
CASES = [(C1, "", 100), (C2, "Nx=64 Ny=24 Nz=12", 40), (C3, "Nx=256 Ny=96", 60)]
IDS = ["c1_2d", "c2_3d_small", "c3_2d_adiabatic_small"]

# bounds (L1, Linf) on the relative errors, about 3-5x the values measured on
# MI355X (printed by the tests).  FP32 storage of the velocities limits the
# transport to ~6e-8 absolute per step, which the residual (orbital) advection
# turns into density errors of ~1e-7 of the density per step.
PAIRED = {"Density": (1e-4, 1e-4), "Energy": (1e-4, 1e-4), "Vx": (2e-5, 5e-5), "Vy": (5e-3, 2e-3),
          "Vz": (5e-3, 2e-3)}
FREE = {"Density": (5e-4, 5e-4), "Energy": (5e-4, 5e-4), "Vx": (2e-4, 5e-4), "Vy": (5e-2, 2e-2),
        "Vz": (5e-2, 2e-2)}


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def active(x, dimz):
    return x[3:-3, 3:-3, :] if dimz else x[:, 3:-3, :]


def errors(a, b):
    d = np.abs(a - b)
    l1 = float(d.sum() / max(np.abs(b).sum(), 1e-300))
    linf = float(d.max() / max(np.abs(b).max(), 1e-300))
    return l1, linf


def compare(tag, s32, s64, bounds):
    fields = ["Density", "Vx", "Vy"] + (["Vz"] if s64.dimz else []) + (["Energy"] if s64.adiabatic else [])
    bad = []
    for f in fields:
        a, b = active(s32.field(f), s64.dimz), active(s64.field(f), s64.dimz)
        l1, linf = errors(a, b)
        say("%s %-7s L1 %.3e  Linf %.3e" % (tag, f, l1, linf))
        if l1 > bounds[f][0] or linf > bounds[f][1]:
            bad.append((f, l1, linf))
    assert not bad, bad


@pytest.mark.parametrize("par,ov,nsteps", CASES, ids=IDS)
def test_fp32_paired_steps(par, ov, nsteps):
    import fargo3d_amd as fa
    s64 = fa.Simulation(par, ov)
    s32 = fa.Simulation(par, (ov + " Precision=32").strip())
    for _ in range(nsteps):
        dt = s64.step()
        s32.force_dt(s64.step_time())
        assert s32.step() == dt
    compare("%s paired %d steps" % (par.split("/")[-1], nsteps), s32, s64, PAIRED)


@pytest.mark.parametrize("par,ov,nsteps", CASES, ids=IDS)
def test_fp32_free_same_time(par, ov, nsteps):
    import fargo3d_amd as fa
    probe = fa.Simulation(par, ov)
    t_end = sum(probe.step() for _ in range(nsteps))
    probe.close()
    ovt = (ov + " DT=%r" % t_end).strip()
    runs = []
    for o in (ovt, ovt + " Precision=32"):
        s = fa.Simulation(par, o)
        n = 0
        while s.scalars()["PhysicalTime"] < t_end * (1 - 1e-12):
            s.step()
            n += 1
            assert n < 4 * nsteps
        runs.append((s, n))
    (s64, n64), (s32, n32) = runs
    say("t = %.6f: %d FP64 steps, %d FP32 steps" % (t_end, n64, n32))
    compare("%s free to t=%.4f" % (par.split("/")[-1], t_end), s32, s64, FREE)


# Ten orbits at r = 1 (t = 20 pi), the precision form against the FP64 path on
# one sequence of time steps (each FP32 step takes the FP64 run's step before
# the output-time clamp, so both clamp alike).  The target of the precision
# form is a relative L1 density error <= 1e-5 over ten orbits; the bound of
# each case is stated with the value measured on MI355X:
#   3-D disk 256x96x32, 3000 steps:              5.0e-7 (target met)
#   adiabatic disk 1024x256, 8734 steps:         2.7e-6 (target met)
#   gas + 4 dust 1024x512, 5 orbits, 5553 steps: 2.0e-6 (target met); over
#     ten orbits the FP64 run itself amplifies a 1e-7 density perturbation to
#     1.5e-6 in the gas and 4.7e-4 in the St = 1 dust after the fifth orbit,
#     and the FP32 form's difference grows with it to 6.3e-5;
#   adiabatic disk 4096x1024, 45505 steps:       2.2e-5 (target not met: the
#     error grows by about 5e-10 per step, a systematic cooling in the inner
#     disk, so the full-resolution horizon has twice the allowed error).
C4 = "fargo3d_amd/setups/c4_multifluid2d.par"
TEN = [(C2, "", 10, 1e-5), (C3, "Nx=1024 Ny=256", 10, 1e-5), (C4, "", 5, 1e-5), (C4, "", 10, 1e-4),
       (C3, "", 10, 3e-5)]
TEN_IDS = ["c2_3d_256x96x32", "c3_adiabatic_1024x256", "c4_gas_4dust_1024x512_5_orbits",
           "c4_gas_4dust_1024x512_10_orbits", "c3_adiabatic_4096x1024"]


@pytest.mark.parametrize("par,ov,orbits,bound", TEN, ids=TEN_IDS)
def test_fp32_ten_orbits(par, ov, orbits, bound):
    import fargo3d_amd as fa
    s64 = fa.Simulation(par, ov)
    s32 = fa.Simulation(par, (ov + " Precision=32").strip())
    t_end = 2.0 * np.pi * orbits
    n, orbit = 0, 1
    while s64.scalars()["PhysicalTime"] < t_end * (1 - 1e-12):
        s64.step()
        s32.force_dt(s64.step_time())
        s32.step()
        n += 1
        if s64.scalars()["PhysicalTime"] >= orbit * 2 * np.pi:
            say("  orbit %d (%d steps)" % (orbit, n))
            orbit += 1
    assert s32.scalars()["PhysicalTime"] == s64.scalars()["PhysicalTime"]
    worst = []
    for k in range(s64.nfluids):
        s64.select_fluid(k)
        s32.select_fluid(k)
        a, b = active(s32.field("Density"), s64.dimz), active(s64.field("Density"), s64.dimz)
        worst.append(errors(a, b)[0])
    s64.select_fluid(0)
    s32.select_fluid(0)
    fields = ["Density", "Vx", "Vy"] + (["Vz"] if s64.dimz else []) + (["Energy"] if s64.adiabatic else [])
    say("%s %s: %g orbits, %d steps" % (par.split("/")[-1], ov or "full size", orbits, n))
    for f in fields:
        l1, linf = errors(active(s32.field(f), s64.dimz), active(s64.field(f), s64.dimz))
        say("  %-8s L1 %.3e  Linf %.3e" % (f, l1, linf))
    say("  density L1 of every fluid: %s (bound %.0e)" % (" ".join("%.2e" % e for e in worst), bound))
    assert worst[0] <= bound, worst
