"""The batched form: independent 2-D disks advanced by one launch per stage.

* Independence: an instance advanced inside a batch of several uneven disks
  is bit-identical to the same instance advanced alone, with its own time
  step (the per-instance CFL and output-time clamp run on the device).
* The 64-disk parameter sweep (Nx in {256,384,768,1536} x Ny in
  {128,192,384}, seeded planet mass, viscosity and aspect ratio): every
  instance inside the batch equals itself advanced alone, and stays within a
  stated bound of the same disk advanced by the plain FP64 path.
* The planets integrated on the device (FP64 RK5, frame rotation, the
  potential from their positions) follow the host's integrator on the same
  time steps, in a rotating and in a fixed frame.
* Agreement with the single-disk FP32 form on the same time steps.
* The autotuner returns one of its candidates and leaves the state unchanged.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

C1 = "fargo3d_amd/setups/c1_fargo2d.par"
SPECS = [(C1, "Nx=256 Ny=128 PlanetMass=1e-5 Nu=1e-6 AspectRatio=0.03"),
         (C1, "Nx=384 Ny=192 PlanetMass=3e-3 Nu=1e-4 AspectRatio=0.08"),
         (C1, "Nx=768 Ny=128 PlanetMass=1e-4 Nu=1e-5 AspectRatio=0.05")]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def state(s):
    return [s.field(f) for f in ("Density", "Vx", "Vy")]


def test_instances_are_independent():
    import fargo3d_amd as fa
    b = fa.Batch(SPECS)
    b.step(20)
    b.sync()
    together = [state(s) for s in b.instances]
    clocks = [b.clock(k) for k in range(b.n)]
    b.close()
    for k, spec in enumerate(SPECS):
        one = fa.Batch([spec])
        one.step(20)
        one.sync()
        alone = state(one.instances[0])
        assert one.clock(0) == clocks[k], (k, one.clock(0), clocks[k])
        for a, c, name in zip(alone, together[k], ("Density", "Vx", "Vy")):
            assert np.array_equal(a.view(np.int64), c.view(np.int64)), (k, name)
        say("instance %d: t = %.6f after 20 steps, identical alone and in the batch" % (k, clocks[k]["PhysicalTime"]))
        one.close()


def rel_l1(a, b):
    return float(np.abs(a - b).sum() / np.abs(b).sum())


def rel_linf(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


# the FP32 form against the FP64 path after SWEEP_STEPS steps of the sweep's
# disks (each on its own time steps): relative L1 and L-infinity bounds
SWEEP_STEPS = 50
BOUND = {"Density": (2e-5, 1e-4), "Vx": (1e-5, 1e-4)}
TIME_BOUND = 2e-5


def test_sweep_of_64_disks():
    import fargo3d_amd as fa
    specs = fa.sweep_specs(64)
    b = fa.Batch(specs)
    assert b.n == 64 and sum(b.cells) > 1e7
    b.step(SWEEP_STEPS)
    b.sync()
    together = [state(s) for s in b.instances]
    clocks = [b.clock(k) for k in range(b.n)]
    planets = [b.planets(k) for k in range(b.n)]
    b.close()
    rows = []
    for k, spec in enumerate(specs):
        one = fa.Batch([spec])
        one.step(SWEEP_STEPS)
        one.sync()
        assert one.clock(0) == clocks[k], (k, one.clock(0), clocks[k])
        pl, of = one.planets(0)
        assert np.array_equal(pl, planets[k][0]) and of == planets[k][1], k
        for a, c, name in zip(state(one.instances[0]), together[k], ("Density", "Vx", "Vy")):
            assert np.array_equal(a.view(np.int64), c.view(np.int64)), (k, name)
        one.close()
        s64 = fa.Simulation(*spec)
        for _ in range(SWEEP_STEPS):
            s64.step()
        t64 = s64.scalars()["PhysicalTime"]
        row = {"k": k, "spec": spec[1], "dt_rel": abs(clocks[k]["PhysicalTime"] - t64) / t64,
               "planet": float(np.max(np.abs(planets[k][0][:, :3] - s64.planets()[:, :3])))}
        for name in BOUND:
            a = together[k][("Density", "Vx", "Vy").index(name)][:, 3:-3, :]
            c = s64.field(name)[:, 3:-3, :]
            row[name] = (rel_l1(a, c), rel_linf(a, c))
        rows.append(row)
        s64.close()
    for r in rows:
        say("%2d %-62s time %.1e planet %.1e %s" % (r["k"], r["spec"], r["dt_rel"], r["planet"],
                                                   " ".join("%s %.2e/%.2e" % (n, r[n][0], r[n][1]) for n in BOUND)))
    worst = {n: (max(r[n][0] for r in rows), max(r[n][1] for r in rows)) for n in BOUND}
    say("64 disks, %d steps: each instance equals itself alone; against the FP64 path, worst relative "
        "L1/Linf: %s; time %.1e" % (SWEEP_STEPS, ", ".join("%s %.2e/%.2e" % (n, v[0], v[1]) for n, v in worst.items()),
                                   max(r["dt_rel"] for r in rows)))
    for r in rows:
        # the clock: the FP32 fields' time steps (the shear limit takes a
        # difference of neighbouring rings' rotation rates, which amplifies
        # the FP32 rounding of vx) against the FP64 path's
        assert r["dt_rel"] <= TIME_BOUND, r
        assert r["planet"] <= 1e-6, r
        for n, (b1, binf) in BOUND.items():
            assert r[n][0] <= b1 and r[n][1] <= binf, (n, r)


@pytest.mark.parametrize("frame", ["G", "F"])
def test_device_planets_follow_the_host_integrator(frame):
    """Batch of one against the single-disk FP32 form driven by the batch's
    time steps: the single disk integrates the planet on the host (glibc,
    the reference's RK5 and frame rotation), the batch on the device."""
    import fargo3d_amd as fa
    # no output-time clamp in these 40 steps (DT beyond the run), so both take
    # exactly the batch's time steps
    ov = "Nx=384 Ny=192 PlanetMass=1e-3 Nu=1e-5 AspectRatio=0.05 DT=100.0"
    if frame == "F":
        ov += " Frame=F OmegaFrame=0.0"
    b = fa.Batch([(C1, ov)])
    s = fa.Simulation(C1, ov + " Precision=32")
    for _ in range(40):
        b.step(1)
        s.force_dt(b.clock(0)["dt"])
        s.step()
    pl, of = b.planets(0)
    ph = s.planets()
    sc = s.scalars()
    dpos = float(np.max(np.abs(pl[:, :6] - ph[:, :6])))
    say("frame %s: planet after 40 steps device %s host %s (max diff %.2e), OMEGAFRAME %.17g vs %.17g"
        % (frame, pl[0, :3], ph[0, :3], dpos, of, sc["OMEGAFRAME"]))
    assert dpos <= 1e-13
    assert abs(of - sc["OMEGAFRAME"]) <= 1e-12 * max(abs(sc["OMEGAFRAME"]), 1.0)
    if frame == "F":
        # a fixed frame: the planet has moved along its orbit (40 steps)
        assert abs(pl[0, 1]) > 1e-3
    for name in ("Density", "Vx", "Vy"):
        a = b.instances[0].field(name)[:, 3:-3, :]
        c = s.field(name)[:, 3:-3, :]
        e = rel_linf(a, c)
        say("  %s batch vs single FP32: rel Linf %.2e" % (name, e))
        assert e <= 1e-4, (name, e)
    b.close()
    s.close()


def test_batch_matches_single_disk_fp32():
    import fargo3d_amd as fa
    par, ov = SPECS[1]
    b = fa.Batch([(par, ov)])
    s = fa.Simulation(par, ov + " Precision=32")
    for _ in range(30):
        dt = s.step()
        b.step(1)
    b.sync()
    assert abs(b.clock(0)["PhysicalTime"] - s.scalars()["PhysicalTime"]) <= 1e-6 * s.scalars()["PhysicalTime"]
    for name in ("Density", "Vx"):
        a = b.instances[0].field(name)[:, 3:-3, :]
        c = s.field(name)[:, 3:-3, :]
        rel = float(np.max(np.abs(a - c)) / np.max(np.abs(c)))
        say("batch vs single FP32 after 30 steps: %s rel %.3e" % (name, rel))
        assert rel <= 1e-4, (name, rel)
    b.close()
    s.close()


def test_tuner_keeps_state():
    import fargo3d_amd as fa
    b = fa.Batch(SPECS)
    b.step(3)
    b.sync()
    before = [state(s) for s in b.instances]
    r, t = b.tune(reps=2)
    assert r in (1, 2, 4, 8) and t in (64, 128, 256)
    after = [state(s) for s in b.instances]
    for x, y in zip(before, after):
        for a, c in zip(x, y):
            assert np.array_equal(a, c)
    say("tuner picked rows_y=%d tile=%d" % (r, t))
    b.close()
