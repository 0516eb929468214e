"""The batched form: independent 2-D disks advanced by one launch per stage.

* Independence: an instance advanced inside a batch of several uneven disks
  is bit-identical to the same instance advanced alone, with its own time
  step (the per-instance CFL and output-time clamp run on the device).
* Agreement with the single-disk FP32 form, which integrates the planet on
  the host every step where the batch holds it on its orbit: the relative
  difference is printed and bounded.
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
