"""GPU parity at the workloads' own horizons, and the disk force.

Every field of every active cell is compared bit for bit with the CPU
restatement (0 differing bits, not a tolerance), the time step of every step
and the integer shift counts exactly, over: the 2-D drop-in disk for its 200
steps (384x128), the 3-D disk (256x96x32) for 50 steps, the 2-D adiabatic
disk (4096x1024) for 10 steps and the gas + 4 dust disk (1024x512) for 20
steps.  The disk force (_ComputeForce, src/compute_force.c:16-219) is checked
on its own and through a planet that feels the disk (AdvanceSystemFromDisk,
src/planets.c:86-113).
"""
import ctypes
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

C1 = "fargo3d_amd/setups/c1_fargo2d.par"
C2 = "fargo3d_amd/setups/c2_p3d.par"
C3 = "fargo3d_amd/setups/c3_adiabatic2d.par"
C4 = "fargo3d_amd/setups/c4_multifluid2d.par"
FEEL = "PlanetConfig=fargo3d_amd/setups/planets/jupiter_feeldisk.cfg"


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def nbits(a, b):
    return int(np.count_nonzero(a.view(np.int64) != b.view(np.int64)))


@pytest.fixture(scope="module")
def fa():
    import fargo3d_amd
    return fargo3d_amd


@pytest.fixture(scope="module")
def orc():
    import oracle
    return oracle


def pair(fa, orc, par, ov="", arch=None):
    o = orc.Oracle(par, ov)
    s = fa.Simulation(par, ov, arch=arch or {})
    return o, s


def fields(s):
    f = ["Density", "Vx", "Vy"] + (["Vz"] if s.dimz else [])
    if s.adiabatic:
        f.append("Energy")
    return f


def compare_bits(o, s, tag):
    nf = s.nfluids
    for q in range(nf):
        if nf > 1:
            o.select_fluid(q)
            s.select_fluid(q)
        for f in fields(s):
            a, b = s.field(f), o.field(f)
            n = nbits(a, b)
            assert n == 0, "%s fluid %d %s: %d bits differ (max rel %.3e)" % (
                tag, q, f, n, float(np.max(np.abs(a - b)) / np.max(np.abs(b))))
    if nf > 1:
        o.select_fluid(0)
        s.select_fluid(0)
    sl = (slice(3, -3), slice(3, -3)) if s.dimz else (slice(None), slice(3, -3))
    assert np.array_equal(s.nshift()[sl], o.nshift()[sl]), "%s: Nshift differs" % tag


@pytest.mark.parametrize("par,ov,nsteps", [(C1, "", 200), (C2, "", 50), (C3, "", 10), (C4, "", 20)],
                         ids=["c1_2d_384x128_200_steps", "c2_3d_256x96x32_50_steps",
                              "c3_2d_adiabatic_4096x1024_10_steps", "c4_gas_4dust_1024x512_20_steps"])
def test_long_parity_zero_bits(fa, orc, par, ov, nsteps):
    say("long parity", par, nsteps)
    o, s = pair(fa, orc, par, ov)
    for n in range(1, nsteps + 1):
        dto, dts = o.step(), s.step()
        assert dto == dts, "step %d: dt %r vs %r" % (n, dts, dto)
        if n % max(1, nsteps // 10) == 0 or n == nsteps:
            say("  step %d dt %.17g" % (n, dts))
            compare_bits(o, s, "step %d" % n)
    assert np.array_equal(o.planets(), s.planets())
    s.close()


@pytest.mark.parametrize("par,ov", [(C1, "Nx=96 Ny=40"), (C2, "Nx=32 Ny=24 Nz=10")], ids=["2d", "3d"])
def test_disk_force_matches_oracle(fa, orc, par, ov):
    """f3d_ComputeForce on a random total density: the 12 sums bit for bit,
    for a body inside the mesh, one inside the inner edge, one outside."""
    o, s = pair(fa, orc, par, ov)
    rng = np.random.default_rng(17)
    td = rng.random(o.field("Total_Density").shape)
    o.field("Total_Density")[...] = td
    s.set_field("Total_Density", td)
    for x, y, z, rsm in ((0.6, 0.8, 0.0, 0.03), (0.1, -0.2, 0.0, 0.01), (2.0, 2.0, 0.01, 0.05)):
        o.call("ComputeForce", x, y, z, rsm, 1e-3)
        lf = (ctypes.c_double * 12)()
        s.call("ComputeForce", x, y, z, rsm, 1e-3, lf)
        got, ref = np.array(lf), o.localforce()
        say("force", x, y, z, got[:3] + got[6:9])
        assert np.array_equal(got, ref), (got, ref)
    s.close()


@pytest.mark.parametrize("arch", [{}, {"FusedCfl": "no", "FusedSources": "no", "FusedTransport": "no"}],
                         ids=["fused", "per_routine"])
def test_whole_step_planet_feels_disk(fa, orc, arch):
    """A planet that feels the disk (FeelsDisk YES): its velocity takes the
    disk's force every step, so its orbit, the potential and the gas all
    depend on _ComputeForce; 30 steps, 0 bits, planet state exactly equal."""
    o, s = pair(fa, orc, C1, "Nx=192 Ny=64 " + FEEL, arch=arch)
    p0 = o.planets().copy()
    for n in range(1, 31):
        dto, dts = o.step(), s.step()
        assert dto == dts, (n, dts, dto)
        assert np.array_equal(o.planets(), s.planets()), n
    compare_bits(o, s, "feel disk")
    # the disk did act on the planet: the orbit differs from the one a
    # planet that does not feel the disk follows
    o2 = orc.Oracle(C1, "Nx=192 Ny=64")
    for _ in range(30):
        o2.step()
    assert not np.array_equal(o.planets()[0, 3:5], o2.planets()[0, 3:5])
    assert np.abs(o.planets()[0, :2] - p0[0, :2]).max() > 0
    s.close()
