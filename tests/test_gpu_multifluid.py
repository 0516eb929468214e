"""GPU parity of the multifluid step (one gas + dust fluids) against the oracle.

The order is src/main.c:366-434: ComputeVmed and cfl of every fluid (each
fluid's cfl reads the last fluid's VxMed), Sources of every fluid, the
implicit drag of all fluids at once (Collisions(dt, 1), src/collisions.c:156),
dust diffusion (src/dust_diffusion_main.c:3), then Transport, damping and
ghost fill of every fluid.  Same bar as the single-fluid parity tests: 1e-12
relative per field, equal time steps, equal shift counts.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

C4 = "fargo3d_amd/setups/c4_multifluid2d.par"
C4_SMALL = "Nx=128 Ny=64"
RTOL = 1e-12
FIELDS = ["Density", "Vx", "Vy"]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def relerr(a, b):
    den = float(np.max(np.abs(b)))
    d = float(np.max(np.abs(a - b)))
    return d / den if den else d


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


def make_pair(fa, orc, ov, arch=None):
    o = orc.Oracle(C4, ov)
    s = fa.Simulation(C4, ov, arch=arch or {})
    o.set_shape(s.nx, s.ny, s.nz, s.dimz)
    return o, s


def compare_fluids(o, s, tag, ghosts=True):
    worst = 0.0
    for k in range(s.nfluids):
        o.select_fluid(k)
        s.select_fluid(k)
        for f in FIELDS:
            a, b = s.field(f), o.field(f)
            if not ghosts:
                a, b = a[:, 3:-3, :], b[:, 3:-3, :]
            e = relerr(a, b)
            say("%s fluid %d %s rel %.3e bits %d" % (tag, k, f, e, nbits(a, b)))
            assert e <= RTOL, (tag, k, f, e)
            worst = max(worst, e)
    o.select_fluid(0)
    s.select_fluid(0)
    return worst


def test_multifluid_initial_state(fa, orc):
    """Every fluid's initial fields, ghost fill and damping references agree bit for bit."""
    say("multifluid initial state", C4_SMALL)
    o, s = make_pair(fa, orc, C4_SMALL)
    assert s.nfluids == 5
    for k in range(5):
        o.select_fluid(k)
        s.select_fluid(k)
        for f in ["Density", "Energy", "Vx", "Vy", "Density0", "Vx0", "Vy0"]:
            a, b = s.field(f), o.field(f)
            assert nbits(a, b) == 0, (k, f)
    with pytest.raises(Exception):
        s.select_fluid(5)


@pytest.mark.parametrize("ov,nsteps", [(C4_SMALL, 12),
                                       (C4_SMALL + " DustDiffusion=no", 6),
                                       (C4_SMALL + " Feedback=no DustDiffusion=no", 6),
                                       ("", 2)],
                         ids=["drag_diffusion", "drag_only", "drag_no_feedback", "full_1024x512"])
def test_multifluid_fused(fa, orc, ov, nsteps):
    """The fused stages plus the drag and diffusion kernels reproduce the oracle's multifluid run."""
    say("multifluid fused", ov, nsteps)
    o, s = make_pair(fa, orc, ov)
    for n in range(1, nsteps + 1):
        dto, dts = o.step(), s.step()
        assert dto == dts, "step %d: dt %r vs %r" % (n, dts, dto)
        if n in (1, 2, nsteps):
            compare_fluids(o, s, "step %d" % n)
            assert np.array_equal(s.nshift()[:, 3:-3], o.nshift()[:, 3:-3])


@pytest.mark.parametrize("form", [0, 1, 2], ids=["sum_in_ring_kernel", "pass_on_side_stream", "pass_before_sweep"])
def test_multifluid_xvmed_forms(fa, orc, form):
    """Every fluid's X-sweep VxMed from the ring kernel's own sum or from one
    pass over all fluids (second stream or step stream) gives the oracle's run."""
    L = fa.lib()
    assert L.f3d_fused_set_xvmed(form) == 0
    try:
        say("multifluid xvmed form", form)
        o, s = make_pair(fa, orc, C4_SMALL)
        for n in range(1, 5):
            dto, dts = o.step(), s.step()
            assert dto == dts, "step %d: dt %r vs %r" % (n, dts, dto)
        compare_fluids(o, s, "xvmed form %d step 4" % form)
        assert np.array_equal(s.nshift()[:, 3:-3], o.nshift()[:, 3:-3])
    finally:
        L.f3d_fused_set_xvmed(-1)


def test_multifluid_routines(fa, orc):
    """The per-routine kernels with the same drag and diffusion kernels agree as well."""
    say("multifluid routines", C4_SMALL)
    o, s = make_pair(fa, orc, C4_SMALL,
                     arch={"FusedCfl": "no", "FusedSources": "no", "FusedTransport": "no"})
    for n in range(1, 5):
        dto, dts = o.step(), s.step()
        assert dto == dts, "step %d: dt %r vs %r" % (n, dts, dto)
    compare_fluids(o, s, "routines step 4", ghosts=False)


def test_multifluid_fp32(fa, orc):
    """The reduced-precision form runs the multifluid step; on one dt sequence
    it stays within the FP32 error of the FP64 run (bounds from measurement:
    about 1e-6 relative after 10 steps)."""
    say("multifluid fp32", C4_SMALL)
    s64 = fa.Simulation(C4, C4_SMALL)
    s32 = fa.Simulation(C4, C4_SMALL + " Precision=32")
    for n in range(10):
        s64.step()
        s32.force_dt(s64.step_time())
        s32.step()
    for k in range(5):
        s64.select_fluid(k)
        s32.select_fluid(k)
        for f in FIELDS:
            a, b = s32.field(f)[:, 3:-3, :], s64.field(f)[:, 3:-3, :]
            e = relerr(a, b)
            say("fp32 fluid %d %s rel %.3e" % (k, f, e))
            assert np.all(np.isfinite(a))
            assert e <= 1e-4, (k, f, e)
