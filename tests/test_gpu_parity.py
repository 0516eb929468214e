"""GPU parity of the HIP path against the CPU restatement (oracle).

Two kinds of check, as FARGO3D's own debugging harness does (FARGO_DEBUG,
src/define.h:523-535): the whole step from the same initial conditions, and
each routine separately from an identical state.  The bar is 1e-12 relative
(max |a-b| / max |b| per field) for floating point and exact equality for the
integer shift counts; the number of bit-level differences is reported too.
"""
import ctypes
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

C1 = "fargo3d_amd/setups/c1_fargo2d.par"
C2 = "fargo3d_amd/setups/c2_p3d.par"
C2_SMALL = "Nx=64 Ny=24 Nz=12"
C3 = "fargo3d_amd/setups/c3_adiabatic2d.par"
C3_SMALL = "Nx=256 Ny=96"
C4 = "fargo3d_amd/setups/c4_multifluid2d.par"
# the 3-D disk with the adiabatic equation of state (energy transported)
C2_AD = C2_SMALL + " Eos=adiabatic EnergyBoundary=symmetric Gamma=1.4"
RTOL = 1e-12

STATE3D = ["Density", "Energy", "Vx", "Vy", "Vx_temp", "Vy_temp", "Mpx", "Mmx", "Mpy", "Mmy",
           "Pot", "Slope", "DivRho", "DensStar", "Qs", "Pressure", "Total_Density", "QL", "QR"]
STATE3D_Z = ["Vz", "Vz_temp", "Mpz", "Mmz"]
STATE2D = ["Density0", "Energy0", "Vx0", "Vy0", "Vz0", "Reduction2D", "VxMed"]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def relerr(a, b):
    den = float(np.max(np.abs(b))) if b.size else 0.0
    d = float(np.max(np.abs(a - b))) if a.size else 0.0
    if den == 0.0:
        return d
    return d / den


def nbits(a, b):
    return int(np.count_nonzero(a.view(np.int64) != b.view(np.int64)))


def fields_of(s):
    """The primitive fields that carry the state (the energy only when it evolves)."""
    f = ["Density", "Vx", "Vy"] + (["Vz"] if s.dimz else [])
    if s.adiabatic:
        f.append("Energy")
    return f


def active_rings(x2d, dimz):
    """Active (j, k) rings of a YZ array of shape (sz, sy)."""
    return x2d[3:-3, 3:-3] if dimz else x2d[:, 3:-3]


def active(x, dimz):
    if dimz:
        return x[3:-3, 3:-3, :]
    return x[:, 3:-3, :]


@pytest.fixture(scope="module")
def fa():
    import fargo3d_amd
    return fargo3d_amd


@pytest.fixture(scope="module")
def orc():
    import oracle
    return oracle


def make_pair(fa, orc, par, ov="", arch=None):
    o = orc.Oracle(par, ov)
    s = fa.Simulation(par, ov, arch=arch or {})
    o.set_shape(s.nx, s.ny, s.nz, s.dimz)
    return o, s


def copy_state(o, s):
    names = STATE3D + (STATE3D_Z if s.dimz else [])
    for n in names:
        s.set_field(n, o.field(n))
    for n in STATE2D:
        s.set_field(n, o.field(n))
    s.set_nshift(o.nshift())
    s.set_planets(o.planets())
    s.set_scalars(o.scalars())


def compare_all(o, s, tag, exact=False):
    names = STATE3D + (STATE3D_Z if s.dimz else [])
    worst = 0.0
    for n in names + STATE2D:
        a, b = s.field(n), o.field(n)
        e = relerr(a, b)
        if e > RTOL or (exact and nbits(a, b)):
            raise AssertionError("%s: field %s rel err %.3e, %d bits differ" % (tag, n, e, nbits(a, b)))
        worst = max(worst, e)
    ns, no = s.nshift(), o.nshift()
    assert np.array_equal(ns, no), "%s: Nshift differs" % tag
    return worst


@pytest.mark.parametrize("par,ov", [(C1, ""), (C2, C2_SMALL), (C3, C3_SMALL)],
                         ids=["c1_2d", "c2_3d_small", "c3_2d_adiabatic_small"])
def test_initial_state(fa, orc, par, ov):
    """Initial conditions + first ghost fill + Stockholm references agree everywhere."""
    say("initial state", par, ov)
    o, s = make_pair(fa, orc, par, ov)
    for n in ["Density", "Energy", "Vx", "Vy"] + (["Vz"] if s.dimz else []):
        a, b = s.field(n), o.field(n)
        assert nbits(a, b) == 0, "%s differs after the initial ghost fill" % n
    for n in ["Density0", "Vx0", "Vy0"]:
        assert nbits(s.field(n), o.field(n)) == 0, n


ROUTINES_2D = [
    ("ComputePressureFieldIso", ()),
    ("Potential", ()),
    ("CorrectVtheta", (1.0e-3,)),
    ("SubStep1_x", ("dt",)),
    ("SubStep1_y", ("dt",)),
    ("visctensor_cyl", ()),
    ("addviscosity_cyl", ("dt",)),
    ("SubStep2_a", ("dt",)),
    ("SubStep2_b", ("dt",)),
    ("boundaries", ()),
    ("momenta_x", ()),
    ("momenta_y", ()),
    ("DivideByRho", ("Mpx",)),
    ("VanLeerY_a", ("DivRho",)),
    ("VanLeerY_b", ("dt", "DivRho", "Qs")),
    ("UpdateY", ("dt", "Mpx", "Qs")),
    ("UpdateDensityY", ("dt", "Density")),
    ("ComputeVmed", ("Vx_temp",)),
    ("ComputeResidual", ("dt",)),
    ("VanLeerX_a", ("Density",)),
    ("VanLeerX_b", ("dt", "Density", "DensStar", "Vx")),
    ("VanLeerX_PPA_a", ("DivRho",)),
    ("VanLeerX_PPA_b", ("DivRho",)),
    ("VanLeerX_PPA_c", ("DivRho",)),
    ("VanLeerX_PPA_d", ("dt", "DivRho", "Qs", "Vx_temp")),
    ("UpdateX", ("dt", "Mpx", "Qs", "Vx")),
    ("UpdateDensityX", ("dt", "Density", "Vx")),
    ("AdvectSHIFT", ("Mpx", "Nshift")),
    ("NewVelocity_x", ()),
    ("NewVelocity_y", ()),
    ("StockholmBoundary", ("dt",)),
    ("ComputeTotalDensity", ()),
]

ROUTINES_3D = [
    ("SubStep1_z", ("dt",)),
    ("visctensor_sph", ()),
    ("addviscosity_sph", ("dt",)),
    ("momenta_z", ()),
    ("VanLeerZ_a", ("Density",)),
    ("VanLeerZ_b", ("dt", "Density", "DensStar")),
    ("UpdateZ", ("dt", "Mpz", "Qs")),
    ("UpdateDensityZ", ("dt", "Density")),
    ("NewVelocity_z", ()),
]


def _args(o, s, args, dt):
    oa, sa = [], []
    for a in args:
        if a == "dt":
            oa.append(dt)
            sa.append(dt)
        elif a == "Nshift":
            oa.append(o.L.orc_nshift(o._s))
            sa.append(ctypes.c_void_p(s.device_ptr("Nshift")))
        else:
            oa.append(a)
            sa.append(a)
    return oa, sa


def _routine_sweep(fa, orc, par, ov, routines, nwarm):
    o, s = make_pair(fa, orc, par, ov)
    dt = 0.0
    for _ in range(nwarm):
        dt = o.step()
    failures = []
    for name, args in routines:
        copy_state(o, s)
        oa, sa = _args(o, s, args, dt)
        say("routine", name, args)
        o.call(name, *oa)
        s.call(name, *sa)
        s.sync()
        try:
            compare_all(o, s, name, exact=False)
        except AssertionError as e:
            failures.append(str(e))
    assert not failures, "\n".join(failures)


def test_routines_2d(fa, orc):
    _routine_sweep(fa, orc, C1, "", ROUTINES_2D, nwarm=3)


def test_routines_3d(fa, orc):
    _routine_sweep(fa, orc, C2, C2_SMALL, ROUTINES_2D[:5] + [("SubStep2_a", ("dt",)), ("SubStep2_b", ("dt",)),
                                                            ("boundaries", ()), ("momenta_x", ()),
                                                            ("ComputeResidual", ("dt",)),
                                                            ("AdvectSHIFT", ("Mpx", "Nshift")),
                                                            ("NewVelocity_x", ()), ("NewVelocity_y", ()),
                                                            ("StockholmBoundary", ("dt",))] + ROUTINES_3D,
                   nwarm=2)


ROUTINES_ADIABATIC = [
    ("ComputePressureFieldAd", ()),
    ("SubStep1_x", ("dt",)),
    ("SubStep1_y", ("dt",)),
    ("SubStep2_a", ("dt",)),
    ("SubStep2_b", ("dt",)),
    ("SubStep3", ("dt",)),
    ("boundaries", ()),
    ("DivideByRho", ("Energy",)),
    ("UpdateY", ("dt", "Energy", "Qs")),
    ("AdvectSHIFT", ("Energy", "Nshift")),
    ("StockholmBoundary", ("dt",)),
]


def test_routines_adiabatic(fa, orc):
    """The energy-equation routines (pressure, heating, compression work, energy ghosts)."""
    _routine_sweep(fa, orc, C3, C3_SMALL, ROUTINES_ADIABATIC, nwarm=2)


def test_reductions_and_cfl(fa, orc):
    o, s = make_pair(fa, orc, C1)
    for _ in range(2):
        o.step()
    copy_state(o, s)
    # cfl (per-cell dt + MIN + shear limit): order-independent MIN, exact
    o.call("ComputeVmed", "Vx")
    s.call("ComputeVmed", "Vx")
    o.call("cfl")
    m = ctypes.c_double()
    s.call("cfl", ctypes.byref(m))
    omin = o.L.orc_field  # noqa: F841 (keeps the symbol table warm)
    # the oracle stores its per-fluid Min in the state; recompute via StepTime
    o.L.orc_CflFluidsMin.restype = ctypes.c_int
    o.L.orc_CflFluidsMin.argtypes = [ctypes.c_void_p]
    o.L.orc_CflFluidsMin(o._s)
    assert m.value == o.scalars()["StepTime"], (m.value, o.scalars()["StepTime"])
    # full reductions in sequential ring order: bit-identical
    o.L.orc_reduction_full_SUM.restype = ctypes.c_double
    o.L.orc_reduction_full_SUM.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)] + [ctypes.c_int] * 4
    ref = o.L.orc_reduction_full_SUM(o._s, o.L.orc_field(o._s, b"Density"), 3, s.ny + 3, 0, 1)
    got = ctypes.c_double()
    s.call("reduction_full_SUM", "Density", 3, s.ny + 3, 0, 1, ctypes.byref(got))
    assert got.value == ref


@pytest.mark.parametrize("par,ov,nsteps", [(C1, "", 20), (C2, C2_SMALL, 6), (C3, C3_SMALL, 10)],
                         ids=["c1_2d", "c2_3d_small", "c3_2d_adiabatic_small"])
def test_whole_step_routines(fa, orc, par, ov, nsteps):
    """The C driver over the per-routine kernels reproduces the CPU run step by step."""
    o, s = make_pair(fa, orc, par, ov, arch={"FusedCfl": "no", "FusedSources": "no", "FusedTransport": "no"})
    for n in range(1, nsteps + 1):
        dto, dts = o.step(), s.step()
        assert dto == dts, "step %d: dt %r vs %r" % (n, dts, dto)
        if n in (1, 2, 5, nsteps):
            worst = 0.0
            for f in fields_of(s):
                a, b = active(s.field(f), s.dimz), active(o.field(f), s.dimz)
                e = relerr(a, b)
                say("step %d %s rel %.3e bits %d" % (n, f, e, nbits(a, b)))
                assert e <= RTOL, (n, f, e)
                worst = max(worst, e)
            assert np.array_equal(s.nshift(), o.nshift())
    po, ps = o.planets(), s.planets()
    assert np.array_equal(po, ps)


def _fused_run(fa, orc, par, ov, nsteps):
    o, s = make_pair(fa, orc, par, ov)
    for n in range(1, nsteps + 1):
        dto, dts = o.step(), s.step()
        assert dto == dts, "step %d: dt %r vs %r" % (n, dts, dto)
        if n in (1, 2, nsteps):
            for f in fields_of(s):
                a, b = active(s.field(f), s.dimz), active(o.field(f), s.dimz)
                e = relerr(a, b)
                say("fused step %d %s rel %.3e bits %d" % (n, f, e, nbits(a, b)))
                assert e <= RTOL, (n, f, e)
            # ghost cells are refilled from active ones at the end of a step
            for f in fields_of(s):
                e = relerr(s.field(f), o.field(f))
                assert e <= RTOL, ("ghosts", n, f, e)
            # the shift counts of the active rings (ghost rings are never shifted
            # into active cells) are integers: exactly equal
            assert np.array_equal(active_rings(s.nshift(), s.dimz), active_rings(o.nshift(), s.dimz)), \
                "step %d: Nshift differs" % n
    return o, s


@pytest.mark.parametrize("par,ov,nsteps", [(C1, "", 30), (C2, C2_SMALL, 8), (C2, "", 3), (C3, C3_SMALL, 20),
                                           (C2, C2_AD, 8), (C3, "", 2)],
                         ids=["c1_2d", "c2_3d_small", "c2_3d_full", "c3_2d_adiabatic_small", "c2_3d_adiabatic_small",
                              "c3_2d_adiabatic_full"])
def test_whole_step_fused(fa, orc, par, ov, nsteps):
    """The fused kernels reproduce the CPU run; the active cells carry the state.
    The full 4096-cell rings take the tiled X sweep."""
    _fused_run(fa, orc, par, ov, nsteps)


@pytest.mark.parametrize("par,ov,nsteps", [(C1, "", 12), (C2, C2_SMALL, 6), (C3, C3_SMALL, 10), (C2, C2_AD, 4)],
                         ids=["c1_2d", "c2_3d_small", "c3_2d_adiabatic_small", "c2_3d_adiabatic_small"])
def test_whole_step_fused_tiled_x(fa, orc, par, ov, nsteps):
    """The tiled X sweep (tiles of 128 cells with halos, shifted writes) on
    rings the LDS ring kernel also handles gives the same results."""
    L = fa.lib()
    assert L.f3d_fused_set_xmode(2) == 0
    try:
        _fused_run(fa, orc, par, ov, nsteps)
    finally:
        L.f3d_fused_set_xmode(0)


@pytest.mark.parametrize("par,ov,nsteps", [(C1, "Nx=256 Ny=64", 8), (C2, "Nx=256 Ny=24 Nz=12", 6),
                                           (C3, C3_SMALL, 10),
                                           (C2, "Nx=256 Ny=24 Nz=12" + C2_AD[len(C2_SMALL):], 4)],
                         ids=["c1_2d", "c2_3d", "c3_2d_adiabatic", "c2_3d_adiabatic"])
def test_whole_step_fused_one_wave_x(fa, orc, par, ov, nsteps):
    """The one-wave ring X sweep (cells in registers, neighbours across lanes
    by wave rotations, rings of 256 cells) gives the oracle's results."""
    L = fa.lib()
    assert L.f3d_fused_set_xmode(3) == 0
    try:
        _fused_run(fa, orc, par, ov, nsteps)
    finally:
        L.f3d_fused_set_xmode(0)


@pytest.mark.parametrize("par,ov,nsteps", [(C3, C3_SMALL, 10), (C3, "Nx=1000 Ny=24", 6), (C3, "Nx=2500 Ny=48", 4),
                                           (C1, "", 8), (C2, C2_SMALL, 4), (C2, C2_AD, 4), (C4, "Nx=128 Ny=64", 4)],
                         ids=["c3_2d_adiabatic_256", "c3_2d_adiabatic_1000", "c3_2d_adiabatic_2500", "c1_2d_384",
                              "c2_3d_64", "c2_3d_adiabatic_64", "c4_five_fluids_128"])
def test_whole_step_fused_one_wave_tiles(fa, orc, par, ov, nsteps):
    """The X sweep by one-wave tiles (256 cells in registers, 244 updated,
    neighbours across lanes by wave rotations, the last tile of a ring
    partial, rings shorter than a tile covered more than once) gives the
    oracle's results bit for bit."""
    L = fa.lib()
    assert L.f3d_fused_set_xmode(4) == 0
    try:
        o, s = _fused_run(fa, orc, par, ov, nsteps)
        for f in fields_of(s):
            assert nbits(active(s.field(f), s.dimz), active(o.field(f), s.dimz)) == 0, (par, ov, f)
        s.close()
    finally:
        L.f3d_fused_set_xmode(0)


@pytest.mark.parametrize("y2v", [-1, 0], ids=["pass_in_y_sweep_launch", "pass_on_second_stream"])
@pytest.mark.parametrize("par,ov,nsteps", [(C3, C3_SMALL, 8), (C3, "Nx=1000 Ny=24", 5), (C4, "Nx=128 Ny=64", 4),
                                           (C1, "Nx=100 Ny=40", 5)],
                         ids=["c3_2d_adiabatic_256", "c3_2d_adiabatic_1000", "c4_five_fluids_128", "c1_2d_ragged"])
def test_whole_step_fused_vmed_in_y_sweep(fa, orc, par, ov, nsteps, y2v):
    """The X sweep's VxMed pass run by workgroups of the 2-D Y sweep's launch
    (before the sweep's own workgroups, every fluid) or on the second stream
    gives the oracle's results bit for bit (the tiled X sweep, which takes
    VxMed from the pass)."""
    L = fa.lib()
    assert L.f3d_fused_set_xmode(4) == 0
    assert L.f3d_fused_set_y2v(y2v) == 0
    try:
        o, s = _fused_run(fa, orc, par, ov, nsteps)
        for f in fields_of(s):
            assert nbits(active(s.field(f), s.dimz), active(o.field(f), s.dimz)) == 0, (par, ov, y2v, f)
        s.close()
    finally:
        L.f3d_fused_set_xmode(0)
        L.f3d_fused_set_y2v(-1)


@pytest.mark.parametrize("fuse", [1, 0], ids=["one_pass", "stress_arrays"])
@pytest.mark.parametrize("par,ov,nsteps", [(C1, "", 12), (C1, "ViscosityModel=alpha Alpha=1e-3", 8),
                                           (C3, C3_SMALL, 10), (C3, C3_SMALL + " ViscosityModel=alpha Alpha=1e-3", 6),
                                           (C1, "Nx=100 Ny=40", 6)],
                         ids=["c1_2d_nu", "c1_2d_alpha", "c3_2d_adiabatic_nu", "c3_2d_adiabatic_alpha",
                              "c1_2d_ragged"])
def test_whole_step_fused_sources_one_pass(fa, orc, par, ov, nsteps, fuse):
    """The 2-D sources of a viscous gas in one pass (the stress computed where
    it is used, x neighbours across lanes, waves overlapping by two columns,
    rings not a multiple of the 62 cells a wave updates) and through the
    stress arrays give the oracle's results bit for bit."""
    L = fa.lib()
    assert L.f3d_fused_set_srcfuse(fuse) == 0
    try:
        o, s = _fused_run(fa, orc, par, ov, nsteps)
        for f in fields_of(s):
            assert nbits(active(s.field(f), s.dimz), active(o.field(f), s.dimz)) == 0, (par, ov, f)
        s.close()
    finally:
        L.f3d_fused_set_srcfuse(-1)


@pytest.mark.parametrize("par,ov,nsteps", [(C2, C2_SMALL, 6), (C2, C2_AD, 4), (C2, "Nx=64 Ny=24 Nz=6", 5),
                                           (C2, "Nx=64 Ny=24 Nz=5", 3), (C1, "", 8), (C3, C3_SMALL, 6),
                                           (C4, "Nx=128 Ny=64", 4), (C1, "Nx=96 Ny=6 Ymin=1.0 Ymax=1.3", 4),
                                           (C1, "Nx=96 Ny=5 Ymin=1.0 Ymax=1.3", 3)],
                         ids=["c2_3d_small", "c2_3d_adiabatic_small", "c2_nz_6", "c2_nz_5_own_launch", "c1_2d",
                              "c3_2d_adiabatic_small", "c4_five_fluids_128", "c1_ny_6", "c1_ny_5_own_launch"])
def test_whole_step_fused_z_fill_in_finish(fa, orc, par, ov, nsteps):
    """The end-of-step ghost fill done by the finishing kernel (3-D: the Z
    sides; 2-D: the Y sides; each thread filling the ghost cells its plane or
    row is the source of) leaves every cell, ghosts included, as the fill of
    its own launch does, and both give the oracle's active cells bit for bit
    (with 5 planes or rows, fewer than the six the merged fill needs, both
    settings take the separate launch)."""
    L = fa.lib()
    full = {}
    for zf in (-1, 0):
        assert L.f3d_fused_set_zfill(zf) == 0
        try:
            o, s = _fused_run(fa, orc, par, ov, nsteps)
            for f in fields_of(s):
                assert nbits(active(s.field(f), s.dimz), active(o.field(f), s.dimz)) == 0, (ov, zf, f)
            full[zf] = {f: s.field(f) for f in fields_of(s) + ["Energy"]}
            s.close()
        finally:
            L.f3d_fused_set_zfill(-1)
    for f in full[0]:
        assert nbits(full[-1][f], full[0][f]) == 0, (ov, f)


@pytest.mark.parametrize("mode", [0, 1, 2, 3, 6, 7], ids=["atomic", "two_pass", "vmed_pass_atomic",
                                                     "vmed_pass_two_pass", "vmed_pass_atomic_one_wave",
                                                     "vmed_pass_two_pass_one_wave"])
def test_cfl_forms(fa, orc, mode):
    """Every form of the fused CFL kernel (atomic or two-pass MIN, VxMed summed
    inside the ring's workgroup or in a pass of its own, four waves or one per
    workgroup) gives the oracle's time step bit for bit."""
    L = fa.lib()
    assert L.f3d_fused_set_cflmode(mode) == 0
    try:
        for par, ov in [(C1, ""), (C2, C2_SMALL), (C3, C3_SMALL), (C3, "Nx=1000 Ny=24")]:
            o, s = make_pair(fa, orc, par, ov)
            for n in range(3):
                dto, dts = o.step(), s.step()
                say("cfl form %d %s step %d dt %r" % (mode, par, n, dts))
                assert dto == dts, (mode, par, n, dts, dto)
            s.close()
    finally:
        L.f3d_fused_set_cflmode(-1)


@pytest.mark.parametrize("split", [0, 1], ids=["one_workgroup_per_ring", "two_workgroups_per_ring"])
@pytest.mark.parametrize("par,ov,nsteps", [(C1, "", 12), (C2, C2_SMALL, 6), (C3, C3_SMALL, 10), (C2, C2_AD, 4)],
                         ids=["c1_2d", "c2_3d_small", "c3_2d_adiabatic_small", "c2_3d_adiabatic_small"])
def test_whole_step_fused_ring_split(fa, orc, par, ov, nsteps, split):
    """The ring X sweep with every quantity in one workgroup, or split over
    two workgroups per ring, gives the oracle's results."""
    L = fa.lib()
    assert L.f3d_fused_set_xsplit(split) == 0
    try:
        _fused_run(fa, orc, par, ov, nsteps)
    finally:
        L.f3d_fused_set_xsplit(-1)


@pytest.mark.parametrize("form", [0, 1, 2], ids=["sum_in_ring_kernel", "pass_on_side_stream", "pass_before_sweep"])
@pytest.mark.parametrize("par,ov,nsteps", [(C1, "", 8), (C2, C2_SMALL, 4), (C3, C3_SMALL, 6)],
                         ids=["c1_2d", "c2_3d_small", "c3_2d_adiabatic_small"])
def test_whole_step_fused_xvmed_forms(fa, orc, par, ov, nsteps, form):
    """The X sweep's VxMed summed inside the ring workgroup, or by a pass of
    its own on a second stream or on the step's stream, gives the oracle's
    results."""
    L = fa.lib()
    assert L.f3d_fused_set_xvmed(form) == 0
    try:
        _fused_run(fa, orc, par, ov, nsteps)
    finally:
        L.f3d_fused_set_xvmed(-1)


@pytest.mark.parametrize("terms", [4, 8, 16])
def test_vmed_pass_terms_per_lane(fa, orc, terms):
    """The one-wave ring sums of VxMed (the CFL's pass over Vx and the X
    sweep's pass over Vx_temp) with 4, 8 or 16 terms per lane give the
    oracle's time step and ring means bit for bit, on rings that fill whole
    chunks of 64 x terms, part of one, and several with a partial last one."""
    L = fa.lib()
    L.f3d_fused_cfl_fluids.restype = ctypes.c_int
    L.f3d_fused_cfl_fluids.argtypes = [ctypes.c_void_p, ctypes.c_int] + [ctypes.c_void_p] * 6 + \
        [ctypes.POINTER(ctypes.c_double)]
    assert L.f3d_fused_set_vmedterms(terms) == 0
    assert L.f3d_fused_set_cflmode(3) == 0      # VxMed of Vx from the pass
    assert L.f3d_fused_set_xvmed(2) == 0        # VxMed of Vx_temp from the pass
    P = ctypes.c_void_p * 8
    try:
        for par, ov in [(C1, ""), (C2, C2_SMALL), (C3, C3_SMALL), (C3, "Nx=1000 Ny=24"), (C3, "Nx=2500 Ny=16")]:
            o, s = make_pair(fa, orc, par, ov)
            for n in range(3):
                # the CFL's pass alone: ComputeVmed(Vx) (src/main.c:366)
                tab = [P(s.device_ptr(nm)) for nm in ("Energy", "Density", "Vx", "Vy", "Vz" if s.dimz else "Vy")]
                got = ctypes.c_double()
                assert L.f3d_fused_cfl_fluids(s.ctx, 1, *tab, s.device_ptr("Vx"), ctypes.byref(got)) == 0
                o.call("ComputeVmed", "Vx")
                vm_o, vm_s = o.field("VxMed"), s.field("VxMed")
                say("vmed terms %d %s %s step %d: CFL pass VxMed bits %d" % (terms, par, ov, n, nbits(vm_s, vm_o)))
                assert nbits(active_rings(vm_s, s.dimz), active_rings(vm_o, s.dimz)) == 0, (terms, par, ov, n)
                dto, dts = o.step(), s.step()
                vm_o, vm_s = o.field("VxMed"), s.field("VxMed")
                say("vmed terms %d %s %s step %d dt %r VxMed bits %d" % (terms, par, ov, n, dts, nbits(vm_s, vm_o)))
                assert dto == dts, (terms, par, ov, n, dts, dto)
                assert nbits(active_rings(vm_s, s.dimz), active_rings(vm_o, s.dimz)) == 0, (terms, par, ov, n)
                assert np.array_equal(active_rings(s.nshift(), s.dimz), active_rings(o.nshift(), s.dimz))
            for f in fields_of(s):
                assert nbits(active(s.field(f), s.dimz), active(o.field(f), s.dimz)) == 0, (terms, par, ov, f)
            s.close()
    finally:
        L.f3d_fused_set_vmedterms(-1)
        L.f3d_fused_set_cflmode(-1)
        L.f3d_fused_set_xvmed(-1)


@pytest.mark.parametrize("ov", [C2_SMALL, C2_AD], ids=["c2_3d_small", "c2_3d_adiabatic_small"])
def test_fused_ghost_fill_random_interior(fa, orc, ov):
    """The fused ghost fill (Z then Y, src/boundary.c:2-23) on a random
    interior: every cell, ghosts and corners included, equals the oracle's
    boundaries() bit for bit."""
    o, s = make_pair(fa, orc, C2, ov)
    rng = np.random.default_rng(11)
    names = ["Density", "Energy", "Vx", "Vy", "Vz"]
    for n in names:
        a = o.field(n)
        pos = n in ("Density", "Energy")
        r = (0.5 + rng.random(a.shape)) * np.abs(a).mean() if pos else rng.standard_normal(a.shape)
        o.field(n)[...] = r
        s.set_field(n, r)
    o.call("boundaries")
    s.call("fused_boundaries")
    for n in names:
        a, b = s.field(n), o.field(n)
        say("ghost fill %s bits %d" % (n, nbits(a, b)))
        assert nbits(a, b) == 0, n
    s.close()


@pytest.mark.parametrize("split", [0, 1], ids=["one_thread_per_column", "two_threads_per_column"])
@pytest.mark.parametrize("par,ov,nsteps", [(C1, "", 8), (C2, C2_SMALL, 4), (C2, C2_AD, 4), (C3, C3_SMALL, 6)],
                         ids=["c1_2d", "c2_3d_small", "c2_3d_adiabatic_small", "c3_2d_adiabatic_small"])
def test_whole_step_fused_column_split(fa, orc, par, ov, nsteps, split):
    """The Y/Z column sweeps with every quantity in one thread, or split over
    two threads per column, give the oracle's results."""
    L = fa.lib()
    assert L.f3d_fused_set_colsplit(split) == 0
    try:
        _fused_run(fa, orc, par, ov, nsteps)
    finally:
        L.f3d_fused_set_colsplit(-1)
