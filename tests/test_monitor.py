"""Monitoring: the monitored quantities (src/mon_dens.c, src/mon_momx.c,
src/mon_momy.c, src/mon_momz.c, src/mon_torque.c) and MonitorGlobal's sums
and files (src/monitor.c).

The CPU restatement is pinned against closed forms on the cylindrical and
spherical meshes:
cell volumes (r_{j+1}^2 - r_j^2) dphi / 2 from the mesh edges, the frame term
of the angular momentum, the face average of vy, and the point-mass torque of
each cell computed from its Cartesian position; a periodic-neighbour slip, a
dropped frame term or a swapped torque sign fails.  The GPU kernels match the
restatement bit for bit on random states (2-D, 3-D spherical, a dust fluid,
the precision form's FP32 fields), and the driver writes the files the
set-ups ask for.
"""
import ctypes
import os

import numpy as np
import pytest

import oracle

C1 = "fargo3d_amd/setups/c1_fargo2d.par"
C2 = "fargo3d_amd/setups/c2_p3d.par"
C4 = "fargo3d_amd/setups/c4_multifluid2d.par"
QNAME = ["dens", "momx", "momy", "momz", "torq"]


def _state(rng, shape):
    return (0.5 + rng.random(shape), rng.standard_normal(shape), rng.standard_normal(shape),
            rng.standard_normal(shape))


def test_oracle_monitors_closed_forms():
    nx, ny = 64, 32
    o = oracle.Oracle(C1, "Nx=%d Ny=%d" % (nx, ny))
    o.set_shape(nx, ny, 1, 0)
    rng = np.random.default_rng(5)
    rho, vx, vy, _ = _state(rng, (1, ny + 6, nx))
    o.field("Density")[...] = rho
    o.field("Vx")[...] = vx
    o.field("Vy")[...] = vy
    om = o.scalars()["OMEGAFRAME"]
    dy = (2.5 - 0.4) / ny
    ymin = 0.4 + dy * (np.arange(ny + 7) - 3)
    r = 0.5 * (ymin[1:] + ymin[:-1])
    dphi = 2 * np.pi / nx
    phi = -np.pi + (np.arange(nx) + 0.5) * dphi
    vol = 0.5 * (ymin[1:] ** 2 - ymin[:-1] ** 2)[:, None] * dphi
    act = slice(3, 3 + ny)

    def got(q, *pl):
        o.call("monitor", q, *(pl or (0.0, 0.0, 0.0, 0.0)))
        return o.field("Slope")[0, act, :].copy()

    def close(a, b):
        assert np.max(np.abs(a - b)) <= 1e-13 * np.max(np.abs(b))

    m = (rho[0] * vol)[act]
    close(got(0), m)
    vxm = 0.5 * (vx[0] + np.roll(vx[0], -1, axis=1))          # i+1 wraps to 0
    close(got(1), ((vxm + r[:, None] * om) * r[:, None] * rho[0] * vol)[act])
    vym = 0.5 * (vy[0, :-1] + vy[0, 1:])
    close(got(2), (rho[0, :-1] * vym * vol[:-1])[act])
    xp, yp, rs = 1.0, 0.05, 0.03
    cx, cy = r[:, None] * np.cos(phi), r[:, None] * np.sin(phi)
    dx, dy_ = cx - xp, cy - yp
    d2 = dx * dx + dy_ * dy_ + rs * rs
    f = rho[0] * vol / (d2 * np.sqrt(d2))
    close(got(4, xp, yp, 0.0, rs), (xp * dy_ * f - yp * dx * f)[act])
    # a mass just ahead of the planet (larger azimuth) pulls it forward
    o.field("Density")[...] = 1e-12
    o.field("Density")[0, 3 + ny // 2, nx // 2 + 2] = 1.0
    t = got(4, r[3 + ny // 2], 0.0, 0.0, rs)
    assert t.sum() > 0
    o.close()


def test_oracle_monitors_spherical_closed_forms():
    """Spherical branches: volumes r^3/3 dphi dcos(theta), the cylindrical
    radius r sin(theta) of the angular momentum, the r factor of momz and the
    torque of cells placed at (r cos phi sin theta, r sin phi sin theta, r cos theta)."""
    nx, ny, nz = 16, 8, 6
    o = oracle.Oracle(C2, "Nx=%d Ny=%d Nz=%d" % (nx, ny, nz))
    o.set_shape(nx, ny, nz, 1)
    rng = np.random.default_rng(9)
    rho, vx, _, vz = _state(rng, (nz + 6, ny + 6, nx))
    o.field("Density")[...] = rho
    o.field("Vx")[...] = vx
    o.field("Vz")[...] = vz
    om = o.scalars()["OMEGAFRAME"]
    dr = (2.5 - 0.4) / ny
    rmin = 0.4 + dr * (np.arange(ny + 7) - 3)
    th0, th1 = 1.42079632679489661922, 1.72079632679489661922
    dth = (th1 - th0) / nz
    tmin = th0 + dth * (np.arange(nz + 7) - 3)
    r = 0.5 * (rmin[1:] + rmin[:-1])
    th = 0.5 * (tmin[1:] + tmin[:-1])
    dphi = 2 * np.pi / nx
    phi = -np.pi + (np.arange(nx) + 0.5) * dphi
    vol = ((rmin[1:] ** 3 - rmin[:-1] ** 3) / 3)[None, :, None] * dphi * \
        (np.cos(tmin[:-1]) - np.cos(tmin[1:]))[:, None, None]
    act = (slice(3, 3 + nz), slice(3, 3 + ny))

    def got(q, *pl):
        o.call("monitor", q, *(pl or (0.0, 0.0, 0.0, 0.0)))
        return o.field("Slope")[act].copy()

    def close(a, b):
        assert np.max(np.abs(a - b)) <= 1e-12 * np.max(np.abs(b))

    close(got(0), (rho * vol)[act])
    rc = r[None, :, None] * np.sin(th)[:, None, None]
    vxm = 0.5 * (vx + np.roll(vx, -1, axis=2))
    close(got(1), ((vxm + rc * om) * rc * rho * vol)[act])
    vzm = np.zeros_like(vz)
    vzm[:-1] = 0.5 * (vz[:-1] + vz[1:])
    close(got(3), (rho * vzm * vol * r[None, :, None])[act])
    xp, yp, zp, rs = 1.0, 0.03, 0.02, 0.05
    st = np.sin(th)[:, None, None]
    cx = r[None, :, None] * np.cos(phi)[None, None, :] * st
    cy = r[None, :, None] * np.sin(phi)[None, None, :] * st
    cz = r[None, :, None] * np.cos(th)[:, None, None] + 0 * phi
    dx, dy, dz = cx - xp, cy - yp, cz - zp
    d2 = dx * dx + dy * dy + dz * dz + rs * rs
    f = rho * vol / (d2 * np.sqrt(d2))
    close(got(4, xp, yp, zp, rs), (xp * dy * f - yp * dx * f)[act])
    o.close()


def _sim_call(s, fp32, q, xp=0.0, yp=0.0, zp=0.0, rs=0.0):
    fn = getattr(s.L, ("f3d32_mon_" if fp32 else "f3d_mon_") + QNAME[q])
    fn.restype = ctypes.c_int
    out = ctypes.c_void_p(s.device_ptr("Pot"))
    args = [ctypes.c_void_p(s.ctx)]
    if q == 4:
        args += [ctypes.c_double(xp), ctypes.c_double(yp), ctypes.c_double(zp), ctypes.c_double(rs)]
    assert fn(*(args + [out])) == 0, s.L.f3d_hip_last_error(s.ctx).decode()


PARITY = {
    "c1_2d": (C1, "Nx=64 Ny=32", 0, 0),
    "c2_3d_sph": (C2, "Nx=32 Ny=16 Nz=8", 0, 0),
    "c4_dust_fluid": (C4, "Nx=64 Ny=32", 2, 0),
    "c2_precision_form": (C2, "Nx=32 Ny=16 Nz=8 Precision=32", 0, 1),
}


@pytest.mark.gpu
@pytest.mark.parametrize("case", list(PARITY))
def test_monitor_parity(case):
    import fargo3d_amd as fa
    par, ov, fluid, fp32 = PARITY[case]
    print("monitor parity", case, flush=True)
    s = fa.Simulation(par, ov)
    o = oracle.Oracle(par, ov.replace(" Precision=32", ""))
    o.set_shape(s.nx, s.ny, s.nz, s.dimz)
    s.select_fluid(fluid)
    o.select_fluid(fluid)
    rng = np.random.default_rng(11)
    names = ["Density", "Vx", "Vy"] + (["Vz"] if s.dimz else [])
    for n, a in zip(names, _state(rng, s.shape)):
        s.set_field(n, a)
        if fp32:
            # the values the precision form holds (FP32; the azimuthal
            # velocity as the FP32 deviation from its FP64 ring profile)
            a = s.field(n)
        o.field(n)[...] = a
    pl = o.planets()[0]
    for q in range(5):
        if q == 3 and not s.dimz:
            continue
        args = (float(pl[0]), float(pl[1]), float(pl[2]), 0.021) if q == 4 else ()
        o.call("monitor", q, *(args or (0.0, 0.0, 0.0, 0.0)))
        _sim_call(s, fp32, q, *args)
        want = o.field("Slope")
        gotf = s.field("Pot")
        sl = (slice(3, -3), slice(3, -3)) if s.dimz else (slice(None), slice(3, -3))
        a, b = gotf[sl], want[sl]
        nb = int(np.count_nonzero(a.view(np.int64) != b.view(np.int64)))
        print("  ", QNAME[q], "bits differ:", nb, flush=True)
        assert nb == 0
    s.close()
    o.close()


@pytest.mark.gpu
@pytest.mark.parametrize("par,ov", [(C1, "Nx=64 Ny=32 Ntot=2 Ninterm=1"), (C2, "Nx=32 Ny=16 Nz=8 Ntot=2 Ninterm=1")])
def test_monitor_files(tmp_path, par, ov):
    import fargo3d_amd as fa
    d = str(tmp_path)
    s = fa.Simulation(par, ov)
    s.write_init(d)
    s.main_loop(0, d)
    mon = os.path.join(d, "monitor", "gas")
    torq = np.loadtxt(os.path.join(mon, "torq_planet_0.dat"))
    assert torq.shape == (2, 2)
    assert torq[-1, 0] == pytest.approx(s.scalars()["PhysicalTime"], rel=1e-11)
    big = open(os.path.join(d, "bigplanet0.dat")).read().splitlines()
    assert [int(x.split("\t")[0]) for x in big] == [0, 1]
    if s.dimz:
        # p3disof: MASS and MOM_X as 2-D maps per pass, TORQ as a scalar
        m2 = np.fromfile(os.path.join(mon, "FG000001", "mass_2d_0000001.dat"))
        assert m2.size == s.ny * s.nz and np.all(m2 > 0)
    else:
        # fargo: MASS, MOM_X, TORQ scalars and the raw Y profile of TORQ
        mass = np.loadtxt(os.path.join(mon, "mass.dat"))
        prof = np.fromfile(os.path.join(mon, "torq_1d_Y_raw_planet_0.dat"))
        assert prof.size == 2 * s.ny
        # the profile sums to the scalar (both from the same ring sums)
        assert np.sum(prof[-s.ny:]) == pytest.approx(torq[-1, 1], rel=1e-9, abs=1e-15)
        # the scalar mass is the sum of rho * V over the final state
        dy = (2.5 - 0.4) / s.ny
        ymin = 0.4 + dy * (np.arange(s.ny + 1))
        vol = 0.5 * (ymin[1:] ** 2 - ymin[:-1] ** 2) * (2 * np.pi / s.nx)
        rho = s.field("Density")[0, 3:-3, :]
        assert mass[-1, 1] == pytest.approx(float(np.sum(rho * vol[:, None])), rel=1e-11)
    s.close()
