"""Mutation check of the oracle pins.

Each entry below is a plausible one-line slip in the CPU restatement
(oracle/oracle_kernels.c) or in the host layer it shares with the product
(mesh, initial conditions, boundary factors): a wrong constant, sign, index
or operand.  The test applies it to a copy of the sources in a temporary
directory, builds that copy, and runs the CPU pins (tests/test_oracle.py and
tests/test_oracle_pins.py) against it through F3D_ORACLE_LIB; at least one
pin must fail for every slip.  Nothing here touches the product library or
the GPU.
"""
import concurrent.futures
import os
import shutil
import subprocess
import sys
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
KER = "oracle/oracle_kernels.c"
MESH = "fargo3d_amd/csrc/host/f3d_mesh.c"
PLAN = "fargo3d_amd/csrc/host/f3d_planets.c"
SETUP = "fargo3d_amd/csrc/host/f3d_setup.c"

# (file, text as in the source, replacement, which occurrence (0-based), label)
MUTANTS = [
    (KER, "- 1.0 / 6.0 * (slope[llxp] - slope[ll])", "- 1.0 / 8.0 * (slope[llxp] - slope[ll])", 0, "PPA_b 1/6 -> 1/8"),
    (KER, "if (6.0 * diff * cord > diff * diff)", "if (3.0 * diff * cord > diff * diff)", 0, "PPA_c 6 diff cord -> 3"),
    (KER, "qs[ll] = q[llxm] + 0.5 * (zone_size_x(IXM(i), j, k) - vx[ll] * dt) * slope[llxm];",
     "qs[ll] = q[llxm] + 0.5 * (zone_size_x(IXM(i), j, k) + vx[ll] * dt) * slope[llxm];", 0, "VanLeerX_b upwind sign"),
    (KER, "else slope[ll] = (2. * dqp * dqm) / (dqm + dqp);", "else slope[ll] = 0.5 * (dqm + dqp);", 0,
     "VanLeerX_a harmonic -> arithmetic"),
    (KER, "mpx[ll] = (vx[llxp] + ymed(j) * OMEGAFRAME) * ymed(j) * rho[ll];",
     "mpx[ll] = (vx[llxp] + ymed(j) * OMEGAFRAME) * ymin(j) * rho[ll];", 0, "momenta_x Mpx ymed -> ymin"),
    (KER, "vy[ll] = (mmy[ll] * Vol(i, j, k) + mpy[llym] * Vol(i, j - 1, k)) /",
     "vy[ll] = (mmy[ll] * Vol(i, j, k) + mpy[llym] * Vol(i, j, k)) /", 0, "NewVelocity_y Vol(j-1) -> Vol(j)"),
    (KER, "vy[l] -= (tauxx[l] + tauxx[lym]) / (ymin(j) * (rho[l] + rho[lym])) * dt;",
     "vy[l] += (tauxx[l] + tauxx[lym]) / (ymin(j) * (rho[l] + rho[lym])) * dt;", 0, "addviscosity_cyl hoop sign"),
    (KER, "vz[l] -= (tauxx[l] + tauxx[lzm]) * cos(zmin(k))", "vz[l] += (tauxx[l] + tauxx[lzm]) * cos(zmin(k))", 0,
     "addviscosity_sph vz cot sign"),
    (KER, "double cfl7 = 4.0 * viscosity", "double cfl7 = 2.0 * viscosity", 0, "cfl viscous factor 4 -> 2"),
    (KER, "double cfl5 = max3(cfl5_a, cfl5_b, cfl5_c) * 4.0 * CVNR;", "double cfl5 = max3(cfl5_a, cfl5_b, cfl5_c) * 4.0;",
     0, "cfl artificial viscosity CVNR dropped"),
    (KER, "double dqm = (q[ll] - q[llym]) / zone_size_y(j, k);", "double dqm = (q[ll] - q[llym]) / zone_size_y(j - 1, k);",
     0, "VanLeerY_a zone size j -> j-1"),
    (KER, "qs[ll] = q[llym] + 0.5 * (zone_size_y(j - 1, k) - vy[ll] * dt) * slope[llym];",
     "qs[ll] = q[llym] + 0.5 * (zone_size_y(j, k) - vy[ll] * dt) * slope[llym];", 0, "VanLeerY_b zone size j-1 -> j"),
    (KER, "double dqp = (q[llzp] - q[ll]) / zone_size_z(j, k + 1);", "double dqp = (q[llzp] - q[ll]) / zone_size_y(j, k + 1);",
     0, "VanLeerZ_a zone size of the other direction"),
    (KER, "qs[ll] = q[llzm] + 0.5 * (zone_size_z(j, k - 1) - vz[ll] * dt) * slope[llzm];",
     "qs[ll] = q[llzm] + 0.5 * (zone_size_z(j, k - 1) + vz[ll] * dt) * slope[llzm];", 0, "VanLeerZ_b upwind sign"),
    (KER, "qs[ll] += ksi * (1.0 - ksi) * (2.0 * q[llxm] - qR[llxm] - qL[llxm]);",
     "qs[ll] += ksi * (1.0 - ksi) * (2.0 * q[llxm] - qR[llxm] - qL[ll]);", 0, "PPA_d qL index"),
    (KER, "double ksi = -vx[ll] * dt / zone_size_x(i, j, k);", "double ksi = vx[ll] * dt / zone_size_x(i, j, k);", 0,
     "PPA_d negative-velocity xi sign"),
    (KER, "if (2.0 * fabs(dqp) < work) work = 2.0 * fabs(dqp);", "if (fabs(dqp) < work) work = fabs(dqp);", 0,
     "PPA_a MC factor 2 dropped on one side"),
    (KER, "qb[ll] += (vy[ll] * qs[ll] * rho_s[ll] * SurfY(i, j, k) - vy[llyp] * qs[llyp] * rho_s[llyp] * SurfY(i, j + 1, k)) *",
     "qb[ll] += (vy[ll] * qs[ll] * rho_s[ll] * SurfY(i, j, k) - vy[llyp] * qs[llyp] * rho_s[llyp] * SurfY(i, j, k)) *", 0,
     "UpdateY SurfY(j+1) -> SurfY(j)"),
    (KER, "qb[ll] += (vz[ll] * qs[ll] * rho_s[ll] * SurfZ(i, j, k) - vz[llzp] * qs[llzp] * rho_s[llzp] * SurfZ(i, j, k + 1)) *",
     "qb[ll] += (vz[ll] * qs[ll] * rho_s[ll] * SurfZ(i, j, k) - vz[llzp] * qs[llzp] * rho_s[llzp] * SurfZ(i, j, k)) *", 0,
     "UpdateZ SurfZ(k+1) -> SurfZ(k)"),
    (KER, "qb[ll] += ((vx[ll] * rho_s[ll] - vx[llxp] * rho_s[llxp]) * SurfX(j, k) * dt * InvVol(i, j, k));",
     "qb[ll] += ((vx[ll] * rho_s[ll] - vx[llxp] * rho_s[ll]) * SurfX(j, k) * dt * InvVol(i, j, k));", 0,
     "UpdateDensityX operand"),
    (KER, "vz[ll] = (mmz[ll] * Vol(i, j, k) + mpz[llzm] * Vol(i, j, k - 1)) /",
     "vz[ll] = (mmz[ll] * Vol(i, j, k) + mpz[llzm] * Vol(i, j, k)) /", 0, "NewVelocity_z Vol(k-1) -> Vol(k)"),
    (KER, "          mpz[ll] *= ymed(j);", "          mpz[ll] *= ymin(j);", 0, "momenta_z r factor"),
    (KER, "tauyx[l] = viscositym * .25 * (rho[l] + rho[lxm] + rho[lym] + rho[lxm - pitch]) *\n"
          "                   ((vy[l] - vy[lxm]) * InvDiffXmed[i] / ymin(j) + (vx[l] - vx[lym]) / (ymed(j) - ymed(j - 1)) -\n"
          "                    .5 * (vx[l] + vx[lym]) / ymin(j));",
     "tauyx[l] = viscositym * .25 * (rho[l] + rho[lxm] + rho[lym] + rho[lxm - pitch]) *\n"
     "                   ((vy[l] - vy[lxm]) * InvDiffXmed[i] / ymin(j) + (vx[l] - vx[lym]) / (ymed(j) - ymed(j - 1)) +\n"
     "                    .5 * (vx[l] + vx[lym]) / ymin(j));", 0, "visctensor_cyl tau_rphi sign"),
    (KER, "        tauxx[l] += viscosity * rho[l] * (vy[lyp] + vy[l]) / ymed(j);\n        tauyy[l] = viscosity * rho[l] * (2.0 * (vy[lyp] - vy[l]) / (ymin(j + 1) - ymin(j)) - div_v);\n        if (DIMZ) {\n          tauzz",
     "        tauxx[l] += viscosity * rho[l] * (vy[lyp] + vy[l]) / ymin(j);\n        tauyy[l] = viscosity * rho[l] * (2.0 * (vy[lyp] - vy[l]) / (ymin(j + 1) - ymin(j)) - div_v);\n        if (DIMZ) {\n          tauzz",
     0, "visctensor_cyl hoop ymed -> ymin"),
    (KER, "tauzz[l] += viscosity * rho[l] * (vy[l] + vy[lyp]) / ymed(j);", "tauzz[l] += viscosity * rho[l] * (vy[l] - vy[lyp]) / ymed(j);",
     0, "visctensor_sph tau_thth hoop sign"),
    (KER, "vx[l] += (tauyx[lyp] + tauyx[l]) / ((rho[lxm] + rho[l]) * ymed(j)) * dt;",
     "vx[l] -= (tauyx[lyp] + tauyx[l]) / ((rho[lxm] + rho[l]) * ymed(j)) * dt;", 0, "addviscosity_sph vx r term sign"),
    (KER, "vy[l] -= (tauzz[l] + tauzz[lym]) / (ymin(j) * (rho[l] + rho[lym])) * dt;",
     "vy[l] += (tauzz[l] + tauzz[lym]) / (ymin(j) * (rho[l] + rho[lym])) * dt;", 0, "addviscosity_sph vy tau_thth sign"),
    (KER, "vy_temp[ll] += -2.0 * (pres_y[ll] - pres_y[llym]) / (rho[ll] * dyrho1 + rho[llym] * dyrho2) * dt;",
     "vy_temp[ll] += -2.0 * (pres_y[ll] - pres_y[llym]) / (rho[ll] * dyrho1 + rho[llym] * dyrho1) * dt;", 0,
     "SubStep2_b Y zone size"),
    (KER, "vz_temp[ll] += -2.0 * dzmed * (pres_z[ll] - pres_z[llzm]) / (rho[ll] * dzrho1 + rho[llzm] * dzrho2) * dt /\n"
          "                           (ymed(j) * (zmed(k) - zmed(k - 1)));",
     "vz_temp[ll] += -2.0 * dzmed * (pres_z[ll] - pres_z[llzm]) / (rho[ll] * dzrho1 + rho[llzm] * dzrho2) * dt /\n"
     "                           (zmed(k) - zmed(k - 1));", 0, "SubStep2_b Z spherical r dropped"),
    (KER, "if (adiabatic) e[ll] += -dt * (pres_y[ll] * (vy[llyp] - vy[ll]) / zone_size_y(j, k));",
     "if (adiabatic) e[ll] += -dt * (pres_y[ll] * (vy[llyp] - vy[ll]) / zone_size_y(j + 1, k));", 0,
     "SubStep2_b Y heating zone size"),
    (KER, "smoothing = S->p.aspectratio * pow(planetdistance / R0, S->p.flaringindex) * planetdistance *\n"
          "                        S->p.thicknesssmoothing;",
     "smoothing = S->p.aspectratio * pow(planetdistance / R0, S->p.flaringindex) *\n"
     "                        S->p.thicknesssmoothing;", 0, "Potential smoothing r_p dropped"),
    (KER, "pot[l] += G * mp * (XC * sys->x[n] + YC * sys->y[n] + ZC * sys->z[n]) /",
     "pot[l] -= G * mp * (XC * sys->x[n] + YC * sys->y[n] + ZC * sys->z[n]) /", 0, "Potential indirect sign"),
    (KER, "double rroche = planetdistance * pow((1.0 / 3.0 * mp / MSTAR), 1.0 / 3.0);",
     "double rroche = planetdistance * pow((1.0 / 3.0 * mp / MSTAR), 1.0 / 2.0);", 0, "Potential Roche exponent"),
    (KER, "vt[l] -= domega * ymed(j) * sin(zmed(k));", "vt[l] -= domega * ymed(j);", 0, "CorrectVtheta spherical sin"),
    (KER, "double tau = ds * sqrt(ymed(j) * ymed(j) * ymed(j) / G / MSTAR);", "double tau = ds * sqrt(ymed(j) * ymed(j) / G / MSTAR);",
     0, "Stockholm tau exponent"),
    (KER, "vx0_target -= (of - of0) * radius;", "vx0_target += (of - of0) * radius;", 0, "Stockholm frame correction sign"),
    (KER, "if (ymed(j) < Y_inf) rampy = (Y_inf - ymed(j)) / (Y_inf - y_min);",
     "if (ymed(j) < Y_inf) rampy = (Y_inf - ymed(j)) / (y_max - Y_inf);", 0, "Stockholm inner ramp width"),
    (KER, "yczc[ll] = S->p.alpha * soundspeed2 / sqrt(G * MSTAR / r3yczc);", "yczc[ll] = S->p.alpha * soundspeed2 * sqrt(G * MSTAR / r3yczc);",
     0, "dust diffusion D = alpha cs^2 Omega"),
    (KER, "double cellmass = Vol(i, j, k) * dens[ll];\n        double dx, dy, dz;",
     "double cellmass = dens[ll];\n        double dx, dy, dz;", 0, "ComputeForce cell mass without volume"),
    (KER, "dist2 += rsmoothing * rsmoothing;\n        double distance = sqrt(dist2);\n        double InvDist3 = 1.0 / (dist2 * distance);\n        InvDist3 *= G * cellmass;\n        fxi",
     "dist2 += rsmoothing;\n        double distance = sqrt(dist2);\n        double InvDist3 = 1.0 / (dist2 * distance);\n        InvDist3 *= G * cellmass;\n        fxi",
     0, "ComputeForce smoothing not squared"),
    (KER, "double shearlimit = CFL * Mindx / fabs(vxmed[l2] / Ymed[j] - vxmed[l2 + 1] / Ymed[j + 1]);",
     "double shearlimit = CFL * Mindx / fabs(vxmed[l2] / Ymed[j] - vxmed[l2 + 1] / Ymed[j]);", 0, "cfl_b shear radius index"),
    (KER, "if (iso) soundspeed2 = cs[ll] * cs[ll];", "if (iso) soundspeed2 = cs[ll];", 0, "cfl isothermal cs^2"),
    (KER, "if (!iso) soundspeed2 = GAMMA * (GAMMA - 1) * cs[ll] / rho[ll];", "if (!iso) soundspeed2 = GAMMA * cs[ll] / rho[ll];", 0,
     "cfl adiabatic sound speed"),
    (KER, "viscosity = ALPHA * cs[ll] * cs[ll] * sqrt(ymed(j) * ymed(j) * ymed(j) / (G * MSTAR));",
     "viscosity = ALPHA * cs[ll] * sqrt(ymed(j) * ymed(j) * ymed(j) / (G * MSTAR));", 0, "cfl alpha viscosity cs^2"),
    (KER, "double cfl3 = (max2(fabs(vy[ll]), fabs(vy[llyp]))) / zone_size_y(j, k);",
     "double cfl3 = (max2(fabs(vy[ll]), fabs(vy[llyp]))) / zone_size_x(i, j, k);", 0, "cfl vy over dx"),
    (PLAN, "o.a = h * h / (G * m * (1. - e * e));", "o.a = h * h / (G * m * (1. + e * e));", 0,
     "guiding centre semi-major axis"),
    (PLAN, "return sqrt(G * m / pow(o.a, 3.0));", "return sqrt(G * m / pow(o.a, 2.0));", 0, "frame mean motion exponent"),
    (PLAN, "sqrt((1.0 - p->eccentricity) / (1.0 + p->eccentricity)) * cos(p->inclination);",
     "sqrt((1.0 + p->eccentricity) / (1.0 - p->eccentricity)) * cos(p->inclination);", 0, "planet apocentre speed"),
    (MESH, "g->szj[j] = 0.5 * (g->ymin[j + 1] * g->ymin[j + 1] - g->ymin[j] * g->ymin[j]);\n      g->invvj[j] = 1.0 / g->szj[j];",
     "g->szj[j] = 0.5 * (g->ymin[j + 1] * g->ymin[j + 1] - g->ymin[j] * g->ymin[j]);\n      g->invvj[j] = 2.0 / g->szj[j];",
     0, "mesh cylindrical InvVj factor"),
    (MESH, "g->invvj[j] = 3. / ((g->ymin[j + 1] * g->ymin[j + 1] * g->ymin[j + 1] -",
     "g->invvj[j] = 2. / ((g->ymin[j + 1] * g->ymin[j + 1] * g->ymin[j + 1] -", 0, "mesh spherical InvVj factor"),
    (MESH, "g->syk[k] = cos(g->zmin[k]) - cos(g->zmin[k + 1]);", "g->syk[k] = sin(g->zmin[k + 1]) - sin(g->zmin[k]);", 0,
     "mesh spherical Syk"),
    (MESH, "g->szk[k] = sin(g->zmin[k]);", "g->szk[k] = sin(g->zmed[k]);", 0, "mesh spherical Szk at the centre"),
    (MESH, "for (int j = 0; j < sy + 1; j++) g->ymin[j] = exp(log(p->ymin) + dy * (j - g->nghy));",
     "for (int j = 0; j < sy + 1; j++) g->ymin[j] = exp(log(p->ymin) + dy * (j - g->nghy + 1));", 0, "mesh log offset"),
    (MESH, "g->syj[j] = g->ymin[j];", "g->syj[j] = g->ymed[j];", 0, "mesh cylindrical Syj at the centre"),
    (SETUP, "x = Vv * (1. + d->aspectratio * d->noise * (u[1] - .5));", "x = Vv * (1. + d->aspectratio * d->noise * (u[1] - .6));",
     0, "IC fargo azimuthal noise offset"),
    (SETUP, "r = Rv * (1.0 + d->noise * (u[0] - .5));\n      break;\n    case F3D_IC_FARGO:",
     "r = Rv * (1.0 + d->noise * (u[0] - .4));\n      break;\n    case F3D_IC_FARGO:", 0,
     "IC Keplerian noise offset"),
    (SETUP, "(2.0 * p->flaringindex - 1.0 - p->sigmaslope));\n      v -= OMEGAFRAME * r;",
     "(2.0 * p->flaringindex - 1.0 + p->sigmaslope));\n      v -= OMEGAFRAME * r;", 0,
     "IC fargo pressure support sign"),
    (SETUP, "pow(sz, -beta - xi + 1. / (h * h));", "pow(sz, -beta + xi + 1. / (h * h));", 0, "IC 3-D colatitude exponent"),
    (SETUP, "r *= exp(d->noise * (sqrt(-2.0 * log(u1)) * cos(2.0 * M_PI * u2)));",
     "r *= 1.0 + d->noise * (sqrt(-2.0 * log(u1)) * cos(2.0 * M_PI * u2));", 0, "IC adiabatic noise not log-normal"),
    (SETUP, "t->ymin_vsq[m] = sqrt(Ymed[jact] / Ymed[jgh]);", "t->ymin_vsq[m] = sqrt(Ymed[jgh] / Ymed[jact]);", 0,
     "boundary 2-D inner Keplerian factor inverted"),
    (SETUP, "double pw = pow(Ymed[jact] / Ymed[jgh], SS + FI + 1.);", "double pw = pow(Ymed[jact] / Ymed[jgh], SS + FI);", 0,
     "boundary 3-D radial density exponent"),
    (SETUP, "FI - 2. - SS + 1. / (AR * AR) * pow(Ymed[j] / R0, -2. * FI));", "FI - 2. + SS + 1. / (AR * AR) * pow(Ymed[j] / R0, -2. * FI));",
     0, "boundary 3-D colatitude density exponent"),
    (SETUP, "double e = exp(-pow(cos(Zmed[k]), 2.) / (AR * AR) *", "double e = exp(pow(cos(Zmed[k]), 2.) / (AR * AR) *", 0,
     "boundary 3-D flared radial exp sign"),
    (SETUP, "d->Y_sup = y_max * pow(p->dampingzone, -2.0 / 3.0);", "d->Y_sup = y_max * pow(p->dampingzone, -1.0 / 3.0);", 0,
     "damping outer zone exponent"),
]

SOURCES = ["oracle/oracle_kernels.c", "oracle/oracle_step.c", "oracle/oracle.h", "oracle/oracle_metric.h",
           "fargo3d_amd/csrc/host/f3d_mesh.c", "fargo3d_amd/csrc/host/f3d_param.c",
           "fargo3d_amd/csrc/host/f3d_planets.c", "fargo3d_amd/csrc/host/f3d_setup.c",
           "fargo3d_amd/csrc/host/f3d_host.h"]
PINS = ["tests/test_oracle.py", "tests/test_oracle_pins.py"]


def apply(text, old, new, nth):
    pos = -1
    for _ in range(nth + 1):
        pos = text.find(old, pos + 1)
        if pos < 0:
            raise AssertionError("mutation site not found: %r" % old[:60])
    return text[:pos] + new + text[pos + len(old):]


def build_copy(tmp, mutant):
    for rel in SOURCES:
        dst = os.path.join(tmp, rel)
        os.makedirs(os.path.dirname(dst), exist_ok=True)
        shutil.copy(os.path.join(ROOT, rel), dst)
    if mutant is not None:
        rel, old, new, nth, _ = mutant
        path = os.path.join(tmp, rel)
        text = open(path).read()
        out = apply(text, old, new, nth)
        assert out != text
        open(path, "w").write(out)
    lib = os.path.join(tmp, "liboracle.so")
    srcs = [os.path.join(tmp, r) for r in SOURCES if r.endswith(".c")]
    subprocess.run(["gcc", "-O1", "-fPIC", "-std=gnu11", "-ffp-contract=off", "-fno-fast-math", "-w", "-shared",
                    "-I" + os.path.join(ROOT, "include"), "-o", lib] + srcs + ["-lm"], check=True)
    return lib


def run_pins(lib):
    env = dict(os.environ, F3D_ORACLE_LIB=lib)
    r = subprocess.run([sys.executable, "-m", "pytest", "-x", "-q", "-p", "no:cacheprovider", "-m", "not gpu"] + PINS,
                       cwd=ROOT, env=env, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=300)
    return r.returncode, r.stdout


def test_mutation_sites_are_unique_enough():
    """Every mutation site exists in the current sources."""
    for rel, old, new, nth, label in MUTANTS:
        text = open(os.path.join(ROOT, rel)).read()
        apply(text, old, new, nth)


def test_unmutated_copy_passes(tmp_path):
    """The copy-and-build recipe itself yields an oracle that passes the pins."""
    code, out = run_pins(build_copy(str(tmp_path), None))
    assert code == 0, out[-3000:]


def _survives(mutant):
    with tempfile.TemporaryDirectory() as tmp:
        code, out = run_pins(build_copy(tmp, mutant))
    return None if code == 1 else "%s (exit %d): %s" % (mutant[4], code, out[-600:])


def test_every_mutant_fails_a_pin():
    """All slips, a few builds and pin runs at a time."""
    workers = max(1, min(8, os.cpu_count() or 1))
    with concurrent.futures.ThreadPoolExecutor(workers) as pool:
        survivors = [r for r in pool.map(_survives, MUTANTS) if r]
    assert not survivors, "slips that passed every pin:\n" + "\n".join(survivors)
