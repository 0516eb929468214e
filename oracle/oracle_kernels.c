/*
 * oracle_kernels.c -- CPU restatement of the hot-path routines.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Each function restates the
 * reference's *_cpu routine named in its comment, with the same loop ranges,
 * scratch-field use and expression association.  "#ifdef Z" in the reference
 * becomes the run-time flag DIMZ: without Z the k loop is absent (k = 0).
 *
 * Parity status of each function (how it is pinned) is listed in the module
 * docstring of tests/test_oracle.py.
 */
#include <math.h>
#include <stdio.h>
#include <string.h>

#include "oracle.h"
#include "oracle_metric.h"

#define G F3D_G
#define MSTAR F3D_MSTAR
#define R0 F3D_R0
#define CVNR F3D_CVNR

#define KRANGE(kb_, size_z_) \
  const int kbeg = DIMZ ? (kb_) : 0; const int kend = DIMZ ? (size_z_) : 1

/* src/CompPresIso.c:10-64 */
void orc_ComputePressureFieldIso(orc_state *S) {
  ORC_MESH(S);
  const double *dens = S->Density, *cs = S->Energy;
  double *pres = S->Pressure;
  int size_y = Ny + 2 * NGHY, size_z = Nz + 2 * NGHZ;
  KRANGE(0, size_z);
  for (int k = kbeg; k < kend; k++)
    for (int j = 0; j < size_y; j++)
      for (int i = 0; i < Nx; i++) {
        int ll = lidx(i, j, k);
        pres[ll] = dens[ll] * cs[ll] * cs[ll];
      }
}

/* src/CompPresAd.c:10-66 */
void orc_ComputePressureFieldAd(orc_state *S) {
  ORC_MESH(S);
  const double *e = S->Energy;
  double *pres = S->Pressure;
  const double GAMMA = S->p.gamma;
  int size_y = Ny + 2 * NGHY, size_z = Nz + 2 * NGHZ;
  KRANGE(0, size_z);
  for (int k = kbeg; k < kend; k++)
    for (int j = 0; j < size_y; j++)
      for (int i = 0; i < Nx; i++) {
        int ll = lidx(i, j, k);
        pres[ll] = (GAMMA - 1.0) * e[ll];
      }
}

/* src/compute_vmed.c:42-97 */
void orc_ComputeVweight(orc_state *S, const double *v, double *q) {
  ORC_MESH(S);
  int size_y = Ny + 2 * NGHY, size_z = Nz + 2 * NGHZ;
  KRANGE(0, size_z);
  for (int k = kbeg; k < kend; k++)
    for (int j = 0; j < size_y; j++)
      for (int i = 0; i < Nx; i++) {
        int ll = lidx(i, j, k);
        q[ll] = v[ll] * (xmin(i + 1) - xmin(i));
      }
}

/* src/reduction_generic.c:5-23 with macro SUM */
void orc_reduction_SUM(orc_state *S, const double *f, int ymin_, int ymax_, int zmin_, int zmax_) {
  ORC_MESH(S);
  double *reduc2d = S->Reduction2D;
  for (int k = zmin_; k < zmax_; k++)
    for (int j = ymin_; j < ymax_; j++) {
      reduc2d[L2D(j, k)] = 0.0;
      for (int i = 0; i < Nx; i++) reduc2d[L2D(j, k)] = reduc2d[L2D(j, k)] + f[lidx(i, j, k)];
    }
}

static void orc_reduction_MIN(orc_state *S, const double *f, int ymin_, int ymax_, int zmin_, int zmax_) {
  ORC_MESH(S);
  double *reduc2d = S->Reduction2D;
  for (int k = zmin_; k < zmax_; k++)
    for (int j = ymin_; j < ymax_; j++) {
      reduc2d[L2D(j, k)] = 1e20;
      for (int i = 0; i < Nx; i++) {
        double a = reduc2d[L2D(j, k)], b = f[lidx(i, j, k)];
        reduc2d[L2D(j, k)] = ((a) < (b) ? (a) : (b));
      }
    }
}

/* src/reduction_full_generic.c:1-18 */
double orc_reduction_full_SUM(orc_state *S, const double *F, int ymin_, int ymax_, int zmin_, int zmax_) {
  ORC_MESH(S);
  orc_reduction_SUM(S, F, ymin_, ymax_, zmin_, zmax_);
  double result = 0.0;
  for (int k = zmin_; k < zmax_; k++)
    for (int j = ymin_; j < ymax_; j++) result = S->Reduction2D[L2D(j, k)] + result;
  return result;
}

double orc_reduction_full_MIN(orc_state *S, const double *F, int ymin_, int ymax_, int zmin_, int zmax_) {
  ORC_MESH(S);
  orc_reduction_MIN(S, F, ymin_, ymax_, zmin_, zmax_);
  double result = 1e20;
  for (int k = zmin_; k < zmax_; k++)
    for (int j = ymin_; j < ymax_; j++) {
      double a = S->Reduction2D[L2D(j, k)];
      result = ((a) < (result) ? (a) : (result));
    }
  return result;
}

/* src/compute_vmed.c:10-39 */
void orc_ComputeVmed(orc_state *S, const double *V) {
  ORC_MESH(S);
  orc_ComputeVweight(S, V, S->Qs);
  orc_reduction_SUM(S, S->Qs, 0, Ny + 2 * NGHY, 0, Nz + 2 * NGHZ);
  for (int k = 0; k < (DIMZ ? Nz + 2 * NGHZ : 1); k++)
    for (int j = 0; j < Ny + 2 * NGHY; j++)
      S->VxMed[L2D(j, k)] = S->Reduction2D[L2D(j, k)] / (GR->XMAX - GR->XMIN);
}

/* src/cfl.c:10-270 (hydro terms; no MHD / non-ideal terms) */
void orc_cfl(orc_state *S) {
  ORC_MESH(S);
  const double *cs = S->Energy, *rho = S->Density, *vx = S->Vx, *vy = S->Vy, *vz = S->Vz;
  const double *vxmed = S->VxMed;
  double *dtime = S->DensStar;
  const double CFL = S->p.cfl, NU = S->p.nu, ALPHA = S->p.alpha, GAMMA = S->p.gamma;
  const int fargo = S->p.fargo, iso = (S->p.eos == F3D_ISOTHERMAL);
  int size_y = Ny + NGHY, size_z = Nz + NGHZ;
  const int kbeg = DIMZ ? NGHZ : 0, kend = DIMZ ? size_z : 1;
  for (int k = kbeg; k < kend; k++)
    for (int j = NGHY; j < size_y; j++)
      for (int i = 0; i < Nx; i++) {
        int ll = lidx(i, j, k);
        int llxp = LXP(ll, i), llyp = LYP(ll), llzp = LZP(ll);
        double vxx, vxxp, soundspeed2, viscosity;
        if (!fargo) {
          vxx = vx[ll];
          vxxp = vx[llxp];
        } else {
          vxx = vx[ll] - vxmed[L2D(j, k)];
          vxxp = vx[llxp] - vxmed[L2D(j, k)];
        }
        if (iso) soundspeed2 = cs[ll] * cs[ll];
        if (S->p.visc == F3D_VISC_ALPHA) {
          if (iso)
            viscosity = ALPHA * cs[ll] * cs[ll] * sqrt(ymed(j) * ymed(j) * ymed(j) / (G * MSTAR));
          else
            viscosity = ALPHA * GAMMA * (GAMMA - 1.0) * cs[ll] / rho[ll] *
                        sqrt(ymed(j) * ymed(j) * ymed(j) / (G * MSTAR));
        } else {
          viscosity = NU;
        }
        if (!iso) soundspeed2 = GAMMA * (GAMMA - 1) * cs[ll] / rho[ll];
        double soundspeed = sqrt(soundspeed2);
        double cfl1_a = soundspeed / zone_size_x(i, j, k);
        double cfl1_b = soundspeed / zone_size_y(j, k);
        double cfl1_c = DIMZ ? soundspeed / zone_size_z(j, k) : 0.0;
        double cfl1 = max3(cfl1_a, cfl1_b, cfl1_c);
        double cfl2 = (max2(fabs(vxx), fabs(vxxp))) / zone_size_x(i, j, k);
        double cfl3 = (max2(fabs(vy[ll]), fabs(vy[llyp]))) / zone_size_y(j, k);
        double cfl4 = DIMZ ? (max2(fabs(vz[ll]), fabs(vz[llzp]))) / zone_size_z(j, k) : 0.0;
        double cfl5_a = fabs(vx[llxp] - vx[ll]) / zone_size_x(i, j, k);
        double cfl5_b = fabs(vy[llyp] - vy[ll]) / zone_size_y(j, k);
        double cfl5_c = DIMZ ? fabs(vz[llzp] - vz[ll]) / zone_size_z(j, k) : 0.0;
        double cfl5 = max3(cfl5_a, cfl5_b, cfl5_c) * 4.0 * CVNR;
        double cfl6 = 0.0;
        double cfl7_a = 1.0 / zone_size_x(i, j, k);
        double cfl7_b = 1.0 / zone_size_y(j, k);
        double cfl7_c = DIMZ ? 1.0 / zone_size_z(j, k) : 0.0;
        double m7 = max3(cfl7_a, cfl7_b, cfl7_c);
        double cfl7 = 4.0 * viscosity * (m7 * m7); /* pow(x,2) is folded to x*x by gcc */
        double cfl8 = 0.0, cfl9 = 0.0, cfl10 = 0.0;
        dtime[ll] = CFL / sqrt(cfl1 * cfl1 + cfl2 * cfl2 + cfl3 * cfl3 + cfl4 * cfl4 + cfl5 * cfl5 +
                               cfl6 * cfl6 + cfl7 * cfl7 + cfl8 * cfl8 + cfl9 * cfl9 + cfl10 * cfl10);
      }
  orc_cfl_b(S);
}

/* src/cfl_b.c:3-41 */
void orc_cfl_b(orc_state *S) {
  ORC_MESH(S);
  double m = orc_reduction_full_MIN(S, S->DensStar, NGHY, Ny + NGHY, NGHZ, Nz + NGHZ);
  if (S->p.fargo) {
    const double *vxmed = S->VxMed;
    const double CFL = S->p.cfl;
    const double Mindx = GR->mindx;
    const int kbeg = DIMZ ? NGHZ : 0, kend = DIMZ ? Nz + NGHZ : 1;
    for (int k = kbeg; k < kend; k++)
      for (int j = NGHY; j < Ny + NGHY; j++) {
        int l2 = L2D(j, k);
        double shearlimit = CFL * Mindx / fabs(vxmed[l2] / Ymed[j] - vxmed[l2 + 1] / Ymed[j + 1]);
        if (shearlimit < m) m = shearlimit;
      }
  }
  S->fl[S->fluid_index].Min = m;
}

/* src/cfl_fluids_min.c:3-22; returns -1 on a null time step */
int orc_CflFluidsMin(orc_state *S) {
  double step = 1e30;
  for (int i = 0; i < S->nfluids; i++)
    if (step > S->fl[i].Min) step = S->fl[i].Min;
  S->StepTime = step;
  double SMALLTIME = (1e-10 * sqrt(R0 * R0 * R0 / (G * MSTAR)));
  if (S->StepTime <= SMALLTIME) return -1;
  return 0;
}

/* src/resetfield.c:10 */
void orc_Reset_field(orc_state *S, double *F) { memset(F, 0, sizeof(double) * S->g.ncell); }

/* src/totaldens.c:10-61 */
void orc_ComputeTotalDensity(orc_state *S) {
  ORC_MESH(S);
  const double *dens = S->Density;
  double *totaldens = S->Total_Density;
  int size_y = Ny + 2 * NGHY, size_z = Nz + 2 * NGHZ;
  KRANGE(0, size_z);
  for (int k = kbeg; k < kend; k++)
    for (int j = 0; j < size_y; j++)
      for (int i = 0; i < Nx; i++) {
        int ll = lidx(i, j, k);
        totaldens[ll] += dens[ll];
      }
}

/* src/compute_force.c:16-219 (no HILLCUT / HALFDISK) + src/planets.c:20-84 */
void orc_ComputeForce(orc_state *S, double x, double y, double z, double rsmoothing, double mass) {
  ORC_MESH(S);
  (void)mass;
  const double *dens = S->Total_Density;
  double *fxi = S->Mmx, *fyi = S->Mpx, *fzi = S->Mmy, *fxhi = S->Mpy, *fyhi = S->Vx_temp,
         *fzhi = S->Vy_temp;
  int size_y = Ny + NGHY, size_z = Nz + NGHZ;
  const int kbeg = DIMZ ? NGHZ : 0, kend = DIMZ ? size_z : 1;
  for (int k = kbeg; k < kend; k++)
    for (int j = NGHY; j < size_y; j++)
      for (int i = 0; i < Nx; i++) {
        int ll = lidx(i, j, k);
        double cellmass = Vol(i, j, k) * dens[ll];
        double dx, dy, dz;
        if (!SPH) {
          dx = ymed(j) * cos(xmed(i)) - x;
          dy = ymed(j) * sin(xmed(i)) - y;
          dz = zmed(k) - z;
        } else {
          dx = ymed(j) * cos(xmed(i)) * sin(zmed(k)) - x;
          dy = ymed(j) * sin(xmed(i)) * sin(zmed(k)) - y;
          dz = ymed(j) * cos(zmed(k)) - z;
        }
        double dist2 = dx * dx + dy * dy + dz * dz;
        dist2 += rsmoothing * rsmoothing;
        double distance = sqrt(dist2);
        double InvDist3 = 1.0 / (dist2 * distance);
        InvDist3 *= G * cellmass;
        fxi[ll] = dx * InvDist3;
        fyi[ll] = dy * InvDist3;
        fzi[ll] = dz * InvDist3;
        fxhi[ll] = 0.0;
        fyhi[ll] = 0.0;
        fzhi[ll] = 0.0;
      }
  double aa = sqrt(x * x + y * y + z * z);
  int index = (int)(((aa - Ymin[NGHY]) / (Ymin[Ny + NGHY + 1] - Ymin[NGHY])) * (double)Ny) + NGHY;
  double *lf = S->localforce;
  double *flds[6] = {S->Mmx, S->Mpx, S->Mmy, S->Mpy, S->Vx_temp, S->Vy_temp};
  int zb = NGHZ, ze = Nz + NGHZ;
  if (index >= NGHY) {
    if (index < Ny + NGHY) {
      for (int c = 0; c < 6; c++) lf[c] = orc_reduction_full_SUM(S, flds[c], NGHY, index, zb, ze);
      for (int c = 0; c < 6; c++) lf[6 + c] = orc_reduction_full_SUM(S, flds[c], index, Ny + NGHY, zb, ze);
    } else {
      for (int c = 0; c < 6; c++) lf[c] = orc_reduction_full_SUM(S, flds[c], NGHY, Ny + NGHY, zb, ze);
      for (int c = 0; c < 6; c++) lf[6 + c] = 0.0;
    }
  } else {
    for (int c = 0; c < 6; c++) lf[c] = 0.0;
    for (int c = 0; c < 6; c++) lf[6 + c] = orc_reduction_full_SUM(S, flds[c], NGHY, Ny + NGHY, zb, ze);
  }
}

/* src/potential.c:51-208 (default star, no GASINDIRECTTERM, no binary) */
void orc_Potential(orc_state *S) {
  ORC_MESH(S);
  double planetmass_taper = f3d_planet_taper(&S->p, S->clk.PhysicalTime);
  double *pot = S->Pot;
  const f3d_psys *sys = &S->sys;
  int size_y = Ny + 2 * NGHY, size_z = Nz + 2 * NGHZ;
  int indirect_term = S->p.indirectterm;
  double taper = planetmass_taper;
  KRANGE(0, size_z);
  for (int k = kbeg; k < kend; k++)
    for (int j = 0; j < size_y; j++)
      for (int i = 0; i < Nx; i++) {
        int l = lidx(i, j, k);
        double XC, YC, ZC;
        if (SPH) {
          XC = ymed(j) * cos(xmed(i)) * sin(zmed(k));
          YC = ymed(j) * sin(xmed(i)) * sin(zmed(k));
          ZC = ymed(j) * cos(zmed(k));
          pot[l] = -G * MSTAR / ymed(j);
        } else {
          XC = ymed(j) * cos(xmed(i));
          YC = ymed(j) * sin(xmed(i));
          ZC = zmed(k);
          pot[l] = -G * MSTAR / sqrt(ymed(j) * ymed(j) + ZC * ZC);
        }
        for (int n = 0; n < sys->nb; n++) {
          double mp = sys->mass[n] * taper;
          double planetdistance = sqrt(sys->x[n] * sys->x[n] + sys->y[n] * sys->y[n] + sys->z[n] * sys->z[n]);
          double rroche = planetdistance * pow((1.0 / 3.0 * mp / MSTAR), 1.0 / 3.0);
          double smoothing;
          if (S->p.rochesmoothing != 0)
            smoothing = rroche * S->p.rochesmoothing;
          else
            smoothing = S->p.aspectratio * pow(planetdistance / R0, S->p.flaringindex) * planetdistance *
                        S->p.thicknesssmoothing;
          smoothing *= smoothing;
          double dist = ((XC - sys->x[n]) * (XC - sys->x[n]) + (YC - sys->y[n]) * (YC - sys->y[n]) +
                         (ZC - sys->z[n]) * (ZC - sys->z[n]));
          if (indirect_term == 1) {
            pot[l] += G * mp * (XC * sys->x[n] + YC * sys->y[n] + ZC * sys->z[n]) /
                      (planetdistance * planetdistance * planetdistance);
          }
          pot[l] += -G * mp / sqrt(dist + smoothing);
        }
      }
}

/* src/correct_vtheta.c:10-68 */
void orc_CorrectVtheta(orc_state *S, double domega) {
  ORC_MESH(S);
  double *vt = S->Vx;
  int size_y = Ny + 2 * NGHY, size_z = Nz + 2 * NGHZ;
  KRANGE(0, size_z);
  for (int k = kbeg; k < kend; k++)
    for (int j = 0; j < size_y; j++)
      for (int i = 0; i < Nx; i++) {
        int l = lidx(i, j, k);
        if (SPH)
          vt[l] -= domega * ymed(j) * sin(zmed(k));
        else
          vt[l] -= domega * ymed(j);
      }
}

/* src/substep1_x.c:10-166 (no MHD) */
void orc_SubStep1_x(orc_state *S, double dt) {
  ORC_MESH(S);
  const double *p = S->Pressure, *pot = S->Pot, *rho = S->Density, *vx = S->Vx;
  double *vx_temp = S->Vx_temp;
  int size_y = Ny + 2 * NGHY - 1, size_z = Nz + 2 * NGHZ - 1;
  int fluidtype = S->Fluidtype, potential = S->p.potential;
  KRANGE(1, size_z);
  for (int k = kbeg; k < kend; k++)
    for (int j = 1; j < size_y; j++)
      for (int i = 0; i < Nx; i++) {
        int ll = lidx(i, j, k);
        int llxm = LXM(ll, i);
        double dxmed = 0.5 * (Sxi[i] + Sxi[IXM(i)]);
        double dtOVERrhom = 2. * dt / (rho[ll] * Sxi[i] + rho[llxm] * Sxi[IXM(i)]) * dxmed;
        vx_temp[ll] = vx[ll];
        if (fluidtype != F3D_DUST) vx_temp[ll] -= dtOVERrhom * (p[ll] - p[llxm]) * Inv_zone_size_xmed(i, j, k);
        if (potential) vx_temp[ll] -= (pot[ll] - pot[llxm]) * dt * Inv_zone_size_xmed(i, j, k);
      }
}

/* src/substep1_y.c:10-260 (no MHD, no shearing box) */
void orc_SubStep1_y(orc_state *S, double dt) {
  ORC_MESH(S);
  const double *p = S->Pressure, *pot = S->Pot, *rho = S->Density;
  const double *vx_half = S->Vx, *vy = S->Vy, *vz_half = S->Vz;
  double *vy_temp = S->Vy_temp;
  int size_y = Ny + 2 * NGHY - 1, size_z = Nz + 2 * NGHZ - 1;
  int fluidtype = S->Fluidtype, potential = S->p.potential;
  const double OMEGAFRAME = S->clk.OMEGAFRAME;
  KRANGE(1, size_z);
  for (int k = kbeg; k < kend; k++)
    for (int j = 1; j < size_y; j++)
      for (int i = 0; i < Nx; i++) {
        int ll = lidx(i, j, k);
        int llxp = LXP(ll, i), llym = LYM(ll), llzp = LZP(ll);
        double dtOVERrhom = 2.0 * dt / (rho[ll] * (ymin(j + 1) - ymin(j)) + rho[llym] * (ymin(j) - ymin(j - 1))) *
                            (ymed(j) - ymed(j - 1));
        vy_temp[ll] = vy[ll];
        if (fluidtype != F3D_DUST) vy_temp[ll] -= dtOVERrhom * (p[ll] - p[llym]) / (ymed(j) - ymed(j - 1));
        if (!SPH) {
          double vphi = .25 * (vx_half[ll] + vx_half[llxp] + vx_half[llym] + vx_half[llxp - pitch]);
          vphi += ymin(j) * OMEGAFRAME;
          vy_temp[ll] += vphi * vphi / ymin(j) * dt;
        } else {
          double vphi = .25 * (vx_half[ll] + vx_half[llxp] + vx_half[llym] + vx_half[llxp - pitch]);
          vphi += ymin(j) * sin(zmed(k)) * OMEGAFRAME;
          double vzz = .25 * (vz_half[ll] + vz_half[llzp] + vz_half[llym] + vz_half[llzp - pitch]);
          vy_temp[ll] += (vphi * vphi + vzz * vzz) / ymin(j) * dt;
        }
        if (potential) vy_temp[ll] -= (pot[ll] - pot[llym]) * dt / (ymed(j) - ymed(j - 1));
      }
}

/* src/substep1_z.c:10-223 (CYLINDRICAL and SPHERICAL, no MHD) */
void orc_SubStep1_z(orc_state *S, double dt) {
  ORC_MESH(S);
  if (!DIMZ) return;
  const double *p = S->Pressure, *pot = S->Pot, *rho = S->Density;
  const double *vx_half = S->Vx, *vz = S->Vz;
  double *vz_temp = S->Vz_temp;
  int size_y = Ny + 2 * NGHY - 1, size_z = Nz + 2 * NGHZ - 1;
  int fluidtype = S->Fluidtype, potential = S->p.potential;
  const double OMEGAFRAME = S->clk.OMEGAFRAME;
  for (int k = 1; k < size_z; k++)
    for (int j = 1; j < size_y; j++)
      for (int i = 0; i < Nx; i++) {
        int ll = lidx(i, j, k);
        int llxp = LXP(ll, i), llzm = LZM(ll);
        double dtOVERrhom = 2.0 * dt / (rho[ll] * (zmin(k + 1) - zmin(k)) + rho[llzm] * (zmin(k) - zmin(k - 1))) *
                            (zmed(k) - zmed(k - 1));
        if (!SPH) {
          vz_temp[ll] = vz[ll];
          if (fluidtype != F3D_DUST) vz_temp[ll] -= dtOVERrhom * (p[ll] - p[llzm]) / (zmed(k) - zmed(k - 1));
          if (potential) vz_temp[ll] -= (pot[ll] - pot[llzm]) * dt / (zmed(k) - zmed(k - 1));
        } else {
          double vphi = .25 * (vx_half[ll] + vx_half[llxp] + vx_half[llzm] + vx_half[llxp - stride]);
          vphi += ymed(j) * sin(zmin(k)) * OMEGAFRAME;
          vz_temp[ll] = vz[ll];
          if (fluidtype != F3D_DUST)
            vz_temp[ll] -= dtOVERrhom * (p[ll] - p[llzm]) / (ymed(j) * (zmed(k) - zmed(k - 1)));
          vz_temp[ll] += vphi * vphi * cos(zmin(k)) / (sin(zmin(k)) * ymed(j)) * dt;
          if (potential) vz_temp[ll] -= (pot[ll] - pot[llzm]) * dt / (ymed(j) * (zmed(k) - zmed(k - 1)));
        }
      }
}

/* viscosity coefficients shared by visctensor_cyl/sph (src/visctensor_cyl.c:141-154) */
static void visc_coefs(orc_state *S, int l, int lxm_, int lym_, int lzm_, int j, double *v, double *vm,
                       double *vzm, double *vzmym, int sph_order) {
  ORC_MESH(S);
  const double *energy = S->Energy, *rho = S->Density;
  const double ALPHA = S->p.alpha, GAMMA = S->p.gamma;
  if (S->p.visc != F3D_VISC_ALPHA) {
    *v = *vm = *vzm = *vzmym = S->p.nu;
    return;
  }
  int lxmp = lxm_ - pitch, lyms = lym_ - stride, lxms = lxm_ - stride;
  if (!DIMZ) {
    /* the z-edge coefficients feed only the tau_xz / tau_zy terms, which
       exist only with Z; without Z their stencil would leave the array */
    lzm_ = lyms = lxms = l;
  }
  if (S->p.eos == F3D_ISOTHERMAL) {
    *v = ALPHA * energy[l] * energy[l] * sqrt(ymed(j) * ymed(j) * ymed(j) / (G * MSTAR));
    *vm = ALPHA * 0.0625 * (energy[l] + energy[lxm_] + energy[lym_] + energy[lxmp]) *
          (energy[l] + energy[lxm_] + energy[lym_] + energy[lxmp]) * sqrt(ymin(j) * ymin(j) * ymin(j) / (G * MSTAR));
    if (!sph_order) {
      *vzmym = ALPHA * 0.0625 * (energy[l] + energy[lym_] + energy[lzm_] + energy[lyms]) *
               (energy[l] + energy[lym_] + energy[lzm_] + energy[lyms]) * sqrt(ymin(j) * ymin(j) * ymin(j) / (G * MSTAR));
      *vzm = ALPHA * 0.0625 * (energy[l] + energy[lzm_] + energy[lxm_] + energy[lxms]) *
             (energy[l] + energy[lzm_] + energy[lxm_] + energy[lxms]) * sqrt(ymed(j) * ymed(j) * ymed(j) / (G * MSTAR));
    } else {
      /* src/visctensor_sph.c:131-134 */
      *vzm = ALPHA * 0.0625 * (energy[l] + energy[lzm_] + energy[lxm_] + energy[lxms]) *
             (energy[l] + energy[lzm_] + energy[lxm_] + energy[lxms]) * sqrt(ymed(j) * ymed(j) * ymed(j) / (G * MSTAR));
      *vzmym = ALPHA * 0.0625 * (energy[l] + energy[lzm_] + energy[lym_] + energy[lyms]) *
               (energy[l] + energy[lzm_] + energy[lym_] + energy[lyms]) * sqrt(ymin(j) * ymin(j) * ymin(j) / (G * MSTAR));
    }
  } else {
    *v = ALPHA * GAMMA * (GAMMA - 1.0) * energy[l] / rho[l] * sqrt(ymed(j) * ymed(j) * ymed(j) / (G * MSTAR));
    *vm = ALPHA * GAMMA * (GAMMA - 1.0) * (energy[l] + energy[lxm_] + energy[lym_] + energy[lxmp]) /
          (rho[l] + rho[lxm_] + rho[lym_] + rho[lxmp]) * sqrt(ymin(j) * ymin(j) * ymin(j) / (G * MSTAR));
    if (!sph_order) {
      *vzmym = ALPHA * GAMMA * (GAMMA - 1.0) * (energy[l] + energy[lym_] + energy[lzm_] + energy[lyms]) /
               (rho[l] + rho[lym_] + rho[lzm_] + rho[lyms]) * sqrt(ymin(j) * ymin(j) * ymin(j) / (G * MSTAR));
      *vzm = ALPHA * GAMMA * (GAMMA - 1.0) * (energy[l] + energy[lzm_] + energy[lxm_] + energy[lxms]) /
             (rho[l] + rho[lzm_] + rho[lxm_] + rho[lxms]) * sqrt(ymed(j) * ymed(j) * ymed(j) / (G * MSTAR));
    } else {
      *vzm = ALPHA * GAMMA * (GAMMA - 1.0) * (energy[l] + energy[lzm_] + energy[lxm_] + energy[lxms]) /
             (rho[l] + rho[lzm_] + rho[lxm_] + rho[lxms]) * sqrt(ymed(j) * ymed(j) * ymed(j) / (G * MSTAR));
      *vzmym = ALPHA * GAMMA * (GAMMA - 1.0) * (energy[l] + energy[lym_] + energy[lzm_] + energy[lyms]) /
               (rho[l] + rho[lym_] + rho[lzm_] + rho[lyms]) * sqrt(ymin(j) * ymin(j) * ymin(j) / (G * MSTAR));
    }
  }
}

/* src/visctensor_cyl.c:10-203 */
void orc_visctensor_cyl(orc_state *S) {
  ORC_MESH(S);
  const double *rho = S->Density, *vx = S->Vx, *vy = S->Vy, *vz = S->Vz;
  double *tauxx = S->Mmx, *tauyy = S->Mmy, *tauzz = S->Mmz, *tauxz = S->Mpx, *tauyx = S->Mpy, *tauzy = S->Mpz;
  int size_y = Ny + 2 * NGHY - 1, size_z = Nz + 2 * NGHZ - 1;
  KRANGE(1, size_z);
  for (int k = kbeg; k < kend; k++)
    for (int j = 1; j < size_y; j++)
      for (int i = 0; i < Nx; i++) {
        int l = lidx(i, j, k);
        int lxp = LXP(l, i), lxm = LXM(l, i), lyp = LYP(l), lym = LYM(l), lzp = LZP(l), lzm = LZM(l);
        double viscosity, viscositym, viscosityzm, viscosityzmym;
        visc_coefs(S, l, lxm, lym, lzm, j, &viscosity, &viscositym, &viscosityzm, &viscosityzmym, 0);
        double div_v = 0.0;
        div_v += (vx[lxp] - vx[l]) * SurfX(j, k);
        div_v += (vy[lyp] * SurfY(i, j + 1, k) - vy[l] * SurfY(i, j, k));
        if (DIMZ) div_v += (vz[lzp] * SurfZ(i, j, k + 1) - vz[l] * SurfZ(i, j, k));
        div_v *= 2.0 / 3.0 * InvVol(i, j, k);
        tauxx[l] = viscosity * rho[l] * (2.0 * (vx[lxp] - vx[l]) / zone_size_x(i, j, k) - div_v);
        tauxx[l] += viscosity * rho[l] * (vy[lyp] + vy[l]) / ymed(j);
        tauyy[l] = viscosity * rho[l] * (2.0 * (vy[lyp] - vy[l]) / (ymin(j + 1) - ymin(j)) - div_v);
        if (DIMZ) {
          tauzz[l] = viscosity * rho[l] * (2.0 * (vz[lzp] - vz[l]) / (zmin(k + 1) - zmin(k)) - div_v);
          tauxz[l] = viscosityzm * .25 * (rho[l] + rho[lzm] + rho[lxm] + rho[lxm - stride]) *
                     ((vx[l] - vx[lzm]) / (zmed(k) - zmed(k - 1)) + (vz[l] - vz[lxm]) * InvDiffXmed[i] / ymed(j));
        }
        tauyx[l] = viscositym * .25 * (rho[l] + rho[lxm] + rho[lym] + rho[lxm - pitch]) *
                   ((vy[l] - vy[lxm]) * InvDiffXmed[i] / ymin(j) + (vx[l] - vx[lym]) / (ymed(j) - ymed(j - 1)) -
                    .5 * (vx[l] + vx[lym]) / ymin(j));
        if (DIMZ)
          tauzy[l] = viscosityzmym * .25 * (rho[l] + rho[lym] + rho[lzm] + rho[lym - stride]) *
                     ((vz[l] - vz[lym]) / (ymed(j) - ymed(j - 1)) + (vy[l] - vy[lzm]) / (zmed(k) - zmed(k - 1)));
      }
}

/* src/addviscosity_cyl.c:10-142 */
void orc_addviscosity_cyl(orc_state *S, double dt) {
  ORC_MESH(S);
  const double *rho = S->Density;
  double *vx = S->Vx_temp, *vy = S->Vy_temp, *vz = S->Vz_temp;
  const double *tauxx = S->Mmx, *tauyy = S->Mmy, *tauzz = S->Mmz, *tauxz = S->Mpx, *tauyx = S->Mpy, *tauzy = S->Mpz;
  int size_y = Ny + 2 * NGHY - 2, size_z = Nz + 2 * NGHZ - 2;
  KRANGE(1, size_z);
  for (int k = kbeg; k < kend; k++)
    for (int j = 1; j < size_y; j++)
      for (int i = 0; i < Nx; i++) {
        int l = lidx(i, j, k);
        int lxp = LXP(l, i), lxm = LXM(l, i), lyp = LYP(l), lym = LYM(l), lzp = LZP(l), lzm = LZM(l);
        vx[l] += 2.0 * (tauxx[l] - tauxx[lxm]) * Inv_zone_size_xmed(i, j, k) / ((rho[l] + rho[lxm])) * dt;
        vx[l] += 2.0 * (ymin(j + 1) * ymin(j + 1) * tauyx[lyp] - ymin(j) * ymin(j) * tauyx[l]) /
                 ((ymin(j + 1) - ymin(j)) * ymed(j) * ymed(j) * (rho[lxm] + rho[l])) * dt;
        if (DIMZ)
          vx[l] += 2.0 * (tauxz[lzp] - tauxz[l]) / ((zmin(k + 1) - zmin(k)) * (rho[lxm] + rho[l])) * dt;
        vy[l] += 2.0 * (ymed(j) * tauyy[l] - ymed(j - 1) * tauyy[lym]) /
                 ((ymed(j) - ymed(j - 1)) * (rho[l] + rho[lym]) * ymin(j)) * dt;
        vy[l] += 2.0 * (tauyx[lxp] - tauyx[l]) / ((xmin(i + 1) - xmin(i)) * ymin(j) * (rho[l] + rho[lym])) * dt;
        vy[l] -= (tauxx[l] + tauxx[lym]) / (ymin(j) * (rho[l] + rho[lym])) * dt;
        if (DIMZ) {
          vy[l] += 2.0 * (tauzy[lzp] - tauzy[l]) / ((zmin(k + 1) - zmin(k)) * (rho[l] + rho[lym])) * dt;
          vz[l] += 2.0 * (tauzz[l] - tauzz[lzm]) / ((zmed(k) - zmed(k - 1)) * (rho[l] + rho[lzm])) * dt;
          vz[l] += 2.0 * (tauxz[lxp] - tauxz[l]) / (zone_size_x(i, j, k) * (rho[l] + rho[lzm])) * dt;
          vz[l] += 2.0 * (ymin(j + 1) * tauzy[lyp] - ymin(j) * tauzy[l]) /
                   ((ymin(j + 1) - ymin(j)) * ymed(j) * (rho[l] + rho[lzm])) * dt;
        }
      }
}

/* src/visctensor_sph.c:10-211 */
void orc_visctensor_sph(orc_state *S) {
  ORC_MESH(S);
  const double *rho = S->Density, *vx = S->Vx, *vy = S->Vy, *vz = S->Vz;
  double *tauxx = S->Mmx, *tauyy = S->Mmy, *tauzz = S->Mmz, *tauxz = S->Mpx, *tauyx = S->Mpy, *tauzy = S->Mpz;
  int size_y = Ny + 2 * NGHY - 1, size_z = Nz + 2 * NGHZ - 1;
  KRANGE(1, size_z);
  for (int k = kbeg; k < kend; k++)
    for (int j = 1; j < size_y; j++)
      for (int i = 0; i < Nx; i++) {
        int l = lidx(i, j, k);
        int lxp = LXP(l, i), lxm = LXM(l, i), lyp = LYP(l), lym = LYM(l), lzp = LZP(l), lzm = LZM(l);
        double viscosity, viscositym, viscosityzm, viscosityzmym;
        visc_coefs(S, l, lxm, lym, lzm, j, &viscosity, &viscositym, &viscosityzm, &viscosityzmym, 1);
        double div_v = 0.0;
        div_v += (vx[lxp] - vx[l]) * SurfX(j, k);
        div_v += (vy[lyp] * SurfY(i, j + 1, k) - vy[l] * SurfY(i, j, k));
        if (DIMZ) div_v += (vz[lzp] * SurfZ(i, j, k + 1) - vz[l] * SurfZ(i, j, k));
        div_v *= 2.0 / 3.0 * InvVol(i, j, k);
        tauxx[l] = viscosity * rho[l] * (2.0 * (vx[lxp] - vx[l]) / zone_size_x(i, j, k) - div_v);
        tauxx[l] += viscosity * rho[l] * (vy[lyp] + vy[l]) / ymed(j);
        if (DIMZ) tauxx[l] += viscosity * rho[l] * (vz[lzp] + vz[l]) * cos(zmed(k)) / (sin(zmed(k)) * ymed(j));
        tauyy[l] = viscosity * rho[l] * (2.0 * (vy[lyp] - vy[l]) / (ymin(j + 1) - ymin(j)) - div_v);
        if (DIMZ) {
          tauzz[l] = viscosity * rho[l] * (2.0 * (vz[lzp] - vz[l]) / (ymed(j) * (zmin(k + 1) - zmin(k))) - div_v);
          tauzz[l] += viscosity * rho[l] * (vy[l] + vy[lyp]) / ymed(j);
          tauxz[l] = viscosityzm * .25 * (rho[l] + rho[lzm] + rho[lxm] + rho[lxm - stride]) *
                     ((vx[l] / sin(zmed(k)) - vx[lzm] / sin(zmed(k - 1))) * sin(zmin(k)) / (ymed(j) * (zmed(k) - zmed(k - 1))) +
                      ((vz[l] - vz[lxm]) * InvDiffXmed[i] / (sin(zmin(k)) * ymed(j))));
        }
        tauyx[l] = viscositym * .25 * (rho[l] + rho[lxm] + rho[lym] + rho[lxm - pitch]) *
                   ((vy[l] - vy[lxm]) * InvDiffXmed[i] / (ymin(j) * sin(zmed(k))) +
                    (vx[l] - vx[lym]) / (ymed(j) - ymed(j - 1)) - .5 * (vx[l] + vx[lym]) / ymin(j));
        if (DIMZ)
          tauzy[l] = viscosityzmym * .25 * (rho[l] + rho[lym] + rho[lzm] + rho[lym - stride]) *
                     ((vz[l] - vz[lym]) / (ymed(j) - ymed(j - 1)) - .5 * (vz[l] + vz[lym]) / ymin(j) +
                      (vy[l] - vy[lzm]) / (ymin(j) * (zmed(k) - zmed(k - 1))));
      }
}

/* src/addviscosity_sph.c:10-150 */
void orc_addviscosity_sph(orc_state *S, double dt) {
  ORC_MESH(S);
  const double *rho = S->Density;
  double *vx = S->Vx_temp, *vy = S->Vy_temp, *vz = S->Vz_temp;
  const double *tauxx = S->Mmx, *tauyy = S->Mmy, *tauzz = S->Mmz, *tauxz = S->Mpx, *tauyx = S->Mpy, *tauzy = S->Mpz;
  int size_y = Ny + 2 * NGHY - 2, size_z = Nz + 2 * NGHZ - 2;
  KRANGE(1, size_z);
  for (int k = kbeg; k < kend; k++)
    for (int j = 1; j < size_y; j++)
      for (int i = 0; i < Nx; i++) {
        int l = lidx(i, j, k);
        int lxp = LXP(l, i), lxm = LXM(l, i), lyp = LYP(l), lym = LYM(l), lzp = LZP(l), lzm = LZM(l);
        vx[l] += 2.0 * (tauxx[l] - tauxx[lxm]) * Inv_zone_size_xmed(i, j, k) / ((rho[l] + rho[lxm])) * dt;
        vx[l] += 2.0 * (ymin(j + 1) * ymin(j + 1) * tauyx[lyp] - ymin(j) * ymin(j) * tauyx[l]) /
                 ((ymin(j + 1) - ymin(j)) * ymed(j) * ymed(j) * (rho[lxm] + rho[l])) * dt;
        vx[l] += (tauyx[lyp] + tauyx[l]) / ((rho[lxm] + rho[l]) * ymed(j)) * dt;
        if (DIMZ) {
          vx[l] += 2.0 * (tauxz[lzp] - tauxz[l]) / (zone_size_z(j, k) * (rho[lxm] + rho[l])) * dt;
          vx[l] += 2.0 * (tauxz[lzp] + tauxz[l]) * cos(zmed(k)) / (sin(zmed(k)) * ymed(j) * (rho[lxm] + rho[l])) * dt;
        }
        vy[l] += 2.0 * (ymed(j) * ymed(j) * tauyy[l] - ymed(j - 1) * ymed(j - 1) * tauyy[lym]) /
                 ((ymed(j) - ymed(j - 1)) * (rho[l] + rho[lym]) * ymin(j) * ymin(j)) * dt;
        vy[l] += 2.0 * (tauyx[lxp] - tauyx[l]) / ((xmin(i + 1) - xmin(i)) * ymin(j) * sin(zmed(k)) * (rho[l] + rho[lym])) * dt;
        vy[l] -= (tauxx[l] + tauxx[lym]) / (ymin(j) * (rho[l] + rho[lym])) * dt;
        if (DIMZ) {
          vy[l] -= (tauzz[l] + tauzz[lym]) / (ymin(j) * (rho[l] + rho[lym])) * dt;
          vy[l] += 2.0 * (tauzy[lzp] * sin(zmin(k + 1)) - tauzy[l] * sin(zmin(k))) /
                   ((zmin(k + 1) - zmin(k)) * ymin(j) * sin(zmed(k)) * (rho[l] + rho[lym])) * dt;
          vz[l] += 2.0 * (tauzz[l] * sin(zmed(k)) - tauzz[lzm] * sin(zmed(k - 1))) /
                   ((zmed(k) - zmed(k - 1)) * ymed(j) * sin(zmin(k)) * (rho[l] + rho[lzm])) * dt;
          vz[l] += 2.0 * (tauxz[lxp] - tauxz[l]) / (edge_size_x_middley_lowz(i, j, k) * (rho[l] + rho[lzm])) * dt;
          vz[l] += 2.0 * (ymin(j + 1) * ymin(j + 1) * ymin(j + 1) * tauzy[lyp] - ymin(j) * ymin(j) * ymin(j) * tauzy[l]) /
                   ((ymin(j + 1) - ymin(j)) * ymed(j) * ymed(j) * ymed(j) * (rho[l] + rho[lzm])) * dt;
          vz[l] -= (tauxx[l] + tauxx[lzm]) * cos(zmin(k)) / (ymed(j) * sin(zmin(k)) * (rho[l] + rho[lzm])) * dt;
        }
      }
}

/* src/substep2_a.c:10-171 (no STRONG_SHOCK) */
void orc_SubStep2_a(orc_state *S, double dt) {
  ORC_MESH(S);
  (void)dt;
  const double *rho = S->Density, *vx = S->Vx, *vy = S->Vy, *vz = S->Vz;
  double *pres_x = S->Mpx, *pres_y = S->Mpy, *pres_z = S->Mpz;
  int size_y = Ny + 2 * NGHY - 1, size_z = Nz + 2 * NGHZ - 1;
  KRANGE(0, size_z);
  for (int k = kbeg; k < kend; k++)
    for (int j = 0; j < size_y; j++)
      for (int i = 0; i < Nx; i++) {
        int ll = lidx(i, j, k);
        int llxp = LXP(ll, i), llyp = LYP(ll), llzp = LZP(ll);
        double dvx = vx[llxp] - vx[ll];
        if (dvx < 0.0) pres_x[ll] = CVNR * CVNR * rho[ll] * dvx * dvx;
        else pres_x[ll] = 0.0;
        double dvy = vy[llyp] - vy[ll];
        if (dvy < 0.0) pres_y[ll] = CVNR * CVNR * rho[ll] * dvy * dvy;
        else pres_y[ll] = 0.0;
        if (DIMZ) {
          double dvz = vz[llzp] - vz[ll];
          if (dvz < 0.0) pres_z[ll] = CVNR * CVNR * rho[ll] * dvz * dvz;
          else pres_z[ll] = 0.0;
        }
      }
}

/* src/substep2_b.c:10-210 */
void orc_SubStep2_b(orc_state *S, double dt) {
  ORC_MESH(S);
  const double *rho = S->Density, *vx = S->Vx, *vy = S->Vy, *vz = S->Vz;
  double *vx_temp = S->Vx_temp, *vy_temp = S->Vy_temp, *vz_temp = S->Vz_temp;
  const double *pres_x = S->Mpx, *pres_y = S->Mpy, *pres_z = S->Mpz;
  double *e = S->Energy;
  const int adiabatic = (S->p.eos == F3D_ADIABATIC);
  int size_y = Ny + 2 * NGHY - 1, size_z = Nz + 2 * NGHZ - 1;
  KRANGE(1, size_z);
  for (int k = kbeg; k < kend; k++)
    for (int j = 1; j < size_y; j++)
      for (int i = 0; i < Nx; i++) {
        int ll = lidx(i, j, k);
        int llxm = LXM(ll, i), llxp = LXP(ll, i), llym = LYM(ll), llyp = LYP(ll), llzm = LZM(ll), llzp = LZP(ll);
        double dxrho1 = Sxi[i];
        double dxrho2 = Sxi[IXM(i)];
        double dxmed = 0.5 * (Sxi[i] + Sxi[IXM(i)]);
        vx_temp[ll] += -2.0 * dt * dxmed * (pres_x[ll] - pres_x[llxm]) / (rho[ll] * dxrho1 + rho[llxm] * dxrho2) *
                       Inv_zone_size_xmed(i, j, k);
        if (adiabatic) e[ll] += -dt * (pres_x[ll] * (vx[llxp] - vx[ll]) / zone_size_x(i, j, k));
        double dyrho1 = ymin(j + 1) - ymin(j);
        double dyrho2 = ymin(j) - ymin(j - 1);
        vy_temp[ll] += -2.0 * (pres_y[ll] - pres_y[llym]) / (rho[ll] * dyrho1 + rho[llym] * dyrho2) * dt;
        if (adiabatic) e[ll] += -dt * (pres_y[ll] * (vy[llyp] - vy[ll]) / zone_size_y(j, k));
        if (DIMZ) {
          double dzmed = zmed(k) - zmed(k - 1);
          double dzrho1 = zmin(k + 1) - zmin(k);
          double dzrho2 = zmin(k) - zmin(k - 1);
          if (SPH)
            vz_temp[ll] += -2.0 * dzmed * (pres_z[ll] - pres_z[llzm]) / (rho[ll] * dzrho1 + rho[llzm] * dzrho2) * dt /
                           (ymed(j) * (zmed(k) - zmed(k - 1)));
          else
            vz_temp[ll] += -2.0 * (pres_z[ll] - pres_z[llzm]) / (rho[ll] * dzrho1 + rho[llzm] * dzrho2) * dt;
          if (adiabatic) e[ll] += -dt * (pres_z[ll] * (vz[llzp] - vz[ll]) / zone_size_z(j, k));
        }
      }
}

/* src/substep3.c:10-122 */
void orc_SubStep3(orc_state *S, double dt) {
  ORC_MESH(S);
  double *e = S->Energy;
  const double *vx = S->Vx_temp, *vy = S->Vy_temp, *vz = S->Vz_temp;
  const double GAMMA = S->p.gamma;
  int size_y = Ny + 2 * NGHY - 1, size_z = Nz + 2 * NGHZ - 1;
  KRANGE(0, size_z);
  for (int k = kbeg; k < kend; k++)
    for (int j = 0; j < size_y; j++)
      for (int i = 0; i < Nx; i++) {
        int ll = lidx(i, j, k);
        int llxp = LXP(ll, i), llyp = LYP(ll), llzp = LZP(ll);
        double div_v = 0.0;
        div_v += (vx[llxp] - vx[ll]) * SurfX(j, k);
        div_v += (vy[llyp] * SurfY(i, j + 1, k) - vy[ll] * SurfY(i, j, k));
        if (DIMZ) div_v += (vz[llzp] * SurfZ(i, j, k + 1) - vz[ll] * SurfZ(i, j, k));
        double term = 0.5 * dt * (GAMMA - 1.) * div_v * InvVol(i, j, k);
        e[ll] *= (1.0 - term) / (1.0 + term);
      }
}

/* src/cpvel.c:3-49 */
void orc_copy_velocities(orc_state *S, int option) {
  size_t n = sizeof(double) * S->g.ncell;
  if (option == 1) { /* VTEMP2V */
    memcpy(S->Vx, S->Vx_temp, n);
    memcpy(S->Vy, S->Vy_temp, n);
    if (S->g.dimz) memcpy(S->Vz, S->Vz_temp, n);
  } else {
    memcpy(S->Vx_temp, S->Vx, n);
    memcpy(S->Vy_temp, S->Vy, n);
    if (S->g.dimz) memcpy(S->Vz_temp, S->Vz, n);
  }
}

/* generated boundary kernels: std/boundary_template.c + scripts/boundparser.py
 * :237-363 with the formulas of std/boundaries.txt:35-102 and the field lists
 * of setups/fargo/fargo.bound.0 and setups/p3disof/p3disof.bound.0.
 * Order as boundaries() (src/boundary.c:2-23): zmin, zmax, ymin, ymax. */
static void bound_z(orc_state *S, int side) {
  ORC_MESH(S);
  double *dens = S->Density, *vx = S->Vx, *vy = S->Vy, *vz = S->Vz;
  const double OF = S->clk.OMEGAFRAME;
  const int sy = Ny + 2 * NGHY;
  for (int m = 0; m < NGHZ; m++)
    for (int j = 0; j < sy; j++)
      for (int i = 0; i < Nx; i++) {
        int lgh, lghs, lact, lacts, lnull, kgh, kact;
        if (!side) {
          lgh = lghs = lidx(i, j, m);
          lact = i + j * pitch + (2 * NGHZ - m - 1) * stride;
          lacts = i + j * pitch + (2 * NGHZ - m) * stride;
          lnull = i + j * pitch + NGHZ * stride;
          kgh = m;
          kact = 2 * NGHZ - m - 1;
        } else {
          lgh = i + j * pitch + (Nz + NGHZ + m) * stride;
          lghs = i + j * pitch + (Nz + NGHZ + 1 + m) * stride;
          lact = lacts = i + j * pitch + (Nz + NGHZ - 1 - m) * stride;
          lnull = i + j * pitch + (Nz + NGHZ) * stride;
          kgh = Nz + NGHZ + m;
          kact = Nz + NGHZ - 1 - m;
        }
        const double *zd = side ? S->bct.zmax_dens : S->bct.zmin_dens;
        double vf = side ? S->bct.zmax_vfac[m] : S->bct.zmin_vfac[m];
        dens[lgh] = dens[lact] * zd[m * sy + j];
        vx[lgh] = (vx[lact] + OF * Ymed[j] * sin(Zmed[kact])) * vf - Ymed[j] * OF * sin(Zmed[kgh]);
        vy[lgh] = vy[lact];
        if (!side || m < NGHZ - 1) vz[lghs] = -vz[lacts];
        vz[lnull] = 0;
        if (S->p.energy_bc) S->Energy[lgh] = S->Energy[lact];
      }
}

static void bound_y(orc_state *S, int side) {
  ORC_MESH(S);
  double *dens = S->Density, *vx = S->Vx, *vy = S->Vy, *vz = S->Vz;
  const double OF = S->clk.OMEGAFRAME;
  const int p3d = (S->p.bc == F3D_BC_P3D);
  const int sz = Nz + 2 * NGHZ;
  for (int k = 0; k < (DIMZ ? sz : 1); k++)
    for (int m = 0; m < NGHY; m++)
      for (int i = 0; i < Nx; i++) {
        int lgh, lghs, lact, lacts, lnull, jgh, jact;
        if (!side) {
          lgh = lghs = lidx(i, m, k);
          lact = i + (2 * NGHY - m - 1) * pitch + k * stride;
          lacts = i + (2 * NGHY - m) * pitch + k * stride;
          lnull = i + NGHY * pitch + k * stride;
          jgh = m;
          jact = 2 * NGHY - m - 1;
        } else {
          lgh = i + (Ny + NGHY + m) * pitch + k * stride;
          lghs = i + (Ny + NGHY + 1 + m) * pitch + k * stride;
          lact = lacts = i + (Ny + NGHY - 1 - m) * pitch + k * stride;
          lnull = i + (Ny + NGHY) * pitch + k * stride;
          jgh = Ny + NGHY + m;
          jact = Ny + NGHY - 1 - m;
        }
        const double dfac = side ? S->bct.ymax_dens[m] : S->bct.ymin_dens[m];
        const double vsq = side ? S->bct.ymax_vsq[m] : S->bct.ymin_vsq[m];
        if (S->fill_density_only) {
          /* FillGhosts(DENS) (src/dust_diffusion_main.c:25) */
          dens[lgh] = p3d ? dens[lact] * dfac * (side ? S->bct.ymax_dexp : S->bct.ymin_dexp)[m * sz + k]
                          : dens[lact] * dfac;
          continue;
        }
        if (!p3d) {
          /* KEPLERIAN2DDENS, KEPLERIAN2DVAZIM, ANTISYMMETRIC */
          dens[lgh] = dens[lact] * dfac;
          vx[lgh] = (vx[lact] + Ymed[jact] * OF) * vsq - Ymed[jgh] * OF;
        } else {
          /* KEPLERIAN3DSPHDENSRAD, KEPLERIAN3DSPHVAZIMRAD, ANTISYMMETRIC, SYMMETRIC */
          const double *dexp = side ? S->bct.ymax_dexp : S->bct.ymin_dexp;
          const double vfac = side ? S->bct.ymax_vfac[m] : S->bct.ymin_vfac[m];
          dens[lgh] = dens[lact] * dfac * dexp[m * sz + k];
          vx[lgh] = (vx[lact] + OF * Ymed[jact] * sin(Zmed[k])) * vsq * vfac - Ymed[jgh] * OF * sin(Zmed[k]);
          vz[lgh] = vz[lact];
        }
        if (!side || m < NGHY - 1) vy[lghs] = -vy[lacts];
        vy[lnull] = 0;
        /* SYMMETRIC, centered |a|a| (std/boundaries.txt:55-57) */
        if (S->p.energy_bc) S->Energy[lgh] = S->Energy[lact];
      }
}

/* src/boundary.c:2-23 */
void orc_boundaries(orc_state *S) {
  if (!S->p.periodicz && S->g.dimz) {
    bound_z(S, 0);
    bound_z(S, 1);
  }
  if (!S->p.periodicy) {
    bound_y(S, 0);
    bound_y(S, 1);
  }
}

/* src/algogas.c:11-27 (sequential: comm is a no-op) */
void orc_FillGhosts(orc_state *S) { orc_boundaries(S); }

/* src/momenta_x.c:10-98 */
void orc_momenta_x(orc_state *S) {
  ORC_MESH(S);
  const double *rho = S->Density, *vx = S->Vx_temp;
  double *mmx = S->Mmx, *mpx = S->Mpx;
  const double OMEGAFRAME = S->clk.OMEGAFRAME;
  int size_y = Ny + 2 * NGHY, size_z = Nz + 2 * NGHZ;
  KRANGE(0, size_z);
  for (int k = kbeg; k < kend; k++)
    for (int j = 0; j < size_y; j++)
      for (int i = 0; i < Nx; i++) {
        int ll = lidx(i, j, k), llxp = LXP(ll, i);
        if (!SPH) {
          mmx[ll] = (vx[ll] + ymed(j) * OMEGAFRAME) * ymed(j) * rho[ll];
          mpx[ll] = (vx[llxp] + ymed(j) * OMEGAFRAME) * ymed(j) * rho[ll];
        } else {
          double rcyl = ymed(j) * sin(zmed(k));
          mmx[ll] = (vx[ll] + rcyl * OMEGAFRAME) * rcyl * rho[ll];
          mpx[ll] = (vx[llxp] + rcyl * OMEGAFRAME) * rcyl * rho[ll];
        }
      }
}

/* src/momenta_y.c:10-74 */
void orc_momenta_y(orc_state *S) {
  ORC_MESH(S);
  const double *rho = S->Density, *vy = S->Vy_temp;
  double *mmy = S->Mmy, *mpy = S->Mpy;
  int size_y = Ny + 2 * NGHY - 1, size_z = Nz + 2 * NGHZ;
  KRANGE(0, size_z);
  for (int k = kbeg; k < kend; k++)
    for (int j = 0; j < size_y; j++)
      for (int i = 0; i < Nx; i++) {
        int ll = lidx(i, j, k), llyp = LYP(ll);
        mmy[ll] = vy[ll] * rho[ll];
        mpy[ll] = vy[llyp] * rho[ll];
      }
}

/* src/momenta_z.c:10-84 */
void orc_momenta_z(orc_state *S) {
  ORC_MESH(S);
  if (!DIMZ) return;
  const double *rho = S->Density, *vz = S->Vz_temp;
  double *mmz = S->Mmz, *mpz = S->Mpz;
  int size_y = Ny + 2 * NGHY, size_z = Nz + 2 * NGHZ - 1;
  for (int k = 0; k < size_z; k++)
    for (int j = 0; j < size_y; j++)
      for (int i = 0; i < Nx; i++) {
        int ll = lidx(i, j, k), llzp = LZP(ll);
        mpz[ll] = vz[llzp] * rho[ll];
        mmz[ll] = vz[ll] * rho[ll];
        if (SPH) {
          mpz[ll] *= ymed(j);
          mmz[ll] *= ymed(j);
        }
      }
}

/* src/divrho.c:10-64 */
void orc_DivideByRho(orc_state *S, const double *q) {
  ORC_MESH(S);
  const double *rho = S->Density;
  double *divrho = S->DivRho;
  int size_y = Ny + 2 * NGHY, size_z = Nz + 2 * NGHZ;
  KRANGE(0, size_z);
  for (int k = kbeg; k < kend; k++)
    for (int j = 0; j < size_y; j++)
      for (int i = 0; i < Nx; i++) {
        int ll = lidx(i, j, k);
        divrho[ll] = q[ll] / rho[ll];
      }
}

/* src/vanleer_y_a.c:10-80 */
void orc_VanLeerY_a(orc_state *S, const double *q) {
  ORC_MESH(S);
  double *slope = S->Slope;
  int size_y = Ny + 2 * NGHY - 1, size_z = Nz + 2 * NGHZ;
  KRANGE(0, size_z);
  for (int k = kbeg; k < kend; k++)
    for (int j = 1; j < size_y; j++)
      for (int i = 0; i < Nx; i++) {
        int ll = lidx(i, j, k), llym = LYM(ll), llyp = LYP(ll);
        double dqm = (q[ll] - q[llym]) / zone_size_y(j, k);
        double dqp = (q[llyp] - q[ll]) / zone_size_y(j + 1, k);
        if (dqp * dqm <= 0) slope[ll] = 0;
        else slope[ll] = 2. * dqp * dqm / (dqm + dqp);
      }
}

/* src/vanleer_y_b.c:10-77 */
void orc_VanLeerY_b(orc_state *S, double dt, const double *q, double *qs) {
  ORC_MESH(S);
  const double *slope = S->Slope, *vy = S->Vy_temp;
  int size_y = Ny + 2 * NGHY - 1, size_z = Nz + 2 * NGHZ;
  KRANGE(0, size_z);
  for (int k = kbeg; k < kend; k++)
    for (int j = 1; j < size_y; j++)
      for (int i = 0; i < Nx; i++) {
        int ll = lidx(i, j, k), llym = LYM(ll);
        if (vy[ll] > 0.)
          qs[ll] = q[llym] + 0.5 * (zone_size_y(j - 1, k) - vy[ll] * dt) * slope[llym];
        else
          qs[ll] = q[ll] - 0.5 * (zone_size_y(j, k) + vy[ll] * dt) * slope[ll];
      }
}

/* src/vanleer_z_a.c:10-81 */
void orc_VanLeerZ_a(orc_state *S, const double *q) {
  ORC_MESH(S);
  if (!DIMZ) return;
  double *slope = S->Slope;
  int size_y = Ny + 2 * NGHY, size_z = Nz + 2 * NGHZ - 1;
  for (int k = 1; k < size_z; k++)
    for (int j = 0; j < size_y; j++)
      for (int i = 0; i < Nx; i++) {
        int ll = lidx(i, j, k), llzm = LZM(ll), llzp = LZP(ll);
        double dqm = (q[ll] - q[llzm]) / zone_size_z(j, k);
        double dqp = (q[llzp] - q[ll]) / zone_size_z(j, k + 1);
        if (dqp * dqm <= 0) slope[ll] = 0;
        else slope[ll] = 2. * dqp * dqm / (dqm + dqp);
      }
}

/* src/vanleer_z_b.c:10-78 */
void orc_VanLeerZ_b(orc_state *S, double dt, const double *q, double *qs) {
  ORC_MESH(S);
  if (!DIMZ) return;
  const double *slope = S->Slope, *vz = S->Vz_temp;
  int size_y = Ny + 2 * NGHY, size_z = Nz + 2 * NGHZ - 1;
  for (int k = 1; k < size_z; k++)
    for (int j = 0; j < size_y; j++)
      for (int i = 0; i < Nx; i++) {
        int ll = lidx(i, j, k), llzm = LZM(ll);
        if (vz[ll] > 0.)
          qs[ll] = q[llzm] + 0.5 * (zone_size_z(j, k - 1) - vz[ll] * dt) * slope[llzm];
        else
          qs[ll] = q[ll] - 0.5 * (zone_size_z(j, k) + vz[ll] * dt) * slope[ll];
      }
}

/* src/vanleer_x_a.c:10-83 */
void orc_VanLeerX_a(orc_state *S, const double *q) {
  ORC_MESH(S);
  double *slope = S->Slope;
  int size_y = Ny + 2 * NGHY, size_z = Nz + 2 * NGHZ;
  KRANGE(0, size_z);
  for (int k = kbeg; k < kend; k++)
    for (int j = 0; j < size_y; j++)
      for (int i = 0; i < Nx; i++) {
        int ll = lidx(i, j, k), llxm = LXM(ll, i), llxp = LXP(ll, i);
        double dqm = (q[ll] - q[llxm]) / zone_size_x(i, j, k);
        double dqp = (q[llxp] - q[ll]) / zone_size_x(IXP(i), j, k);
        if (dqp * dqm <= 0.0) slope[ll] = 0.0;
        else slope[ll] = (2. * dqp * dqm) / (dqm + dqp);
      }
}

/* src/vanleer_x_b.c:10-77 */
void orc_VanLeerX_b(orc_state *S, double dt, const double *q, double *qs, const double *vx) {
  ORC_MESH(S);
  const double *slope = S->Slope;
  int size_y = Ny + 2 * NGHY, size_z = Nz + 2 * NGHZ;
  KRANGE(0, size_z);
  for (int k = kbeg; k < kend; k++)
    for (int j = 0; j < size_y; j++)
      for (int i = 0; i < Nx; i++) {
        int ll = lidx(i, j, k), llxm = LXM(ll, i);
        if (vx[ll] > 0.0)
          qs[ll] = q[llxm] + 0.5 * (zone_size_x(IXM(i), j, k) - vx[ll] * dt) * slope[llxm];
        else
          qs[ll] = q[ll] - 0.5 * (zone_size_x(i, j, k) + vx[ll] * dt) * slope[ll];
      }
}

/* src/fargo_ppa_a.c:10-77 (no PPA_STEEPENER) */
void orc_VanLeerX_PPA_a(orc_state *S, const double *q) {
  ORC_MESH(S);
  double *slope = S->Slope;
  int size_y = Ny + 2 * NGHY, size_z = Nz + 2 * NGHZ;
  KRANGE(0, size_z);
  for (int k = kbeg; k < kend; k++)
    for (int j = 0; j < size_y; j++)
      for (int i = 0; i < Nx; i++) {
        int l = lidx(i, j, k), lxm = LXM(l, i), lxp = LXP(l, i);
        double dqm = (q[l] - q[lxm]);
        double dqp = (q[lxp] - q[l]);
        if (dqp * dqm <= 0.0) slope[l] = 0.0;
        else {
          slope[l] = 0.5 * (q[lxp] - q[lxm]);
          double work = fabs(slope[l]);
          if (2.0 * fabs(dqm) < work) work = 2.0 * fabs(dqm);
          if (2.0 * fabs(dqp) < work) work = 2.0 * fabs(dqp);
          if (slope[l] < 0) slope[l] = -work;
          else slope[l] = work;
        }
      }
}

/* src/fargo_ppa_b.c:10-67 */
void orc_VanLeerX_PPA_b(orc_state *S, const double *q) {
  ORC_MESH(S);
  const double *slope = S->Slope;
  double *qL = S->QL, *qR = S->QR;
  int size_y = Ny + 2 * NGHY, size_z = Nz + 2 * NGHZ;
  KRANGE(0, size_z);
  for (int k = kbeg; k < kend; k++)
    for (int j = 0; j < size_y; j++)
      for (int i = 0; i < Nx; i++) {
        int ll = lidx(i, j, k), llxp = LXP(ll, i);
        double temp = q[ll] + 0.5 * (q[llxp] - q[ll]) - 1.0 / 6.0 * (slope[llxp] - slope[ll]);
        qR[ll] = temp;
        qL[llxp] = temp;
      }
}

/* src/fargo_ppa_c.c:10-73 */
void orc_VanLeerX_PPA_c(orc_state *S, const double *q) {
  ORC_MESH(S);
  double *qL = S->QL, *qR = S->QR;
  int size_y = Ny + 2 * NGHY, size_z = Nz + 2 * NGHZ;
  KRANGE(0, size_z);
  for (int k = kbeg; k < kend; k++)
    for (int j = 0; j < size_y; j++)
      for (int i = 0; i < Nx; i++) {
        int ll = lidx(i, j, k);
        if ((qR[ll] - q[ll]) * (q[ll] - qL[ll]) < 0.0) {
          qL[ll] = q[ll];
          qR[ll] = q[ll];
        }
        double diff = qR[ll] - qL[ll];
        double cord = q[ll] - 0.5 * (qL[ll] + qR[ll]);
        if (6.0 * diff * cord > diff * diff) qL[ll] = 3.0 * q[ll] - 2.0 * qR[ll];
        if (-diff * diff > 6.0 * diff * cord) qR[ll] = 3.0 * q[ll] - 2.0 * qL[ll];
      }
}

/* src/fargo_ppa_d.c:10-80 */
void orc_VanLeerX_PPA_d(orc_state *S, double dt, const double *q, double *qs, const double *vx) {
  ORC_MESH(S);
  const double *qL = S->QL, *qR = S->QR;
  int size_y = Ny + 2 * NGHY, size_z = Nz + 2 * NGHZ;
  KRANGE(0, size_z);
  for (int k = kbeg; k < kend; k++)
    for (int j = 0; j < size_y; j++)
      for (int i = 0; i < Nx; i++) {
        int ll = lidx(i, j, k), llxm = LXM(ll, i);
        if (vx[ll] > 0.0) {
          double ksi = vx[ll] * dt / zone_size_x(i, j, k);
          qs[ll] = qR[llxm] + ksi * (q[llxm] - qR[llxm]);
          qs[ll] += ksi * (1.0 - ksi) * (2.0 * q[llxm] - qR[llxm] - qL[llxm]);
        } else {
          double ksi = -vx[ll] * dt / zone_size_x(i, j, k);
          qs[ll] = qL[ll] + ksi * (q[ll] - qL[ll]);
          qs[ll] += ksi * (1.0 - ksi) * (2.0 * q[ll] - qR[ll] - qL[ll]);
        }
      }
}

/* src/update_x.c:10-82 */
void orc_UpdateX(orc_state *S, double dt, double *qb, const double *qs, const double *vx) {
  ORC_MESH(S);
  const double *rho_s = S->DensStar;
  int size_y = Ny + 2 * NGHY, size_z = Nz + 2 * NGHZ;
  KRANGE(0, size_z);
  for (int k = kbeg; k < kend; k++)
    for (int j = 0; j < size_y; j++)
      for (int i = 0; i < Nx; i++) {
        int ll = lidx(i, j, k), llxp = LXP(ll, i);
        qb[ll] += ((vx[ll] * qs[ll] * rho_s[ll] - vx[llxp] * qs[llxp] * rho_s[llxp]) * SurfX(j, k) * dt *
                   InvVol(i, j, k));
      }
}

/* src/update_y.c:10-82 */
void orc_UpdateY(orc_state *S, double dt, double *qb, const double *qs) {
  ORC_MESH(S);
  const double *vy = S->Vy_temp, *rho_s = S->DensStar;
  int size_y = Ny + 2 * NGHY - 1, size_z = Nz + 2 * NGHZ;
  KRANGE(0, size_z);
  for (int k = kbeg; k < kend; k++)
    for (int j = 0; j < size_y; j++)
      for (int i = 0; i < Nx; i++) {
        int ll = lidx(i, j, k), llyp = LYP(ll);
        qb[ll] += (vy[ll] * qs[ll] * rho_s[ll] * SurfY(i, j, k) - vy[llyp] * qs[llyp] * rho_s[llyp] * SurfY(i, j + 1, k)) *
                  dt * InvVol(i, j, k);
      }
}

/* src/update_z.c */
void orc_UpdateZ(orc_state *S, double dt, double *qb, const double *qs) {
  ORC_MESH(S);
  if (!DIMZ) return;
  const double *vz = S->Vz_temp, *rho_s = S->DensStar;
  int size_y = Ny + 2 * NGHY, size_z = Nz + 2 * NGHZ - 1;
  for (int k = 0; k < size_z; k++)
    for (int j = 0; j < size_y; j++)
      for (int i = 0; i < Nx; i++) {
        int ll = lidx(i, j, k), llzp = LZP(ll);
        qb[ll] += (vz[ll] * qs[ll] * rho_s[ll] * SurfZ(i, j, k) - vz[llzp] * qs[llzp] * rho_s[llzp] * SurfZ(i, j, k + 1)) *
                  dt * InvVol(i, j, k);
      }
}

/* src/update_density_x.c:10-80 */
void orc_UpdateDensityX(orc_state *S, double dt, double *qb, const double *vx) {
  ORC_MESH(S);
  const double *rho_s = S->DensStar;
  int size_y = Ny + 2 * NGHY, size_z = Nz + 2 * NGHZ;
  KRANGE(0, size_z);
  for (int k = kbeg; k < kend; k++)
    for (int j = 0; j < size_y; j++)
      for (int i = 0; i < Nx; i++) {
        int ll = lidx(i, j, k), llxp = LXP(ll, i);
        qb[ll] += ((vx[ll] * rho_s[ll] - vx[llxp] * rho_s[llxp]) * SurfX(j, k) * dt * InvVol(i, j, k));
      }
}

/* src/update_density_y.c:10-80 */
void orc_UpdateDensityY(orc_state *S, double dt, double *qb) {
  ORC_MESH(S);
  const double *vy = S->Vy_temp, *rho_s = S->DensStar;
  int size_y = Ny + 2 * NGHY - 1, size_z = Nz + 2 * NGHZ;
  KRANGE(0, size_z);
  for (int k = kbeg; k < kend; k++)
    for (int j = 0; j < size_y; j++)
      for (int i = 0; i < Nx; i++) {
        int ll = lidx(i, j, k), llyp = LYP(ll);
        qb[ll] += (vy[ll] * rho_s[ll] * SurfY(i, j, k) - vy[llyp] * rho_s[llyp] * SurfY(i, j + 1, k)) * dt *
                  InvVol(i, j, k);
      }
}

/* src/update_density_z.c */
void orc_UpdateDensityZ(orc_state *S, double dt, double *qb) {
  ORC_MESH(S);
  if (!DIMZ) return;
  const double *vz = S->Vz_temp, *rho_s = S->DensStar;
  int size_y = Ny + 2 * NGHY, size_z = Nz + 2 * NGHZ - 1;
  for (int k = 0; k < size_z; k++)
    for (int j = 0; j < size_y; j++)
      for (int i = 0; i < Nx; i++) {
        int ll = lidx(i, j, k), llzp = LZP(ll);
        qb[ll] += (vz[ll] * rho_s[ll] * SurfZ(i, j, k) - vz[llzp] * rho_s[llzp] * SurfZ(i, j, k + 1)) * dt *
                  InvVol(i, j, k);
      }
}

/* src/compute_residual.c:10-88 */
void orc_ComputeResidual(orc_state *S, double dt) {
  ORC_MESH(S);
  double *vx = S->Vx_temp, *vxr = S->Vx;
  const double *vxmed = S->VxMed;
  int *nshift = S->Nshift;
  int size_y = Ny + 2 * NGHY, size_z = Nz + 2 * NGHZ;
  KRANGE(0, size_z);
  for (int k = kbeg; k < kend; k++)
    for (int j = 0; j < size_y; j++)
      for (int i = 0; i < Nx; i++) {
        int ll = lidx(i, j, k), ll2D = L2D(j, k);
        double ntilde = vxmed[ll2D] * dt / zone_size_x(i, j, k);
        double nround = floor(ntilde + 0.5);
        if (i == 0) nshift[ll2D] = (int)nround;
        vxr[ll] = vx[ll] - vxmed[ll2D];
        vx[ll] = (ntilde - nround) * zone_size_x(i, j, k) / dt;
      }
}

/* src/advect_shift.c:9-68 */
void orc_AdvectSHIFT(orc_state *S, double *f, const int *nshift) {
  ORC_MESH(S);
  double *aux = S->Pressure;
  int size_y = Ny + 2 * NGHY, size_z = Nz + 2 * NGHZ;
  for (int k = 0; k < size_z; k++)
    for (int j = 0; j < size_y; j++)
      for (int i = 0; i < Nx; i++) {
        int ll = lidx(i, j, k);
        int itarget = i - nshift[L2D(j, k)];
        while (itarget < 0) itarget += Nx;
        while (itarget >= Nx) itarget -= Nx;
        int ltarget = ll - i + itarget;
        aux[ll] = f[ltarget];
      }
  memcpy(f, aux, sizeof(double) * S->g.ncell);
}

/* src/newvel_x.c:10-99 */
void orc_NewVelocity_x(orc_state *S) {
  ORC_MESH(S);
  const double *rho = S->Density, *mmx = S->Mmx, *mpx = S->Mpx;
  double *vx = S->Vx;
  const double OMEGAFRAME = S->clk.OMEGAFRAME;
  int size_y = Ny + 2 * NGHY, size_z = Nz + 2 * NGHZ;
  KRANGE(1, size_z);
  for (int k = kbeg; k < kend; k++)
    for (int j = 1; j < size_y; j++)
      for (int i = 0; i < Nx; i++) {
        int ll = lidx(i, j, k), llxm = LXM(ll, i);
        if (!SPH) {
          vx[ll] = (mmx[ll] * Vol(i, j, k) + mpx[llxm] * Vol(IXM(i), j, k)) /
                       ((rho[ll] * Vol(i, j, k) + rho[llxm] * Vol(IXM(i), j, k)) * ymed(j)) -
                   OMEGAFRAME * ymed(j);
        } else {
          double rcyl = ymed(j) * sin(zmed(k));
          vx[ll] = (mmx[ll] * Vol(i, j, k) + mpx[llxm] * Vol(IXM(i), j, k)) /
                       ((rho[ll] * Vol(i, j, k) + rho[llxm] * Vol(IXM(i), j, k)) * rcyl) -
                   OMEGAFRAME * rcyl;
        }
      }
}

/* src/newvel_y.c:10-80 */
void orc_NewVelocity_y(orc_state *S) {
  ORC_MESH(S);
  const double *rho = S->Density, *mmy = S->Mmy, *mpy = S->Mpy;
  double *vy = S->Vy;
  int size_y = Ny + 2 * NGHY, size_z = Nz + 2 * NGHZ;
  KRANGE(1, size_z);
  for (int k = kbeg; k < kend; k++)
    for (int j = 1; j < size_y; j++)
      for (int i = 0; i < Nx; i++) {
        int ll = lidx(i, j, k), llym = LYM(ll);
        vy[ll] = (mmy[ll] * Vol(i, j, k) + mpy[llym] * Vol(i, j - 1, k)) /
                 (rho[ll] * Vol(i, j, k) + rho[llym] * Vol(i, j - 1, k));
      }
}

/* src/newvel_z.c:10-84 */
void orc_NewVelocity_z(orc_state *S) {
  ORC_MESH(S);
  if (!DIMZ) return;
  const double *rho = S->Density, *mmz = S->Mmz, *mpz = S->Mpz;
  double *vz = S->Vz;
  int size_y = Ny + 2 * NGHY, size_z = Nz + 2 * NGHZ;
  for (int k = 1; k < size_z; k++)
    for (int j = 1; j < size_y; j++)
      for (int i = 0; i < Nx; i++) {
        int ll = lidx(i, j, k), llzm = LZM(ll);
        vz[ll] = (mmz[ll] * Vol(i, j, k) + mpz[llzm] * Vol(i, j, k - 1)) /
                 (rho[ll] * Vol(i, j, k) + rho[llzm] * Vol(i, j, k - 1));
        if (SPH) vz[ll] /= ymed(j);
      }
}

/* src/stockholm.c:10-197 */
void orc_StockholmBoundary(orc_state *S, double dt) {
  ORC_MESH(S);
  double *rho = S->Density, *vx = S->Vx, *vy = S->Vy, *vz = S->Vz, *e = S->Energy;
  const double *rho0 = S->Density0, *vx0 = S->Vx0, *vy0 = S->Vy0, *vz0 = S->Vz0, *e0 = S->Energy0;
  (void)e; (void)e0;
  const double y_max = S->p.ymax, y_min = S->p.ymin, z_max = S->p.zmax, z_min = S->p.zmin;
  const double of = S->clk.OMEGAFRAME, of0 = S->clk.OMEGAFRAME0;
  const double ds = S->p.taudamp;
  const int periodic_z = S->p.periodicz;
  const double Y_inf = S->damp.Y_inf, Y_sup = S->damp.Y_sup, Z_inf = S->damp.Z_inf, Z_sup = S->damp.Z_sup;
  int size_y = Ny + 2 * NGHY, size_z = Nz + 2 * NGHZ;
  KRANGE(0, size_z);
  for (int k = kbeg; k < kend; k++)
    for (int j = 0; j < size_y; j++)
      for (int i = 0; i < Nx; i++) {
        int l = lidx(i, j, k);
        double rampy = 0.0, rampz = 0.0, rampzz = 0.0;
        if (ymed(j) > Y_sup) rampy = (ymed(j) - Y_sup) / (y_max - Y_sup);
        if (ymed(j) < Y_inf) rampy = (Y_inf - ymed(j)) / (Y_inf - y_min);
        rampy *= rampy;
        if (DIMZ) {
          if (zmed(k) > Z_sup) rampz = (zmed(k) - Z_sup) / (z_max - Z_sup);
          if (zmed(k) < Z_inf) rampz = (Z_inf - zmed(k)) / (Z_inf - z_min);
          rampz = rampz * rampz;
          if (zmin(k) > Z_sup) rampzz = (zmin(k) - Z_sup) / (z_max - Z_sup);
          if (zmin(k) < Z_inf) rampzz = (Z_inf - zmin(k)) / (Z_inf - z_min);
          rampzz = rampzz * rampzz;
        }
        if (periodic_z) {
          rampz = 0.0;
          rampzz = 0.0;
        }
        double ramp = rampy + rampz;
        double rampi = rampy + rampzz;
        double tau = ds * sqrt(ymed(j) * ymed(j) * ymed(j) / G / MSTAR);
        if (ramp > 0.0) {
          double taud = tau / ramp;
          rho[l] = (rho[l] * taud + rho0[L2D(j, k)] * dt) / (dt + taud);
          double vx0_target = vx0[L2D(j, k)];
          double radius = ymed(j);
          if (SPH) radius *= sin(zmed(k));
          vx0_target -= (of - of0) * radius;
          vx[l] = (vx[l] * taud + vx0_target * dt) / (dt + taud);
          vy[l] = (vy[l] * taud + vy0[L2D(j, k)] * dt) / (dt + taud);
        }
        if (DIMZ) {
          if (rampi > 0.0) {
            double taud = tau / rampi;
            vz[l] = (vz[l] * taud + vz0[L2D(j, k)] * dt) / (dt + taud);
          }
        }
      }
}

/* src/init_stockholm.c:3-126 (fresh start) */
void orc_init_stockholm(orc_state *S) {
  ORC_MESH(S);
  for (int k = 0; k < (DIMZ ? Nz + 2 * NGHZ : 1); k++)
    for (int j = 0; j < Ny + 2 * NGHY; j++) {
      int l = lidx(0, j, k), l2 = L2D(j, k);
      if (S->p.eos == F3D_ADIABATIC) S->Energy0[l2] = S->Energy[l];
      S->Vx0[l2] = S->Vx[l];
      S->Vy0[l2] = S->Vy[l];
      if (DIMZ) S->Vz0[l2] = S->Vz[l];
      S->Density0[l2] = S->Density[l];
    }
}

/* ------------------------------------------------------------------ multifluid */

/* _collisions_cpu (src/collisions.c:10-154) with the STOKESNUMBER collision
   matrix (src/collision_kernel.h:14-94) and the partial-pivot Gauss solver
   (src/gauss.h:1-45), option 1 (V_temp in and out), one direction
   (id1, id2, id3) per call. */
static void collisions_dir(orc_state *S, double dt, int id1, int id2, int id3) {
  ORC_MESH(S);
  const int NF = S->nfluids;
  double *rho[F3D_MAX_FLUIDS], *vel[F3D_MAX_FLUIDS];
  for (int f = 0; f < NF; f++) {
    rho[f] = S->fl[f].Density;
    vel[f] = id1 ? S->fl[f].Vx_temp : id2 ? S->fl[f].Vy_temp : S->fl[f].Vz_temp;
  }
  const double *alpha = S->alpha;
  const int size_y = Ny + 2 * NGHY, size_z = Nz + 2 * NGHZ;
  KRANGE(1, size_z);
  double b[F3D_MAX_FLUIDS], m[F3D_MAX_FLUIDS * F3D_MAX_FLUIDS];
  for (int k = kbeg; k < kend; k++)
    for (int j = 1; j < size_y; j++)
      for (int i = 0; i < Nx; i++) {
        const int l = lidx(i, j, k);
        const int idm = LXM(l, i) * id1 + LYM(l) * id2 + LZM(l) * id3;
        const double omega = (id1 + id3) * sqrt(G * MSTAR / (ymed(j) * ymed(j) * ymed(j))) +
                             id2 * sqrt(G * MSTAR / (ymin(j) * ymin(j) * ymin(j)));
        for (int o = 0; o < NF; o++) {
          for (int p = 0; p < NF; p++) {
            const double rho_p = 0.5 * (rho[p][l] + rho[p][idm]);
            if (p != o) {
              const double rho_o = 0.5 * (rho[o][l] + rho[o][idm]);
              if (p > o) m[p + o * NF] = -dt * omega * alpha[p + o * NF] * rho_p / rho_o;
              else m[p + o * NF] = -dt * omega * alpha[p + o * NF];
            } else {
              double sum = 0.0;
              for (int q = 0; q < NF; q++) {
                if (q != p) {
                  const double rho_q = 0.5 * (rho[q][l] + rho[q][idm]);
                  if (q > p) sum += alpha[q + p * NF] * rho_q / rho_p;
                  else sum += alpha[q + p * NF];
                }
              }
              m[p + p * NF] = 1.0 + dt * omega * sum;
            }
          }
          b[o] = vel[o][l];
        }
        for (int ir = 0; ir < NF - 1; ir++) {
          /* pivot: as src/gauss.h:8-14 (the stored maximum is the signed value) */
          double max_value = -1;
          int ir_max = ir;
          for (int ir2 = ir; ir2 < NF; ir2++) {
            if (fabs(m[ir + ir2 * NF]) >= max_value) {
              max_value = m[ir + ir2 * NF];
              ir_max = ir2;
            }
          }
          if (ir_max != ir) {
            for (int ic = 0; ic < NF; ic++) {
              double temp = m[ic + ir * NF];
              m[ic + ir * NF] = m[ic + ir_max * NF];
              m[ic + ir_max * NF] = temp;
            }
            double temp = b[ir];
            b[ir] = b[ir_max];
            b[ir_max] = temp;
          }
          for (int ir2 = ir + 1; ir2 < NF; ir2++) {
            const double factor = m[ir + ir2 * NF] / m[ir + ir * NF];
            for (int ic = 0; ic < NF; ic++) m[ic + ir2 * NF] -= m[ic + ir * NF] * factor;
            b[ir2] -= b[ir] * factor;
          }
        }
        for (int ir = NF - 1; ir > -1; ir--) {
          for (int ic = NF - 1; ic > ir; ic--) b[ir] -= m[ic + ir * NF] * b[ic];
          b[ir] /= m[ir + ir * NF];
        }
        for (int o = 0; o < NF; o++) vel[o][l] = b[o];
      }
}

/* Collisions(dt, 1) (src/collisions.c:156-189): X, then Y, then Z */
void orc_Collisions(orc_state *S, double dt) {
  collisions_dir(S, dt, 1, 0, 0);
  collisions_dir(S, dt, 0, 1, 0);
  if (S->g.dimz) collisions_dir(S, dt, 0, 0, 1);
}

/* DustDiffusion_Coefficients_cpu (src/dust_diffusion_coefficients.c:10-143):
   the diffusion coefficient at cell centres and at lower Y faces, from the
   gas (Fluids[0]); 2-D isothermal forms */
void orc_DustDiffusion_Coefficients(orc_state *S) {
  ORC_MESH(S);
  const double *cs = S->fl[0].Energy;
  double *yczc = S->Sdiffyczc, *yfzc = S->Sdiffyfzc;
  const int size_y = Ny + 2 * NGHY, size_z = Nz + 2 * NGHZ;
  KRANGE(0, size_z);
  for (int k = kbeg; k < kend; k++)
    for (int j = 0; j < size_y; j++)
      for (int i = 0; i < Nx; i++) {
        const int ll = lidx(i, j, k), llym = LYM(ll);
        if (S->p.visc == F3D_VISC_ALPHA) {
          const double soundspeed2 = cs[ll] * cs[ll];
          double soundspeedf2;
          if (j == 0) soundspeedf2 = soundspeed2;
          else soundspeedf2 = 0.5 * (cs[ll] + cs[llym]) * 0.5 * (cs[ll] + cs[llym]);
          const double r3yczc = ymed(j) * ymed(j) * ymed(j);
          const double r3yfzc = ymin(j) * ymin(j) * ymin(j);
          yczc[ll] = S->p.alpha * soundspeed2 / sqrt(G * MSTAR / r3yczc);
          yfzc[ll] = S->p.alpha * soundspeedf2 / sqrt(G * MSTAR / r3yfzc);
        } else {
          yczc[ll] = S->p.nu;
          yfzc[ll] = S->p.nu;
        }
      }
}

/* DustDiffusion_Core_cpu (src/dust_diffusion_core.c:10-192), CYLINDRICAL 2-D:
   the selected (dust) fluid's updated density into Pressure */
void orc_DustDiffusion_Core(orc_state *S, double dt) {
  ORC_MESH(S);
  const double *yfzc = S->Sdiffyfzc, *yczc = S->Sdiffyczc;
  const double *rhod = S->Density, *rhog = S->fl[0].Density;
  double *temp = S->Pressure;
  const int size_y = Ny + 2 * NGHY - 1, size_z = Nz + 2 * NGHZ - 1;
  KRANGE(1, size_z);
  for (int k = kbeg; k < kend; k++)
    for (int j = 1; j < size_y; j++)
      for (int i = 0; i < Nx; i++) {
        const int ll = lidx(i, j, k), llxm = LXM(ll, i), llxp = LXP(ll, i), llyp = LYP(ll), llym = LYM(ll);
        double update = 0.0;
        const double c = rhod[ll] / (rhod[ll] + rhog[ll]);
        const double dxmin = xmin(i + 1) - xmin(i);
        const double dxmedp = 1.0 / InvDiffXmed[IXP(i)];
        const double dxmedm = 1.0 / InvDiffXmed[i];
        const double d1 = 0.25 * (rhod[ll] + rhog[ll] + rhod[llxp] + rhog[llxp]) * (yczc[llxp] + yczc[ll]);
        const double d2 = 0.25 * (rhod[ll] + rhog[ll] + rhod[llxm] + rhog[llxm]) * (yczc[llxm] + yczc[ll]);
        const double cxp = rhod[llxp] / (rhod[llxp] + rhog[llxp]);
        const double cxm = rhod[llxm] / (rhod[llxm] + rhog[llxm]);
        update += 1.0 / ymed(j) / ymed(j) / (dxmin) * (d1 * (cxp - c) / (dxmedp) - (d2 * (c - cxm)) / (dxmedm));
        const double d3 = 0.5 * (rhod[ll] + rhog[ll] + rhod[llyp] + rhog[llyp]) * yfzc[llyp];
        const double d4 = 0.5 * (rhod[ll] + rhog[ll] + rhod[llym] + rhog[llym]) * yfzc[ll];
        const double cyp = rhod[llyp] / (rhod[llyp] + rhog[llyp]);
        const double cym = rhod[llym] / (rhod[llym] + rhog[llym]);
        update += 1.0 / ymed(j) / (ymin(j + 1) - ymin(j)) *
                  (ymin(j + 1) * d3 * (cyp - c) / (ymed(j + 1) - ymed(j)) -
                   (ymin(j) * d4 * (c - cym)) / (ymed(j) - ymed(j - 1)));
        temp[ll] = rhod[ll] + dt * update;
      }
}

/* DustDiffusion_Main (src/dust_diffusion_main.c:3-28): coefficients, then for
   each dust fluid the core, copy_field(Density, Pressure) (src/copy_field.c:3-12)
   and the density ghost fill */
void orc_DustDiffusion_Main(orc_state *S, double dt) {
  orc_DustDiffusion_Coefficients(S);
  for (int f = 0; f < S->nfluids; f++) {
    orc_select_fluid(S, f);
    if (S->Fluidtype != F3D_DUST) continue;
    orc_DustDiffusion_Core(S, dt);
    memcpy(S->Density, S->Pressure, S->g.ncell * sizeof(double));
    S->fill_density_only = 1;
    orc_FillGhosts(S);
    S->fill_density_only = 0;
  }
}

/* src/mon_dens.c:10-68, src/mon_momx.c, src/mon_momy.c, src/mon_momz.c,
   src/mon_torque.c: the monitored quantity into Slope over the active cells,
   the cells the monitor reductions read (src/monitor.c:232-233).
   q: 0 mass, 1 momx, 2 momy, 3 momz, 4 torq. */
void orc_monitor(orc_state *S, int q, double Xplanet, double Yplanet, double Zplanet, double rsmoothing) {
  ORC_MESH(S);
  const double OMEGAFRAME = S->clk.OMEGAFRAME;
  const double *dens = S->Density, *vx = S->Vx, *vy = S->Vy, *vz = S->Vz;
  double *interm = S->Slope;
  const double rsm2 = rsmoothing * rsmoothing;
  const int kbeg = DIMZ ? NGHZ : 0, kend = DIMZ ? Nz + NGHZ : 1;
  for (int k = kbeg; k < kend; k++)
    for (int j = NGHY; j < Ny + NGHY; j++)
      for (int i = 0; i < Nx; i++) {
        int ll = lidx(i, j, k);
        int lxp = (i < Nx - 1) ? ll + 1 : ll - (Nx - 1);
        int lyp = ll + pitch, lzp = ll + stride;
        double rcyl, dx, dy, dz = 0.0;
        switch (q) {
          case 0:
            interm[ll] = dens[ll] * Vol(i, j, k);
            break;
          case 1:
            if (SPH) {
              rcyl = ymed(j) * sin(zmed(k));
              interm[ll] = (.5 * (vx[ll] + vx[lxp]) + rcyl * OMEGAFRAME) * rcyl * dens[ll] * Vol(i, j, k);
            } else {
              interm[ll] = (.5 * (vx[ll] + vx[lxp]) + ymed(j) * OMEGAFRAME) * ymed(j) * dens[ll] * Vol(i, j, k);
            }
            break;
          case 2:
            interm[ll] = dens[ll] * .5 * (vy[ll] + vy[lyp]) * Vol(i, j, k);
            break;
          case 3:
            interm[ll] = dens[ll] * .5 * (vz[ll] + vz[lzp]) * Vol(i, j, k);
            if (SPH) interm[ll] *= ymed(j);
            break;
          default: {
            double cellmass = Vol(i, j, k) * dens[ll];
            if (SPH) {
              dx = ymed(j) * cos(xmed(i)) * sin(zmed(k)) - Xplanet;
              dy = ymed(j) * sin(xmed(i)) * sin(zmed(k)) - Yplanet;
              if (DIMZ) dz = ymed(j) * cos(zmed(k)) - Zplanet;
            } else {
              dx = ymed(j) * cos(xmed(i)) - Xplanet;
              dy = ymed(j) * sin(xmed(i)) - Yplanet;
              if (DIMZ) dz = zmed(k) - Zplanet;
            }
            double dist2 = dx * dx + dy * dy + dz * dz;
            dist2 += rsm2;
            double distance = sqrt(dist2);
            double InvDist3 = 1.0 / (dist2 * distance);
            InvDist3 *= G * cellmass;
            double fxi = dx * InvDist3;
            double fyi = dy * InvDist3;
            interm[ll] = Xplanet * fyi - Yplanet * fxi;
          }
        }
      }
}
