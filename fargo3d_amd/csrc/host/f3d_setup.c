/*
 * f3d_setup.c -- initial conditions, boundary-factor tables, damping set-up.
 *
 * Initial conditions restate the setups' condinit.c (setups/fargo/condinit.c
 * :25-50, setups/p3disof/condinit.c:14-52).  Boundary tables hold the host
 * (glibc) evaluation of every transcendental factor that the generated
 * boundary kernels (std/boundaries.txt:35-102 through scripts/boundparser.py
 * :237-363) compute from mesh coordinates.
 */
#include "f3d_host.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define G F3D_G
#define MSTAR F3D_MSTAR
#define R0 F3D_R0

void f3d_seed(uint64_t seed) {
  if (seed) srand48((long)seed);
}

/* Standard normal deviate from two drand48 draws (Box-Muller). */
static double gauss48(void) {
  double u1 = drand48(), u2 = drand48();
  if (u1 < 1e-300) u1 = 1e-300;
  return sqrt(-2.0 * log(u1)) * cos(2.0 * M_PI * u2);
}

int f3d_initial_conditions(const f3d_grid *g, const f3d_params *p, const f3d_clock *c,
                           double *dens, double *energy, double *vx, double *vy, double *vz) {
  return f3d_initial_conditions_fluid(g, p, c, F3D_GAS, dens, energy, vx, vy, vz);
}

/* setups/fargo_multifluid/condinit.c:8-52: gas in pressure-supported
   rotation, each dust fluid Keplerian and pressureless with EPSILON times the
   gas density; no noise */
static void multifluid_ic(const f3d_grid *g, const f3d_params *p, double OMEGAFRAME, int fluidtype, double *dens,
                          double *energy, double *vx, double *vy) {
  for (int j = 0; j < g->sy; j++) {
    for (int i = 0; i < g->sx; i++) {
      size_t l = (size_t)i + (size_t)j * g->pitch;
      double r = g->ymed[j];
      double omega = sqrt(G * MSTAR / r / r / r);
      double rhog = p->sigma0 * pow(r / R0, -p->sigmaslope);
      double rhod = rhog * p->epsilon;
      if (fluidtype == F3D_GAS) {
        dens[l] = rhog;
        vx[l] = omega * r * sqrt(1.0 + pow(p->aspectratio, 2) * pow(r / R0, 2 * p->flaringindex) *
                                           (2.0 * p->flaringindex - 1.0 - p->sigmaslope));
        vy[l] = 0.0;
        energy[l] = p->aspectratio * pow(r / R0, p->flaringindex) * omega * r;
      } else {
        dens[l] = rhod;
        vx[l] = omega * r;
        vy[l] = 0.0;
        energy[l] = 0.0;
      }
      vx[l] -= OMEGAFRAME * r;
    }
  }
}

int f3d_initial_conditions_fluid(const f3d_grid *g, const f3d_params *p, const f3d_clock *c, int fluidtype,
                                 double *dens, double *energy, double *vx, double *vy, double *vz) {
  const double OMEGAFRAME = c->OMEGAFRAME;
  f3d_seed(p->seed);
  size_t n = g->ncell;
  memset(dens, 0, n * sizeof(double));
  memset(energy, 0, n * sizeof(double));
  memset(vx, 0, n * sizeof(double));
  memset(vy, 0, n * sizeof(double));
  if (vz) memset(vz, 0, n * sizeof(double));
  if (strcmp(p->setup, "fargo_multifluid") == 0) {
    if (g->dimz) return -1;
    multifluid_ic(g, p, OMEGAFRAME, fluidtype, dens, energy, vx, vy);
    return 0;
  }
  if (!g->dimz) {
    /* 2-D cylindrical disks: setups/fargo/condinit.c:25-50 */
    for (int j = 0; j < g->sy; j++) {
      for (int i = 0; i < g->sx; i++) {
        size_t l = (size_t)i + (size_t)j * g->pitch;
        double r = g->ymed[j];
        double omega = sqrt(G * MSTAR / r / r / r);
        double soundspeed;
        if (p->ic_keplerian) {
          /* v_phi = r^-1/2 - Omega_f r, v_r = 0, uniform noise on density only */
          dens[l] = p->sigma0 * pow(r / R0, -p->sigmaslope) * (1.0 + p->noise * (drand48() - .5));
          soundspeed = p->aspectratio * pow(r / R0, p->flaringindex) * omega * r;
          vx[l] = omega * r;
          vx[l] -= OMEGAFRAME * r;
          vy[l] = 0.0;
        } else if (p->eos == F3D_ADIABATIC) {
          /* log-normal density perturbations, pressure-supported rotation */
          dens[l] = p->sigma0 * pow(r / R0, -p->sigmaslope);
          if (p->noise != 0.0) dens[l] *= exp(p->noise * gauss48());
          soundspeed = p->aspectratio * pow(r / R0, p->flaringindex) * omega * r;
          vx[l] = omega * r * sqrt(1.0 + pow(p->aspectratio, 2) * pow(r / R0, 2 * p->flaringindex) *
                                             (2.0 * p->flaringindex - 1.0 - p->sigmaslope));
          vx[l] -= OMEGAFRAME * r;
          vy[l] = 0.0;
        } else {
          dens[l] = p->sigma0 * pow(r / R0, -p->sigmaslope) * (1.0 + p->noise * (drand48() - .5));
          soundspeed = p->aspectratio * pow(r / R0, p->flaringindex) * omega * r;
          vx[l] = omega * r * sqrt(1.0 + pow(p->aspectratio, 2) * pow(r / R0, 2 * p->flaringindex) *
                                             (2.0 * p->flaringindex - 1.0 - p->sigmaslope));
          vx[l] -= OMEGAFRAME * r;
          vx[l] *= (1. + p->aspectratio * p->noise * (drand48() - .5));
          vy[l] = soundspeed * p->noise * (drand48() - .5);
        }
        if (p->eos == F3D_ISOTHERMAL)
          energy[l] = soundspeed;
        else
          energy[l] = pow(soundspeed, 2) * dens[l] / (p->gamma - 1.0);
      }
    }
    return 0;
  }
  if (g->geometry != F3D_SPHERICAL) return -1;
  /* 3-D spherical disk: setups/p3disof/condinit.c:14-52 */
  for (int k = 0; k < g->sz; k++) {
    for (int j = 0; j < g->sy; j++) {
      double r = g->ymed[j];
      double r3 = r * r * r;
      double omega = sqrt(G * MSTAR / (r3));
      for (int i = 0; i < g->nx; i++) {
        size_t l = (size_t)i + (size_t)j * g->pitch + (size_t)k * g->stride;
        vy[l] = vz[l] = 0.0;
        vx[l] = omega * r;
        double xi = p->sigmaslope + 1. + p->flaringindex;
        double beta = 1. - 2 * p->flaringindex;
        double h = p->aspectratio * pow(r / R0, p->flaringindex);
        double sz = sin(g->zmed[k]);
        if (p->flaringindex == 0.0) {
          dens[l] = p->sigma0 / sqrt(2.0 * M_PI) / (R0 * p->aspectratio) * pow(r / R0, -xi) *
                    pow(sz, -beta - xi + 1. / (h * h));
        } else {
          dens[l] = p->sigma0 / sqrt(2.0 * M_PI) / (R0 * p->aspectratio) * pow(r / R0, -xi) *
                    pow(sz, -xi - beta) *
                    exp((1. - pow(sz, -2. * p->flaringindex)) / 2. / p->flaringindex / (h * h));
        }
        vx[l] *= sqrt(pow(sz, -2. * p->flaringindex) - (beta + xi) * h * h);
        vx[l] -= OMEGAFRAME * r * sz;
        if (p->eos == F3D_ISOTHERMAL)
          energy[l] = h * sqrt(G * MSTAR / r);
        else
          energy[l] = dens[l] * h * h * G * MSTAR / r / (p->gamma - 1.0);
      }
    }
  }
  return 0;
}

int f3d_bctab_init(f3d_bctab *t, const f3d_grid *g, const f3d_params *p) {
  memset(t, 0, sizeof(*t));
  const int ny = g->ny, nghy = g->nghy, nz = g->nz, nghz = g->nghz;
  const double *Ymed = g->ymed, *Zmed = g->zmed;
  const double SS = p->sigmaslope, FI = p->flaringindex, AR = p->aspectratio;
  for (int m = 0; m < nghy; m++) {
    /* ymin side: jgh = m, jact = 2*nghy-m-1 (scripts/boundparser.py:323-331) */
    int jgh = m, jact = 2 * nghy - m - 1;
    t->ymin_dens[m] = pow(Ymed[jact] / Ymed[jgh], SS);
    t->ymin_vsq[m] = sqrt(Ymed[jact] / Ymed[jgh]);
    t->ymin_vfac[m] = (1. + (2. + SS - FI) * (Ymed[jact] - Ymed[jgh]) / R0 * FI * AR * AR *
                                pow((Ymed[jgh] + Ymed[jact]) / (2. * R0), 2. * FI - 1.));
    /* ymax side: jgh = ny+nghy+m, jact = ny+nghy-1-m (scripts/boundparser.py:332-340) */
    jgh = ny + nghy + m;
    jact = ny + nghy - 1 - m;
    t->ymax_dens[m] = pow(Ymed[jact] / Ymed[jgh], SS);
    t->ymax_vsq[m] = sqrt(Ymed[jact] / Ymed[jgh]);
    t->ymax_vfac[m] = (1. + (2. + SS - FI) * (Ymed[jact] - Ymed[jgh]) / R0 * FI * AR * AR *
                                pow((Ymed[jgh] + Ymed[jact]) / (2. * R0), 2. * FI - 1.));
  }
  if (g->dimz && p->bc == F3D_BC_P3D) {
    /* 3-D density factors: KEPLERIAN3DSPHDENSRAD (radial) uses
       ds*pow(Ymed(jact)/Ymed(jgh),SS+FI+1.)*exp(...); the pow part is
       stored in ymin_dens/ymax_dens, the exp part per (m,k). */
    t->ymin_dexp = (double *)calloc((size_t)nghy * g->sz, sizeof(double));
    t->ymax_dexp = (double *)calloc((size_t)nghy * g->sz, sizeof(double));
    t->zmin_dens = (double *)calloc((size_t)nghz * g->sy, sizeof(double));
    t->zmax_dens = (double *)calloc((size_t)nghz * g->sy, sizeof(double));
    for (int m = 0; m < nghy; m++) {
      for (int side = 0; side < 2; side++) {
        int jgh = side ? ny + nghy + m : m;
        int jact = side ? ny + nghy - 1 - m : 2 * nghy - m - 1;
        double pw = pow(Ymed[jact] / Ymed[jgh], SS + FI + 1.);
        if (side) t->ymax_dens[m] = pw; else t->ymin_dens[m] = pw;
        for (int k = 0; k < g->sz; k++) {
          double e = exp(-pow(cos(Zmed[k]), 2.) / (AR * AR) *
                         (1. - Ymed[jgh] / (.5 * (Ymed[jact] + Ymed[jgh]))) * FI *
                         pow(.5 * (Ymed[jgh] + Ymed[jact]) / R0, -2. * FI - 1.));
          if (side) t->ymax_dexp[(size_t)m * g->sz + k] = e;
          else t->ymin_dexp[(size_t)m * g->sz + k] = e;
        }
      }
    }
    for (int m = 0; m < nghz; m++) {
      for (int side = 0; side < 2; side++) {
        /* zmin: kgh = m, kact = 2*nghz-m-1; zmax: kgh = nz+nghz+m, kact = nz+nghz-1-m */
        int kgh = side ? nz + nghz + m : m;
        int kact = side ? nz + nghz - 1 - m : 2 * nghz - m - 1;
        double vf = (1. + FI * cos(.5 * (Zmed[kgh] + Zmed[kact])) * (Zmed[kact] - Zmed[kgh]));
        if (side) t->zmax_vfac[m] = vf; else t->zmin_vfac[m] = vf;
        for (int j = 0; j < g->sy; j++) {
          double d = pow(sin(Zmed[kgh]) / sin(Zmed[kact]),
                         FI - 2. - SS + 1. / (AR * AR) * pow(Ymed[j] / R0, -2. * FI));
          if (side) t->zmax_dens[(size_t)m * g->sy + j] = d;
          else t->zmin_dens[(size_t)m * g->sy + j] = d;
        }
      }
    }
  }
  return 0;
}

void f3d_bctab_free(f3d_bctab *t) {
  free(t->ymin_dexp); free(t->ymax_dexp); free(t->zmin_dens); free(t->zmax_dens);
  t->ymin_dexp = t->ymax_dexp = t->zmin_dens = t->zmax_dens = NULL;
}

void f3d_damp_init(f3d_damp *d, const f3d_params *p) {
  /* src/stockholm.c:76-90 */
  double y_min = p->ymin, y_max = p->ymax, z_min = p->zmin, z_max = p->zmax;
  d->Y_inf = y_min * pow(p->dampingzone, 2.0 / 3.0);
  d->Y_sup = y_max * pow(p->dampingzone, -2.0 / 3.0);
  d->Z_inf = z_min - (z_max - z_min);
  d->Z_sup = z_max + (z_max - z_min);
  if (p->geometry == F3D_CYLINDRICAL) {
    d->Z_inf = z_min + (z_max - z_min) * 0.1;
    d->Z_sup = z_max - (z_max - z_min) * 0.1;
    if (p->periodicz) {
      d->Z_inf = z_min - R0;
      d->Z_sup = z_max + R0;
    }
  } else {
    d->Z_inf = M_PI / 2.0 - (M_PI / 2.0 - z_min) * (1.0 - p->killingbccolatitude);
    d->Z_sup = M_PI / 2.0 + (M_PI / 2.0 - z_min) * (1.0 - p->killingbccolatitude);
  }
}
