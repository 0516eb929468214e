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

/* ---- glibc's drand48 generator: X' = (a X + c) mod 2^48, a draw is X'/2^48 ---- */
#define LCG_A 0x5DEECE66DULL
#define LCG_C 0xBULL
#define LCG_MASK ((1ULL << 48) - 1)

/* the n-fold composition of X -> (a X + c) mod 2^48, as X -> (A X + C):
   binary powering (every power of one affine map commutes with the others) */
static void lcg_pow(unsigned long long a, unsigned long long c, unsigned long long n, unsigned long long *A,
                    unsigned long long *C) {
  unsigned long long ra = 1, rc = 0;
  while (n) {
    if (n & 1) {
      rc = (a * rc + c) & LCG_MASK;
      ra = (a * ra) & LCG_MASK;
    }
    c = (a * c + c) & LCG_MASK;
    a = (a * a) & LCG_MASK;
    n >>= 1;
  }
  *A = ra;
  *C = rc;
}

unsigned long long f3d_drand48_skip(unsigned long long x0, unsigned long long n) {
  unsigned long long A, C;
  lcg_pow(LCG_A, LCG_C, n, &A, &C);
  return (A * (x0 & LCG_MASK) + C) & LCG_MASK;
}

/* the state srand48(seed) sets (glibc keeps the low 32 bits of the seed
   above 0x330E), or glibc's initial state without srand48 */
unsigned long long f3d_drand48_state(uint64_t seed) {
  if (!seed) return 0x1234ABCD330EULL;
  return ((((unsigned long long)seed) & 0xffffffffULL) << 16) | 0x330EULL;
}

void f3d_drand48_jumps(int ndraw, unsigned long long A[48], unsigned long long C[48]) {
  unsigned long long a, c;
  lcg_pow(LCG_A, LCG_C, (unsigned long long)ndraw, &a, &c);
  for (int b = 0; b < 48; b++) {
    A[b] = a;
    C[b] = c;
    c = (a * c + c) & LCG_MASK;
    a = (a * a) & LCG_MASK;
  }
}

/* ---- initial conditions: per-ring profiles and per-cell noise ----
 * Every set-up below is a profile of the rings (j,k) -- the density factor R,
 * the azimuthal velocity V in the frame, the sound speed CS and E, the
 * energy (PLAIN) or cs^2 (adiabatic) -- times noise drawn per cell in the
 * set-up's drand48 order.  f3d_ic_profiles evaluates the profiles with the
 * set-up's expressions (glibc transcendentals); f3d_ic_cell is the per-cell
 * arithmetic, which the device CondInit repeats operation for operation. */
int f3d_ic_profiles(const f3d_grid *g, const f3d_params *p, const f3d_clock *c, int fluidtype, f3d_ic_desc *d,
                    double *R, double *V, double *CS, double *E) {
  const double OMEGAFRAME = c->OMEGAFRAME;
  memset(d, 0, sizeof(*d));
  d->noise = p->noise;
  d->aspectratio = p->aspectratio;
  d->gamma = p->gamma;
  d->adiabatic = (p->eos == F3D_ADIABATIC);
  d->x0 = f3d_drand48_state(p->seed);
  size_t n2 = g->n2d;
  memset(R, 0, n2 * sizeof(double));
  memset(V, 0, n2 * sizeof(double));
  memset(CS, 0, n2 * sizeof(double));
  memset(E, 0, n2 * sizeof(double));
  if (strcmp(p->setup, "fargo_multifluid") == 0) {
    /* setups/fargo_multifluid/condinit.c:8-52: gas in pressure-supported
       rotation, each dust fluid Keplerian and pressureless with EPSILON times
       the gas density; no noise */
    if (g->dimz) return -1;
    d->mode = F3D_IC_PLAIN;
    d->adiabatic = 0;
    for (int j = 0; j < g->sy; j++) {
      double r = g->ymed[j];
      double omega = sqrt(G * MSTAR / r / r / r);
      double rhog = p->sigma0 * pow(r / R0, -p->sigmaslope);
      double v;
      if (fluidtype == F3D_GAS) {
        R[j] = rhog;
        v = omega * r * sqrt(1.0 + pow(p->aspectratio, 2) * pow(r / R0, 2 * p->flaringindex) *
                                       (2.0 * p->flaringindex - 1.0 - p->sigmaslope));
        E[j] = p->aspectratio * pow(r / R0, p->flaringindex) * omega * r;
      } else {
        R[j] = rhog * p->epsilon;
        v = omega * r;
        E[j] = 0.0;
      }
      v -= OMEGAFRAME * r;
      V[j] = v;
    }
    return 0;
  }
  if (!g->dimz) {
    /* 2-D cylindrical disks: setups/fargo/condinit.c:25-50 */
    if (p->ic_keplerian) {
      /* v_phi = r^-1/2 - Omega_f r, v_r = 0, uniform noise on the density only */
      d->mode = F3D_IC_KEPLER;
      d->ndraw = 1;
    } else if (p->eos == F3D_ADIABATIC) {
      /* log-normal density perturbations, pressure-supported rotation */
      d->mode = F3D_IC_LOGNORMAL;
      d->ndraw = p->noise != 0.0 ? 2 : 0;
    } else {
      d->mode = F3D_IC_FARGO;
      d->ndraw = 3;
    }
    for (int j = 0; j < g->sy; j++) {
      double r = g->ymed[j];
      double omega = sqrt(G * MSTAR / r / r / r);
      double soundspeed = p->aspectratio * pow(r / R0, p->flaringindex) * omega * r;
      double v;
      R[j] = p->sigma0 * pow(r / R0, -p->sigmaslope);
      if (p->ic_keplerian)
        v = omega * r;
      else
        v = omega * r * sqrt(1.0 + pow(p->aspectratio, 2) * pow(r / R0, 2 * p->flaringindex) *
                                       (2.0 * p->flaringindex - 1.0 - p->sigmaslope));
      v -= OMEGAFRAME * r;
      V[j] = v;
      CS[j] = soundspeed;
      E[j] = pow(soundspeed, 2);
    }
    return 0;
  }
  if (g->geometry != F3D_SPHERICAL) return -1;
  /* 3-D spherical disk: setups/p3disof/condinit.c:14-52, no noise */
  d->mode = F3D_IC_PLAIN;
  d->adiabatic = 0;
  for (int k = 0; k < g->sz; k++) {
    for (int j = 0; j < g->sy; j++) {
      size_t l2 = (size_t)j + (size_t)k * g->sy;
      double r = g->ymed[j];
      double r3 = r * r * r;
      double omega = sqrt(G * MSTAR / (r3));
      double v = omega * r;
      double xi = p->sigmaslope + 1. + p->flaringindex;
      double beta = 1. - 2 * p->flaringindex;
      double h = p->aspectratio * pow(r / R0, p->flaringindex);
      double sz = sin(g->zmed[k]);
      double dens;
      if (p->flaringindex == 0.0) {
        dens = p->sigma0 / sqrt(2.0 * M_PI) / (R0 * p->aspectratio) * pow(r / R0, -xi) *
               pow(sz, -beta - xi + 1. / (h * h));
      } else {
        dens = p->sigma0 / sqrt(2.0 * M_PI) / (R0 * p->aspectratio) * pow(r / R0, -xi) * pow(sz, -xi - beta) *
               exp((1. - pow(sz, -2. * p->flaringindex)) / 2. / p->flaringindex / (h * h));
      }
      v *= sqrt(pow(sz, -2. * p->flaringindex) - (beta + xi) * h * h);
      v -= OMEGAFRAME * r * sz;
      R[l2] = dens;
      V[l2] = v;
      if (p->eos == F3D_ISOTHERMAL)
        E[l2] = h * sqrt(G * MSTAR / r);
      else
        E[l2] = dens * h * h * G * MSTAR / r / (p->gamma - 1.0);
    }
  }
  return 0;
}

void f3d_ic_cell(const f3d_ic_desc *d, size_t l2, const double *u, double *rho, double *e, double *vx,
                 double *vy) {
  const double Rv = d->R[l2], Vv = d->V[l2];
  double r = Rv, x = Vv, y = 0.0;
  switch (d->mode) {
    case F3D_IC_PLAIN:
      *rho = Rv;
      *vx = Vv;
      *vy = 0.0;
      *e = d->E[l2];
      return;
    case F3D_IC_KEPLER:
      r = Rv * (1.0 + d->noise * (u[0] - .5));
      break;
    case F3D_IC_FARGO:
      r = Rv * (1.0 + d->noise * (u[0] - .5));
      x = Vv * (1. + d->aspectratio * d->noise * (u[1] - .5));
      y = d->CS[l2] * d->noise * (u[2] - .5);
      break;
    default: /* F3D_IC_LOGNORMAL */
      if (d->noise != 0.0) {
        double u1 = u[0], u2 = u[1];
        if (u1 < 1e-300) u1 = 1e-300;
        r *= exp(d->noise * (sqrt(-2.0 * log(u1)) * cos(2.0 * M_PI * u2)));
      }
      break;
  }
  *rho = r;
  *vx = x;
  *vy = y;
  *e = d->adiabatic ? d->E[l2] * r / (d->gamma - 1.0) : d->CS[l2];
}

int f3d_initial_conditions(const f3d_grid *g, const f3d_params *p, const f3d_clock *c,
                           double *dens, double *energy, double *vx, double *vy, double *vz) {
  return f3d_initial_conditions_fluid(g, p, c, F3D_GAS, dens, energy, vx, vy, vz);
}

int f3d_initial_conditions_fluid(const f3d_grid *g, const f3d_params *p, const f3d_clock *c, int fluidtype,
                                 double *dens, double *energy, double *vx, double *vy, double *vz) {
  size_t n = g->ncell, n2 = g->n2d;
  double *buf = (double *)malloc(4 * n2 * sizeof(double));
  if (!buf) return -1;
  f3d_ic_desc d;
  if (f3d_ic_profiles(g, p, c, fluidtype, &d, buf, buf + n2, buf + 2 * n2, buf + 3 * n2) != 0) {
    free(buf);
    return -1;
  }
  d.R = buf;
  d.V = buf + n2;
  d.CS = buf + 2 * n2;
  d.E = buf + 3 * n2;
  f3d_seed(p->seed);
  memset(dens, 0, n * sizeof(double));
  memset(energy, 0, n * sizeof(double));
  memset(vx, 0, n * sizeof(double));
  memset(vy, 0, n * sizeof(double));
  if (vz) memset(vz, 0, n * sizeof(double));
  /* cells in the set-ups' loop order: x fastest, then y, then z; the draws
     of one cell in the order its condinit.c makes them */
  for (int k = 0; k < g->sz; k++)
    for (int j = 0; j < g->sy; j++)
      for (int i = 0; i < g->nx; i++) {
        size_t l = (size_t)i + (size_t)j * g->pitch + (size_t)k * g->stride;
        double u[3] = {0.0, 0.0, 0.0};
        for (int q = 0; q < d.ndraw; q++) u[q] = drand48();
        f3d_ic_cell(&d, (size_t)j + (size_t)k * g->sy, u, dens + l, energy + l, vx + l, vy + l);
      }
  free(buf);
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
