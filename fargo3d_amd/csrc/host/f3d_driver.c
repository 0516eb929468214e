/*
 * f3d_driver.c -- host step driver of the MI355X build.
 *
 * The order of operations restates src/main.c:361-448, src/algogas.c:30-159
 * and src/transport.c:69-193; initialisation restates src/main.c:176-316.
 * Each routine is called through a function-pointer table (the reference's
 * src/global.h:303-416 table), bound by default to the per-routine GPU entry
 * points of include/fargo3d_hip.h.  The fused stage kernels replace whole
 * groups of routines when the corresponding "Fused*" switch is on.
 */
#include "f3d_driver.h"

#include <ctype.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <errno.h>

typedef struct {
  int fused_cfl, fused_sources, fused_transport;
} f3d_arch;

struct f3d_sim {
  f3d_params p;
  f3d_grid g;
  f3d_psys sys;
  f3d_clock clk;
  f3d_bctab bct;
  f3d_damp damp;
  f3d_hip_ctx *ctx;
  f3d_fields f;
  f3d_phys ph;
  f3d_arch arch;
  double dtemp, StepTime, dt_last;
  double DiskOnPrimaryAcceleration[3];
  long alreadycalculated;
  int prec32;    /* FP32 fields and the f3d32_fused_* stages */
  double dt_force; /* > 0: the next step's time step (measurement of two runs on one dt sequence) */
  size_t es;     /* bytes per element of the 3-D fields and the damping references */
  /* fluids (src/allocate.c:45 CreateFluid: each fluid owns its primitive fields and damping
     references; the scratch fields are shared) */
  int nfl, cur;
  int ftype[F3D_MAX_FLUIDS];
  double alpha[F3D_MAX_FLUIDS * F3D_MAX_FLUIDS]; /* ColRate matrix, [i + j*nfl] (src/colrate.c:10-23) */
  f3d_fields fl[F3D_MAX_FLUIDS];                 /* the per-fluid members of each fluid */
  long pot_step;                                 /* Timestepcount of the last Potential launch */
  double *monbuf;                                /* FP64 cell values of a monitored quantity */
  int mon_counter;                               /* MonCounter (src/monitor.c:14) */
  double *vb; /* precision form: the ring profile V0[j + k*sy] its azimuthal velocities deviate from */
  char err[512];
};

/* the members of f3d_fields each fluid owns */
#define F3D_FLUID_MEMBERS(X) \
  X(Density) X(Energy) X(Vx) X(Vy) X(Vz) X(Vx_temp) X(Vy_temp) X(Vz_temp) X(Density0) X(Energy0) X(Vx0) X(Vy0) X(Vz0)

/* SelectFluid (src/LowTasks.c:565): point the bound field set at fluid n */
static void select_fluid(f3d_sim *S, int n) {
#define SEL(m) S->f.m = S->fl[n].m;
  F3D_FLUID_MEMBERS(SEL)
#undef SEL
  S->cur = n;
  S->ph.fluidtype = S->ftype[n];
  f3d_hip_bind(S->ctx, &S->f);
  f3d_hip_set_phys(S->ctx, &S->ph);
}


#define ERR(S, ...)                                        \
  do {                                                     \
    snprintf((S)->err, sizeof((S)->err), __VA_ARGS__);     \
    return -1;                                             \
  } while (0)
#define CALL(S, expr)                                                                    \
  do {                                                                                   \
    if ((expr) != 0) ERR(S, "%s: %s", #expr, f3d_hip_last_error((S)->ctx));             \
  } while (0)

const char *f3d_sim_error(f3d_sim *S) { return S ? S->err : "no simulation"; }
int f3d_sim_nfluids(f3d_sim *S) { return S->nfl; }

int f3d_sim_select_fluid(f3d_sim *S, int n) {
  if (n < 0 || n >= S->nfl) ERR(S, "no fluid %d (%d fluids)", n, S->nfl);
  select_fluid(S, n);
  return 0;
}
f3d_hip_ctx *f3d_sim_ctx(f3d_sim *S) { return S->ctx; }
const f3d_params *f3d_sim_params(f3d_sim *S) { return &S->p; }
const f3d_psys *f3d_sim_psys(f3d_sim *S) { return &S->sys; }
const f3d_clock *f3d_sim_clock(f3d_sim *S) { return &S->clk; }

int f3d_sim_force_dt(f3d_sim *S, double dt) {
  S->dt_force = dt;
  return 0;
}

double f3d_sim_dt_out(f3d_sim *S) { return S->p.dt_out; }

int f3d_sim_eos(f3d_sim *S) { return S->p.eos; }

void f3d_sim_shape(f3d_sim *S, int out[4]) {
  out[0] = S->g.nx;
  out[1] = S->g.ny;
  out[2] = S->g.nz;
  out[3] = S->g.dimz;
}

static double *dalloc_bytes(size_t bytes, int *fail) {
  void *p = NULL;
  if (f3d_hip_malloc(&p, bytes) != 0) {
    *fail = 1;
    return NULL;
  }
  return (double *)p;
}
#define dalloc(n, fail) dalloc_bytes((n) * S->es, fail)

/* The precision form stores every fluid's azimuthal velocity (Vx, Vx_temp)
   and its damping reference (Vx0) as the FP32 deviation from the FP64 ring
   profile S->vb: 1 for such a 3-D array, 2 for the 2-D reference, 0 for
   every other array. */
static int vx_kind(f3d_sim *S, const void *dev) {
  if (!S->prec32 || !S->vb) return 0;
  for (int k = 0; k < S->nfl; k++) {
    if (dev == S->fl[k].Vx || dev == S->fl[k].Vx_temp) return 1;
    if (dev == S->fl[k].Vx0) return 2;
  }
  return 0;
}

/* the profile value of element idx of an array of that kind */
static double vb_at(const f3d_sim *S, size_t idx, int kind) {
  if (kind == 2) return S->vb[idx];
  size_t j = (idx / (size_t)S->g.pitch) % (size_t)S->g.sy, k = idx / (size_t)S->g.stride;
  return S->vb[j + k * (size_t)S->g.sy];
}

/* host buffer of n doubles -> device array of the simulation's element type */
static int upload(f3d_sim *S, void *dev, const double *h, size_t n) {
  if (!S->prec32) return f3d_hip_memcpy_h2d(S->ctx, dev, h, n * sizeof(double));
  float *t = (float *)malloc(n * sizeof(float));
  if (!t) return -1;
  const int kind = vx_kind(S, dev);
  for (size_t i = 0; i < n; i++) t[i] = (float)(kind ? h[i] - vb_at(S, i, kind) : h[i]);
  int rc = f3d_hip_memcpy_h2d(S->ctx, dev, t, n * sizeof(float));
  if (rc == 0) rc = f3d_hip_sync(S->ctx);
  free(t);
  return rc;
}

static int download(f3d_sim *S, double *h, const void *dev, size_t n) {
  if (!S->prec32) return f3d_hip_memcpy_d2h(S->ctx, h, dev, n * sizeof(double));
  float *t = (float *)malloc(n * sizeof(float));
  if (!t) return -1;
  int rc = f3d_hip_memcpy_d2h(S->ctx, t, dev, n * sizeof(float));
  if (rc == 0) rc = f3d_hip_sync(S->ctx);
  const int kind = vx_kind(S, dev);
  for (size_t i = 0; rc == 0 && i < n; i++) h[i] = kind ? (double)t[i] + vb_at(S, i, kind) : (double)t[i];
  free(t);
  return rc;
}

/* run parameters and host scalars as the kernels see them */
int f3d_sim_push_phys(f3d_sim *S) {
  f3d_phys *ph = &S->ph;
  memset(ph, 0, sizeof(*ph));
  ph->fluidtype = S->ftype[S->cur];
  ph->eos = S->p.eos;
  ph->visc = S->p.visc;
  ph->potential = S->p.potential;
  ph->fargo = S->p.fargo;
  ph->stockholm = S->p.stockholm;
  ph->indirectterm = S->p.indirectterm;
  ph->periodicz = S->p.periodicz;
  ph->periodicy = S->p.periodicy;
  ph->bc = S->p.bc;
  ph->energy_bc = S->p.energy_bc;
  ph->nu = S->p.nu;
  ph->alpha = S->p.alpha;
  ph->gamma = S->p.gamma;
  ph->cfl = S->p.cfl;
  ph->OMEGAFRAME = S->clk.OMEGAFRAME;
  ph->OMEGAFRAME0 = S->clk.OMEGAFRAME0;
  ph->PhysicalTime = S->clk.PhysicalTime;
  ph->ymin_par = S->p.ymin;
  ph->ymax_par = S->p.ymax;
  ph->zmin_par = S->p.zmin;
  ph->zmax_par = S->p.zmax;
  ph->taudamp = S->p.taudamp;
  ph->Y_inf = S->damp.Y_inf;
  ph->Y_sup = S->damp.Y_sup;
  ph->Z_inf = S->damp.Z_inf;
  ph->Z_sup = S->damp.Z_sup;
  for (int q = 0; q < 3; q++) {
    ph->ymin_dens[q] = S->bct.ymin_dens[q];
    ph->ymax_dens[q] = S->bct.ymax_dens[q];
    ph->ymin_vsq[q] = S->bct.ymin_vsq[q];
    ph->ymax_vsq[q] = S->bct.ymax_vsq[q];
    ph->ymin_vfac[q] = S->bct.ymin_vfac[q];
    ph->ymax_vfac[q] = S->bct.ymax_vfac[q];
    ph->zmin_vfac[q] = S->bct.zmin_vfac[q];
    ph->zmax_vfac[q] = S->bct.zmax_vfac[q];
  }
  ph->ymin_dexp = S->bct.ymin_dexp;
  ph->ymax_dexp = S->bct.ymax_dexp;
  ph->zmin_dens = S->bct.zmin_dens;
  ph->zmax_dens = S->bct.zmax_dens;
  /* planets at the current positions (Potential_cpu, src/potential.c:51-131) */
  double taper = f3d_planet_taper(&S->p, S->clk.PhysicalTime);
  ph->nplanets = S->sys.nb;
  for (int n = 0; n < S->sys.nb && n < 16; n++) {
    ph->px[n] = S->sys.x[n];
    ph->py[n] = S->sys.y[n];
    ph->pz[n] = S->sys.z[n];
    ph->pmass[n] = S->sys.mass[n] * taper;
    ph->psmooth2[n] = f3d_planet_smoothing2(&S->sys, &S->p, n, taper);
  }
  CALL(S, f3d_hip_set_phys(S->ctx, ph));
  return 0;
}

static int apply_overrides(f3d_params *p, const char *ov, char *err, int errlen) {
  if (!ov) return 0;
  char buf[4096];
  snprintf(buf, sizeof(buf), "%s", ov);
  for (char *tok = strtok(buf, " \t\n"); tok; tok = strtok(NULL, " \t\n")) {
    char *eq = strchr(tok, '=');
    if (!eq) continue;
    *eq = 0;
    if (f3d_params_set(p, tok, eq + 1) != 0) {
      if (err) snprintf(err, errlen, "unknown parameter %s", tok);
      return -1;
    }
  }
  return 0;
}

f3d_sim *f3d_sim_create(const char *parfile, const char *overrides, char *err, int errlen) {
  f3d_sim *S = (f3d_sim *)calloc(1, sizeof(f3d_sim));
  f3d_params_defaults(&S->p);
  S->arch.fused_cfl = S->arch.fused_sources = S->arch.fused_transport = 1;
  if (parfile && parfile[0] && f3d_params_read(&S->p, parfile, err, (size_t)errlen) != 0) goto fail;
  if (apply_overrides(&S->p, overrides, err, errlen) != 0) goto fail;
  f3d_params_finalize(&S->p);
  if (S->p.nsnap != 0) {
    /* NSNAP > 0 makes the reference write a SPECIFIC output every NSNAP
       steps and renumber TimeStep (src/main.c:347-353), which the main loop
       here does not do */
    if (err) snprintf(err, errlen, "NSNAP=%d: snapshots every NSNAP steps are not supported (use NSNAP 0)",
                      S->p.nsnap);
    goto fail;
  }
  if (f3d_grid_init(&S->g, &S->p) != 0) {
    if (err) snprintf(err, errlen, "bad mesh parameters");
    goto fail;
  }
  if (S->p.nfluids < 1 || S->p.nfluids > F3D_MAX_FLUIDS) {
    if (err) snprintf(err, errlen, "NFLUIDS=%d: 1 to %d fluids", S->p.nfluids, F3D_MAX_FLUIDS);
    goto fail;
  }
  S->nfl = S->p.nfluids;
  /* fluid 0 is the gas, the others dust (setups/fargo_multifluid/condinit.c:58-76) */
  for (int k = 0; k < S->nfl; k++) S->ftype[k] = k == 0 ? F3D_GAS : F3D_DUST;
  /* ColRate(INVSTOKESk, 0, k, feedback) (src/colrate.c:10-23) */
  for (int k = 1; k < S->nfl; k++) {
    S->alpha[0 + k * S->nfl] = S->p.invstokes[k];
    S->alpha[k + 0 * S->nfl] = S->p.feedback ? S->p.invstokes[k] : 0.0;
  }
  S->prec32 = (S->p.precision == 32);
  S->es = S->prec32 ? sizeof(float) : sizeof(double);
  f3d_bctab_init(&S->bct, &S->g, &S->p);
  f3d_damp_init(&S->damp, &S->p);
  if (f3d_psys_init(&S->sys, &S->p, err, (size_t)errlen) != 0) goto fail;
  if (S->prec32) {
    /* the FP32 form runs the fused stages only; the disk force on planets
       (FP64 per-routine reductions) is not part of it */
    for (int k = 0; k < S->sys.nb; k++)
      if (S->sys.feeldisk[k]) {
        if (err) snprintf(err, errlen, "the FP32 form has no disk force on planets (FeelDisk)");
        goto fail;
      }
  }
  S->clk.OMEGAFRAME = S->p.omegaframe;
  if (S->p.corotating) S->clk.OMEGAFRAME = f3d_psys_info(&S->sys, &S->p, F3D_PSYS_FREQUENCY);
  S->clk.OMEGAFRAME0 = S->clk.OMEGAFRAME;
  S->alreadycalculated = -1;
  S->pot_step = -1;
  {
    f3d_mesh_desc md;
    memset(&md, 0, sizeof(md));
    const f3d_grid *g = &S->g;
    md.nx = g->nx; md.ny = g->ny; md.nz = g->nz; md.nghy = g->nghy; md.nghz = g->nghz;
    md.dimz = g->dimz; md.geometry = g->geometry; md.pitch = g->pitch; md.stride = g->stride;
    md.XMIN = g->XMIN; md.XMAX = g->XMAX; md.mindx = g->mindx;
    md.xmin = g->xmin; md.ymin = g->ymin; md.zmin = g->zmin;
    md.xmed = g->xmed; md.ymed = g->ymed; md.zmed = g->zmed;
    md.sxi = g->sxi; md.sxj = g->sxj; md.syj = g->syj; md.szj = g->szj;
    md.sxk = g->sxk; md.syk = g->syk; md.szk = g->szk; md.invvj = g->invvj; md.invdiffxmed = g->invdiffxmed;
    md.cosxmed = g->cosxmed; md.sinxmed = g->sinxmed; md.sinzmed = g->sinzmed; md.coszmed = g->coszmed;
    md.sinzmin = g->sinzmin; md.coszmin = g->coszmin;
    if (f3d_hip_create(&md, &S->ctx) != 0) {
      if (err) snprintf(err, errlen, "%s", f3d_hip_last_error(NULL));
      goto fail;
    }
  }
  {
    size_t n = S->g.ncell, n2 = S->g.n2d;
    int bad = 0;
    f3d_fields *f = &S->f;
    for (int k = 0; k < S->nfl; k++) {
      f3d_fields *F = &S->fl[k];
      F->Density = dalloc(n, &bad); F->Energy = dalloc(n, &bad);
      F->Vx = dalloc(n, &bad); F->Vy = dalloc(n, &bad);
      F->Vx_temp = dalloc(n, &bad); F->Vy_temp = dalloc(n, &bad);
      if (S->g.dimz) { F->Vz = dalloc(n, &bad); F->Vz_temp = dalloc(n, &bad); }
      F->Density0 = dalloc(n2, &bad); F->Energy0 = dalloc(n2, &bad); F->Vx0 = dalloc(n2, &bad);
      F->Vy0 = dalloc(n2, &bad); F->Vz0 = dalloc(n2, &bad);
    }
    f->Mpx = dalloc(n, &bad); f->Mmx = dalloc(n, &bad); f->Mpy = dalloc(n, &bad); f->Mmy = dalloc(n, &bad);
    if (S->g.dimz) { f->Mpz = dalloc(n, &bad); f->Mmz = dalloc(n, &bad); }
    /* the potential is FP64 in both forms (its differences balance the centrifugal terms) */
    f->Pot = dalloc_bytes(n * sizeof(double), &bad); f->Slope = dalloc(n, &bad); f->DivRho = dalloc(n, &bad);
    f->DensStar = dalloc(n, &bad); f->Qs = dalloc(n, &bad); f->Pressure = dalloc(n, &bad);
    f->Total_Density = dalloc(n, &bad); f->QL = dalloc(n, &bad); f->QR = dalloc(n, &bad);
    /* ring sums stay FP64 in both forms */
    f->Reduction2D = dalloc_bytes(n2 * sizeof(double), &bad); f->VxMed = dalloc_bytes(n2 * sizeof(double), &bad);
    {
      void *pn = NULL;
      if (f3d_hip_malloc(&pn, n2 * sizeof(int)) != 0) bad = 1;
      f->Nshift = (int *)pn;
    }
    if (bad) {
      if (err) snprintf(err, errlen, "device allocation failed");
      goto fail;
    }
    /* fields start at zero (src/allocate.c:140-158) */
    for (int k = 0; k < S->nfl; k++) {
      f3d_fields *F = &S->fl[k];
      double *own3[] = {F->Density, F->Energy, F->Vx, F->Vy, F->Vz, F->Vx_temp, F->Vy_temp, F->Vz_temp};
      for (size_t a = 0; a < sizeof(own3) / sizeof(own3[0]); a++)
        if (own3[a]) f3d_hip_memset(S->ctx, own3[a], 0, n * S->es);
      double *own2[] = {F->Density0, F->Energy0, F->Vx0, F->Vy0, F->Vz0};
      for (size_t a = 0; a < sizeof(own2) / sizeof(own2[0]); a++) f3d_hip_memset(S->ctx, own2[a], 0, n2 * S->es);
    }
    f3d_hip_memset(S->ctx, f->Pot, 0, n * sizeof(double));
    double **all[] = {&f->Mpx, &f->Mmx, &f->Mpy, &f->Mmy, &f->Mpz, &f->Mmz, &f->Slope, &f->DivRho,
                      &f->DensStar, &f->Qs, &f->Pressure, &f->Total_Density, &f->QL, &f->QR};
    for (size_t a = 0; a < sizeof(all) / sizeof(all[0]); a++)
      if (*all[a]) f3d_hip_memset(S->ctx, *all[a], 0, n * S->es);
    f3d_hip_memset(S->ctx, f->Reduction2D, 0, n2 * sizeof(double));
    f3d_hip_memset(S->ctx, f->VxMed, 0, n2 * sizeof(double));
    f3d_hip_memset(S->ctx, f->Nshift, 0, n2 * sizeof(int));
    select_fluid(S, 0);
  }
  {
    /* CondInit of every fluid (src/main.c:265): on the host and uploaded, or
       (CondInitOnDevice) on the GPU from the same per-ring profiles and the
       same drand48 stream */
    size_t n = S->g.ncell, n2 = S->g.n2d;
    double *prof = (double *)malloc(4 * n2 * sizeof(double));
    double *h = S->p.ic_device ? NULL : (double *)calloc(5 * n, sizeof(double));
    int bad = !prof || (!S->p.ic_device && !h);
    for (int k = 0; k < S->nfl && !bad; k++) {
      const f3d_fields *F = &S->fl[k];
      f3d_ic_desc d;
      if (f3d_ic_profiles(&S->g, &S->p, &S->clk, S->ftype[k], &d, prof, prof + n2, prof + 2 * n2, prof + 3 * n2)) {
        bad = 1;
        break;
      }
      d.R = prof;
      d.V = prof + n2;
      d.CS = prof + 2 * n2;
      d.E = prof + 3 * n2;
      if (k == 0 && S->prec32) {
        /* the profile the precision form's azimuthal velocities deviate
           from: the gas's initial V(j,k), FP64 */
        S->vb = (double *)malloc(n2 * sizeof(double));
        if (!S->vb || f3d_hip_set_vbase(S->ctx, d.V) != 0) {
          bad = 1;
          break;
        }
        memcpy(S->vb, d.V, n2 * sizeof(double));
      }
      if (S->p.ic_device) {
        bad = (S->prec32 ? f3d32_CondInit : f3d_CondInit)(S->ctx, &d, F->Density, F->Energy, F->Vx, F->Vy,
                                                           S->g.dimz ? F->Vz : NULL) != 0;
        continue;
      }
      f3d_initial_conditions_fluid(&S->g, &S->p, &S->clk, S->ftype[k], h, h + n, h + 2 * n, h + 3 * n,
                                   S->g.dimz ? h + 4 * n : NULL);
      bad = upload(S, F->Density, h, n) | upload(S, F->Energy, h + n, n) | upload(S, F->Vx, h + 2 * n, n) |
            upload(S, F->Vy, h + 3 * n, n);
      if (S->g.dimz) bad |= upload(S, F->Vz, h + 4 * n, n);
    }
    free(h);
    free(prof);
    if (bad) {
      if (err) snprintf(err, errlen, "the initial conditions failed%s%s", S->p.ic_device ? ": " : "",
                        S->p.ic_device ? f3d_hip_last_error(S->ctx) : "");
      goto fail;
    }
  }
  /* FillGhosts of every fluid, then init_stockholm of every fluid (src/main.c:311-316) */
  for (int pass = 0; pass < 2; pass++) {
    if (pass == 1 && !S->p.stockholm) break;
    for (int k = 0; k < S->nfl; k++) {
      select_fluid(S, k);
      if (f3d_sim_push_phys(S) != 0) {
        if (err) snprintf(err, errlen, "%s", S->err);
        goto fail;
      }
      int rc = pass == 0 ? (S->prec32 ? f3d32_fused_boundaries(S->ctx) : f3d_boundaries(S->ctx))
                         : (S->prec32 ? f3d32_fused_init_stockholm(S->ctx) : f3d_init_stockholm(S->ctx));
      if (rc != 0) {
        if (err) snprintf(err, errlen, "%s", f3d_hip_last_error(S->ctx));
        goto fail;
      }
    }
  }
  select_fluid(S, 0);
  if (f3d_sim_push_phys(S) != 0 || f3d_hip_sync(S->ctx) != 0) {
    if (err) snprintf(err, errlen, "%s", S->err[0] ? S->err : f3d_hip_last_error(S->ctx));
    goto fail;
  }
  return S;
fail:
  f3d_sim_destroy(S);
  return NULL;
}

void f3d_sim_destroy(f3d_sim *S) {
  if (!S) return;
  if (S->ctx) f3d_hip_sync(S->ctx);
  for (int k = 0; k < S->nfl; k++) {
#define FREE1(m) \
  if (S->fl[k].m) f3d_hip_free(S->fl[k].m);
    F3D_FLUID_MEMBERS(FREE1)
#undef FREE1
  }
  f3d_fields *f = &S->f;
  void *all[] = {f->Mpx, f->Mmx, f->Mpy, f->Mmy, f->Mpz, f->Mmz, f->Pot, f->Slope, f->DivRho, f->DensStar, f->Qs,
                 f->Pressure, f->Total_Density, f->QL, f->QR, f->Reduction2D, f->VxMed, f->Nshift};
  for (size_t a = 0; a < sizeof(all) / sizeof(all[0]); a++)
    if (all[a]) f3d_hip_free(all[a]);
  if (S->ctx) f3d_hip_destroy(S->ctx);
  if (S->monbuf) f3d_hip_free(S->monbuf);
  free(S->vb);
  f3d_bctab_free(&S->bct);
  f3d_grid_free(&S->g);
  free(S);
}

/* ---- func_arch table ---- */
static const char *GPU_GROUPS[] = {
    "computepressurefieldiso", "computepressurefieldad", "computetotaldensity", "substep1", "substep2",
    "substep3", "dividebyrho", "vanleer", "momenta", "update", "updatedensity", "newvelocity", "reduction",
    "advectshift", "computeresidual", "computevweight", "computevmed", "potential", "correctvtheta", "cfl",
    "computeforce", "vanleerppa", "copyvelocities", "viscoustensor", "addviscosity", "boundaries",
    "stockholmboundary", "resetfield", "communications", "checkmute", "setuphook", "copyfield",
    "changeframe", "fillghostsx", "monitor", NULL};

int f3d_sim_set_arch(f3d_sim *S, const char *group, const char *mode) {
  char g[64], m[16];
  snprintf(g, sizeof(g), "%s", group);
  snprintf(m, sizeof(m), "%s", mode);
  for (char *c = g; *c; c++) *c = (char)tolower((unsigned char)*c);
  for (char *c = m; *c; c++) *c = (char)tolower((unsigned char)*c);
  int on = (m[0] == 'y' || m[0] == '1' || m[0] == 'f' || m[0] == 't');
  if (strcmp(g, "fusedcfl") == 0) { S->arch.fused_cfl = on; return 0; }
  if (strcmp(g, "fusedsources") == 0) { S->arch.fused_sources = on; return 0; }
  if (strcmp(g, "fusedtransport") == 0) { S->arch.fused_transport = on; return 0; }
  for (int i = 0; GPU_GROUPS[i]; i++) {
    if (strcmp(g, GPU_GROUPS[i]) != 0) continue;
    if (m[0] == 'c') ERR(S, "group %s: no CPU implementation is built into this library", group);
    return 0; /* GPU / HIP: the per-routine kernels */
  }
  /* groups of subsystems outside the hydro step (MHD, RAM, collisions, ...) */
  return 0;
}

int f3d_sim_read_arch(f3d_sim *S, const char *path) {
  FILE *in = fopen(path, "r");
  if (!in) ERR(S, "cannot open %s", path);
  char s[512], name[128], val[128];
  while (fgets(s, sizeof(s), in)) {
    if (sscanf(s, "%127s", name) != 1 || name[0] == '#') continue;
    char *s1 = s + strlen(name);
    if (sscanf(s1 + strspn(s1, "\t :=>"), "%127s", val) != 1) continue;
    if (f3d_sim_set_arch(S, name, val) != 0) {
      fclose(in);
      return -1;
    }
  }
  fclose(in);
  return 0;
}

/* ---- the step ---- */

/* the Potential kernel of this step: the planets at their current positions
   (src/potential.c:32); it does not need dt, so the step can launch it ahead
   of the CFL read-back */
static int launch_potential(f3d_sim *S) {
  CALL(S, f3d_sim_push_phys(S));
  CALL(S, S->prec32 ? f3d32_fused_potential(S->ctx) : f3d_Potential(S->ctx));
  S->pot_step = S->clk.Timestepcount;
  return 0;
}

/* compute_potential (src/potential.c:10-49) */
static int compute_potential(f3d_sim *S, double dt) {
  if (S->alreadycalculated == S->clk.Timestepcount) return 0;
  if (S->p.corotating) f3d_psys_info(&S->sys, &S->p, F3D_PSYS_MARK);
  /* The disk's force on the star (ComputeAccel at the origin) feeds only the
     gas indirect term, which no supported setup enables; it is evaluated when
     a planet feels the disk (AdvanceSystemFromDisk), where it is consumed. */
  if (S->pot_step != S->clk.Timestepcount && launch_potential(S)) return -1;
  for (int k = 0; k < S->sys.nb; k++) {
    if (!S->sys.feeldisk[k]) continue;
    if (S->prec32) ERR(S, "the FP32 form has no disk force on planets (FeelDisk)");
    double lf[12];
    double m = S->sys.mass[k], x = S->sys.x[k], y = S->sys.y[k], z = S->sys.z[k];
    double r = sqrt(x * x + y * y + z * z), smoothing;
    if (S->p.rochesmoothing != 0)
      smoothing = r * pow(m / 3. / F3D_MSTAR, 1. / 3.) * S->p.rochesmoothing;
    else
      smoothing = S->p.aspectratio * pow(r / F3D_R0, S->p.flaringindex) * r * S->p.thicknesssmoothing;
    CALL(S, f3d_ComputeForce(S->ctx, x, y, z, smoothing, m, lf));
    double gam[3];
    if (S->p.excludehill) {
      gam[0] = lf[3] + lf[9]; gam[1] = lf[4] + lf[10]; gam[2] = lf[5] + lf[11];
    } else {
      gam[0] = lf[0] + lf[6]; gam[1] = lf[1] + lf[7]; gam[2] = lf[2] + lf[8];
    }
    S->sys.vx[k] += dt * gam[0];
    S->sys.vy[k] += dt * gam[1];
    S->sys.vz[k] += dt * gam[2];
  }
  f3d_host_advance_planets(&S->sys, &S->p, &S->clk, dt);
  S->alreadycalculated = S->clk.Timestepcount;
  return 0;
}

static int fill_ghosts(f3d_sim *S) {
  CALL(S, f3d_boundaries(S->ctx));
  return 0;
}

/* Sources (src/algogas.c:30-133), one kernel per routine */
static int sources_routines(f3d_sim *S, double dt) {
  f3d_hip_ctx *c = S->ctx;
  if (S->p.eos == F3D_ADIABATIC)
    CALL(S, f3d_ComputePressureFieldAd(c));
  else
    CALL(S, f3d_ComputePressureFieldIso(c));
  if (S->p.potential) {
    if (compute_potential(S, dt)) return -1;
    CALL(S, f3d_sim_push_phys(S)); /* OMEGAFRAME after the frame update */
    if (S->p.corotating) CALL(S, f3d_CorrectVtheta(c, S->clk.Domega));
  }
  CALL(S, f3d_SubStep1_x(c, dt));
  CALL(S, f3d_SubStep1_y(c, dt));
  if (S->g.dimz) CALL(S, f3d_SubStep1_z(c, dt));
  if (S->p.visc != F3D_VISC_NONE) {
    if (S->g.geometry == F3D_SPHERICAL) {
      CALL(S, f3d_visctensor_sph(c));
      CALL(S, f3d_addviscosity_sph(c, dt));
    } else {
      CALL(S, f3d_visctensor_cyl(c));
      CALL(S, f3d_addviscosity_cyl(c, dt));
    }
  }
  CALL(S, f3d_SubStep2_a(c, dt));
  CALL(S, f3d_SubStep2_b(c, dt));
  if (S->p.eos == F3D_ADIABATIC) CALL(S, f3d_SubStep3(c, dt));
  CALL(S, f3d_copy_velocities(c, 1));
  if (fill_ghosts(S)) return -1;
  CALL(S, f3d_copy_velocities(c, 2));
  return 0;
}

static int transport_y(f3d_sim *S, double *Q, double dt) {
  f3d_hip_ctx *c = S->ctx;
  if (Q != S->f.Density) {
    CALL(S, f3d_DivideByRho(c, Q));
    CALL(S, f3d_VanLeerY_a(c, S->f.DivRho));
    CALL(S, f3d_VanLeerY_b(c, dt, S->f.DivRho, S->f.Qs));
    CALL(S, f3d_UpdateY(c, dt, Q, S->f.Qs));
  } else {
    CALL(S, f3d_UpdateDensityY(c, dt, Q));
  }
  return 0;
}

static int transport_z(f3d_sim *S, double *Q, double dt) {
  f3d_hip_ctx *c = S->ctx;
  if (Q != S->f.Density) {
    CALL(S, f3d_DivideByRho(c, Q));
    CALL(S, f3d_VanLeerZ_a(c, S->f.DivRho));
    CALL(S, f3d_VanLeerZ_b(c, dt, S->f.DivRho, S->f.Qs));
    CALL(S, f3d_UpdateZ(c, dt, Q, S->f.Qs));
  } else {
    CALL(S, f3d_UpdateDensityZ(c, dt, Q));
  }
  return 0;
}

static int vanleerx(f3d_sim *S, int ppa, const double *Q, double *Qs, const double *Vx_t, double dt) {
  f3d_hip_ctx *c = S->ctx;
  if (!ppa) {
    CALL(S, f3d_VanLeerX_a(c, Q));
    CALL(S, f3d_VanLeerX_b(c, dt, Q, Qs, Vx_t));
  } else {
    CALL(S, f3d_VanLeerX_PPA_a(c, Q));
    CALL(S, f3d_VanLeerX_PPA_b(c, Q));
    CALL(S, f3d_VanLeerX_PPA_c(c, Q));
    CALL(S, f3d_VanLeerX_PPA_d(c, dt, Q, Qs, Vx_t));
  }
  return 0;
}

static int transport_x(f3d_sim *S, int ppa, double *Q, const double *Vx_t, double dt) {
  f3d_hip_ctx *c = S->ctx;
  if (Q != S->f.Density) {
    CALL(S, f3d_DivideByRho(c, Q));
    if (vanleerx(S, ppa, S->f.DivRho, S->f.Qs, Vx_t, dt)) return -1;
    CALL(S, f3d_UpdateX(c, dt, Q, S->f.Qs, Vx_t));
  } else {
    CALL(S, f3d_UpdateDensityX(c, dt, Q, Vx_t));
  }
  return 0;
}

static int x_advection(f3d_sim *S, int ppa, const double *Vx_t, double dt) {
  f3d_fields *f = &S->f;
  if (vanleerx(S, ppa, f->Density, f->DensStar, Vx_t, dt)) return -1;
  double *qs[8];
  int n = 0;
  qs[n++] = f->Mpx; qs[n++] = f->Mmx; qs[n++] = f->Mpy; qs[n++] = f->Mmy;
  if (S->g.dimz) { qs[n++] = f->Mpz; qs[n++] = f->Mmz; }
  if (S->p.eos == F3D_ADIABATIC) qs[n++] = f->Energy;
  qs[n++] = f->Density;
  for (int q = 0; q < n; q++)
    if (transport_x(S, ppa, qs[q], Vx_t, dt)) return -1;
  return 0;
}

/* transport() (src/transport.c:69-193), one kernel per routine */
static int transport_routines(f3d_sim *S, double dt) {
  f3d_hip_ctx *c = S->ctx;
  f3d_fields *f = &S->f;
  if (S->p.fargo) CALL(S, f3d_ComputeVmed(c, f->Vx_temp));
  CALL(S, f3d_momenta_x(c));
  CALL(S, f3d_momenta_y(c));
  if (S->g.dimz) CALL(S, f3d_momenta_z(c));
  double *qs[8];
  int n = 0;
  qs[n++] = f->Mpx; qs[n++] = f->Mmx; qs[n++] = f->Mpy; qs[n++] = f->Mmy;
  if (S->g.dimz) { qs[n++] = f->Mpz; qs[n++] = f->Mmz; }
  if (S->p.eos == F3D_ADIABATIC) qs[n++] = f->Energy;
  qs[n++] = f->Density;
  if (S->g.dimz && S->g.nz > 1) {
    CALL(S, f3d_VanLeerZ_a(c, f->Density));
    CALL(S, f3d_VanLeerZ_b(c, dt, f->Density, f->DensStar));
    for (int q = 0; q < n; q++)
      if (transport_z(S, qs[q], dt)) return -1;
  }
  if (S->g.ny > 1) {
    CALL(S, f3d_VanLeerY_a(c, f->Density));
    CALL(S, f3d_VanLeerY_b(c, dt, f->Density, f->DensStar));
    for (int q = 0; q < n; q++)
      if (transport_y(S, qs[q], dt)) return -1;
  }
  if (S->g.nx > 1) {
    if (!S->p.fargo) {
      if (x_advection(S, 0, f->Vx_temp, dt)) return -1;
    } else {
      CALL(S, f3d_ComputeResidual(c, dt));
      if (x_advection(S, 0, f->Vx, dt)) return -1;
      if (x_advection(S, 1, f->Vx_temp, dt)) return -1;
      for (int q = 0; q < n; q++) CALL(S, f3d_AdvectSHIFT(c, qs[q], f->Nshift));
    }
  }
  CALL(S, f3d_NewVelocity_x(c));
  CALL(S, f3d_NewVelocity_y(c));
  if (S->g.dimz) CALL(S, f3d_NewVelocity_z(c));
  return 0;
}

/* Collisions(dt, 1) and DustDiffusion_Main(dt) (src/main.c:412-418) */
static int fluid_coupling(f3d_sim *S, double dt) {
  if (S->nfl < 2) return 0;
  f3d_hip_ctx *c = S->ctx;
  double *rho[F3D_MAX_FLUIDS], *vxt[F3D_MAX_FLUIDS], *vyt[F3D_MAX_FLUIDS], *vzt[F3D_MAX_FLUIDS];
  for (int k = 0; k < S->nfl; k++) {
    rho[k] = S->fl[k].Density;
    vxt[k] = S->fl[k].Vx_temp;
    vyt[k] = S->fl[k].Vy_temp;
    vzt[k] = S->fl[k].Vz_temp;
  }
  if (S->p.dragforce)
    CALL(S, S->prec32 ? f3d32_fused_collisions(c, dt, S->nfl, rho, vxt, vyt, vzt, S->alpha)
                      : f3d_fused_collisions(c, dt, S->nfl, rho, vxt, vyt, vzt, S->alpha));
  if (S->p.dustdiffusion)
    CALL(S, S->prec32 ? f3d32_fused_dust_diffusion(c, dt, S->nfl, rho, S->ftype, S->fl[0].Energy)
                      : f3d_fused_dust_diffusion(c, dt, S->nfl, rho, S->ftype, S->fl[0].Energy));
  return 0;
}

int f3d_sim_step(f3d_sim *S, double *dt_out) {
  f3d_hip_ctx *c = S->ctx;
  const double DT = S->p.dt_out;
  const int nfl = S->nfl;
  if (S->dtemp >= DT) S->dtemp = 0.0;
  if (S->prec32 && !(S->arch.fused_cfl && S->arch.fused_sources && S->arch.fused_transport))
    ERR(S, "the FP32 form runs the fused stages only");
  int feel = 0;
  for (int k = 0; k < S->sys.nb; k++) feel |= S->sys.feeldisk[k];
  /* the Potential kernel needs no dt: it is queued behind the CFL and runs
     while the host waits for and handles dt */
  const int pot_early =
      S->arch.fused_sources && S->p.potential && !feel && S->alreadycalculated != S->clk.Timestepcount;
  /* MULTIFLUID(ComputeVmed(Vx)); MULTIFLUID(cfl()) (src/main.c:366-381): every
     fluid's CFL reads the VxMed of the last fluid */
  double step = 1e30;
  if (!S->arch.fused_cfl && S->p.fargo)
    for (int k = 0; k < nfl; k++) {
      select_fluid(S, k);
      CALL(S, f3d_ComputeVmed(c, S->f.Vx));
    }
  if (S->arch.fused_cfl || S->prec32) {
    /* every fluid's cfl into one device MIN, one read-back */
    double *cs[F3D_MAX_FLUIDS], *rho[F3D_MAX_FLUIDS], *vx[F3D_MAX_FLUIDS], *vy[F3D_MAX_FLUIDS], *vz[F3D_MAX_FLUIDS];
    for (int k = 0; k < nfl; k++) {
      cs[k] = S->fl[k].Energy;
      rho[k] = S->fl[k].Density;
      vx[k] = S->fl[k].Vx;
      vy[k] = S->fl[k].Vy;
      vz[k] = S->fl[k].Vz;
    }
    if (f3d_sim_push_phys(S)) return -1;
    double fluids_min;
    if (S->prec32)
      CALL(S, f3d32_fused_cfl_fluids_async(c, nfl, cs, rho, vx, vy, vz, S->fl[nfl - 1].Vx));
    else
      CALL(S, f3d_fused_cfl_fluids_async(c, nfl, cs, rho, vx, vy, vz, S->fl[nfl - 1].Vx));
    if (pot_early && launch_potential(S)) return -1;
    CALL(S, S->prec32 ? f3d32_fused_cfl_wait(c, &fluids_min) : f3d_fused_cfl_wait(c, &fluids_min));
    step = fluids_min;
  } else {
    if (pot_early && launch_potential(S)) return -1;
    for (int k = 0; k < nfl; k++) {
      select_fluid(S, k);
      if (f3d_sim_push_phys(S)) return -1;
      double fluid_min;
      CALL(S, f3d_cfl(c, &fluid_min));
      /* CflFluidsMin (src/cfl_fluids_min.c:3-22) */
      if (step > fluid_min) step = fluid_min;
    }
  }
  if (S->dt_force > 0.0) {
    step = S->dt_force;
    S->dt_force = 0.0;
  }
  S->StepTime = step;
  if (S->StepTime <= (1e-10 * sqrt(F3D_R0 * F3D_R0 * F3D_R0 / (F3D_G * F3D_MSTAR))))
    ERR(S, "Error!!!--> Null dt");
  double dt = S->StepTime;
  S->dtemp += dt;
  if (S->dtemp > DT) dt = DT - (S->dtemp - dt);
  /* Total_Density (src/main.c:399-403) feeds only the disk force; the fused
     path skips it when no planet feels the disk */
  if (!S->prec32 && (!S->arch.fused_sources || (S->p.potential && feel))) {
    CALL(S, f3d_Reset_field(c, S->f.Total_Density));
    for (int k = 0; k < nfl; k++) {
      select_fluid(S, k);
      CALL(S, f3d_ComputeTotalDensity(c));
    }
  }
  /* MULTIFLUID(Sources(dt)) (src/main.c:410): isothermal runs with at most
     one viscous fluid take every fluid in one pass */
  int nvisc = 0;
  for (int k = 0; k < nfl; k++) nvisc += (S->p.visc != 0 && S->ftype[k] == F3D_GAS);
  const int srcs_all = nfl > 1 && S->arch.fused_sources && S->p.eos != F3D_ADIABATIC && nvisc <= 1;
  if (srcs_all) {
    select_fluid(S, 0);
    if (S->p.potential && compute_potential(S, dt)) return -1;
    if (f3d_sim_push_phys(S)) return -1;
    f3d_fields all[F3D_MAX_FLUIDS];
    for (int k = 0; k < nfl; k++) {
      all[k] = S->f;
#define SELS(m) all[k].m = S->fl[k].m;
      F3D_FLUID_MEMBERS(SELS)
#undef SELS
    }
    const double dom = (S->p.potential && S->p.corotating) ? S->clk.Domega : 0.0;
    CALL(S, S->prec32 ? f3d32_fused_sources_fluids(c, dt, dom, nfl, all, S->ftype)
                      : f3d_fused_sources_fluids(c, dt, dom, nfl, all, S->ftype));
  }
  for (int k = 0; k < nfl && !srcs_all; k++) {
    select_fluid(S, k);
    if (S->arch.fused_sources) {
      if (S->p.potential) {
        if (compute_potential(S, dt)) return -1;
      }
      if (f3d_sim_push_phys(S)) return -1;
      const double dom = (S->p.potential && S->p.corotating) ? S->clk.Domega : 0.0;
      CALL(S, S->prec32 ? f3d32_fused_sources(c, dt, dom) : f3d_fused_sources(c, dt, dom));
    } else {
      if (f3d_sim_push_phys(S)) return -1;
      if (sources_routines(S, dt)) return -1;
    }
  }
  if (fluid_coupling(S, dt)) return -1;
  /* MULTIFLUID(Transport(dt)), then MULTIFLUID(StockholmBoundary(dt)) and
     MULTIFLUID(FillGhosts) (src/main.c:420-434); the fused transport ends with
     the damping and the ghost fill of its own fluid */
  if (S->arch.fused_transport && nfl > 1) {
    /* every fluid's sweeps, one X-sweep launch over all fluids */
    f3d_fields all[F3D_MAX_FLUIDS];
    for (int k = 0; k < nfl; k++) {
      all[k] = S->f;
#define SEL(m) all[k].m = S->fl[k].m;
      F3D_FLUID_MEMBERS(SEL)
#undef SEL
    }
    if (f3d_sim_push_phys(S)) return -1;
    CALL(S, S->prec32 ? f3d32_fused_transport_fluids(c, dt, nfl, all) : f3d_fused_transport_fluids(c, dt, nfl, all));
  } else {
    for (int k = 0; k < nfl; k++) {
      select_fluid(S, k);
      if (S->arch.fused_transport) {
        CALL(S, S->prec32 ? f3d32_fused_transport(c, dt, 1) : f3d_fused_transport(c, dt, 1));
      } else {
        if (transport_routines(S, dt)) return -1;
      }
    }
  }
  S->clk.PhysicalTime += dt;
  S->clk.Timestepcount++;
  if (!S->arch.fused_transport) {
    if (f3d_sim_push_phys(S)) return -1;
    for (int k = 0; k < nfl; k++) {
      select_fluid(S, k);
      if (S->p.stockholm) CALL(S, f3d_StockholmBoundary(c, dt));
    }
    for (int k = 0; k < nfl; k++) {
      select_fluid(S, k);
      if (fill_ghosts(S)) return -1;
    }
  }
  select_fluid(S, 0);
  S->dt_last = dt;
  if (dt_out) *dt_out = dt;
  return 0;
}

int f3d_sim_run(f3d_sim *S, int nsteps) {
  for (int s = 0; s < nsteps; s++)
    if (f3d_sim_step(S, NULL)) return -1;
  CALL(S, f3d_hip_sync(S->ctx));
  return 0;
}

/* ---- field access ---- */
/* is2d: 1 for YZ arrays; the element type is the simulation's except for the
   FP64 ring sums (Reduction2D, VxMed) and the FP64 potential */
static double *field_ptr(f3d_sim *S, const char *name, int *is2d) {
  f3d_fields *f = &S->f;
  struct { const char *n; double *p; int d2; } t[] = {
      {"Density", f->Density, 0}, {"Energy", f->Energy, 0}, {"Vx", f->Vx, 0}, {"Vy", f->Vy, 0},
      {"Vz", f->Vz, 0}, {"Vx_temp", f->Vx_temp, 0}, {"Vy_temp", f->Vy_temp, 0}, {"Vz_temp", f->Vz_temp, 0},
      {"Mpx", f->Mpx, 0}, {"Mmx", f->Mmx, 0}, {"Mpy", f->Mpy, 0}, {"Mmy", f->Mmy, 0}, {"Mpz", f->Mpz, 0},
      {"Mmz", f->Mmz, 0}, {"Pot", f->Pot, 0}, {"Slope", f->Slope, 0}, {"DivRho", f->DivRho, 0},
      {"DensStar", f->DensStar, 0}, {"Qs", f->Qs, 0}, {"Pressure", f->Pressure, 0},
      {"Total_Density", f->Total_Density, 0}, {"QL", f->QL, 0}, {"QR", f->QR, 0},
      {"Density0", f->Density0, 1}, {"Energy0", f->Energy0, 1}, {"Vx0", f->Vx0, 1}, {"Vy0", f->Vy0, 1},
      {"Vz0", f->Vz0, 1}, {"Reduction2D", f->Reduction2D, 1}, {"VxMed", f->VxMed, 1}};
  for (size_t i = 0; i < sizeof(t) / sizeof(t[0]); i++)
    if (strcmp(t[i].n, name) == 0) {
      if (is2d) *is2d = t[i].d2;
      return t[i].p;
    }
  return NULL;
}

double *f3d_sim_device_field(f3d_sim *S, const char *name) {
  if (strcmp(name, "Nshift") == 0) return (double *)S->f.Nshift;
  return field_ptr(S, name, NULL);
}

int f3d_sim_set_nshift(f3d_sim *S, const int *host) {
  CALL(S, f3d_hip_memcpy_h2d(S->ctx, S->f.Nshift, host, S->g.n2d * sizeof(int)));
  return 0;
}

int f3d_sim_set_planets(f3d_sim *S, const double *in, int n) {
  if (n != S->sys.nb) ERR(S, "planet count mismatch (%d vs %d)", n, S->sys.nb);
  for (int i = 0; i < n; i++) {
    S->sys.x[i] = in[8 * i + 0]; S->sys.y[i] = in[8 * i + 1]; S->sys.z[i] = in[8 * i + 2];
    S->sys.vx[i] = in[8 * i + 3]; S->sys.vy[i] = in[8 * i + 4]; S->sys.vz[i] = in[8 * i + 5];
    S->sys.mass[i] = in[8 * i + 6];
  }
  return f3d_sim_push_phys(S);
}

long f3d_sim_field_size(f3d_sim *S, const char *name) {
  int d2 = 0;
  if (!field_ptr(S, name, &d2)) return -1;
  return d2 ? (long)S->g.n2d : (long)S->g.ncell;
}

static int is_fp64_array(f3d_sim *S, const double *p) {
  return p == S->f.Reduction2D || p == S->f.VxMed || p == S->f.Pot;
}

int f3d_sim_get_field(f3d_sim *S, const char *name, double *host) {
  int d2 = 0;
  double *p = field_ptr(S, name, &d2);
  if (!p) ERR(S, "no field %s", name);
  size_t n = d2 ? S->g.n2d : S->g.ncell;
  if (is_fp64_array(S, p)) CALL(S, f3d_hip_memcpy_d2h(S->ctx, host, p, n * sizeof(double)));
  else CALL(S, download(S, host, p, n));
  return 0;
}

int f3d_sim_set_field(f3d_sim *S, const char *name, const double *host) {
  int d2 = 0;
  double *p = field_ptr(S, name, &d2);
  if (!p) ERR(S, "no field %s", name);
  size_t n = d2 ? S->g.n2d : S->g.ncell;
  if (is_fp64_array(S, p)) CALL(S, f3d_hip_memcpy_h2d(S->ctx, p, host, n * sizeof(double)));
  else CALL(S, upload(S, p, host, n));
  return 0;
}

int f3d_sim_get_nshift(f3d_sim *S, int *host) {
  CALL(S, f3d_hip_memcpy_d2h(S->ctx, host, S->f.Nshift, S->g.n2d * sizeof(int)));
  return 0;
}

void f3d_sim_get_scalars(f3d_sim *S, double o[8]) {
  o[0] = S->clk.PhysicalTime; o[1] = S->clk.OMEGAFRAME; o[2] = S->clk.OMEGAFRAME0;
  o[3] = S->clk.Domega; o[4] = S->dtemp; o[5] = S->StepTime; o[6] = S->clk.XAxisRotationAngle;
  o[7] = (double)S->clk.Timestepcount;
}

void f3d_sim_set_scalars(f3d_sim *S, const double in[8]) {
  S->clk.PhysicalTime = in[0]; S->clk.OMEGAFRAME = in[1]; S->clk.OMEGAFRAME0 = in[2];
  S->clk.Domega = in[3]; S->dtemp = in[4]; S->StepTime = in[5]; S->clk.XAxisRotationAngle = in[6];
  S->clk.Timestepcount = (long)in[7];
}

void f3d_sim_get_planets(f3d_sim *S, double *out, int maxn) {
  for (int i = 0; i < S->sys.nb && i < maxn; i++) {
    out[8 * i + 0] = S->sys.x[i]; out[8 * i + 1] = S->sys.y[i]; out[8 * i + 2] = S->sys.z[i];
    out[8 * i + 3] = S->sys.vx[i]; out[8 * i + 4] = S->sys.vy[i]; out[8 * i + 5] = S->sys.vz[i];
    out[8 * i + 6] = S->sys.mass[i]; out[8 * i + 7] = (double)S->sys.nb;
  }
}

/* ---- outputs and restart ---- */

/* WriteMerging's loop (src/output.c:283-315) on one host array: for each
   active k, each active j, the nx active values of the row; x fastest,
   native-endian doubles, no header */
int f3d_write_active(const double *h, int nx, int ny, int nz, int nghy, int nghz, long pitch, long stride,
                     const char *path) {
  FILE *out = fopen(path, "wb");
  if (!out) return -1;
  size_t rows = 0;
  for (int k = nghz; k < nz + nghz; k++)
    for (int j = nghy; j < ny + nghy; j++)
      rows += fwrite(h + (size_t)j * pitch + (size_t)k * stride, sizeof(double), (size_t)nx, out) == (size_t)nx;
  int rc = fclose(out);
  return (rows == (size_t)ny * nz && rc == 0) ? 0 : -1;
}

/* RestartDat with Restart_Full (src/LowTasks.c:1025-1040) on one host array:
   fills the active cells, leaves the ghosts; -1 if the file is missing or
   does not hold exactly nx*ny*nz doubles */
int f3d_read_active(double *h, int nx, int ny, int nz, int nghy, int nghz, long pitch, long stride,
                    const char *path) {
  FILE *in = fopen(path, "rb");
  if (!in) return -1;
  fseek(in, 0, SEEK_END);
  long sz = ftell(in);
  fseek(in, 0, SEEK_SET);
  int ok = sz == (long)sizeof(double) * nx * ny * nz;
  for (int k = nghz; ok && k < nz + nghz; k++)
    for (int j = nghy; ok && j < ny + nghy; j++)
      ok = fread(h + (size_t)j * pitch + (size_t)k * stride, sizeof(double), (size_t)nx, in) == (size_t)nx;
  fclose(in);
  return ok ? 0 : -1;
}

/* active-cell extent of the grid in the (nz, nghz) of WriteMerging: 2-D runs
   have one k row and no Z ghosts */
static void active_extent(const f3d_grid *g, int *nz, int *nghz) {
  *nz = g->dimz ? g->nz : 1;
  *nghz = g->dimz ? g->nghz : 0;
}

static int write_dev_field(f3d_sim *S, const double *dev, const char *path) {
  const f3d_grid *g = &S->g;
  size_t n = g->ncell;
  double *h = (double *)malloc(n * sizeof(double));
  if (!h) ERR(S, "out of host memory");
  if (download(S, h, dev, n) != 0) {
    free(h);
    ERR(S, "copy of a field for %s failed", path);
  }
  int nz, nghz;
  active_extent(g, &nz, &nghz);
  int rc = f3d_write_active(h, g->nx, g->ny, nz, g->nghy, nghz, g->pitch, g->stride, path);
  free(h);
  if (rc) ERR(S, "cannot write %s", path);
  return 0;
}

int f3d_sim_write_field(f3d_sim *S, const char *name, const char *path) {
  int d2 = 0;
  double *p = field_ptr(S, name, &d2);
  if (!p || d2) ERR(S, "no 3-D field %s", name);
  return write_dev_field(S, p, path);
}

/* fluid names of the set-ups: the gas is "gas" (setups/fargo/condinit.c:54),
   the dust fluids "dust1", "dust2", ... (setups/fargo_multifluid/condinit.c:54-69) */
static void fluid_name(int k, char *buf, size_t n) {
  if (k == 0) snprintf(buf, n, "gas");
  else snprintf(buf, n, "dust%d", k);
}

/* the primitive fields WriteOutputs(ALL) writes for fluid k, in its order
   (src/output.c:662-690): density, energy (not for dust), vx, vy, vz (3-D) */
static int output_fields(f3d_sim *S, int k, const char **suffix, double **dev) {
  const f3d_fields *F = &S->fl[k];
  int n = 0;
  suffix[n] = "dens"; dev[n++] = F->Density;
  if (S->ftype[k] != F3D_DUST) { suffix[n] = "energy"; dev[n++] = F->Energy; }
  suffix[n] = "vx"; dev[n++] = F->Vx;
  suffix[n] = "vy"; dev[n++] = F->Vy;
  if (S->g.dimz) { suffix[n] = "vz"; dev[n++] = F->Vz; }
  return n;
}

/* WritePlanetSystemFile's planet<k>.dat (big: bigplanet<k>.dat) line
   (src/output.c:46-75, 235-250) */
static int write_planet_files(f3d_sim *S, int n, const char *dir, int big) {
  char path[1024];
  for (int i = 0; i < S->sys.nb; i++) {
    snprintf(path, sizeof(path), big ? "%s/bigplanet%d.dat" : "%s/planet%d.dat", dir, i);
    FILE *out = fopen(path, "a");
    if (!out) ERR(S, "cannot write %s", path);
    fprintf(out, "%d\t%#.18g\t%#.18g\t%#.18g\t%#.18g\t%#.18g\t%#.18g\t%#.18g\t%#.18g\t%#.18g\n", n, S->sys.x[i],
            S->sys.y[i], S->sys.z[i], S->sys.vx[i], S->sys.vy[i], S->sys.vz[i], S->sys.mass[i],
            S->clk.PhysicalTime, S->clk.OMEGAFRAME);
    if (fclose(out) != 0) ERR(S, "cannot write %s", path);
  }
  return 0;
}

int f3d_sim_write_outputs(f3d_sim *S, int n, const char *dir) {
  if (write_planet_files(S, n, dir, 0)) return -1;
  char name[32], path[1024];
  for (int k = 0; k < S->nfl; k++) {
    const char *suffix[5];
    double *dev[5];
    int m = output_fields(S, k, suffix, dev);
    fluid_name(k, name, sizeof(name));
    for (int q = 0; q < m; q++) {
      snprintf(path, sizeof(path), "%s/%s%s%d.dat", dir, name, suffix[q], n);
      if (write_dev_field(S, dev[q], path)) return -1;
    }
  }
  return 0;
}

/* the damping references and the file names init_stockholm gives them
   (src/init_stockholm.c:30-50, 106-124): density, vx, vy, vz (3-D), energy
   (adiabatic) */
static int stockholm_arrays(f3d_sim *S, int k, const char **suffix, double **dev) {
  const f3d_fields *F = &S->fl[k];
  int n = 0;
  suffix[n] = "dens"; dev[n++] = F->Density0;
  suffix[n] = "vx"; dev[n++] = F->Vx0;
  suffix[n] = "vy"; dev[n++] = F->Vy0;
  if (S->g.dimz) { suffix[n] = "vz"; dev[n++] = F->Vz0; }
  if (S->p.eos == F3D_ADIABATIC) { suffix[n] = "energy"; dev[n++] = F->Energy0; }
  return n;
}

/* Write2D with GHOSTINC of one YZ array (src/output2d.c:3-59): on one rank
   every (j, k) of the array, ghosts and corners included, j fastest */
static int write_stockholm_files(f3d_sim *S, const char *dir) {
  const size_t n2 = S->g.n2d;
  double *h = (double *)malloc(n2 * sizeof(double));
  if (!h) ERR(S, "out of host memory");
  char name[32], path[1024];
  for (int k = 0; k < S->nfl; k++) {
    const char *suffix[5];
    double *dev[5];
    int m = stockholm_arrays(S, k, suffix, dev);
    fluid_name(k, name, sizeof(name));
    for (int q = 0; q < m; q++) {
      snprintf(path, sizeof(path), "%s/%s%s0_2d.dat", dir, name, suffix[q]);
      FILE *out = NULL;
      if (download(S, h, dev[q], n2) != 0 || !(out = fopen(path, "wb")) ||
          fwrite(h, sizeof(double), n2, out) != n2) {
        if (out) fclose(out);
        free(h);
        ERR(S, "cannot write %s", path);
      }
      if (fclose(out) != 0) {
        free(h);
        ERR(S, "cannot write %s", path);
      }
    }
  }
  free(h);
  return 0;
}

int f3d_sim_write_init(f3d_sim *S, const char *dir) {
  /* EmptyPlanetSystemFiles (src/output.c:33-44): planet<k>.dat only;
     bigplanet<k>.dat is appended to and never emptied */
  char path[1024];
  for (int i = 0; i < S->sys.nb; i++) {
    snprintf(path, sizeof(path), "%s/planet%d.dat", dir, i);
    FILE *out = fopen(path, "w");
    if (!out) ERR(S, "cannot write %s", path);
    if (fclose(out) != 0) ERR(S, "cannot write %s", path);
  }
  return S->p.stockholm ? write_stockholm_files(S, dir) : 0;
}

/* GetfromPlanetFile (src/output.c:77-110): the values of the line of
   planet<k>.dat whose first column is n; v[0..8] = columns 2..10 */
static int planet_line(f3d_sim *S, int k, int n, const char *dir, double v[9]) {
  char path[1024], line[1024];
  snprintf(path, sizeof(path), "%s/planet%d.dat", dir, k);
  FILE *in = fopen(path, "r");
  if (!in) ERR(S, "cannot read %s", path);
  int found = 0;
  while (!found && fgets(line, sizeof(line), in)) {
    char *e;
    long t = strtol(line, &e, 10);
    if (e == line || t != n) continue;
    char *q = e;
    int c = 0;
    for (; c < 9; c++) {
      v[c] = strtod(q, &e);
      if (e == q) break;
      q = e;
    }
    found = c == 9;
  }
  fclose(in);
  if (!found) ERR(S, "no entry %d in %s", n, path);
  return 0;
}

int f3d_sim_restart(f3d_sim *S, int n, const char *dir) {
  /* RestartSimulation of every fluid (src/LowTasks.c:816-881), merged files */
  const f3d_grid *g = &S->g;
  const size_t nc = g->ncell, n2 = g->n2d;
  int nz, nghz;
  active_extent(g, &nz, &nghz);
  double *h = (double *)malloc((nc > n2 ? nc : n2) * sizeof(double));
  if (!h) ERR(S, "out of host memory");
  char name[32], path[1024];
  for (int k = 0; k < S->nfl; k++) {
    const char *suffix[5];
    double *dev[5];
    int m = output_fields(S, k, suffix, dev);
    fluid_name(k, name, sizeof(name));
    for (int q = 0; q < m; q++) {
      snprintf(path, sizeof(path), "%s/%s%s%d.dat", dir, name, suffix[q], n);
      /* the ghosts keep the initial condition's values until the fill below */
      if (download(S, h, dev[q], nc) != 0 ||
          f3d_read_active(h, g->nx, g->ny, nz, g->nghy, nghz, g->pitch, g->stride, path) != 0 ||
          upload(S, dev[q], h, nc) != 0) {
        free(h);
        ERR(S, "cannot restart from %s (missing, or not %d x %d x %d doubles)", path, g->nx, g->ny, nz);
      }
    }
  }
  /* the clock and the planets from planet0.dat (src/main.c:256-260) */
  if (S->sys.nb > 0) {
    double v[9];
    for (int i = 0; i < S->sys.nb; i++) {
      if (planet_line(S, i, n, dir, v)) {
        free(h);
        return -1;
      }
      if (i == 0) {
        S->clk.PhysicalTime = v[7];
        S->clk.OMEGAFRAME = v[8];
      }
      S->sys.x[i] = v[0]; S->sys.y[i] = v[1]; S->sys.z[i] = v[2];
      S->sys.vx[i] = v[3]; S->sys.vy[i] = v[4]; S->sys.vz[i] = v[5];
      S->sys.mass[i] = v[6];
    }
  }
  S->dtemp = 0.0;
  /* FillGhosts, then init_stockholm of every fluid (src/main.c:311-316): a
     restart reads each damping reference from its 0_2d file when the file is
     there with the full size, else takes it from the restarted fields
     (src/init_stockholm.c:30-105), and writes the files again */
  for (int pass = 0; pass < 2; pass++) {
    if (pass == 1 && !S->p.stockholm) break;
    for (int k = 0; k < S->nfl; k++) {
      select_fluid(S, k);
      if (f3d_sim_push_phys(S)) {
        free(h);
        return -1;
      }
      int rc = pass == 0 ? (S->prec32 ? f3d32_fused_boundaries(S->ctx) : f3d_boundaries(S->ctx))
                         : (S->prec32 ? f3d32_fused_init_stockholm(S->ctx) : f3d_init_stockholm(S->ctx));
      if (rc != 0) {
        free(h);
        ERR(S, "%s", f3d_hip_last_error(S->ctx));
      }
      if (pass == 0) continue;
      const char *suffix[5];
      double *dev[5];
      int m = stockholm_arrays(S, k, suffix, dev);
      fluid_name(k, name, sizeof(name));
      for (int q = 0; q < m; q++) {
        snprintf(path, sizeof(path), "%s/%s%s0_2d.dat", dir, name, suffix[q]);
        FILE *in = fopen(path, "rb");
        if (!in) continue;
        fseek(in, 0, SEEK_END);
        long sz = ftell(in);
        fseek(in, 0, SEEK_SET);
        int ok = sz == (long)(n2 * sizeof(double)) && fread(h, sizeof(double), n2, in) == n2;
        fclose(in);
        if (ok && upload(S, dev[q], h, n2) != 0) {
          free(h);
          ERR(S, "upload of %s failed", path);
        }
      }
    }
  }
  free(h);
  select_fluid(S, 0);
  if (f3d_sim_push_phys(S)) return -1;
  if (S->p.stockholm && write_stockholm_files(S, dir)) return -1;
  CALL(S, f3d_hip_sync(S->ctx));
  return n * S->p.ninterm;
}

/* one pass of the DT loop (src/main.c:361-448): steps until the output
   interval DT is reached; returns the number of steps */
int f3d_sim_advance(f3d_sim *S) {
  int steps = 0;
  do {
    if (f3d_sim_step(S, NULL)) return -1;
    steps++;
  } while (S->dtemp < S->p.dt_out);
  CALL(S, f3d_hip_sync(S->ctx));
  return steps;
}

/* ---- monitoring (src/monitor.c) ---- */

static int make_dir(f3d_sim *S, const char *path) {
  if (mkdir(path, 0755) != 0 && errno != EEXIST) ERR(S, "cannot create %s", path);
  return 0;
}

/* MonitorFunction (src/monitor.c:67-211) for the quantity in Reduction2D:
   the 2-D map (NOGHOSTINC), the raw Y profile and the scalar, each if the
   quantity is in that kind's mask */
static int monitor_function(f3d_sim *S, int bit, const char *mname, int plnb, int timestep, const char *fdir,
                            int m2d, int myraw, int mscalar) {
  const f3d_grid *g = &S->g;
  const int sy = g->ny + 2 * g->nghy;
  const int kb = g->dimz ? g->nghz : 0, ke = g->dimz ? g->nz + g->nghz : 1;
  const size_t n2 = g->n2d;
  char pl[32], path[1200];
  if (plnb < 0) pl[0] = 0;
  else snprintf(pl, sizeof(pl), "_planet_%d", plnb);
  double *R = (double *)malloc(n2 * sizeof(double));
  if (!R) ERR(S, "out of host memory");
  if (f3d_hip_memcpy_d2h(S->ctx, R, S->f.Reduction2D, n2 * sizeof(double)) != 0) {
    free(R);
    ERR(S, "copy of Reduction2D failed: %s", f3d_hip_last_error(S->ctx));
  }
  int rc = 0;
  FILE *out = NULL;
  if (m2d & bit) {
    /* Write2D NOGHOSTINC into monitor/<fluid>/FG<timestep>/ (src/output2d.c:3-59) */
    char fg[1100];
    snprintf(fg, sizeof(fg), "%s/FG%06d", fdir, timestep);
    if (make_dir(S, fg) == 0) {
      snprintf(path, sizeof(path), "%s/%s_2d_%07d%s.dat", fg, mname, S->mon_counter, pl);
      if ((out = fopen(path, "wb")) != NULL) {
        for (int k = kb; k < ke; k++) fwrite(R + (size_t)k * sy + g->nghy, sizeof(double), (size_t)g->ny, out);
        fclose(out);
      } else rc = -1;
    } else rc = -1;
  }
  if (rc == 0 && (myraw & bit)) {
    double *prof = (double *)calloc((size_t)g->ny, sizeof(double));
    for (int j = g->nghy; prof && j < g->ny + g->nghy; j++)
      for (int k = kb; k < ke; k++) prof[j - g->nghy] += R[(size_t)k * sy + j];
    snprintf(path, sizeof(path), "%s/%s_1d_Y_raw%s.dat", fdir, mname, pl);
    if (prof && (out = fopen(path, "ab")) != NULL) {
      fwrite(prof, sizeof(double), (size_t)g->ny, out);
      fclose(out);
    } else rc = -1;
    free(prof);
  }
  if (rc == 0 && (mscalar & bit)) {
    double lsum = 0.0;
    for (int k = kb; k < ke; k++)
      for (int j = g->nghy; j < g->ny + g->nghy; j++) lsum += R[(size_t)k * sy + j];
    snprintf(path, sizeof(path), "%s/%s%s.dat", fdir, mname, pl);
    if ((out = fopen(path, "a")) != NULL) {
      fprintf(out, "%.12g\t%.12g\n", S->clk.PhysicalTime, lsum);
      fclose(out);
    } else rc = -1;
  }
  free(R);
  if (rc) ERR(S, "cannot write %s", path);
  return 0;
}

int f3d_sim_monitor(f3d_sim *S, int m2d, int myraw, int mscalar, int timestep, const char *dir) {
  /* MonitorGlobal of every fluid (src/monitor.c:236-254): each monitored
     quantity in increasing bit order, per planet for the torque */
  static const struct { int bit; const char *name; int q; int perplanet; } Q[] = {
      {F3D_MON_MASS, "mass", 0, 0}, {F3D_MON_MOM_X, "momx", 1, 0}, {F3D_MON_MOM_Y, "momy", 2, 0},
      {F3D_MON_MOM_Z, "momz", 3, 0}, {F3D_MON_TORQ, "torq", 4, 1}};
  const int all = m2d | myraw | mscalar;
  if (all & ~(F3D_MON_MASS | F3D_MON_MOM_X | F3D_MON_MOM_Y | F3D_MON_MOM_Z | F3D_MON_TORQ))
    ERR(S, "monitor: only MASS, MOM_X, MOM_Y, MOM_Z and TORQ are monitored");
  if ((all & F3D_MON_MOM_Z) && !S->g.dimz) ERR(S, "monitor: MOM_Z needs a Z dimension");
  const f3d_grid *g = &S->g;
  if (!S->monbuf) {
    void *p = NULL;
    if (f3d_hip_malloc(&p, g->ncell * sizeof(double)) != 0) ERR(S, "monitor: device allocation failed");
    S->monbuf = (double *)p;
  }
  char mdir[1024], fdir[1100], name[32];
  snprintf(mdir, sizeof(mdir), "%s/monitor", dir);
  if (make_dir(S, mdir)) return -1;
  const int kb = g->dimz ? g->nghz : 0, ke = g->dimz ? g->nz + g->nghz : 1;
  for (int f = 0; f < S->nfl; f++) {
    select_fluid(S, f);
    if (f3d_sim_push_phys(S)) return -1;
    fluid_name(f, name, sizeof(name));
    snprintf(fdir, sizeof(fdir), "%s/%s", mdir, name);
    if (make_dir(S, fdir)) return -1;
    for (size_t a = 0; a < sizeof(Q) / sizeof(Q[0]); a++) {
      if (!(all & Q[a].bit)) continue;
      const int nb = Q[a].perplanet ? S->sys.nb : 1;
      for (int i = 0; i < nb; i++) {
        double xp = 0, yp = 0, zp = 0, rsm = 0;
        if (Q[a].perplanet) {
          /* src/mon_torque.c:15-16: the smoothing length of the planet */
          xp = S->sys.x[i]; yp = S->sys.y[i]; zp = S->sys.z[i];
          const double rp = sqrt(xp * xp + yp * yp + zp * zp);
          rsm = S->p.thicknesssmoothing * S->p.aspectratio * pow(rp / F3D_R0, S->p.flaringindex) * rp;
        }
        double *b = S->monbuf;
        int rc;
        switch (Q[a].q) {
          case 0: rc = S->prec32 ? f3d32_mon_dens(S->ctx, b) : f3d_mon_dens(S->ctx, b); break;
          case 1: rc = S->prec32 ? f3d32_mon_momx(S->ctx, b) : f3d_mon_momx(S->ctx, b); break;
          case 2: rc = S->prec32 ? f3d32_mon_momy(S->ctx, b) : f3d_mon_momy(S->ctx, b); break;
          case 3: rc = S->prec32 ? f3d32_mon_momz(S->ctx, b) : f3d_mon_momz(S->ctx, b); break;
          default:
            rc = S->prec32 ? f3d32_mon_torq(S->ctx, xp, yp, zp, rsm, b) : f3d_mon_torq(S->ctx, xp, yp, zp, rsm, b);
        }
        if (rc) ERR(S, "monitor %s: %s", Q[a].name, f3d_hip_last_error(S->ctx));
        CALL(S, f3d_reduction_SUM(S->ctx, b, g->nghy, g->ny + g->nghy, kb, ke));
        if (monitor_function(S, Q[a].bit, Q[a].name, Q[a].perplanet ? i : -1, timestep, fdir, m2d, myraw,
                             mscalar))
          return -1;
      }
    }
  }
  select_fluid(S, 0);
  if (f3d_sim_push_phys(S)) return -1;
  S->mon_counter++;
  return 0;
}

/* the main loop with its outputs (src/main.c:324-356): from begin_i to NTOT,
   the planet files and WriteOutputs(ALL) of every fluid every NINTERM passes */
int f3d_sim_main_loop(f3d_sim *S, int begin_i, const char *dir) {
  const int ninterm = S->p.ninterm > 0 ? S->p.ninterm : 1;
  for (int i = begin_i; i <= S->p.ntot; i++) {
    const int timestep = i / ninterm;
    if (i % ninterm == 0 && f3d_sim_write_outputs(S, timestep, dir)) return -1;
    if (i == S->p.ntot) break;
    if (f3d_sim_advance(S) < 0) return -1;
    /* after each DT pass: MonitorGlobal of every fluid with the set-up's
       monitors, then the bigplanet line (src/main.c:452-461) */
    if ((S->p.mon_2d | S->p.mon_y_raw | S->p.mon_scalar) &&
        f3d_sim_monitor(S, S->p.mon_2d, S->p.mon_y_raw, S->p.mon_scalar, timestep, dir))
      return -1;
    if (write_planet_files(S, timestep, dir, 1)) return -1;
  }
  return 0;
}
