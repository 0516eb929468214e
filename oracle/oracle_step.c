/*
 * oracle_step.c -- order of operations of one hydro step, CPU restatement.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Restates the DT loop of
 * src/main.c:361-448, Sources/Transport of src/algogas.c:30-159, transport()
 * of src/transport.c:69-193 and compute_potential of src/potential.c:10-49,
 * plus the initialisation sequence of src/main.c:176-316.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

static double *fld(size_t n) { return (double *)calloc(n ? n : 1, sizeof(double)); }

/* parse "name=value name=value" overrides (the -o flag, src/redefoptions.c) */
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

void orc_select_fluid(orc_state *S, int n) {
  orc_fluid *f = &S->fl[n];
  S->fluid_index = n;
  S->Fluidtype = f->type;
  S->Density = f->Density; S->Energy = f->Energy;
  S->Vx = f->Vx; S->Vy = f->Vy; S->Vz = f->Vz;
  S->Vx_temp = f->Vx_temp; S->Vy_temp = f->Vy_temp; S->Vz_temp = f->Vz_temp;
  S->Density0 = f->Density0; S->Energy0 = f->Energy0;
  S->Vx0 = f->Vx0; S->Vy0 = f->Vy0; S->Vz0 = f->Vz0;
}

orc_state *orc_create(const char *parfile, const char *overrides, char *err, int errlen) {
  orc_state *S = (orc_state *)calloc(1, sizeof(orc_state));
  f3d_params_defaults(&S->p);
  if (parfile && parfile[0] && f3d_params_read(&S->p, parfile, err, (size_t)errlen) != 0) {
    free(S);
    return NULL;
  }
  if (apply_overrides(&S->p, overrides, err, errlen) != 0) {
    free(S);
    return NULL;
  }
  f3d_params_finalize(&S->p);
  if (f3d_grid_init(&S->g, &S->p) != 0) {
    if (err) snprintf(err, errlen, "bad mesh");
    free(S);
    return NULL;
  }
  size_t n = S->g.ncell, n2 = S->g.n2d;
  S->nfluids = S->p.nfluids;
  if (S->nfluids > F3D_MAX_FLUIDS) {
    if (err) snprintf(err, errlen, "too many fluids");
    free(S);
    return NULL;
  }
  for (int f = 0; f < S->nfluids; f++) {
    orc_fluid *F = &S->fl[f];
    F->type = f == 0 ? F3D_GAS : F3D_DUST; /* setups/fargo_multifluid/condinit.c:58-76 */
    F->Density = fld(n); F->Energy = fld(n);
    F->Vx = fld(n); F->Vy = fld(n); F->Vz = fld(n);
    F->Vx_temp = fld(n); F->Vy_temp = fld(n); F->Vz_temp = fld(n);
    F->Density0 = fld(n2); F->Energy0 = fld(n2);
    F->Vx0 = fld(n2); F->Vy0 = fld(n2); F->Vz0 = fld(n2);
  }
  S->Mpx = fld(n); S->Mmx = fld(n); S->Mpy = fld(n); S->Mmy = fld(n); S->Mpz = fld(n); S->Mmz = fld(n);
  S->Pot = fld(n); S->Slope = fld(n); S->DivRho = fld(n); S->DensStar = fld(n); S->Qs = fld(n);
  S->Pressure = fld(n); S->Total_Density = fld(n); S->QL = fld(n); S->QR = fld(n);
  S->Reduction2D = fld(n2); S->VxMed = fld(n2);
  S->Nshift = (int *)calloc(n2, sizeof(int));
  S->Sdiffyczc = fld(n); S->Sdiffyfzc = fld(n);
  /* ColRate(INVSTOKESk, gas, k, feedback) (src/colrate.c:10-23) */
  for (int f = 1; f < S->nfluids; f++) {
    S->alpha[0 + f * S->nfluids] = S->p.invstokes[f];
    S->alpha[f + 0 * S->nfluids] = S->p.feedback ? S->p.invstokes[f] : 0.0;
  }
  S->alreadycalculated = -1;
  S->compute_disk_force = 1;
  f3d_bctab_init(&S->bct, &S->g, &S->p);
  f3d_damp_init(&S->damp, &S->p);
  /* src/main.c:238-242 */
  if (f3d_psys_init(&S->sys, &S->p, err, (size_t)errlen) != 0) {
    orc_destroy(S);
    return NULL;
  }
  S->clk.OMEGAFRAME = S->p.omegaframe;
  if (S->p.corotating) S->clk.OMEGAFRAME = f3d_psys_info(&S->sys, &S->p, F3D_PSYS_FREQUENCY);
  S->clk.OMEGAFRAME0 = S->clk.OMEGAFRAME;
  /* CondInit, FillGhosts, init_stockholm (src/main.c:265, 311, 315) */
  for (int f = 0; f < S->nfluids; f++) {
    orc_select_fluid(S, f);
    f3d_initial_conditions_fluid(&S->g, &S->p, &S->clk, S->Fluidtype, S->Density, S->Energy, S->Vx, S->Vy,
                                 S->g.dimz ? S->Vz : NULL);
  }
  for (int f = 0; f < S->nfluids; f++) {
    orc_select_fluid(S, f);
    orc_FillGhosts(S);
  }
  if (S->p.stockholm)
    for (int f = 0; f < S->nfluids; f++) {
      orc_select_fluid(S, f);
      orc_init_stockholm(S);
    }
  orc_select_fluid(S, 0);
  return S;
}

void orc_destroy(orc_state *S) {
  if (!S) return;
  for (int f = 0; f < S->nfluids; f++) {
    orc_fluid *F = &S->fl[f];
    double *a[] = {F->Density, F->Energy, F->Vx, F->Vy, F->Vz, F->Vx_temp, F->Vy_temp, F->Vz_temp,
                   F->Density0, F->Energy0, F->Vx0, F->Vy0, F->Vz0};
    for (size_t i = 0; i < sizeof(a) / sizeof(a[0]); i++) free(a[i]);
  }
  double *b[] = {S->Mpx, S->Mmx, S->Mpy, S->Mmy, S->Mpz, S->Mmz, S->Pot, S->Slope, S->DivRho,
                 S->DensStar, S->Qs, S->Pressure, S->Total_Density, S->QL, S->QR, S->Reduction2D, S->VxMed,
                 S->Sdiffyczc, S->Sdiffyfzc};
  for (size_t i = 0; i < sizeof(b) / sizeof(b[0]); i++) free(b[i]);
  free(S->Nshift);
  f3d_bctab_free(&S->bct);
  f3d_grid_free(&S->g);
  free(S);
}

/* src/planets.c:67-84 */
static void compute_accel(orc_state *S, double x, double y, double z, double rs, double m, double *acc) {
  orc_ComputeForce(S, x, y, z, rs, m);
  const double *lf = S->localforce;
  if (S->p.excludehill) {
    acc[0] = lf[3] + lf[9];
    acc[1] = lf[4] + lf[10];
    acc[2] = lf[5] + lf[11];
  } else {
    acc[0] = lf[0] + lf[6];
    acc[1] = lf[1] + lf[7];
    acc[2] = lf[2] + lf[8];
  }
}

/* src/potential.c:10-49 */
static void compute_potential(orc_state *S, double dt) {
  if (S->alreadycalculated == S->clk.Timestepcount) return;
  if (S->p.corotating) f3d_psys_info(&S->sys, &S->p, F3D_PSYS_MARK);
  int need_force = S->compute_disk_force;
  for (int k = 0; k < S->sys.nb; k++) need_force |= S->sys.feeldisk[k];
  if (need_force) compute_accel(S, 0.0, 0.0, 0.0, 0.0, 0.0, S->DiskOnPrimaryAcceleration);
  for (int c = 0; c < 3; c++) S->IndirectTerm[c] = S->p.indirectterm ? -S->DiskOnPrimaryAcceleration[c] : 0.0;
  orc_Potential(S);
  /* AdvanceSystemFromDisk (src/planets.c:86-113) */
  for (int k = 0; k < S->sys.nb; k++) {
    if (!S->sys.feeldisk[k]) continue;
    double m = S->sys.mass[k], x = S->sys.x[k], y = S->sys.y[k], z = S->sys.z[k];
    double r = sqrt(x * x + y * y + z * z), smoothing;
    if (S->p.rochesmoothing != 0)
      smoothing = r * pow(m / 3. / F3D_MSTAR, 1. / 3.) * S->p.rochesmoothing;
    else
      smoothing = S->p.aspectratio * pow(r / F3D_R0, S->p.flaringindex) * r * S->p.thicknesssmoothing;
    double gam[3];
    compute_accel(S, x, y, z, smoothing, m, gam);
    S->sys.vx[k] += dt * gam[0];
    S->sys.vy[k] += dt * gam[1];
    S->sys.vz[k] += dt * gam[2];
  }
  f3d_host_advance_planets(&S->sys, &S->p, &S->clk, dt);
  S->alreadycalculated = S->clk.Timestepcount;
}

/* src/algogas.c:30-133 */
void orc_Sources(orc_state *S, double dt) {
  if (S->p.eos == F3D_ADIABATIC)
    orc_ComputePressureFieldAd(S);
  else
    orc_ComputePressureFieldIso(S);
  if (S->p.potential) {
    compute_potential(S, dt);
    if (S->p.corotating) orc_CorrectVtheta(S, S->clk.Domega);
  }
  orc_SubStep1_x(S, dt);
  orc_SubStep1_y(S, dt);
  if (S->g.dimz) orc_SubStep1_z(S, dt);
  if (S->p.visc != F3D_VISC_NONE && S->Fluidtype == F3D_GAS) {
    if (S->g.geometry == F3D_SPHERICAL) {
      orc_visctensor_sph(S);
      orc_addviscosity_sph(S, dt);
    } else {
      orc_visctensor_cyl(S);
      orc_addviscosity_cyl(S, dt);
    }
  }
  orc_SubStep2_a(S, dt);
  orc_SubStep2_b(S, dt);
  if (S->p.eos == F3D_ADIABATIC && S->Fluidtype == F3D_GAS) orc_SubStep3(S, dt);
  orc_copy_velocities(S, 1);
  orc_FillGhosts(S);
  orc_copy_velocities(S, 2);
}

/* src/transport.c:170-200 */
static void TransportY(orc_state *S, double *Q, double dt) {
  if (Q != S->Density) {
    orc_DivideByRho(S, Q);
    orc_VanLeerY_a(S, S->DivRho);
    orc_VanLeerY_b(S, dt, S->DivRho, S->Qs);
    orc_UpdateY(S, dt, Q, S->Qs);
  } else {
    orc_UpdateDensityY(S, dt, Q);
  }
}

static void TransportZ(orc_state *S, double *Q, double dt) {
  if (Q != S->Density) {
    orc_DivideByRho(S, Q);
    orc_VanLeerZ_a(S, S->DivRho);
    orc_VanLeerZ_b(S, dt, S->DivRho, S->Qs);
    orc_UpdateZ(S, dt, Q, S->Qs);
  } else {
    orc_UpdateDensityZ(S, dt, Q);
  }
}

/* VanLeerX (src/transport.c:5-8) or VanLeerX_PPA (src/vanleerx_ppa.c:3-12) */
static void vanleerx(orc_state *S, int ppa, const double *Q, double *Qs, const double *Vx_t, double dt) {
  if (!ppa) {
    orc_VanLeerX_a(S, Q);
    orc_VanLeerX_b(S, dt, Q, Qs, Vx_t);
  } else {
    orc_VanLeerX_PPA_a(S, Q);
    orc_VanLeerX_PPA_b(S, Q);
    orc_VanLeerX_PPA_c(S, Q);
    orc_VanLeerX_PPA_d(S, dt, Q, Qs, Vx_t);
  }
}

/* TransportX and X_advection (src/transport.c:10-67) */
static void TransportX(orc_state *S, int ppa, double *Q, const double *Vx_t, double dt) {
  if (Q != S->Density) {
    orc_DivideByRho(S, Q);
    vanleerx(S, ppa, S->DivRho, S->Qs, Vx_t, dt);
    orc_UpdateX(S, dt, Q, S->Qs, Vx_t);
  } else {
    orc_UpdateDensityX(S, dt, Q, Vx_t);
  }
}

static void X_advection(orc_state *S, int ppa, const double *Vx_t, double dt) {
  vanleerx(S, ppa, S->Density, S->DensStar, Vx_t, dt);
  TransportX(S, ppa, S->Mpx, Vx_t, dt);
  TransportX(S, ppa, S->Mmx, Vx_t, dt);
  TransportX(S, ppa, S->Mpy, Vx_t, dt);
  TransportX(S, ppa, S->Mmy, Vx_t, dt);
  if (S->g.dimz) {
    TransportX(S, ppa, S->Mpz, Vx_t, dt);
    TransportX(S, ppa, S->Mmz, Vx_t, dt);
  }
  if (S->p.eos == F3D_ADIABATIC) TransportX(S, ppa, S->Energy, Vx_t, dt);
  TransportX(S, ppa, S->Density, Vx_t, dt);
}

/* transport() (src/transport.c:69-193) */
static void transport(orc_state *S, double dt) {
  orc_momenta_x(S);
  orc_momenta_y(S);
  if (S->g.dimz) orc_momenta_z(S);
  if (S->g.dimz && S->g.nz > 1) {
    orc_VanLeerZ_a(S, S->Density);
    orc_VanLeerZ_b(S, dt, S->Density, S->DensStar);
    TransportZ(S, S->Mpx, dt);
    TransportZ(S, S->Mmx, dt);
    TransportZ(S, S->Mpy, dt);
    TransportZ(S, S->Mmy, dt);
    TransportZ(S, S->Mpz, dt);
    TransportZ(S, S->Mmz, dt);
    if (S->p.eos == F3D_ADIABATIC) TransportZ(S, S->Energy, dt);
    TransportZ(S, S->Density, dt);
  }
  if (S->g.ny > 1) {
    orc_VanLeerY_a(S, S->Density);
    orc_VanLeerY_b(S, dt, S->Density, S->DensStar);
    TransportY(S, S->Mpx, dt);
    TransportY(S, S->Mmx, dt);
    TransportY(S, S->Mpy, dt);
    TransportY(S, S->Mmy, dt);
    if (S->g.dimz) {
      TransportY(S, S->Mpz, dt);
      TransportY(S, S->Mmz, dt);
    }
    if (S->p.eos == F3D_ADIABATIC) TransportY(S, S->Energy, dt);
    TransportY(S, S->Density, dt);
  }
  if (S->g.nx > 1) {
    if (!S->p.fargo) {
      X_advection(S, 0, S->Vx_temp, dt);
    } else {
      orc_ComputeResidual(S, dt);
      X_advection(S, 0, S->Vx, dt);
      X_advection(S, 1, S->Vx_temp, dt);
      orc_AdvectSHIFT(S, S->Mpx, S->Nshift);
      orc_AdvectSHIFT(S, S->Mmx, S->Nshift);
      orc_AdvectSHIFT(S, S->Mpy, S->Nshift);
      orc_AdvectSHIFT(S, S->Mmy, S->Nshift);
      if (S->g.dimz) {
        orc_AdvectSHIFT(S, S->Mpz, S->Nshift);
        orc_AdvectSHIFT(S, S->Mmz, S->Nshift);
      }
      if (S->p.eos == F3D_ADIABATIC) orc_AdvectSHIFT(S, S->Energy, S->Nshift);
      orc_AdvectSHIFT(S, S->Density, S->Nshift);
    }
  }
  orc_NewVelocity_x(S);
  orc_NewVelocity_y(S);
  if (S->g.dimz) orc_NewVelocity_z(S);
}

/* src/algogas.c:135-159 */
void orc_Transport(orc_state *S, double dt) {
  if (S->p.fargo) orc_ComputeVmed(S, S->Vx_temp);
  transport(S, dt);
}

/* one pass of the DT loop body (src/main.c:361-448) */
double orc_step(orc_state *S) {
  const double DT = S->p.dt_out;
  if (S->dtemp >= DT) S->dtemp = 0.0; /* a new DT interval begins (src/main.c:359) */
  if (S->p.fargo)
    for (int f = 0; f < S->nfluids; f++) {
      orc_select_fluid(S, f);
      orc_ComputeVmed(S, S->Vx);
    }
  for (int f = 0; f < S->nfluids; f++) {
    orc_select_fluid(S, f);
    orc_cfl(S);
  }
  if (orc_CflFluidsMin(S) != 0) return -1.0;
  double dt = S->StepTime;
  S->dtemp += dt;
  if (S->dtemp > DT) dt = DT - (S->dtemp - dt);
  orc_Reset_field(S, S->Total_Density);
  for (int f = 0; f < S->nfluids; f++) {
    orc_select_fluid(S, f);
    orc_ComputeTotalDensity(S);
  }
  for (int f = 0; f < S->nfluids; f++) {
    orc_select_fluid(S, f);
    orc_Sources(S, dt);
  }
  if (S->p.dragforce) orc_Collisions(S, dt); /* src/main.c:412-414 */
  if (S->p.dustdiffusion) orc_DustDiffusion_Main(S, dt); /* src/main.c:416-418 */
  for (int f = 0; f < S->nfluids; f++) {
    orc_select_fluid(S, f);
    orc_Transport(S, dt);
  }
  S->clk.PhysicalTime += dt;
  S->clk.Timestepcount++;
  if (S->p.stockholm)
    for (int f = 0; f < S->nfluids; f++) {
      orc_select_fluid(S, f);
      orc_StockholmBoundary(S, dt);
    }
  for (int f = 0; f < S->nfluids; f++) {
    orc_select_fluid(S, f);
    orc_FillGhosts(S);
  }
  orc_select_fluid(S, 0);
  S->dt_last = dt;
  return dt;
}

double *orc_field(orc_state *S, const char *name) {
  struct { const char *n; double *p; } t[] = {
      {"Density", S->Density}, {"Energy", S->Energy}, {"Vx", S->Vx}, {"Vy", S->Vy}, {"Vz", S->Vz},
      {"Vx_temp", S->Vx_temp}, {"Vy_temp", S->Vy_temp}, {"Vz_temp", S->Vz_temp},
      {"Density0", S->Density0}, {"Energy0", S->Energy0}, {"Vx0", S->Vx0}, {"Vy0", S->Vy0}, {"Vz0", S->Vz0},
      {"Mpx", S->Mpx}, {"Mmx", S->Mmx}, {"Mpy", S->Mpy}, {"Mmy", S->Mmy}, {"Mpz", S->Mpz}, {"Mmz", S->Mmz},
      {"Pot", S->Pot}, {"Slope", S->Slope}, {"DivRho", S->DivRho}, {"DensStar", S->DensStar},
      {"Qs", S->Qs}, {"Pressure", S->Pressure}, {"Total_Density", S->Total_Density}, {"QL", S->QL},
      {"QR", S->QR}, {"Reduction2D", S->Reduction2D}, {"VxMed", S->VxMed},
      {"Sdiffyczc", S->Sdiffyczc}, {"Sdiffyfzc", S->Sdiffyfzc}};
  for (size_t i = 0; i < sizeof(t) / sizeof(t[0]); i++)
    if (strcmp(t[i].n, name) == 0) return t[i].p;
  return NULL;
}

int *orc_nshift(orc_state *S) { return S->Nshift; }

void orc_get_scalars(orc_state *S, double *o) {
  o[0] = S->clk.PhysicalTime; o[1] = S->clk.OMEGAFRAME; o[2] = S->clk.OMEGAFRAME0;
  o[3] = S->clk.Domega; o[4] = S->dtemp; o[5] = S->StepTime; o[6] = S->clk.XAxisRotationAngle;
  o[7] = (double)S->clk.Timestepcount;
}

void orc_set_scalars(orc_state *S, const double *in) {
  S->clk.PhysicalTime = in[0]; S->clk.OMEGAFRAME = in[1]; S->clk.OMEGAFRAME0 = in[2];
  S->clk.Domega = in[3]; S->dtemp = in[4]; S->StepTime = in[5]; S->clk.XAxisRotationAngle = in[6];
  S->clk.Timestepcount = (long)in[7];
}

/* mesh shape: nx, ny, nz, dimz */
void orc_shape(orc_state *S, int *out) {
  out[0] = S->g.nx; out[1] = S->g.ny; out[2] = S->g.nz; out[3] = S->g.dimz;
}

/* The host-layer arrays the oracle was built with (mesh, boundary-factor
 * tables, damping-zone limits), so the tests can pin them against closed
 * forms; *len receives the element count, NULL for an unknown name. */
const double *orc_array(orc_state *S, const char *name, int *len) {
  const f3d_grid *g = &S->g;
  const f3d_bctab *b = &S->bct;
  int sx = g->sx, sy = g->sy, sz = g->sz;
  struct { const char *n; const double *p; int len; } t[] = {
      {"xmin", g->xmin, sx + 1}, {"xmed", g->xmed, sx}, {"ymin", g->ymin, sy + 1}, {"ymed", g->ymed, sy},
      {"zmin", g->zmin, sz + 1}, {"zmed", g->zmed, sz}, {"sxi", g->sxi, sx}, {"sxj", g->sxj, sy},
      {"syj", g->syj, sy}, {"szj", g->szj, sy}, {"sxk", g->sxk, sz}, {"syk", g->syk, sz}, {"szk", g->szk, sz},
      {"invvj", g->invvj, sy}, {"invdiffxmed", g->invdiffxmed, sx}, {"mindx", &g->mindx, 1},
      {"ymin_dens", b->ymin_dens, F3D_NGH}, {"ymax_dens", b->ymax_dens, F3D_NGH},
      {"ymin_vsq", b->ymin_vsq, F3D_NGH}, {"ymax_vsq", b->ymax_vsq, F3D_NGH},
      {"ymin_vfac", b->ymin_vfac, F3D_NGH}, {"ymax_vfac", b->ymax_vfac, F3D_NGH},
      {"zmin_vfac", b->zmin_vfac, F3D_NGH}, {"zmax_vfac", b->zmax_vfac, F3D_NGH},
      {"ymin_dexp", b->ymin_dexp, b->ymin_dexp ? F3D_NGH * sz : 0},
      {"ymax_dexp", b->ymax_dexp, b->ymax_dexp ? F3D_NGH * sz : 0},
      {"zmin_dens", b->zmin_dens, b->zmin_dens ? F3D_NGH * sy : 0},
      {"zmax_dens", b->zmax_dens, b->zmax_dens ? F3D_NGH * sy : 0},
      {"damp", &S->damp.Y_inf, 4}};
  for (size_t i = 0; i < sizeof(t) / sizeof(t[0]); i++)
    if (strcmp(t[i].n, name) == 0) {
      if (len) *len = t[i].len;
      return t[i].p;
    }
  if (len) *len = 0;
  return NULL;
}

/* the 12 sums _ComputeForce leaves (src/compute_force.c:160-215) */
void orc_get_localforce(orc_state *S, double *out12) { memcpy(out12, S->localforce, 12 * sizeof(double)); }

/* planet n's state: x y z vx vy vz mass */
void orc_set_planet(orc_state *S, int n, const double *in7) {
  if (n < 0 || n >= S->sys.nb) return;
  S->sys.x[n] = in7[0]; S->sys.y[n] = in7[1]; S->sys.z[n] = in7[2];
  S->sys.vx[n] = in7[3]; S->sys.vy[n] = in7[4]; S->sys.vz[n] = in7[5];
  S->sys.mass[n] = in7[6];
}

void orc_get_planets(orc_state *S, double *out, int maxn) {
  for (int i = 0; i < S->sys.nb && i < maxn; i++) {
    out[8 * i + 0] = S->sys.x[i]; out[8 * i + 1] = S->sys.y[i]; out[8 * i + 2] = S->sys.z[i];
    out[8 * i + 3] = S->sys.vx[i]; out[8 * i + 4] = S->sys.vy[i]; out[8 * i + 5] = S->sys.vz[i];
    out[8 * i + 6] = S->sys.mass[i]; out[8 * i + 7] = (double)S->sys.nb;
  }
}

/* Planet-only integration for the orbit tests: advance the planetary system
 * by time t in nsub RK5 substeps (src/runge_kutta.c), no disk, no frame change. */
void orc_psys_integrate(orc_state *S, double t, int nsub) {
  for (int s = 0; s < nsub; s++) f3d_psys_rk5(&S->sys, &S->p, &S->clk, t / nsub);
}
