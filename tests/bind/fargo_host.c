/*
 * fargo_host.c -- a miniature FARGO3D host for the binding test (TEST
 * INFRASTRUCTURE).  It owns the function-pointer table and the globals of
 * fargo3d_api.h, allocates every field on the host and the device (the
 * CreateField of src/allocate.c:115-198, without pitch padding), sets the
 * run up (src/main.c:176-316 for one fluid: parameters, mesh, planets,
 * CondInit, the first ghost fill, init_stockholm) and runs the DT loop body
 * (src/main.c:361-448, Sources and Transport of src/algogas.c:30-159,
 * transport of src/transport.c:69-193, compute_potential of
 * src/potential.c:10-49) calling every routine through the table, the way
 * the reference's main loop does after ChangeArch.  Host-side parts stay on
 * the host as in the reference: ComputeVmed's division of the ring sums
 * (src/compute_vmed.c:16-35), CflFluidsMin, the planets' RK5 and the frame.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "fargo3d_api.h"
#include "fargo3d_hip.h"

/* ---- the function-pointer table (src/global.h:303-416) ---- */
void (*ComputePressureFieldIso)();
void (*ComputePressureFieldAd)();
void (*SubStep1_x)(real);
void (*SubStep1_y)(real);
void (*SubStep1_z)(real);
void (*SubStep2_a)(real);
void (*SubStep2_b)(real);
void (*SubStep3)(real);
void (*DivideByRho)(Field *);
void (*VanLeerX_a)(Field *);
void (*VanLeerX_b)(real, Field *, Field *, Field *);
void (*VanLeerY_a)(Field *);
void (*VanLeerY_b)(real, Field *, Field *);
void (*VanLeerZ_a)(Field *);
void (*VanLeerZ_b)(real, Field *, Field *);
void (*momenta_x)();
void (*momenta_y)();
void (*momenta_z)();
void (*reduction_SUM)(Field *, int, int, int, int);
void (*reduction_MIN)(Field *, int, int, int, int);
void (*UpdateX)(real, Field *, Field *, Field *);
void (*UpdateY)(real, Field *, Field *);
void (*UpdateZ)(real, Field *, Field *);
void (*UpdateDensityX)(real, Field *, Field *);
void (*UpdateDensityY)(real, Field *);
void (*UpdateDensityZ)(real, Field *);
void (*NewVelocity_x)();
void (*NewVelocity_y)();
void (*NewVelocity_z)();
void (*AdvectSHIFT)(Field *, FieldInt2D *);
void (*ComputeResidual)(real);
void (*ComputeVweight)(Field *, Field *);
void (*Potential)();
void (*CorrectVtheta)(real);
void (*cfl)(void);
void (*_ComputeForce)(real, real, real, real, real);
void (*copy_velocities)(int);
void (*VanLeerX_PPA_a)(Field *);
void (*VanLeerX_PPA_b)(Field *);
void (*VanLeerX_PPA_c)(Field *);
void (*VanLeerX_PPA_d)(real, Field *, Field *, Field *);
void (*Reset_field)(Field *);
void (*ComputeTotalDensity)();
void (*StockholmBoundary)(real);
void (*visctensor_cyl)();
void (*addviscosity_cyl)(real);
void (*visctensor_sph)();
void (*addviscosity_sph)(real);
void (*boundary_ymin[1])();
void (*boundary_ymax[1])();
void (*boundary_zmin[1])();
void (*boundary_zmax[1])();

/* ---- globals ---- */
int Nx, Ny, Nz, NGHY, NGHZ, Pitch_gpu, Stride_gpu, Fluidtype = F3D_GAS, FluidIndex;
real XMIN, XMAX, Mindx, OMEGAFRAME, OMEGAFRAME0, PhysicalTime, Domega;
real *Xmin, *Ymin, *Zmin, *Xmed, *Ymed, *Zmed, *Sxi, *Sxj, *Syj, *Szj, *Sxk, *Syk, *Szk, *InvVj, *InvDiffXmed;
real Min[8];
Field *Density, *Energy, *Vx, *Vy, *Vz, *Vx_temp, *Vy_temp, *Vz_temp;
Field *Mpx, *Mmx, *Mpy, *Mmy, *Mpz, *Mmz, *Pot, *Slope, *DivRho, *DensStar, *Qs, *Pressure, *Total_Density, *QL, *QR;
Field2D *Density0, *Energy0, *Vx0, *Vy0, *Vz0, *Reduction2D, *VxMed;
FieldInt2D *Nshift;
real localforce[12];
f3d_params P;
f3d_grid GRID;
f3d_bctab BCT;
f3d_damp DAMP;
f3d_psys Sys;
static f3d_clock CLK;
static real StepTime, dtemp;
static long TimeStepCount, alreadycalculated = -1;

static Field *CreateField(const char *name, size_t n) {
  Field *f = (Field *)calloc(1, sizeof(Field));
  f->name = (char *)name;
  f->field_cpu = (real *)calloc(n, sizeof(real));
  if (f3d_hip_malloc((void **)&f->field_gpu, n * sizeof(real))) return NULL;
  return f;
}
static Field2D *CreateField2D(const char *name, size_t n) { return (Field2D *)CreateField(name, n); }

static void DestroyField(Field *f) {
  if (!f) return;
  free(f->field_cpu);
  f3d_hip_free(f->field_gpu);
  free(f);
}

static void to_gpu(Field *f, size_t n) { HipToDevice(f->field_gpu, f->field_cpu, n * sizeof(real)); }
static void to_cpu(Field *f, size_t n) { HipToHost(f->field_cpu, f->field_gpu, n * sizeof(real)); }

static void FillGhosts(void) {
  /* boundaries() (src/boundary.c:2-23) */
  if (GRID.dimz && !P.periodicz) {
    boundary_zmin[FluidIndex]();
    boundary_zmax[FluidIndex]();
  }
  if (!P.periodicy) {
    boundary_ymin[FluidIndex]();
    boundary_ymax[FluidIndex]();
  }
}

/* src/compute_vmed.c:10-39: the ring sums on the device, the division on
   the host (INPUT2D(Reduction2D), OUTPUT2D(VxMed)) */
static void ComputeVmed(Field *V) {
  ComputeVweight(V, Qs);
  reduction_SUM(Qs, 0, Ny + 2 * NGHY, 0, Nz + 2 * NGHZ);
  const size_t n2 = GRID.n2d;
  to_cpu((Field *)Reduction2D, n2);
  for (size_t l = 0; l < n2; l++) VxMed->field_cpu[l] = Reduction2D->field_cpu[l] / (XMAX - XMIN);
  to_gpu((Field *)VxMed, n2);
}

/* src/potential.c:10-49 (no planet feels the disk here) */
static void compute_potential(real dt) {
  if (alreadycalculated == TimeStepCount) return;
  if (P.corotating) f3d_psys_info(&Sys, &P, F3D_PSYS_MARK);
  Potential();
  CLK.OMEGAFRAME = OMEGAFRAME;
  CLK.PhysicalTime = PhysicalTime;
  f3d_host_advance_planets(&Sys, &P, &CLK, dt);
  OMEGAFRAME = CLK.OMEGAFRAME;
  Domega = CLK.Domega;
  alreadycalculated = TimeStepCount;
}

/* src/algogas.c:30-133 */
static void Sources(real dt) {
  if (P.eos == F3D_ADIABATIC) ComputePressureFieldAd();
  else ComputePressureFieldIso();
  if (P.potential) {
    compute_potential(dt);
    if (P.corotating) CorrectVtheta(Domega);
  }
  SubStep1_x(dt);
  SubStep1_y(dt);
  if (GRID.dimz) SubStep1_z(dt);
  if (P.visc != F3D_VISC_NONE) {
    if (GRID.geometry == F3D_SPHERICAL) {
      visctensor_sph();
      addviscosity_sph(dt);
    } else {
      visctensor_cyl();
      addviscosity_cyl(dt);
    }
  }
  SubStep2_a(dt);
  SubStep2_b(dt);
  if (P.eos == F3D_ADIABATIC) SubStep3(dt);
  copy_velocities(1); /* VTEMP2V */
  FillGhosts();
  copy_velocities(2); /* V2VTEMP */
}

/* src/transport.c:10-67 */
static void TransportY(Field *Q, real dt) {
  if (Q != Density) {
    DivideByRho(Q);
    VanLeerY_a(DivRho);
    VanLeerY_b(dt, DivRho, Qs);
    UpdateY(dt, Q, Qs);
  } else {
    UpdateDensityY(dt, Q);
  }
}
static void TransportZ(Field *Q, real dt) {
  if (Q != Density) {
    DivideByRho(Q);
    VanLeerZ_a(DivRho);
    VanLeerZ_b(dt, DivRho, Qs);
    UpdateZ(dt, Q, Qs);
  } else {
    UpdateDensityZ(dt, Q);
  }
}
static void VanLeerXs(int ppa, Field *Q, Field *Qs_, Field *Vx_t, real dt) {
  if (!ppa) {
    VanLeerX_a(Q);
    VanLeerX_b(dt, Q, Qs_, Vx_t);
  } else { /* VanLeerX_PPA (src/vanleerx_ppa.c:3-12) */
    VanLeerX_PPA_a(Q);
    VanLeerX_PPA_b(Q);
    VanLeerX_PPA_c(Q);
    VanLeerX_PPA_d(dt, Q, Qs_, Vx_t);
  }
}
static void TransportX(int ppa, Field *Q, Field *Vx_t, real dt) {
  if (Q != Density) {
    DivideByRho(Q);
    VanLeerXs(ppa, DivRho, Qs, Vx_t, dt);
    UpdateX(dt, Q, Qs, Vx_t);
  } else {
    UpdateDensityX(dt, Q, Vx_t);
  }
}
static void X_advection(int ppa, Field *Vx_t, real dt) {
  VanLeerXs(ppa, Density, DensStar, Vx_t, dt);
  TransportX(ppa, Mpx, Vx_t, dt);
  TransportX(ppa, Mmx, Vx_t, dt);
  TransportX(ppa, Mpy, Vx_t, dt);
  TransportX(ppa, Mmy, Vx_t, dt);
  if (GRID.dimz) {
    TransportX(ppa, Mpz, Vx_t, dt);
    TransportX(ppa, Mmz, Vx_t, dt);
  }
  if (P.eos == F3D_ADIABATIC) TransportX(ppa, Energy, Vx_t, dt);
  TransportX(ppa, Density, Vx_t, dt);
}

/* Transport (src/algogas.c:135-159) with transport (src/transport.c:69-193) */
static void Transport(real dt) {
  if (P.fargo) ComputeVmed(Vx_temp);
  momenta_x();
  momenta_y();
  if (GRID.dimz) momenta_z();
  if (GRID.dimz && Nz > 1) {
    VanLeerZ_a(Density);
    VanLeerZ_b(dt, Density, DensStar);
    Field *qz[] = {Mpx, Mmx, Mpy, Mmy, Mpz, Mmz};
    for (int q = 0; q < 6; q++) TransportZ(qz[q], dt);
    if (P.eos == F3D_ADIABATIC) TransportZ(Energy, dt);
    TransportZ(Density, dt);
  }
  VanLeerY_a(Density);
  VanLeerY_b(dt, Density, DensStar);
  TransportY(Mpx, dt);
  TransportY(Mmx, dt);
  TransportY(Mpy, dt);
  TransportY(Mmy, dt);
  if (GRID.dimz) {
    TransportY(Mpz, dt);
    TransportY(Mmz, dt);
  }
  if (P.eos == F3D_ADIABATIC) TransportY(Energy, dt);
  TransportY(Density, dt);
  if (!P.fargo) {
    X_advection(0, Vx_temp, dt);
  } else {
    ComputeResidual(dt);
    X_advection(0, Vx, dt);
    X_advection(1, Vx_temp, dt);
    Field *sh[] = {Mpx, Mmx, Mpy, Mmy, Mpz, Mmz};
    for (int q = 0; q < (GRID.dimz ? 6 : 4); q++) AdvectSHIFT(sh[q], Nshift);
    if (P.eos == F3D_ADIABATIC) AdvectSHIFT(Energy, Nshift);
    AdvectSHIFT(Density, Nshift);
  }
  NewVelocity_x();
  NewVelocity_y();
  if (GRID.dimz) NewVelocity_z();
}

/* ---- the test's interface ---- */

int bind_setup(const char *parfile, const char *overrides, char *err, int errlen) {
  f3d_params_defaults(&P);
  if (f3d_params_read(&P, parfile, err, (size_t)errlen)) return -1;
  char buf[1024];
  snprintf(buf, sizeof(buf), "%s", overrides ? overrides : "");
  for (char *t = strtok(buf, " "); t; t = strtok(NULL, " ")) {
    char *eq = strchr(t, '=');
    if (!eq) continue;
    *eq = 0;
    if (f3d_params_set(&P, t, eq + 1)) {
      snprintf(err, errlen, "unknown parameter %s", t);
      return -1;
    }
  }
  f3d_params_finalize(&P);
  if (P.nfluids != 1 || P.precision != 64) {
    snprintf(err, errlen, "the binding test runs one FP64 fluid");
    return -1;
  }
  if (f3d_grid_init(&GRID, &P)) return -1;
  /* src/global.h sizes and mesh arrays (src/allocate.c:201-213) */
  Nx = GRID.nx; Ny = GRID.ny; Nz = GRID.nz; NGHY = GRID.nghy; NGHZ = GRID.nghz;
  Pitch_gpu = GRID.pitch; Stride_gpu = GRID.stride;
  XMIN = GRID.XMIN; XMAX = GRID.XMAX; Mindx = GRID.mindx;
  Xmin = GRID.xmin; Ymin = GRID.ymin; Zmin = GRID.zmin; Xmed = GRID.xmed; Ymed = GRID.ymed; Zmed = GRID.zmed;
  Sxi = GRID.sxi; Sxj = GRID.sxj; Syj = GRID.syj; Szj = GRID.szj; Sxk = GRID.sxk; Syk = GRID.syk; Szk = GRID.szk;
  InvVj = GRID.invvj; InvDiffXmed = GRID.invdiffxmed;
  f3d_bctab_init(&BCT, &GRID, &P);
  f3d_damp_init(&DAMP, &P);
  if (f3d_psys_init(&Sys, &P, err, (size_t)errlen)) return -1;
  for (int k = 0; k < Sys.nb; k++)
    if (Sys.feeldisk[k]) {
      snprintf(err, errlen, "the binding test's host has no disk force on planets");
      return -1;
    }
  OMEGAFRAME = P.omegaframe;
  if (P.corotating) OMEGAFRAME = f3d_psys_info(&Sys, &P, F3D_PSYS_FREQUENCY);
  OMEGAFRAME0 = OMEGAFRAME;
  CLK.OMEGAFRAME = CLK.OMEGAFRAME0 = OMEGAFRAME;
  HipInit(); /* the mesh on the device (src/light_global_dev.c:5-23) */
  if (HipError[0]) {
    snprintf(err, errlen, "%s", HipError);
    return -1;
  }
  const size_t n = GRID.ncell, n2 = GRID.n2d;
  Field **f3[] = {&Density, &Energy, &Vx, &Vy, &Vz, &Vx_temp, &Vy_temp, &Vz_temp, &Mpx, &Mmx, &Mpy, &Mmy, &Mpz,
                  &Mmz, &Pot, &Slope, &DivRho, &DensStar, &Qs, &Pressure, &Total_Density, &QL, &QR};
  for (size_t a = 0; a < sizeof(f3) / sizeof(f3[0]); a++)
    if (!(*f3[a] = CreateField("f", n))) return -1;
  Field2D **f2[] = {&Density0, &Energy0, &Vx0, &Vy0, &Vz0, &Reduction2D, &VxMed};
  for (size_t a = 0; a < sizeof(f2) / sizeof(f2[0]); a++)
    if (!(*f2[a] = CreateField2D("f2", n2))) return -1;
  Nshift = (FieldInt2D *)CreateField("ns", n2);
  /* CondInit (src/main.c:265) and the upload */
  f3d_initial_conditions(&GRID, &P, &CLK, Density->field_cpu, Energy->field_cpu, Vx->field_cpu, Vy->field_cpu,
                         GRID.dimz ? Vz->field_cpu : NULL);
  Field *up[] = {Density, Energy, Vx, Vy, Vz};
  for (int a = 0; a < 5; a++) to_gpu(up[a], n);
  HipBindFluid();
  ChangeArchHip();
  FillGhosts(); /* src/main.c:311 */
  if (P.stockholm) {
    /* init_stockholm (src/init_stockholm.c:3-126), a host routine: the
       references are the first cell of every ring */
    Field *dn[] = {Density, Energy, Vx, Vy, Vz};
    for (int a = 0; a < 5; a++) to_cpu(dn[a], n);
    for (size_t l2 = 0; l2 < n2; l2++) {
      const size_t j = l2 % GRID.sy, k = l2 / GRID.sy, l = j * GRID.pitch + k * GRID.stride;
      Density0->field_cpu[l2] = Density->field_cpu[l];
      Energy0->field_cpu[l2] = Energy->field_cpu[l];
      Vx0->field_cpu[l2] = Vx->field_cpu[l];
      Vy0->field_cpu[l2] = Vy->field_cpu[l];
      Vz0->field_cpu[l2] = GRID.dimz ? Vz->field_cpu[l] : 0.0;
    }
    Field2D *rf[] = {Density0, Energy0, Vx0, Vy0, Vz0};
    for (int a = 0; a < 5; a++) to_gpu((Field *)rf[a], n2);
  }
  return HipError[0] ? -1 : 0;
}

/* one pass of the DT loop body (src/main.c:361-448); returns dt, or -1 */
double bind_step(void) {
  const real DT = P.dt_out;
  if (dtemp >= DT) dtemp = 0.0;
  if (P.fargo) ComputeVmed(Vx);
  FluidIndex = 0;
  cfl();
  StepTime = Min[0]; /* CflFluidsMin (src/cfl_fluids_min.c:3-22), one fluid */
  if (HipError[0] || StepTime <= 1e-10) return -1.0;
  real dt = StepTime;
  dtemp += dt;
  if (dtemp > DT) dt = DT - (dtemp - dt);
  Reset_field(Total_Density);
  ComputeTotalDensity();
  Sources(dt);
  Transport(dt);
  PhysicalTime += dt;
  TimeStepCount++;
  if (P.stockholm) StockholmBoundary(dt);
  FillGhosts();
  return HipError[0] ? -1.0 : dt;
}

const char *bind_error(void) { return HipError; }

/* releases the run so that bind_setup can start another */
void bind_teardown(void) {
  Field **f3[] = {&Density, &Energy, &Vx, &Vy, &Vz, &Vx_temp, &Vy_temp, &Vz_temp, &Mpx, &Mmx, &Mpy, &Mmy, &Mpz,
                  &Mmz, &Pot, &Slope, &DivRho, &DensStar, &Qs, &Pressure, &Total_Density, &QL, &QR};
  for (size_t a = 0; a < sizeof(f3) / sizeof(f3[0]); a++) {
    DestroyField(*f3[a]);
    *f3[a] = NULL;
  }
  Field2D **f2[] = {&Density0, &Energy0, &Vx0, &Vy0, &Vz0, &Reduction2D, &VxMed};
  for (size_t a = 0; a < sizeof(f2) / sizeof(f2[0]); a++) {
    DestroyField((Field *)*f2[a]);
    *f2[a] = NULL;
  }
  DestroyField((Field *)Nshift);
  Nshift = NULL;
  HipFini();
  f3d_bctab_free(&BCT);
  f3d_grid_free(&GRID);
  PhysicalTime = Domega = 0.0;
  StepTime = dtemp = 0.0;
  TimeStepCount = 0;
  alreadycalculated = -1;
  HipError[0] = 0;
  memset(&Sys, 0, sizeof(Sys));
  memset(&CLK, 0, sizeof(CLK));
}

/* a field's device contents (nx*sy*sz doubles) into host memory */
int bind_field(const char *name, double *out) {
  struct { const char *n; Field *f; } t[] = {{"Density", Density}, {"Energy", Energy}, {"Vx", Vx}, {"Vy", Vy},
                                             {"Vz", Vz}};
  for (size_t a = 0; a < sizeof(t) / sizeof(t[0]); a++)
    if (strcmp(name, t[a].n) == 0) {
      HipToHost(out, t[a].f->field_gpu, GRID.ncell * sizeof(real));
      return HipError[0] ? -1 : 0;
    }
  return -1;
}

void bind_planets(double *out) {
  for (int i = 0; i < Sys.nb; i++) {
    out[7 * i + 0] = Sys.x[i]; out[7 * i + 1] = Sys.y[i]; out[7 * i + 2] = Sys.z[i];
    out[7 * i + 3] = Sys.vx[i]; out[7 * i + 4] = Sys.vy[i]; out[7 * i + 5] = Sys.vz[i];
    out[7 * i + 6] = Sys.mass[i];
  }
}
