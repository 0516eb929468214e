/*
 * fargo3d_hip.h -- C ABI of the MI355X (gfx950) implementation of FARGO3D's
 * hydro time step.
 *
 * Drop-in model.  FARGO3D calls every routine of its step through a global
 * function pointer (src/global.h:303-416) that ChangeArch binds to a *_cpu
 * or *_gpu implementation from std/func_arch.cfg (src/change_arch.c:146-570).
 * Those implementations read the selected fluid's fields, the mesh and the
 * run parameters from globals.  This ABI keeps the same routine names and
 * argument lists and makes the globals explicit in one handle:
 *   - f3d_hip_create   uploads the mesh (the LightGlobalDev step,
 *                      src/light_global_dev.c:5-23);
 *   - f3d_hip_bind     names the device buffers of the selected fluid and the
 *                      shared scratch fields (SelectFluid, src/LowTasks.c:565);
 *   - f3d_hip_set_phys sets the run parameters and host scalars the kernels
 *                      read (.par values, OMEGAFRAME, planets).
 * A Field* argument of the reference becomes the device pointer of that field
 * (Field->field_gpu); everything else is plain C.  Buffers are x-fastest
 * arrays with the pitch/stride given at creation.
 *
 * Errors: every entry point returns 0 on success or a negative code; the
 * message is available from f3d_hip_last_error().  (The reference instead
 * exits the process from check_errors, src/lowtasks.cu:6-15.)
 * Launches are asynchronous on the handle's stream unless a routine returns a
 * host value (the reductions and cfl), which synchronises.
 */
#ifndef FARGO3D_HIP_H
#define FARGO3D_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct f3d_hip_ctx f3d_hip_ctx;

/* Mesh description (host arrays; sizes: x* nx+1 / nx, y* sy+1 / sy, z* sz+1 / sz
 * with sy = ny+2*nghy, sz = nz+2*nghz).  Fields mirror src/global.h:210-242. */
typedef struct f3d_mesh_desc {
  int nx, ny, nz, nghy, nghz, dimz, geometry; /* geometry: 1 cylindrical, 2 spherical */
  int pitch, stride;                          /* in elements */
  double XMIN, XMAX, mindx;
  const double *xmin, *ymin, *zmin, *xmed, *ymed, *zmed;
  const double *sxi, *sxj, *syj, *szj, *sxk, *syk, *szk, *invvj, *invdiffxmed;
  const double *cosxmed, *sinxmed;            /* cos/sin(Xmed), nx */
  const double *sinzmed, *coszmed;            /* sin/cos(Zmed), sz */
  const double *sinzmin, *coszmin;            /* sin/cos(Zmin), sz+1 */
} f3d_mesh_desc;

/* Device buffers of the selected fluid and the shared scratch fields
 * (src/LowTasks.c:565-642).  NULL where a field does not exist (e.g. Vz in 2-D). */
typedef struct f3d_fields {
  double *Density, *Energy, *Vx, *Vy, *Vz, *Vx_temp, *Vy_temp, *Vz_temp;
  double *Mpx, *Mmx, *Mpy, *Mmy, *Mpz, *Mmz;
  double *Pot, *Slope, *DivRho, *DensStar, *Qs, *Pressure, *Total_Density, *QL, *QR;
  double *Density0, *Energy0, *Vx0, *Vy0, *Vz0; /* Stockholm references, YZ arrays */
  double *Reduction2D, *VxMed;                   /* YZ arrays */
  int *Nshift;                                   /* YZ int array */
} f3d_fields;

/* Run parameters and host scalars the kernels read. */
typedef struct f3d_phys {
  int fluidtype;       /* 1 gas, 2 dust (Fluidtype) */
  int eos;             /* 1 isothermal, 2 adiabatic */
  int visc;            /* 0 none, 1 NU, 2 ALPHA */
  int potential, fargo, stockholm, indirectterm, periodicz, periodicy;
  int bc;              /* 1: 2-D Keplerian disk boundaries, 2: 3-D spherical disk boundaries */
  int energy_bc;       /* 1: Energy ghosts SYMMETRIC on the non-periodic sides (setup .bound entry) */
  double nu, alpha, gamma, cfl;
  double OMEGAFRAME, OMEGAFRAME0, PhysicalTime;
  double ymin_par, ymax_par, zmin_par, zmax_par, taudamp;
  double Y_inf, Y_sup, Z_inf, Z_sup;               /* damping zones (src/stockholm.c:76-90) */
  /* boundary factors (host glibc evaluation of std/boundaries.txt formulas) */
  double ymin_dens[3], ymax_dens[3], ymin_vsq[3], ymax_vsq[3], ymin_vfac[3], ymax_vfac[3];
  double zmin_vfac[3], zmax_vfac[3];
  const double *ymin_dexp, *ymax_dexp, *zmin_dens, *zmax_dens; /* host arrays, 3*sz / 3*sy */
  /* planets for Potential (positions at the start of the step) */
  int nplanets;
  double px[16], py[16], pz[16], pmass[16]; /* mass already multiplied by the taper */
  double psmooth2[16];                      /* smoothing^2 (src/potential.c:120-131) */
} f3d_phys;

/* ---- runtime ---- */
int f3d_hip_create(const f3d_mesh_desc *mesh, f3d_hip_ctx **out);
void f3d_hip_destroy(f3d_hip_ctx *ctx);
int f3d_hip_bind(f3d_hip_ctx *ctx, const f3d_fields *fields);
int f3d_hip_set_phys(f3d_hip_ctx *ctx, const f3d_phys *phys);
/* Initial conditions on the device (the setups' CondInit: setups/fargo/
 * condinit.c:25-50, setups/p3disof/condinit.c:14-52, setups/fargo_multifluid/
 * condinit.c:8-52).  A set-up's initial state is a per-ring profile -- every
 * transcendental of the coordinates, which the host evaluates with glibc --
 * times the seeded noise of each cell.  The noise is glibc's drand48 stream
 * in the set-up's call order (cells x fastest, then y, then z; ndraw draws per
 * cell), which the device reproduces by jump-ahead of the 48-bit generator
 * X' = (a X + c) mod 2^48 (a = 0x5DEECE66D, c = 0xB; a draw is X' / 2^48). */
enum { F3D_IC_PLAIN = 0, F3D_IC_KEPLER = 1, F3D_IC_FARGO = 2, F3D_IC_LOGNORMAL = 3 };
typedef struct f3d_ic_desc {
  int mode;      /* F3D_IC_*: PLAIN rho = R, vx = V, vy = 0, e = E (no noise);
                    KEPLER rho = R (1 + noise (U0 - .5)), vx = V, vy = 0;
                    FARGO  rho = R (1 + noise (U0 - .5)), vx = V (1 + h noise (U1 - .5)),
                           vy = CS noise (U2 - .5);
                    LOGNORMAL rho = R exp(noise sqrt(-2 log U0) cos(2 pi U1)) (noise != 0), vx = V, vy = 0;
                    energy (not PLAIN): E rho / (gamma - 1) if adiabatic (E = cs^2), else CS */
  int ndraw;     /* drand48 draws per cell */
  int adiabatic;
  double noise, aspectratio, gamma;
  unsigned long long x0;                 /* generator state before the first draw (48 bits) */
  const double *R, *V, *CS, *E;          /* host arrays of sy*sz per-ring values (index j + k*sy) */
} f3d_ic_desc;
/* The state of every cell (ghosts included) of one fluid, FP64 fields.  No
 * counterpart entry point in the reference (CondInit runs on its host). */
int f3d_CondInit(f3d_hip_ctx *ctx, const f3d_ic_desc *d, double *dens, double *energy, double *vx, double *vy,
                 double *vz);
/* The same into the precision form's FP32 fields (behind the same pointer
 * types), the azimuthal velocity as its deviation from the profile of
 * f3d_hip_set_vbase. */
int f3d32_CondInit(f3d_hip_ctx *ctx, const f3d_ic_desc *d, double *dens, double *energy, double *vx, double *vy,
                   double *vz);
/* drand48's state after n draws from state x0 (the generator's jump-ahead;
 * host only, no device use) */
unsigned long long f3d_drand48_skip(unsigned long long x0, unsigned long long n);

/* Precision form (no counterpart in the reference, whose fields are all of
 * one type): the FP64 ring profile V0[j + k*sy] (sy*sz values) that the FP32
 * fields Vx, Vx_temp and the damping reference Vx0 are stored as deviations
 * from, vx = V0(j,k) + (double)stored.  Set once, before the first f3d32_*
 * stage on the handle and before any FP32 azimuthal velocity is written. */
int f3d_hip_set_vbase(f3d_hip_ctx *ctx, const double *host_v0);
int f3d_hip_sync(f3d_hip_ctx *ctx);
const char *f3d_hip_last_error(f3d_hip_ctx *ctx);
/* device memory helpers for callers without their own allocator */
int f3d_hip_malloc(void **ptr, size_t bytes);
int f3d_hip_free(void *ptr);
int f3d_hip_memcpy_h2d(f3d_hip_ctx *ctx, void *dst, const void *src, size_t bytes);
int f3d_hip_memcpy_d2h(f3d_hip_ctx *ctx, void *dst, const void *src, size_t bytes);
int f3d_hip_memcpy_d2d(f3d_hip_ctx *ctx, void *dst, const void *src, size_t bytes);
int f3d_hip_memset(f3d_hip_ctx *ctx, void *dst, int value, size_t bytes);
int f3d_hip_device_count(void);
/* HIP-event timing of the fused stages and Potential on the handle's stream
 * (no counterpart in the reference; used by the benchmark) */
int f3d_hip_stage_count(void);
const char *f3d_hip_stage_name(int stage);
int f3d_hip_timing_enable(f3d_hip_ctx *ctx, int on);
/* time only the stages whose bit is set (an event pair around a stage puts
 * a few microseconds between launches, so the benchmark times one stage in
 * its timed region); enable(on) is mask ~0 or 0 */
int f3d_hip_timing_mask(f3d_hip_ctx *ctx, unsigned mask);
int f3d_hip_timing_read(f3d_hip_ctx *ctx, double *ms, long *counts);
/* time only every n-th launch of each timed stage (n >= 1; -2 otherwise):
 * the averages then come from a sample of the launches, and the others run
 * without the event pair's gap */
int f3d_hip_timing_every(f3d_hip_ctx *ctx, int n);

/* ---- per-routine entry points (reference pointer table, src/global.h:303-416) ---- */
/* ComputePressureFieldIso/Ad: src/CompPresIso.c:10, src/CompPresAd.c:10 */
int f3d_ComputePressureFieldIso(f3d_hip_ctx *ctx);
int f3d_ComputePressureFieldAd(f3d_hip_ctx *ctx);
/* ComputeTotalDensity: src/totaldens.c:10; Reset_field: src/resetfield.c:10 */
int f3d_ComputeTotalDensity(f3d_hip_ctx *ctx);
int f3d_Reset_field(f3d_hip_ctx *ctx, double *F);
/* SubStep1_x/y/z: src/substep1_x.c:10, src/substep1_y.c:10, src/substep1_z.c:10 */
int f3d_SubStep1_x(f3d_hip_ctx *ctx, double dt);
int f3d_SubStep1_y(f3d_hip_ctx *ctx, double dt);
int f3d_SubStep1_z(f3d_hip_ctx *ctx, double dt);
/* SubStep2_a/b: src/substep2_a.c:10, src/substep2_b.c:10; SubStep3: src/substep3.c:10 */
int f3d_SubStep2_a(f3d_hip_ctx *ctx, double dt);
int f3d_SubStep2_b(f3d_hip_ctx *ctx, double dt);
int f3d_SubStep3(f3d_hip_ctx *ctx, double dt);
/* visctensor_cyl/sph + addviscosity_cyl/sph: src/visctensor_cyl.c:10,
 * src/addviscosity_cyl.c:10, src/visctensor_sph.c:10, src/addviscosity_sph.c:10 */
int f3d_visctensor_cyl(f3d_hip_ctx *ctx);
int f3d_addviscosity_cyl(f3d_hip_ctx *ctx, double dt);
int f3d_visctensor_sph(f3d_hip_ctx *ctx);
int f3d_addviscosity_sph(f3d_hip_ctx *ctx, double dt);
/* Potential: src/potential.c:51; CorrectVtheta: src/correct_vtheta.c:10 */
int f3d_Potential(f3d_hip_ctx *ctx);
int f3d_CorrectVtheta(f3d_hip_ctx *ctx, double domega);
/* _ComputeForce: src/compute_force.c:16 (fills the 12 force sums, host values) */
int f3d_ComputeForce(f3d_hip_ctx *ctx, double x, double y, double z, double rsmoothing, double mass,
                     double localforce_out[12]);
/* copy_velocities: src/cpvel.c:3 (option 1 = VTEMP2V, 2 = V2VTEMP) */
int f3d_copy_velocities(f3d_hip_ctx *ctx, int option);
/* boundaries: src/boundary.c:2 (zmin, zmax, ymin, ymax of the bound family) */
int f3d_boundaries(f3d_hip_ctx *ctx);
/* the same ghost fill on explicit buffers (e.g. V_temp during the mid-step fill) */
int f3d_boundaries_on(f3d_hip_ctx *ctx, double *Density, double *Vx, double *Vy, double *Vz);
/* momenta_x/y/z: src/momenta_x.c:10, src/momenta_y.c:10, src/momenta_z.c:10 */
int f3d_momenta_x(f3d_hip_ctx *ctx);
int f3d_momenta_y(f3d_hip_ctx *ctx);
int f3d_momenta_z(f3d_hip_ctx *ctx);
/* DivideByRho: src/divrho.c:10 */
int f3d_DivideByRho(f3d_hip_ctx *ctx, const double *Q);
/* VanLeerX/Y/Z_a/b: src/vanleer_x_a.c:10, src/vanleer_x_b.c:10, src/vanleer_y_a.c:10,
 * src/vanleer_y_b.c:10, src/vanleer_z_a.c:10, src/vanleer_z_b.c:10 */
int f3d_VanLeerX_a(f3d_hip_ctx *ctx, const double *Q);
int f3d_VanLeerX_b(f3d_hip_ctx *ctx, double dt, const double *Q, double *Qs, const double *Vx_t);
int f3d_VanLeerY_a(f3d_hip_ctx *ctx, const double *Q);
int f3d_VanLeerY_b(f3d_hip_ctx *ctx, double dt, const double *Q, double *Qs);
int f3d_VanLeerZ_a(f3d_hip_ctx *ctx, const double *Q);
int f3d_VanLeerZ_b(f3d_hip_ctx *ctx, double dt, const double *Q, double *Qs);
/* VanLeerX_PPA_a/b/c/d: src/fargo_ppa_a.c:10, src/fargo_ppa_b.c:10, src/fargo_ppa_c.c:10,
 * src/fargo_ppa_d.c:10 */
int f3d_VanLeerX_PPA_a(f3d_hip_ctx *ctx, const double *Q);
int f3d_VanLeerX_PPA_b(f3d_hip_ctx *ctx, const double *Q);
int f3d_VanLeerX_PPA_c(f3d_hip_ctx *ctx, const double *Q);
int f3d_VanLeerX_PPA_d(f3d_hip_ctx *ctx, double dt, const double *Q, double *Qs, const double *Vx_t);
/* UpdateX/Y/Z: src/update_x.c:10, src/update_y.c:10, src/update_z.c;
 * UpdateDensityX/Y/Z: src/update_density_x.c:10, src/update_density_y.c:10, src/update_density_z.c */
int f3d_UpdateX(f3d_hip_ctx *ctx, double dt, double *Q, const double *Qs, const double *Vx_t);
int f3d_UpdateY(f3d_hip_ctx *ctx, double dt, double *Q, const double *Qs);
int f3d_UpdateZ(f3d_hip_ctx *ctx, double dt, double *Q, const double *Qs);
int f3d_UpdateDensityX(f3d_hip_ctx *ctx, double dt, double *Q, const double *Vx_t);
int f3d_UpdateDensityY(f3d_hip_ctx *ctx, double dt, double *Q);
int f3d_UpdateDensityZ(f3d_hip_ctx *ctx, double dt, double *Q);
/* NewVelocity_x/y/z: src/newvel_x.c:10, src/newvel_y.c:10, src/newvel_z.c:10 */
int f3d_NewVelocity_x(f3d_hip_ctx *ctx);
int f3d_NewVelocity_y(f3d_hip_ctx *ctx);
int f3d_NewVelocity_z(f3d_hip_ctx *ctx);
/* ComputeVweight: src/compute_vmed.c:42; ComputeVmed: src/compute_vmed.c:10 */
int f3d_ComputeVweight(f3d_hip_ctx *ctx, const double *V, double *Q);
int f3d_ComputeVmed(f3d_hip_ctx *ctx, const double *V);
/* ComputeResidual: src/compute_residual.c:10; AdvectSHIFT: src/advect_shift.c:9 */
int f3d_ComputeResidual(f3d_hip_ctx *ctx, double dt);
int f3d_AdvectSHIFT(f3d_hip_ctx *ctx, double *F, const int *Nshift);
/* reduction_SUM/MIN: src/reduction_generic.c:5 (per-ring, into Reduction2D, ring sums in
 * sequential i order); reduction_full_*: src/reduction_full_generic.c:1 (host value) */
int f3d_reduction_SUM(f3d_hip_ctx *ctx, const double *F, int ymin, int ymax, int zmin, int zmax);
int f3d_reduction_MIN(f3d_hip_ctx *ctx, const double *F, int ymin, int ymax, int zmin, int zmax);
int f3d_reduction_full_SUM(f3d_hip_ctx *ctx, const double *F, int ymin, int ymax, int zmin, int zmax,
                           double *result);
/* Monitored quantities of the bound fluid into an FP64 array `out` (the
 * reference writes them into Slope), active cells only, the ones the monitor
 * reductions read (src/monitor.c:232-233): mon_dens src/mon_dens.c:10,
 * mon_momx src/mon_momx.c:10, mon_momy src/mon_momy.c:10, mon_momz
 * src/mon_momz.c:10, mon_torq src/mon_torque.c:10 (planet position and
 * smoothing passed in place of the Xplanet.. globals).  The f3d32_ twins read
 * the precision form's FP32 fields and compute in FP64. */
int f3d_mon_dens(f3d_hip_ctx *ctx, double *out);
int f3d_mon_momx(f3d_hip_ctx *ctx, double *out);
int f3d_mon_momy(f3d_hip_ctx *ctx, double *out);
int f3d_mon_momz(f3d_hip_ctx *ctx, double *out);
int f3d_mon_torq(f3d_hip_ctx *ctx, double xp, double yp, double zp, double rsmoothing, double *out);
int f3d32_mon_dens(f3d_hip_ctx *ctx, double *out);
int f3d32_mon_momx(f3d_hip_ctx *ctx, double *out);
int f3d32_mon_momy(f3d_hip_ctx *ctx, double *out);
int f3d32_mon_momz(f3d_hip_ctx *ctx, double *out);
int f3d32_mon_torq(f3d_hip_ctx *ctx, double xp, double yp, double zp, double rsmoothing, double *out);
int f3d_reduction_full_MIN(f3d_hip_ctx *ctx, const double *F, int ymin, int ymax, int zmin, int zmax,
                           double *result);
/* cfl: src/cfl.c:10 + cfl_b: src/cfl_b.c:3 (per-fluid Min, host value) */
int f3d_cfl(f3d_hip_ctx *ctx, double *fluid_min);
/* StockholmBoundary: src/stockholm.c:10; init_stockholm: src/init_stockholm.c:3 */
int f3d_StockholmBoundary(f3d_hip_ctx *ctx, double dt);
int f3d_init_stockholm(f3d_hip_ctx *ctx);

/* ---- fused stages (no counterpart in the reference: the same arithmetic as
 * the routine sequences they replace, in fewer HBM passes) ---- */
/* ComputeVmed(Vx) + cfl + cfl_b in one launch (one workgroup per ring, one
 * atomic MIN per ring); writes the fluid's Min to *fluid_min. */
int f3d_fused_cfl(f3d_hip_ctx *ctx, double *fluid_min);
/* Sources from the pressure field to the mid-step ghost fill (src/algogas.c:34-118),
 * after Potential; domega is the CorrectVtheta argument (0 when not corotating). */
int f3d_fused_sources(f3d_hip_ctx *ctx, double dt, double domega);
/* MULTIFLUID(Sources(dt)) (src/main.c:410) of nf fluids (type[q]: 1 gas,
 * 2 dust) in one pass: the viscous fluid's stresses, one SubStep1/2 launch
 * over every fluid, the mid-step ghost fill.  Isothermal runs with at most
 * one viscous fluid (else -1; call f3d_fused_sources per fluid). */
int f3d_fused_sources_fluids(f3d_hip_ctx *ctx, double dt, double domega, int nf, const f3d_fields *fl,
                             const int *type);
/* Transport (src/algogas.c:135-159) and, when stockholm != 0, StockholmBoundary and
 * the end-of-step ghost fill. */
int f3d_fused_transport(f3d_hip_ctx *ctx, double dt, int end_of_step);
/* MULTIFLUID(Transport(dt)), MULTIFLUID(StockholmBoundary(dt)) and
 * MULTIFLUID(FillGhosts) (src/main.c:420-434) for nf (1..8) fluids whose fields
 * are fl[0..nf-1] (the scratch, potential, VxMed and Nshift pointers are taken
 * from the handle's bound set): every fluid's sweeps, with the X sweep of all
 * fluids in one launch.  VxMed and Nshift end with the last fluid's values. */
int f3d_fused_transport_fluids(f3d_hip_ctx *ctx, double dt, int nf, const f3d_fields *fl);
/* Rows (Y sweep; 0: automatic, the default) and planes (Z sweep) one thread
 * marches through; a tuning knob that does not change results.  Returns -2
 * for rows_y below 0 or rows_z below 1. */
int f3d_fused_set_tiling(int rows_y, int rows_z);
/* X-sweep kernel: 0 automatic (one wave per ring with the cells in
 * registers for rings of 256 cells from 2048 rings x fluids on, else the
 * ring in LDS when it fits, else tiles of 256 cells with halos, one wave
 * each, the cells in registers), 1 ring in LDS, 2 tiles of 128 cells in LDS,
 * 3 one wave per ring (rings of 256 cells; others as 0), 4 one-wave tiles;
 * results do not depend on it.  Returns -2 for other values. */
int f3d_fused_set_xmode(int mode);
/* Ring X sweep: 1 two workgroups per ring, each carrying the density and
 * half of the transported quantities (about half the LDS, so more
 * workgroups per CU), 0 one workgroup per ring, -1 automatic; results do not
 * depend on it.  Returns -2 for other values. */
int f3d_fused_set_xsplit(int split);
/* Cells per workgroup of the tiled X sweep: 64, 128 (default) or 256;
 * results do not depend on it.  Returns -2 for other values. */
int f3d_fused_set_xtile(int cells);
/* Where the ring X sweep takes each ring's VxMed (ComputeVmed of Vx_temp,
 * src/compute_vmed.c:14-30) from: 0 its own sequential sum, 1 a one-wave-
 * per-ring pass on a second stream that overlaps the Z/Y sweeps, 2 that pass
 * on the step's stream just before the X sweep, -1 automatic (1 from 1024
 * rings x fluids on); the tiled sweep always uses the pass.  Results do not
 * depend on it.  Returns -2 for other values. */
int f3d_fused_set_xvmed(int form);
/* Column (Y and Z) sweeps: 1 two threads per column, each with the density
 * and half of the quantities, 0 one thread per column, -1 automatic (split
 * only for sweeps with fewer waves than SIMDs); results do not depend on it.
 * Returns -2 for other values. */
int f3d_fused_set_colsplit(int split);
/* Form of the fused CFL, bits: 1 ring minima to an array and a
 * one-workgroup second pass (else one 64-bit atomic MIN per ring); 2 the
 * rings' VxMed from a pass of its own; 4 (with bit 2) one wave per
 * workgroup.  -1 automatic.  The result is the same value in every form.
 * Returns -2 for other values. */
int f3d_fused_set_cflmode(int mode);
/* Terms per lane of the one-wave sequential ring sums of VxMed (the CFL's
 * and the X sweep's passes): 4, 8 or 16, -1 automatic (by ring length); the
 * sums are the same additions in the same order in every form.  Returns -2
 * for other values. */
int f3d_fused_set_vmedterms(int terms);
/* 2-D sources of a viscous gas with a potential: one pass that computes the
 * viscous stress where it is used (no stress arrays), -1 automatic (on), or
 * 0 the stress pass and the update pass of their own; the same arithmetic
 * in both.  Returns -2 for other values. */
int f3d_fused_set_srcfuse(int on);
/* End-of-step ghost fill: -1 automatic (on): in 3-D the Z sides, in 2-D the
 * Y sides, are filled by the finishing kernel from the values it has just
 * written (in 3-D the Y fill then runs as its own launch); 0: the fills as
 * launches of their own.  The same values in both.  Returns -2 for other
 * values. */
int f3d_fused_set_zfill(int on);
/* 2-D X sweep of long or many rings: the VxMed pass it reads runs in the Y
 * sweep's launch (its own workgroups beside the sweep's), -1 automatic (on),
 * or 0 on the second stream; the same sums in both.  Returns -2 for other
 * values. */
int f3d_fused_set_y2v(int on);
/* Potential, the ghost fill and the damping references on the fused stages'
 * scratch-free kernels (the same formulas as f3d_Potential, f3d_boundaries,
 * f3d_init_stockholm) */
int f3d_fused_potential(f3d_hip_ctx *ctx);
int f3d_fused_boundaries(f3d_hip_ctx *ctx);
int f3d_fused_init_stockholm(f3d_hip_ctx *ctx);
/* f3d_fused_cfl with the ring sums (ComputeVmed, src/compute_vmed.c:14-30)
 * taken over vxm instead of the bound fluid's Vx: with several fluids the
 * reference runs MULTIFLUID(ComputeVmed(Vx)) before MULTIFLUID(cfl())
 * (src/main.c:366-381), so every fluid's cfl reads the last fluid's VxMed.
 * vxm NULL means the bound Vx. */
int f3d_fused_cfl_vm(f3d_hip_ctx *ctx, const double *vxm, double *fluid_min);
/* cfl of nf fluids (cs = each fluid's Energy) and CflFluidsMin
 * (src/cfl_fluids_min.c:3-22) with one device MIN and one read-back; the ring
 * sums are taken over vxm as in f3d_fused_cfl_vm (vz ignored in 2-D). */
int f3d_fused_cfl_fluids(f3d_hip_ctx *ctx, int nf, double *const *cs, double *const *rho, double *const *vx,
                         double *const *vy, double *const *vz, const double *vxm, double *step_min);
/* f3d_fused_cfl_fluids in two halves (no counterpart in the reference, whose
 * cfl returns after its reduction): _async queues the launches and returns;
 * f3d_fused_cfl_wait waits for them alone and returns the step.  A host can
 * queue work that does not need dt (the planets' potential) in between, so
 * that it runs while the host handles dt. */
int f3d_fused_cfl_fluids_async(f3d_hip_ctx *ctx, int nf, double *const *cs, double *const *rho,
                               double *const *vx, double *const *vy, double *const *vz, const double *vxm);
int f3d_fused_cfl_wait(f3d_hip_ctx *ctx, double *step_min);
/* f3d_fused_cfl_fluids with no allocation (no counterpart in the reference;
 * for callers that cannot allocate device memory: code under stream capture,
 * serving loops, harnesses that forbid hipMalloc).  Same arguments and the
 * same time step bit for bit (vxm must not be NULL); the per-ring minima and
 * result slots live in a fixed device array of the library, at most 262144
 * rings x fluids (larger meshes return -2 with a message), shared by every
 * call on a GPU: calls that overlap on one GPU race.  Synchronous. */
int f3d_fused_cfl_fluids_noalloc(f3d_hip_ctx *ctx, int nf, double *const *cs, double *const *rho,
                                 double *const *vx, double *const *vy, double *const *vz, const double *vxm,
                                 double *step_min);

/* ---- coupling of several fluids (src/main.c:412-418) ---- */
/* Collisions(dt, 1): src/collisions.c:156 (+ _collisions_cpu :10, ColRate
 * src/colrate.c:10).  Implicit drag on the V_temp fields of nf (2..8) fluids,
 * one direction after the other; alpha is the ColRate matrix, alpha[i + j*nf]
 * the rate fluid i feels from fluid j (the STOKESNUMBER form: multiplied by
 * the local Keplerian frequency).  rho/vxt/vyt/vzt hold each fluid's device
 * fields (vzt ignored in 2-D). */
int f3d_fused_collisions(f3d_hip_ctx *ctx, double dt, int nf, double *const *rho, double *const *vxt,
                         double *const *vyt, double *const *vzt, const double *alpha);
/* DustDiffusion_Main(dt): src/dust_diffusion_main.c:3 (+ the coefficients,
 * src/dust_diffusion_coefficients.c:10, and the core,
 * src/dust_diffusion_core.c:10).  Diffuses the density of every fluid whose
 * type is 2 (dust) against the gas density rho[0]; cs_gas is the gas
 * Energy (sound speed) field.  2-D cylindrical meshes. */
int f3d_fused_dust_diffusion(f3d_hip_ctx *ctx, double dt, int nf, double *const *rho, const int *type,
                             const double *cs_gas);

/* ---- reduced-precision form (no counterpart in the reference): the fused
 * stages on FP32 fields (every 3-D field and the damping references bound to
 * the handle are FP32 arrays behind the same pointers), with FP32 stencil
 * arithmetic and FP64 kept for the ring sums (VxMed stays an FP64 array), the
 * residual/shift split (Nshift stays int) and the CFL time step. ---- */
int f3d32_fused_cfl(f3d_hip_ctx *ctx, double *fluid_min);
int f3d32_fused_sources(f3d_hip_ctx *ctx, double dt, double domega);
int f3d32_fused_sources_fluids(f3d_hip_ctx *ctx, double dt, double domega, int nf, const f3d_fields *fl,
                               const int *type);
int f3d32_fused_transport(f3d_hip_ctx *ctx, double dt, int end_of_step);
int f3d32_fused_transport_fluids(f3d_hip_ctx *ctx, double dt, int nf, const f3d_fields *fl);
int f3d32_fused_potential(f3d_hip_ctx *ctx);
int f3d32_fused_boundaries(f3d_hip_ctx *ctx);
int f3d32_fused_init_stockholm(f3d_hip_ctx *ctx);
int f3d32_fused_set_tiling(int rows_y, int rows_z);
int f3d32_fused_set_xmode(int mode);
int f3d32_fused_set_xsplit(int split);
int f3d32_fused_set_xtile(int cells);
int f3d32_fused_set_xvmed(int form);
int f3d32_fused_set_colsplit(int split);
int f3d32_fused_set_cflmode(int mode);
int f3d32_fused_set_vmedterms(int terms);
int f3d32_fused_set_srcfuse(int on);
int f3d32_fused_set_zfill(int on);
int f3d32_fused_set_y2v(int on);
int f3d32_fused_cfl_vm(f3d_hip_ctx *ctx, const double *vxm, double *fluid_min);
int f3d32_fused_cfl_fluids(f3d_hip_ctx *ctx, int nf, double *const *cs, double *const *rho, double *const *vx,
                           double *const *vy, double *const *vz, const double *vxm, double *step_min);
int f3d32_fused_cfl_fluids_async(f3d_hip_ctx *ctx, int nf, double *const *cs, double *const *rho,
                                 double *const *vx, double *const *vy, double *const *vz, const double *vxm);
int f3d32_fused_cfl_wait(f3d_hip_ctx *ctx, double *step_min);
int f3d32_fused_collisions(f3d_hip_ctx *ctx, double dt, int nf, double *const *rho, double *const *vxt,
                           double *const *vyt, double *const *vzt, const double *alpha);
int f3d32_fused_dust_diffusion(f3d_hip_ctx *ctx, double dt, int nf, double *const *rho, const int *type,
                               const double *cs_gas);

#ifdef __cplusplus
}
#endif
#endif
