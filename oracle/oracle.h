/*
 * oracle.h -- CPU restatement of FARGO3D's hydro time step.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker for the HIP path: only the
 * tests, the smoke check and the bench's CPU-baseline leg may load it.  The
 * product (fargo3d_amd) never links, imports or calls anything here.
 *
 * Every routine restates the reference's sequential CPU form (the *_cpu
 * function of the cited file) loop by loop: the same loop ranges (ghost rows
 * included), the same scratch-field reuse, and the same expression
 * association (the CPU macro forms of src/define.h:226-271 are kept textually
 * in oracle_metric.h), compiled with -ffp-contract=off so that it is
 * bit-identical to a gcc -O3 x86-64 build of the reference.
 */
#ifndef F3D_ORACLE_H
#define F3D_ORACLE_H

#include "../fargo3d_amd/csrc/host/f3d_host.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_fluid {
  int type; /* F3D_GAS / F3D_DUST */
  double *Density, *Energy, *Vx, *Vy, *Vz, *Vx_temp, *Vy_temp, *Vz_temp;
  double *Density0, *Energy0, *Vx0, *Vy0, *Vz0; /* Stockholm references (YZ) */
  double Min; /* per-fluid CFL result (src/cfl_b.c:6) */
} orc_fluid;

typedef struct orc_state {
  f3d_params p;
  f3d_grid g;
  f3d_psys sys;
  f3d_clock clk;
  f3d_bctab bct;
  f3d_damp damp;
  int nfluids, fluid_index;
  orc_fluid fl[F3D_MAX_FLUIDS];
  /* the selected fluid (SelectFluid, src/LowTasks.c:565-593) */
  int Fluidtype;
  double *Density, *Energy, *Vx, *Vy, *Vz, *Vx_temp, *Vy_temp, *Vz_temp;
  double *Density0, *Energy0, *Vx0, *Vy0, *Vz0;
  /* shared scratch fields (src/LowTasks.c:595-642) */
  double *Mpx, *Mmx, *Mpy, *Mmy, *Mpz, *Mmz, *Pot, *Slope, *DivRho, *DensStar, *Qs;
  double *Pressure, *Total_Density, *QL, *QR;
  double *Reduction2D, *VxMed;
  int *Nshift;
  /* scalars */
  double StepTime, dtemp, dt_last;
  double IndirectTerm[3], DiskOnPrimaryAcceleration[3];
  double localforce[12];
  long alreadycalculated;
  int compute_disk_force; /* 1: evaluate ComputeAccel every step as the reference does */
  /* multifluid: collision rates (Alpha, src/colrate.c:10-23), dust diffusion
     coefficients (Sdiffyczc/Sdiffyfzc), density-only ghost fill flag */
  double alpha[F3D_MAX_FLUIDS * F3D_MAX_FLUIDS];
  double *Sdiffyczc, *Sdiffyfzc;
  int fill_density_only;
} orc_state;

/* lifetime */
orc_state *orc_create(const char *parfile, const char *overrides, char *err, int errlen);
void orc_destroy(orc_state *S);
void orc_select_fluid(orc_state *S, int n);

/* one iteration of the DT loop (src/main.c:361-448); returns dt used, <0 on error */
double orc_step(orc_state *S);

/* routines (same names as the reference's *_cpu functions) */
void orc_ComputePressureFieldIso(orc_state *S);
void orc_ComputePressureFieldAd(orc_state *S);
void orc_ComputeVweight(orc_state *S, const double *V, double *Q);
void orc_reduction_SUM(orc_state *S, const double *F, int ymin, int ymax, int zmin, int zmax);
double orc_reduction_full_SUM(orc_state *S, const double *F, int ymin, int ymax, int zmin, int zmax);
double orc_reduction_full_MIN(orc_state *S, const double *F, int ymin, int ymax, int zmin, int zmax);
void orc_ComputeVmed(orc_state *S, const double *V);
void orc_cfl(orc_state *S);
void orc_cfl_b(orc_state *S);
int orc_CflFluidsMin(orc_state *S);
void orc_ComputeTotalDensity(orc_state *S);
void orc_Reset_field(orc_state *S, double *F);
void orc_ComputeForce(orc_state *S, double x, double y, double z, double rsmoothing, double mass);
void orc_monitor(orc_state *S, int q, double Xplanet, double Yplanet, double Zplanet, double rsmoothing);
void orc_Potential(orc_state *S);
void orc_CorrectVtheta(orc_state *S, double domega);
void orc_SubStep1_x(orc_state *S, double dt);
void orc_SubStep1_y(orc_state *S, double dt);
void orc_SubStep1_z(orc_state *S, double dt);
void orc_visctensor_cyl(orc_state *S);
void orc_addviscosity_cyl(orc_state *S, double dt);
void orc_visctensor_sph(orc_state *S);
void orc_addviscosity_sph(orc_state *S, double dt);
void orc_SubStep2_a(orc_state *S, double dt);
void orc_SubStep2_b(orc_state *S, double dt);
void orc_SubStep3(orc_state *S, double dt);
void orc_copy_velocities(orc_state *S, int option);
void orc_boundaries(orc_state *S);
void orc_FillGhosts(orc_state *S);
void orc_momenta_x(orc_state *S);
void orc_momenta_y(orc_state *S);
void orc_momenta_z(orc_state *S);
void orc_DivideByRho(orc_state *S, const double *Q);
void orc_VanLeerY_a(orc_state *S, const double *Q);
void orc_VanLeerY_b(orc_state *S, double dt, const double *Q, double *Qs);
void orc_VanLeerZ_a(orc_state *S, const double *Q);
void orc_VanLeerZ_b(orc_state *S, double dt, const double *Q, double *Qs);
void orc_VanLeerX_a(orc_state *S, const double *Q);
void orc_VanLeerX_b(orc_state *S, double dt, const double *Q, double *Qs, const double *Vx_t);
void orc_VanLeerX_PPA_a(orc_state *S, const double *Q);
void orc_VanLeerX_PPA_b(orc_state *S, const double *Q);
void orc_VanLeerX_PPA_c(orc_state *S, const double *Q);
void orc_VanLeerX_PPA_d(orc_state *S, double dt, const double *Q, double *Qs, const double *Vx_t);
void orc_UpdateX(orc_state *S, double dt, double *Q, const double *Qs, const double *Vx_t);
void orc_UpdateY(orc_state *S, double dt, double *Q, const double *Qs);
void orc_UpdateZ(orc_state *S, double dt, double *Q, const double *Qs);
void orc_UpdateDensityX(orc_state *S, double dt, double *Q, const double *Vx_t);
void orc_UpdateDensityY(orc_state *S, double dt, double *Q);
void orc_UpdateDensityZ(orc_state *S, double dt, double *Q);
void orc_ComputeResidual(orc_state *S, double dt);
void orc_AdvectSHIFT(orc_state *S, double *F, const int *NS);
void orc_NewVelocity_x(orc_state *S);
void orc_NewVelocity_y(orc_state *S);
void orc_NewVelocity_z(orc_state *S);
void orc_StockholmBoundary(orc_state *S, double dt);
void orc_init_stockholm(orc_state *S);
void orc_Sources(orc_state *S, double dt);
void orc_Collisions(orc_state *S, double dt);
void orc_DustDiffusion_Coefficients(orc_state *S);
void orc_DustDiffusion_Core(orc_state *S, double dt);
void orc_DustDiffusion_Main(orc_state *S, double dt);
void orc_Transport(orc_state *S, double dt);

/* field access for the Python side */
double *orc_field(orc_state *S, const char *name);
int *orc_nshift(orc_state *S);
void orc_get_scalars(orc_state *S, double *out8);
void orc_set_scalars(orc_state *S, const double *in8);
void orc_get_planets(orc_state *S, double *out, int maxn);
void orc_psys_integrate(orc_state *S, double t, int nsub);
void orc_shape(orc_state *S, int *out4);
const double *orc_array(orc_state *S, const char *name, int *len);
void orc_get_localforce(orc_state *S, double *out12);
void orc_set_planet(orc_state *S, int n, const double *in7);

#ifdef __cplusplus
}
#endif
#endif
