/*
 * f3d_host.h -- host-side data model of the FARGO3D hydro step (plain C).
 *
 * This header holds what both the HIP product path and any CPU checker need
 * on the host: the run parameters (the .par file + the compile-time options
 * FARGO3D takes from a setup's .opt file), the mesh and its metric arrays,
 * the planetary system and its RK5 integrator, and the initial conditions of
 * the supported setups.  Nothing here touches a GPU.
 *
 * Reference data model: Field/mesh globals in src/global.h:210-265, mesh
 * construction src/LowTasks.c:198-563, parameters src/param.c:98-217.
 */
#ifndef F3D_HOST_H
#define F3D_HOST_H

#include <stddef.h>
#include <stdint.h>

#include "fargo3d_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Ghost widths (src/define.h:57-63). X has no ghosts (GHOSTSX off). */
#define F3D_NGH 3
#define F3D_MAX_FLUIDS 8
#define F3D_MAX_PLANETS 16

/* Scale-free units (src/fondam.h: G_SF, MSTAR_SF, R0_SF). */
#define F3D_G 1.0
#define F3D_MSTAR 1.0
#define F3D_R0 1.0
/* CVNR (src/define.h:9) */
#define F3D_CVNR 1.41

enum f3d_geometry { F3D_CYLINDRICAL = 1, F3D_SPHERICAL = 2 };
enum f3d_eos { F3D_ISOTHERMAL = 1, F3D_ADIABATIC = 2 };
enum f3d_visc { F3D_VISC_NONE = 0, F3D_VISC_NU = 1, F3D_VISC_ALPHA = 2 };
enum f3d_fluidtype { F3D_GAS = 1, F3D_DUST = 2 };
/* monitored quantities (the bit values of src/define.h:121-132) */
#define F3D_MON_MASS 1
#define F3D_MON_MOM_X 64
#define F3D_MON_MOM_Y 128
#define F3D_MON_MOM_Z 256
#define F3D_MON_TORQ 512
/* Boundary families (std/boundaries.txt:35-102, setups/<name>/<name>.bound.0) */
enum f3d_bc { F3D_BC_FARGO2D = 1, F3D_BC_P3D = 2 };

/* Run parameters: the .par values used on the hydro path plus the
 * compile-time switches FARGO3D reads from a setup's .opt file. */
typedef struct f3d_params {
  char setup[64];
  /* mesh (src/LowTasks.c:198-563) */
  int nx, ny, nz;
  double xmin, xmax, ymin, ymax, zmin, zmax;
  char spacing[16];            /* "Lin" or "Log" */
  /* disk */
  double aspectratio, sigma0, sigmaslope, flaringindex, nu, alpha, gamma;
  double noise;
  /* numerics */
  double cfl;
  /* planets */
  char planetconfig[256];
  double thicknesssmoothing, rochesmoothing, eccentricity, inclination;
  int indirectterm, excludehill;
  double masstaper, releasedate, releaseradius, semimajoraxis, orbitalradius, planetmass;
  /* frame */
  char frame[16];
  double omegaframe;
  /* damping */
  double dampingzone, taudamp, killingbccolatitude;
  int periodicz, periodicy;
  /* outputs / loop */
  double dt_out;               /* DT: physical time between outputs */
  int ninterm, ntot;
  int nsnap;                   /* NSNAP: snapshots every NSNAP steps (0: none) */
  char funcarchfile[256];
  char outputdir[256];
  /* .opt switches (compile-time in FARGO3D, run-time here) */
  int geometry;                /* F3D_CYLINDRICAL / F3D_SPHERICAL */
  int dimz;                    /* -DZ */
  int eos;                     /* -DISOTHERMAL / -DADIABATIC */
  int visc;                    /* -DVISCOSITY / -DALPHAVISCOSITY */
  int potential;               /* -DPOTENTIAL */
  int stockholm;               /* -DSTOCKHOLM */
  int fargo;                   /* 1 = FARGO orbital advection, 0 = -DSTANDARD */
  int bc;                      /* boundary family */
  int energy_bc;               /* 1: Energy ghosts SYMMETRIC (|a|a|) on the Y (and Z) sides */
  int precision;               /* 64: FP64 fields (the drop-in); 32: FP32 fields, FP64 sums/residual/CFL */
  int nfluids;                 /* fluid 0 is the gas, the others dust (setups/fargo_multifluid) */
  double invstokes[F3D_MAX_FLUIDS]; /* Invstokes1.. : inverse Stokes number of dust fluid k */
  double epsilon;              /* dust-to-gas density ratio of each dust fluid */
  int dragforce;               /* -DDRAGFORCE with -DSTOKESNUMBER (implicit drag, src/collisions.c) */
  int dustdiffusion;           /* -DDUSTDIFFUSION (src/dust_diffusion_main.c) */
  int feedback;                /* the gas feels the drag of the dust (ColRate feedback, src/colrate.c) */
  int ic_keplerian;            /* 1: v_phi = r^-1/2 - Omega_f r, v_r = 0, noise on density only */
  int mon_2d, mon_y_raw, mon_scalar; /* MONITOR_2D, MONITOR_Y_RAW, MONITOR_SCALAR of the .opt file */
  uint64_t seed;               /* srand48 seed for the noise (0 = glibc default state) */
  int ic_device;               /* CondInitOnDevice: the initial state computed on the GPU (f3d_CondInit) */
  /* derived (src/param.c:236-240) */
  int corotating, guidingcenter;
} f3d_params;

/* Mesh and metric arrays (src/global.h:210-242; src/LowTasks.c:198-563).
 * Extra tables hold the glibc cos/sin of the 1-D coordinates so that a
 * device never evaluates a transcendental of a coordinate. */
typedef struct f3d_grid {
  int nx, ny, nz;              /* active cells */
  int nghy, nghz;
  int dimz;
  int geometry;
  int sx, sy, sz;              /* nx, ny+2nghy, nz+2nghz */
  int pitch, stride;           /* elements */
  size_t ncell;                /* sx*sy*sz */
  size_t n2d;                  /* sy*sz */
  double XMIN, XMAX, YMIN, YMAX, ZMIN, ZMAX;
  double *xmin, *xmed, *ymin, *ymed, *zmin, *zmed;
  double *sxi, *sxj, *syj, *szj, *sxk, *syk, *szk, *invvj, *invdiffxmed;
  double *cosxmed, *sinxmed;   /* nx */
  double *sinzmed, *coszmed, *sinzmin, *coszmin; /* sz (+1 for zmin) */
  double mindx;
} f3d_grid;

/* Planetary system (src/structs.h PlanetarySystem; src/psys.c). */
typedef struct f3d_psys {
  int nb;
  double x[F3D_MAX_PLANETS], y[F3D_MAX_PLANETS], z[F3D_MAX_PLANETS];
  double vx[F3D_MAX_PLANETS], vy[F3D_MAX_PLANETS], vz[F3D_MAX_PLANETS];
  double mass[F3D_MAX_PLANETS], acc[F3D_MAX_PLANETS];
  int feeldisk[F3D_MAX_PLANETS], feelothers[F3D_MAX_PLANETS];
  /* GetPsysInfo's static MARK state (src/psys.c:370) */
  double X_planet, Y_planet;
  /* RK5 stage vectors k1..k6; static arrays in the reference
     (src/runge_kutta.c:8) that keep their values between calls, where the
     zero coefficients of TranslatePlanetRK5 multiply them */
  double rk[6][6 * F3D_MAX_PLANETS];
} f3d_psys;

/* Host scalars that evolve with the run (src/global.h). */
typedef struct f3d_clock {
  double PhysicalTime;
  double OMEGAFRAME, OMEGAFRAME0, Domega;
  double XAxisRotationAngle;
  long Timestepcount;
} f3d_clock;

/* ---- parameters ---- */
void f3d_params_defaults(f3d_params *p);
/* Apply the .opt-equivalent switches of a named setup ("fargo", "p3disof",
 * "fargo_adiabatic", ...). Returns 0, or -1 for an unknown setup. */
int f3d_params_setup_options(f3d_params *p, const char *setup);
/* Read a .par file (src/param.c:98-217). Returns 0 or -1 (message in err). */
int f3d_params_read(f3d_params *p, const char *path, char *err, size_t errlen);
/* Set one variable by name, case-insensitive (the -o "name=value" path,
 * src/redefoptions.c). Returns 0, or -1 for an unknown name. */
int f3d_params_set(f3d_params *p, const char *name, const char *value);
/* Derived flags (src/param.c:236-240). */
void f3d_params_finalize(f3d_params *p);

/* ---- mesh ---- */
int f3d_grid_init(f3d_grid *g, const f3d_params *p);
void f3d_grid_free(f3d_grid *g);

/* ---- planets (src/psys.c, src/runge_kutta.c) ---- */
int f3d_psys_init(f3d_psys *sys, const f3d_params *p, char *err, size_t errlen);
/* Add one planet as a .cfg line would (dist, mass, accretion, feeldisk, feelothers). */
void f3d_psys_add(f3d_psys *sys, const f3d_params *p, double dist, double mass,
                  double accret, int feeldisk, int feelothers);
enum { F3D_PSYS_GET = 0, F3D_PSYS_MARK = 1, F3D_PSYS_FREQUENCY = 2 };
double f3d_psys_info(f3d_psys *sys, const f3d_params *p, int action);
void f3d_psys_rotate(f3d_psys *sys, f3d_clock *c, double angle);
void f3d_psys_rk5(f3d_psys *sys, const f3d_params *p, const f3d_clock *c, double dt);
/* Planet smoothing length as Potential_cpu computes it (src/potential.c:120-131):
 * returns smoothing^2 for planet n at the current position. */
double f3d_planet_smoothing2(const f3d_psys *sys, const f3d_params *p, int n, double taper);
double f3d_planet_taper(const f3d_params *p, double PhysicalTime);

/* compute_potential's host part after the Potential kernel
 * (src/potential.c:32-48): RK5 sub-cycles, frame update. */
void f3d_host_advance_planets(f3d_psys *sys, const f3d_params *p, f3d_clock *c, double dt);

/* ---- initial conditions ---- */
/* Fill density/energy/vx/vy/vz (size grid->ncell each; vz may be NULL in 2-D)
 * as the setup's condinit.c does. */
int f3d_initial_conditions(const f3d_grid *g, const f3d_params *p, const f3d_clock *c,
                           double *dens, double *energy, double *vx, double *vy, double *vz);
/* the same for fluid type F3D_GAS or F3D_DUST of a multifluid setup */
int f3d_initial_conditions_fluid(const f3d_grid *g, const f3d_params *p, const f3d_clock *c, int fluidtype,
                                 double *dens, double *energy, double *vx, double *vy, double *vz);
/* The two parts of those: the per-ring profiles R, V, CS, E (sy*sz each,
 * caller's arrays) and the scalars of the set-up's per-cell noise into d
 * (its R/V/CS/E pointers are left for the caller to set), and the
 * arithmetic of one cell from its draws u[0..ndraw) -- the device CondInit
 * (f3d_CondInit) repeats it operation for operation. */
int f3d_ic_profiles(const f3d_grid *g, const f3d_params *p, const f3d_clock *c, int fluidtype, f3d_ic_desc *d,
                    double *R, double *V, double *CS, double *E);
void f3d_ic_cell(const f3d_ic_desc *d, size_t l2, const double *u, double *rho, double *e, double *vx,
                 double *vy);
/* drand48's state after srand48(seed) (glibc's initial state for seed 0),
 * and the jump table of ndraw draws: entry b maps a state X to the state
 * 2^b * ndraw draws later, X -> (A[b] X + C[b]) mod 2^48 */
unsigned long long f3d_drand48_state(uint64_t seed);
void f3d_drand48_jumps(int ndraw, unsigned long long A[48], unsigned long long C[48]);

/* ---- boundary tables ----
 * The generated boundary kernels (std/boundaries.txt, scripts/boundparser.py)
 * evaluate pow/exp/sin/cos/sqrt of mesh coordinates. The host evaluates
 * those factors once with glibc; kernels then only multiply. */
typedef struct f3d_bctab {
  /* Y sides, index m = 0..nghy-1 is the loop index j of the generated kernel */
  double ymin_dens[F3D_NGH], ymax_dens[F3D_NGH];      /* 2-D: pow(Ymed(jact)/Ymed(jgh), SIGMASLOPE) */
  double ymin_vsq[F3D_NGH], ymax_vsq[F3D_NGH];        /* sqrt(Ymed(jact)/Ymed(jgh)) */
  double ymin_vfac[F3D_NGH], ymax_vfac[F3D_NGH];      /* 3-D: (1.+...) factor of KEPLERIAN3DSPHVAZIMRAD */
  double *ymin_dexp, *ymax_dexp;  /* 3-D: [m*sz+k] exp(...) factor of KEPLERIAN3DSPHDENSRAD */
  /* Z sides (3-D), index [m*sy + j] */
  double *zmin_dens, *zmax_dens;  /* pow(sin(Zmed(kgh))/sin(Zmed(kact)), ...) */
  double zmin_vfac[F3D_NGH], zmax_vfac[F3D_NGH];      /* (1.+f*cos(..)*(..)) */
} f3d_bctab;
int f3d_bctab_init(f3d_bctab *t, const f3d_grid *g, const f3d_params *p);
void f3d_bctab_free(f3d_bctab *t);

/* Stockholm damping setup (src/stockholm.c:62-82): per-j tau and ramps. */
typedef struct f3d_damp {
  double Y_inf, Y_sup, Z_inf, Z_sup;
} f3d_damp;
void f3d_damp_init(f3d_damp *d, const f3d_params *p);

/* drand48 wrapper used by the setups (glibc; srand48 when seed != 0). */
void f3d_seed(uint64_t seed);

#ifdef __cplusplus
}
#endif
#endif
