/*
 * fargo3d_api.h -- FARGO3D's side of the drop-in boundary, restated for the
 * binding test (TEST INFRASTRUCTURE: built into tests/bind/lib, never into
 * the product).
 *
 * The field types keep the members the binding reads (src/structs.h:67-140:
 * struct field, field2D, fieldint2D), the function-pointer table keeps the
 * hydro step's entries with the reference's pointer types
 * (src/global.h:303-416), and the globals are the ones its wrappers read
 * (src/global.h:210-265 mesh and sizes, the selected fluid's fields of
 * SelectFluid, src/LowTasks.c:565-593).  hip_bind.c binds the table to the
 * library the way INTEGRATION.md describes; fargo_host.c is a miniature
 * FARGO3D host that runs the DT loop body through the table.
 */
#ifndef FARGO3D_API_H
#define FARGO3D_API_H

#include "f3d_host.h"

typedef double real;

typedef struct field {
  char *name;
  real *field_cpu;
  real *field_gpu;
} Field;
typedef struct field2D {
  char *name;
  real *field_cpu;
  real *field_gpu;
} Field2D;
typedef struct fieldint2D {
  char *name;
  int *field_cpu;
  int *field_gpu;
} FieldInt2D;

/* src/global.h:303-416 (the hydro-step entries) */
extern void (*ComputePressureFieldIso)();
extern void (*ComputePressureFieldAd)();
extern void (*SubStep1_x)(real);
extern void (*SubStep1_y)(real);
extern void (*SubStep1_z)(real);
extern void (*SubStep2_a)(real);
extern void (*SubStep2_b)(real);
extern void (*SubStep3)(real);
extern void (*DivideByRho)(Field *);
extern void (*VanLeerX_a)(Field *);
extern void (*VanLeerX_b)(real, Field *, Field *, Field *);
extern void (*VanLeerY_a)(Field *);
extern void (*VanLeerY_b)(real, Field *, Field *);
extern void (*VanLeerZ_a)(Field *);
extern void (*VanLeerZ_b)(real, Field *, Field *);
extern void (*momenta_x)();
extern void (*momenta_y)();
extern void (*momenta_z)();
extern void (*reduction_SUM)(Field *, int, int, int, int);
extern void (*reduction_MIN)(Field *, int, int, int, int);
extern void (*UpdateX)(real, Field *, Field *, Field *);
extern void (*UpdateY)(real, Field *, Field *);
extern void (*UpdateZ)(real, Field *, Field *);
extern void (*UpdateDensityX)(real, Field *, Field *);
extern void (*UpdateDensityY)(real, Field *);
extern void (*UpdateDensityZ)(real, Field *);
extern void (*NewVelocity_x)();
extern void (*NewVelocity_y)();
extern void (*NewVelocity_z)();
extern void (*AdvectSHIFT)(Field *, FieldInt2D *);
extern void (*ComputeResidual)(real);
extern void (*ComputeVweight)(Field *, Field *);
extern void (*Potential)();
extern void (*CorrectVtheta)(real);
extern void (*cfl)(void);
extern void (*_ComputeForce)(real, real, real, real, real);
extern void (*copy_velocities)(int);
extern void (*VanLeerX_PPA_a)(Field *);
extern void (*VanLeerX_PPA_b)(Field *);
extern void (*VanLeerX_PPA_c)(Field *);
extern void (*VanLeerX_PPA_d)(real, Field *, Field *, Field *);
extern void (*Reset_field)(Field *);
extern void (*ComputeTotalDensity)();
extern void (*StockholmBoundary)(real);
extern void (*visctensor_cyl)();
extern void (*addviscosity_cyl)(real);
extern void (*visctensor_sph)();
extern void (*addviscosity_sph)(real);
extern void (*boundary_ymin[1])();
extern void (*boundary_ymax[1])();
extern void (*boundary_zmin[1])();
extern void (*boundary_zmax[1])();

/* globals the wrappers read (src/global.h) */
extern int Nx, Ny, Nz, NGHY, NGHZ, Pitch_gpu, Stride_gpu, Fluidtype, FluidIndex;
extern real XMIN, XMAX, Mindx, OMEGAFRAME, OMEGAFRAME0, PhysicalTime, Domega;
extern real *Xmin, *Ymin, *Zmin, *Xmed, *Ymed, *Zmed, *Sxi, *Sxj, *Syj, *Szj, *Sxk, *Syk, *Szk, *InvVj, *InvDiffXmed;
extern real Min[8];
extern Field *Density, *Energy, *Vx, *Vy, *Vz, *Vx_temp, *Vy_temp, *Vz_temp;
extern Field *Mpx, *Mmx, *Mpy, *Mmy, *Mpz, *Mmz, *Pot, *Slope, *DivRho, *DensStar, *Qs, *Pressure, *Total_Density, *QL,
    *QR;
extern Field2D *Density0, *Energy0, *Vx0, *Vy0, *Vz0, *Reduction2D, *VxMed;
extern FieldInt2D *Nshift;
extern real localforce[12];
/* the .par values and the host-side tables (param.h and friends) */
extern f3d_params P;
extern f3d_grid GRID;
extern f3d_bctab BCT;
extern f3d_damp DAMP;
extern f3d_psys Sys;

/* hip_bind.c */
void HipInit(void);
void HipBindFluid(void);
void HipFini(void);
void HipToDevice(void *dst, const void *src, size_t bytes);
void HipToHost(void *dst, const void *src, size_t bytes);
void ChangeArchHip(void);
extern char HipError[512];

#endif
