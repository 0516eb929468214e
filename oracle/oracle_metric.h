/*
 * oracle_metric.h -- mesh metric forms of the CPU build (TEST INFRASTRUCTURE).
 *
 * The reference defines its metric terms as textual macros without outer
 * parentheses (src/define.h:226-271, CPU branch), so an expression such as
 * "x*dt*InvVol(i,j,k)" associates as ((x*dt)*InvVj[j])/(Syk[k]*Sxi[i]).
 * Those forms are kept textually identical here because the association
 * decides the last bit of every update; the geometry switch is a run-time
 * flag (CYL / SPH) instead of an #ifdef.
 */
#ifndef F3D_ORACLE_METRIC_H
#define F3D_ORACLE_METRIC_H

#include <math.h>

/* Brings the mesh of state S into scope under the reference's names. */
#define ORC_MESH(S)                                                           \
  const f3d_grid *GR = &(S)->g;                                               \
  const int Nx = GR->nx, Ny = GR->ny, Nz = GR->nz;                            \
  const int NGHY = GR->nghy, NGHZ = GR->nghz;                                 \
  const int pitch = GR->pitch, stride = GR->stride;                           \
  const int DIMZ = GR->dimz;                                                  \
  const int SPH = (GR->geometry == F3D_SPHERICAL);                            \
  const double *Xmin = GR->xmin, *Ymin = GR->ymin, *Zmin = GR->zmin;          \
  const double *Xmed = GR->xmed, *Ymed = GR->ymed, *Zmed = GR->zmed;          \
  const double *Sxi = GR->sxi, *Sxj = GR->sxj, *Syj = GR->syj, *Szj = GR->szj; \
  const double *Sxk = GR->sxk, *Syk = GR->syk, *Szk = GR->szk;                \
  const double *InvVj = GR->invvj, *InvDiffXmed = GR->invdiffxmed;            \
  (void)Nx; (void)Ny; (void)Nz; (void)NGHY; (void)NGHZ; (void)pitch;          \
  (void)stride; (void)DIMZ; (void)SPH; (void)Xmin; (void)Ymin; (void)Zmin;    \
  (void)Xmed; (void)Ymed; (void)Zmed; (void)Sxi; (void)Sxj; (void)Syj;        \
  (void)Szj; (void)Sxk; (void)Syk; (void)Szk; (void)InvVj; (void)InvDiffXmed

#define xmin(i) Xmin[(i)]
#define ymin(i) Ymin[(i)]
#define zmin(i) Zmin[(i)]
#define xmed(i) Xmed[(i)]
#define ymed(i) Ymed[(i)]
#define zmed(i) Zmed[(i)]

/* src/define.h:228-236 (CPU) */
#define InvVol(i, j, k) InvVj[(j)] / (Syk[(k)] * Sxi[(i)])
#define Vol(i, j, k) Syk[(k)] * Sxi[(i)] / InvVj[(j)]
#define SurfX(j, k) Sxj[(j)] * Sxk[(k)]
#define SurfY(i, j, k) Syj[(j)] * Syk[(k)] * Sxi[(i)]
#define SurfZ(i, j, k) Szj[(j)] * Szk[(k)] * Sxi[(i)]

/* src/define.h:259-271 (CPU); CYLINDRICAL / SPHERICAL picked at run time */
#define zone_size_x(i, j, k) (SPH ? (Sxi[(i)] * Ymed[(j)] * sin(Zmed[(k)])) : (Sxi[(i)] * Ymed[(j)]))
#define Inv_zone_size_xmed(i, j, k) \
  (SPH ? (InvDiffXmed[(i)] / Ymed[(j)] / sin(Zmed[(k)])) : (InvDiffXmed[(i)] / Ymed[(j)]))
#define zone_size_y(j, k) (Ymin[(j) + 1] - Ymin[(j)])
#define zone_size_z(j, k) (SPH ? ((Zmin[(k) + 1] - Zmin[(k)]) * Ymed[(j)]) : (Zmin[(k) + 1] - Zmin[(k)]))
/* src/define.h:316-321 (SPHERICAL) */
#define edge_size_x_middley_lowz(i, j, k) (Sxi[(i)] * Ymed[(j)] * sin(Zmin[(k)]))

/* index forms, no X ghosts (src/define.h:390-420) */
#define lidx(i, j, k) ((i) + (j) * pitch + (k) * stride)
#define LXP(l, i) (((i) < (Nx - 1)) ? ((l) + 1) : ((l) - (Nx - 1)))
#define LXM(l, i) (((i) > 0) ? ((l) - 1) : ((l) + (Nx - 1)))
#define IXM(i) ((i) > 0 ? ((i) - 1) : Nx - 1)
#define IXP(i) ((i) < Nx - 1 ? ((i) + 1) : 0)
#define LYP(l) ((l) + pitch)
#define LYM(l) ((l) - pitch)
#define LZP(l) (DIMZ ? (l) + stride : (l))
#define LZM(l) (DIMZ ? (l) - stride : (l))
#define L2D(j, k) ((j) + (k) * (Ny + 2 * NGHY))

#define max2(a, b) ((a) > (b) ? (a) : (b))
#define max3(a, b, c) (((a) > (b) ? (a) : (b)) > (c) ? ((a) > (b) ? (a) : (b)) : (c))

#endif
