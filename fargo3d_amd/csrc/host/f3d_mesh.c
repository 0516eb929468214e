/*
 * f3d_mesh.c -- mesh coordinates and metric arrays.
 *
 * Follows InitSpace (src/LowTasks.c:198-463, the linear and logarithmic
 * branches; fixed-file and mesh-density spacing are out of scope) and
 * InitSurfaces (src/LowTasks.c:466-563) for a sequential run (Y0 = Z0 = 0).
 * Transcendental tables (cos/sin of Xmed, Zmed, Zmin) are evaluated here
 * with glibc so that device kernels reproduce the CPU arithmetic exactly.
 */
#include "f3d_host.h"

#include <ctype.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

static double *alloc0(size_t n) { return (double *)calloc(n ? n : 1, sizeof(double)); }

int f3d_grid_init(f3d_grid *g, const f3d_params *p) {
  memset(g, 0, sizeof(*g));
  g->nx = p->nx;
  g->ny = p->ny;
  g->dimz = p->dimz;
  g->nz = p->dimz ? p->nz : 1;
  g->nghy = F3D_NGH;
  g->nghz = p->dimz ? F3D_NGH : 0;
  g->geometry = p->geometry;
  g->sx = g->nx;
  g->sy = g->ny + 2 * g->nghy;
  g->sz = g->nz + 2 * g->nghz;
  g->pitch = g->sx;
  g->stride = g->sx * g->sy;
  g->ncell = (size_t)g->sx * g->sy * g->sz;
  g->n2d = (size_t)g->sy * g->sz;
  g->XMIN = p->xmin; g->XMAX = p->xmax;
  g->YMIN = p->ymin; g->YMAX = p->ymax;
  g->ZMIN = p->zmin; g->ZMAX = p->zmax;
  if (g->nx < 1 || g->ny < 1 || g->nz < 1) return -1;

  const int sx = g->sx, sy = g->sy, sz = g->sz;
  g->xmin = alloc0(sx + 1); g->xmed = alloc0(sx);
  g->ymin = alloc0(sy + 1); g->ymed = alloc0(sy);
  g->zmin = alloc0(sz + 1); g->zmed = alloc0(sz);
  g->sxi = alloc0(sx); g->invdiffxmed = alloc0(sx);
  g->sxj = alloc0(sy); g->syj = alloc0(sy); g->szj = alloc0(sy); g->invvj = alloc0(sy);
  g->sxk = alloc0(sz); g->syk = alloc0(sz); g->szk = alloc0(sz);
  g->cosxmed = alloc0(sx); g->sinxmed = alloc0(sx);
  g->sinzmed = alloc0(sz); g->coszmed = alloc0(sz);
  g->sinzmin = alloc0(sz + 1); g->coszmin = alloc0(sz + 1);

  /* InitSpace, linear / logarithmic branch (src/LowTasks.c:320-370) */
  double dx = (p->xmax - p->xmin) / p->nx;
  double dy = (p->ymax - p->ymin) / p->ny;
  double dz = p->dimz ? (p->zmax - p->zmin) / p->nz : 0.0;
  int logy = (toupper((unsigned char)p->spacing[0]) == 'L' && toupper((unsigned char)p->spacing[1]) == 'O');
  if (logy) {
    dy = (log(p->ymax) - log(p->ymin)) / p->ny;
    for (int j = 0; j < sy + 1; j++) g->ymin[j] = exp(log(p->ymin) + dy * (j - g->nghy));
  } else {
    for (int j = 0; j < sy + 1; j++) g->ymin[j] = p->ymin + dy * (j - g->nghy);
  }
  for (int k = 0; k < sz + 1; k++) g->zmin[k] = p->dimz ? p->zmin + dz * (k - g->nghz) : 0.0;
  for (int i = 0; i < sx + 1; i++) g->xmin[i] = p->xmin + dx * (i - 0);

  for (int i = 0; i < sx; i++) g->xmed[i] = 0.5 * (g->xmin[i + 1] + g->xmin[i]);
  for (int j = 0; j < sy; j++) g->ymed[j] = 0.5 * (g->ymin[j + 1] + g->ymin[j]);
  for (int k = 0; k < sz; k++) g->zmed[k] = 0.5 * (g->zmin[k + 1] + g->zmin[k]);
  for (int i = 1; i < sx; i++) g->invdiffxmed[i] = 1. / (g->xmed[i] - g->xmed[i - 1]);
  g->invdiffxmed[0] = 1. / (g->xmed[0] - (g->xmed[g->nx - 1] - (p->xmax - p->xmin)));

  /* InitSurfaces (src/LowTasks.c:466-563) */
  for (int i = 0; i < sx; i++) g->sxi[i] = g->xmin[i + 1] - g->xmin[i];
  g->mindx = g->sxi[0];
  for (int i = 1; i < sx; i++)
    if (g->sxi[i] < g->mindx) g->mindx = g->sxi[i];
  if (g->geometry == F3D_CYLINDRICAL) {
    for (int j = 0; j < sy; j++) {
      g->sxj[j] = g->ymin[j + 1] - g->ymin[j];
      g->syj[j] = g->ymin[j];
      g->szj[j] = 0.5 * (g->ymin[j + 1] * g->ymin[j + 1] - g->ymin[j] * g->ymin[j]);
      g->invvj[j] = 1.0 / g->szj[j];
    }
    for (int k = 0; k < sz; k++) {
      g->sxk[k] = p->dimz ? g->zmin[k + 1] - g->zmin[k] : 1.0;
      g->syk[k] = g->sxk[k];
      g->szk[k] = 1.0;
    }
  } else {
    for (int j = 0; j < sy; j++) {
      g->sxj[j] = 0.5 * (g->ymin[j + 1] * g->ymin[j + 1] - g->ymin[j] * g->ymin[j]);
      g->syj[j] = g->ymin[j] * g->ymin[j];
      g->szj[j] = 0.5 * (g->ymin[j + 1] * g->ymin[j + 1] - g->ymin[j] * g->ymin[j]);
      g->invvj[j] = 3. / ((g->ymin[j + 1] * g->ymin[j + 1] * g->ymin[j + 1] -
                           g->ymin[j] * g->ymin[j] * g->ymin[j]));
    }
    for (int k = 0; k < sz; k++) {
      if (p->dimz) {
        g->sxk[k] = g->zmin[k + 1] - g->zmin[k];
        g->syk[k] = cos(g->zmin[k]) - cos(g->zmin[k + 1]);
        g->szk[k] = sin(g->zmin[k]);
      } else {
        g->sxk[k] = g->syk[k] = g->szk[k] = 1.0;
      }
    }
  }
  /* glibc tables of the coordinate transcendentals */
  for (int i = 0; i < sx; i++) {
    g->cosxmed[i] = cos(g->xmed[i]);
    g->sinxmed[i] = sin(g->xmed[i]);
  }
  for (int k = 0; k < sz; k++) {
    g->sinzmed[k] = sin(g->zmed[k]);
    g->coszmed[k] = cos(g->zmed[k]);
  }
  for (int k = 0; k < sz + 1; k++) {
    g->sinzmin[k] = sin(g->zmin[k]);
    g->coszmin[k] = cos(g->zmin[k]);
  }
  return 0;
}

void f3d_grid_free(f3d_grid *g) {
  double **arrs[] = {&g->xmin, &g->xmed, &g->ymin, &g->ymed, &g->zmin, &g->zmed,
                     &g->sxi, &g->sxj, &g->syj, &g->szj, &g->sxk, &g->syk, &g->szk,
                     &g->invvj, &g->invdiffxmed, &g->cosxmed, &g->sinxmed,
                     &g->sinzmed, &g->coszmed, &g->sinzmin, &g->coszmin};
  for (size_t i = 0; i < sizeof(arrs) / sizeof(arrs[0]); i++) {
    free(*arrs[i]);
    *arrs[i] = NULL;
  }
}
