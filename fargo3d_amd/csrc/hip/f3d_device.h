// f3d_device.h -- device-side mesh view and metric forms for the gfx950 kernels.
//
// Layout: every 3-D field is a dense x-fastest array of sx*sy*sz reals with
// l = i + j*pitch + k*stride (src/define.h:414-420, no X ghosts); 2-D (YZ)
// fields use l2 = j + k*sy.  Coordinates and metric factors live in one
// device buffer described by f3d_dmesh (the device copy FARGO3D makes in
// LightGlobalDev, src/light_global_dev.c:5-23), plus host-evaluated glibc
// tables of cos/sin of the coordinates so that no kernel evaluates a
// transcendental of a coordinate.
//
// Association: FARGO3D's CPU metric macros have no outer parentheses
// (src/define.h:228-236), so "x*dt*InvVol(i,j,k)" rounds as
// ((x*dt)*InvVj[j])/(Syk[k]*Sxi[i]).  The helpers below take the left
// operand and reproduce that order exactly; kernels are compiled with
// -ffp-contract=off so no multiply-add is fused.
#pragma once

#include <hip/hip_runtime.h>

namespace f3d {

enum { CYL = 1, SPH = 2 };

struct dmesh {
  int nx, ny, nz, nghy, nghz, dimz, geom;
  int sx, sy, sz, pitch, stride;
  double XMIN, XMAX, mindx;
  const double *xmin, *ymin, *zmin, *xmed, *ymed, *zmed;
  const double *sxi, *sxj, *syj, *szj, *sxk, *syk, *szk, *invvj, *invdiffxmed;
  const double *cosxmed, *sinxmed, *sinzmed, *coszmed, *sinzmin, *coszmin;
};

__device__ __forceinline__ int lxp_of(int l, int i, int nx) { return (i < nx - 1) ? l + 1 : l - (nx - 1); }
__device__ __forceinline__ int lxm_of(int l, int i, int nx) { return (i > 0) ? l - 1 : l + (nx - 1); }
__device__ __forceinline__ int ixm_of(int i, int nx) { return i > 0 ? i - 1 : nx - 1; }
__device__ __forceinline__ int ixp_of(int i, int nx) { return i < nx - 1 ? i + 1 : 0; }

// x * InvVol(i,j,k)  ->  (x*InvVj[j]) / (Syk[k]*Sxi[i])
__device__ __forceinline__ double mul_invvol(const dmesh& m, double x, int i, int j, int k) {
  return (x * m.invvj[j]) / (m.syk[k] * m.sxi[i]);
}
// x * Vol(i,j,k)  ->  ((x*Syk[k])*Sxi[i]) / InvVj[j]
__device__ __forceinline__ double mul_vol(const dmesh& m, double x, int i, int j, int k) {
  return ((x * m.syk[k]) * m.sxi[i]) / m.invvj[j];
}
// Vol(i,j,k) standing alone  ->  (Syk[k]*Sxi[i]) / InvVj[j]
__device__ __forceinline__ double vol(const dmesh& m, int i, int j, int k) {
  return (m.syk[k] * m.sxi[i]) / m.invvj[j];
}
// x * SurfX(j,k)
__device__ __forceinline__ double mul_surfx(const dmesh& m, double x, int j, int k) {
  return (x * m.sxj[j]) * m.sxk[k];
}
// x * SurfY(i,j,k)
__device__ __forceinline__ double mul_surfy(const dmesh& m, double x, int i, int j, int k) {
  return ((x * m.syj[j]) * m.syk[k]) * m.sxi[i];
}
// x * SurfZ(i,j,k)
__device__ __forceinline__ double mul_surfz(const dmesh& m, double x, int i, int j, int k) {
  return ((x * m.szj[j]) * m.szk[k]) * m.sxi[i];
}
template <int G>
__device__ __forceinline__ double zone_size_x(const dmesh& m, int i, int j, int k) {
  if (G == SPH) return (m.sxi[i] * m.ymed[j]) * m.sinzmed[k];
  return m.sxi[i] * m.ymed[j];
}
template <int G>
__device__ __forceinline__ double inv_zone_size_xmed(const dmesh& m, int i, int j, int k) {
  if (G == SPH) return (m.invdiffxmed[i] / m.ymed[j]) / m.sinzmed[k];
  return m.invdiffxmed[i] / m.ymed[j];
}
__device__ __forceinline__ double zone_size_y(const dmesh& m, int j) { return m.ymin[j + 1] - m.ymin[j]; }
template <int G>
__device__ __forceinline__ double zone_size_z(const dmesh& m, int j, int k) {
  if (G == SPH) return (m.zmin[k + 1] - m.zmin[k]) * m.ymed[j];
  return m.zmin[k + 1] - m.zmin[k];
}

// van Leer harmonic slope (src/vanleer_y_a.c:61-68): 0 at extrema.
__device__ __forceinline__ double vanleer_slope(double dqm, double dqp) {
  if (dqp * dqm <= 0.0) return 0.0;
  return 2. * dqp * dqm / (dqm + dqp);
}
// upwind interface value (src/vanleer_y_b.c:60-65)
__device__ __forceinline__ double vanleer_face(double v, double dt, double qm, double sm, double dzm,
                                               double q, double s, double dz) {
  if (v > 0.) return qm + 0.5 * (dzm - v * dt) * sm;
  return q - 0.5 * (dz + v * dt) * s;
}

}  // namespace f3d
