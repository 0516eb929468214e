// f3d_device_body.h -- device-side mesh view and metric forms, written once
// for the storage/arithmetic type `real` of a build (included by
// f3d_device.h with real = double in namespace f3d, and by f3d_device32.h
// with real = float in namespace f3d32; no include guard on purpose).
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

namespace F3D_NS {

typedef F3D_REAL real;
#ifdef F3D_MESH_REAL
typedef F3D_MESH_REAL mreal;
#else
typedef F3D_REAL mreal;
#endif

enum { CYL = 1, SPH = 2 };

struct dmesh {
  int nx, ny, nz, nghy, nghz, dimz, geom;
  int sx, sy, sz, pitch, stride;
  double XMIN, XMAX, mindx;
  const mreal *xmin, *ymin, *zmin, *xmed, *ymed, *zmed;
  const mreal *sxi, *sxj, *syj, *szj, *sxk, *syk, *szk, *invvj, *invdiffxmed;
  const mreal *cosxmed, *sinxmed, *sinzmed, *coszmed, *sinzmin, *coszmin;
  // FP64 copies for the parts that stay FP64 in the reduced-precision form
  // (ring sums, residual/shift split, CFL); the same arrays in the FP64 build
  const double *xmin64, *ymin64, *zmin64, *ymed64, *sinzmed64, *sxi64, *zmed64, *sinzmin64, *coszmin64;
  // precision form only: V0(j,k), the FP64 ring profile the azimuthal
  // velocities are stored as deviations from (index j + k*sy); null in FP64
  const double* vb;
};

__device__ __forceinline__ int lxp_of(int l, int i, int nx) { return (i < nx - 1) ? l + 1 : l - (nx - 1); }
__device__ __forceinline__ int lxm_of(int l, int i, int nx) { return (i > 0) ? l - 1 : l + (nx - 1); }
__device__ __forceinline__ int ixm_of(int i, int nx) { return i > 0 ? i - 1 : nx - 1; }
__device__ __forceinline__ int ixp_of(int i, int nx) { return i < nx - 1 ? i + 1 : 0; }

// Division by a divisor that several quotients share, correctly rounded: with
// r = RN(1/d) and q0 = RN(a r), the remainder a - q0 d is exact (one FMA)
// and q0 + r (a - q0 d), rounded once (a second FMA), is RN(a/d) -- the
// correction step of Markstein's division (round to nearest, d finite and
// nonzero, no overflow or underflow).  Checked with exact integer arithmetic
// for every pair of significands of the binary formats of precision 4 to 16
// (1.4e9 pairs; q0 alone is wrong in 27% of them) and against IEEE division
// on 6e8 random and near-midpoint binary64 pairs.  An exact quotient
// (remainder 0, also a = -0) keeps q0.  So each quotient costs a multiply and
// two FMAs instead of the full division sequence, and equals a/d bit for bit.
struct Recip {
  real d, r;
};
__device__ __forceinline__ Recip recip(real d) { return Recip{d, real(1.0) / d}; }
__device__ __forceinline__ real rdiv(real a, const Recip& R) {
  const real q0 = a * R.r;
  const real rem = fma(-q0, R.d, a);
  return rem == real(0.0) ? q0 : fma(rem, R.r, q0);
}
// the same in FP64 in every build (the CFL time step)
struct RecipD {
  double d, r;
};
__device__ __forceinline__ RecipD recipd(double d) { return RecipD{d, 1.0 / d}; }
__device__ __forceinline__ double rdivd(double a, const RecipD& R) {
  const double q0 = a * R.r;
  const double rem = fma(-q0, R.d, a);
  return rem == 0.0 ? q0 : fma(rem, R.r, q0);
}

// x * InvVol(i,j,k)  ->  (x*InvVj[j]) / (Syk[k]*Sxi[i])
__device__ __forceinline__ real mul_invvol(const dmesh& m, real x, int i, int j, int k) {
  return (x * m.invvj[j]) / (m.syk[k] * m.sxi[i]);
}
// its divisor, for rdiv(): several quantities of one cell share it
__device__ __forceinline__ Recip invvol_recip(const dmesh& m, int i, int k) { return recip(m.syk[k] * m.sxi[i]); }
__device__ __forceinline__ real mul_invvol(const dmesh& m, real x, int j, const Recip& V) { return rdiv(x * m.invvj[j], V); }
// x * Vol(i,j,k)  ->  ((x*Syk[k])*Sxi[i]) / InvVj[j]
__device__ __forceinline__ real mul_vol(const dmesh& m, real x, int i, int j, int k) {
  return ((x * m.syk[k]) * m.sxi[i]) / m.invvj[j];
}
// the same with the ring's divisor InvVj[j] shared (rdiv())
__device__ __forceinline__ real mul_vol(const dmesh& m, real x, int i, int k, const Recip& IV) {
  return rdiv((x * m.syk[k]) * m.sxi[i], IV);
}
// Vol(i,j,k) standing alone  ->  (Syk[k]*Sxi[i]) / InvVj[j]
__device__ __forceinline__ real vol(const dmesh& m, int i, int j, int k) {
  return (m.syk[k] * m.sxi[i]) / m.invvj[j];
}
// x * SurfX(j,k)
__device__ __forceinline__ real mul_surfx(const dmesh& m, real x, int j, int k) {
  return (x * m.sxj[j]) * m.sxk[k];
}
// x * SurfY(i,j,k)
__device__ __forceinline__ real mul_surfy(const dmesh& m, real x, int i, int j, int k) {
  return ((x * m.syj[j]) * m.syk[k]) * m.sxi[i];
}
// x * SurfZ(i,j,k)
__device__ __forceinline__ real mul_surfz(const dmesh& m, real x, int i, int j, int k) {
  return ((x * m.szj[j]) * m.szk[k]) * m.sxi[i];
}
template <int G>
__device__ __forceinline__ real zone_size_x(const dmesh& m, int i, int j, int k) {
  if (G == SPH) return (m.sxi[i] * m.ymed[j]) * m.sinzmed[k];
  return m.sxi[i] * m.ymed[j];
}
template <int G>
__device__ __forceinline__ real inv_zone_size_xmed(const dmesh& m, int i, int j, int k) {
  if (G == SPH) return (m.invdiffxmed[i] / m.ymed[j]) / m.sinzmed[k];
  return m.invdiffxmed[i] / m.ymed[j];
}
__device__ __forceinline__ real zone_size_y(const dmesh& m, int j) { return m.ymin[j + 1] - m.ymin[j]; }
template <int G>
__device__ __forceinline__ real zone_size_z(const dmesh& m, int j, int k) {
  if (G == SPH) return (m.zmin[k + 1] - m.zmin[k]) * m.ymed[j];
  return m.zmin[k + 1] - m.zmin[k];
}

// FP64 zone sizes from the FP64 copies (the same operations as above)
template <int G>
__device__ __forceinline__ double zone_size_x64(const dmesh& m, int i, int j, int k) {
  if (G == SPH) return (m.sxi64[i] * m.ymed64[j]) * m.sinzmed64[k];
  return m.sxi64[i] * m.ymed64[j];
}
__device__ __forceinline__ double zone_size_y64(const dmesh& m, int j) { return m.ymin64[j + 1] - m.ymin64[j]; }
template <int G>
__device__ __forceinline__ double zone_size_z64(const dmesh& m, int j, int k) {
  if (G == SPH) return (m.zmin64[k + 1] - m.zmin64[k]) * m.ymed64[j];
  return m.zmin64[k + 1] - m.zmin64[k];
}

// The MIN of the CFL's n ring minima by one workgroup (blockDim.x a multiple
// of 64, at most 1024; MIN is order independent, so this equals one atomic
// MIN per ring) into result[slot], and, when host_out is given, straight into
// pinned host memory: the value, then, after a system-scope fence, the
// sequence number the host polls for.  sh: blockDim.x / 64 doubles of LDS.
__device__ inline void cfl_min_publish(const double* ringmin, int n, unsigned long long* result, int slot,
                                       double* host_out, unsigned long long seq, double* sh) {
  double v = __longlong_as_double(0x7FF0000000000000ll);  // +inf
  // eight loads per thread issued before any is compared; an index past the
  // end reads the last minimum again, which leaves the MIN unchanged
  for (int i0 = threadIdx.x; i0 < n; i0 += 8 * blockDim.x) {
    double x[8];
#pragma unroll
    for (int u = 0; u < 8; u++) x[u] = ringmin[min(i0 + u * (int)blockDim.x, n - 1)];
#pragma unroll
    for (int u = 0; u < 8; u++) v = (x[u] < v) ? x[u] : v;
  }
  for (int o = 32; o > 0; o >>= 1) {
    const double w = __shfl_down(v, o, 64);
    v = (w < v) ? w : v;
  }
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < (int)(blockDim.x >> 6); w++) v = (sh[w] < v) ? sh[w] : v;
    result[slot] = (unsigned long long)__double_as_longlong(v);
    if (host_out) {
      host_out[0] = v;
      __threadfence_system();
      *reinterpret_cast<volatile unsigned long long*>(host_out + 1) = seq;
    }
  }
}

// van Leer harmonic slope (src/vanleer_y_a.c:61-68): 0 at extrema.  The
// quotient is formed in every lane and the extremum case selected after it
// (the same value in both cases): a branch around the division would split
// the sweeps' loops into blocks that the next row's loads cannot be moved
// across.
__device__ __forceinline__ real vanleer_slope(real dqm, real dqp) {
  const real p = dqp * dqm;
  const real s = real(2.) * dqp * dqm / (dqm + dqp);
  return (p <= real(0.0)) ? real(0.0) : s;
}
// upwind interface value (src/vanleer_y_b.c:60-65)
__device__ __forceinline__ real vanleer_face(real v, real dt, real qm, real sm, real dzm,
                                               real q, real s, real dz) {
  if (v > real(0.)) return qm + real(0.5) * (dzm - v * dt) * sm;
  return q - real(0.5) * (dz + v * dt) * s;
}

}  // namespace F3D_NS

