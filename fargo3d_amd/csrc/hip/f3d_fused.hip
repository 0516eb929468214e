// f3d_fused.hip -- the fast path: the hydro step in a handful of launches.
//
// The per-routine entry points (f3d_routines.hip) move every intermediate
// array through HBM, as the reference's one-launch-per-routine GPU build
// does (about 370 array passes per step).  The kernels here compute the same
// values with the same per-cell arithmetic (f3d_cell.h) but keep
// intermediates on chip:
//
//   kf_cfl         ring-owning: VxMed (sequential ring sum), per-cell CFL,
//                  ring minima, and in the last workgroup the global MIN and
//                  shear limit -> dt (one launch, one 8-byte copy to host)
//   kf_src_a       viscous stress, artificial pressure and potential at
//                  every cell the next kernel reads
//   kf_src_b       SubStep1 + addviscosity + SubStep2_b on active cells
//   (boundaries)   mid-step ghost fill on V_temp
//   kf_sweep_z     3-D: momenta + Z sweep, one thread per (i,j) column
//                  marching in k with a register window
//   kf_sweep_y     momenta (2-D) + Y sweep, columns marching in j
//   kf_sweep_x     ring-owning: VxMed, residual, van Leer and PPA X sweeps,
//                  integer shift and NewVelocity_x in LDS
//   kf_finish      NewVelocity_y/z and Stockholm damping
//   (boundaries)   end-of-step ghost fill
//
// Only the active cells of the primitive fields carry information from one
// stage to the next: every ghost value is rewritten by the ghost fill from
// active values (std/boundaries.txt formulas read mirror rows only), so the
// fused kernels evaluate ghost rows only where an active cell's stencil
// reads them, and there with the reference's formulas.
#include <cstring>

#include "f3d_internal.h"
#include "f3d_cell.h"

using namespace f3d;

namespace {

constexpr int WG = 256;
constexpr int BXF = 64;

// fused scratch buffers (allocated on first use, each n3 doubles)
enum { FS_RHOY = 0, FS_RHOX, FS_QX, FS_QY, FS_QZ, FS_Z0, FS_COUNT = FS_Z0 + 7 };

struct Scratch {
  double* f[FS_COUNT];
  double* ringmin;
  unsigned* counter;
  double* result;
};

Scratch g_scr[16];
void* g_owner[16];

int scratch_of(f3d_hip_ctx* ctx, Scratch** out) {
  int slot = -1;
  for (int s = 0; s < 16; s++)
    if (g_owner[s] == ctx) slot = s;
  if (slot < 0) {
    for (int s = 0; s < 16 && slot < 0; s++)
      if (!g_owner[s]) slot = s;
    if (slot < 0) return -1;
    Scratch& S = g_scr[slot];
    memset(&S, 0, sizeof(S));
    for (int a = 0; a < FS_COUNT; a++) {
      if (hipMalloc(&S.f[a], ctx->n3 * sizeof(double)) != hipSuccess) return -1;
      if (hipMemset(S.f[a], 0, ctx->n3 * sizeof(double)) != hipSuccess) return -1;
    }
    if (hipMalloc(&S.ringmin, (ctx->n2 + 64) * sizeof(double)) != hipSuccess) return -1;
    if (hipMalloc(&S.counter, 64 * sizeof(unsigned)) != hipSuccess) return -1;
    if (hipMemset(S.counter, 0, 64 * sizeof(unsigned)) != hipSuccess) return -1;
    if (hipMalloc(&S.result, 64 * sizeof(double)) != hipSuccess) return -1;
    g_owner[slot] = ctx;
  }
  *out = &g_scr[slot];
  return 0;
}

__device__ __forceinline__ double wave_min(double v) {
  for (int o = 32; o > 0; o >>= 1) {
    double w = __shfl_down(v, o, 64);
    v = (w < v) ? w : v;
  }
  return v;
}

// Block-wide min; every thread gets the result.
__device__ double block_min(double v, double* sh) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
  v = wave_min(v);
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  double r = 1e20;
  if (threadIdx.x == 0) {
    for (int w = 0; w < nw; w++) r = (sh[w] < r) ? sh[w] : r;
    sh[0] = r;
  }
  __syncthreads();
  r = sh[0];
  __syncthreads();
  return r;
}

}  // namespace

// ----------------------------------------------------------------- kf_cfl
// One workgroup per ring (j,k), j in [nghy, ny+nghy] (the extra ring feeds the
// shear limit of the last active ring), k active.
template <int G, int DZ>
__global__ void __launch_bounds__(WG) kf_cfl(dmesh m, const double* cs, const double* rho, const double* vx,
                                             const double* vy, const double* vz, double* vxmed, double* ringmin,
                                             unsigned* counter, double* result, cell::CflPar cp) {
  extern __shared__ double lds[];  // nx weights + 8 for reductions
  double* w = lds;
  double* sh = lds + m.nx;
  const int nrj = m.ny + 1;
  const int j = m.nghy + (int)(blockIdx.x % nrj);
  const int k = DZ ? m.nghz + (int)(blockIdx.x / nrj) : 0;
  const int base = j * m.pitch + k * m.stride;
  const int l2 = j + k * m.sy;
  // ComputeVweight + reduction_SUM: Qs = v*(xmin(i+1)-xmin(i)), summed in i order
  for (int i = threadIdx.x; i < m.nx; i += blockDim.x) w[i] = vx[base + i] * (m.xmin[i + 1] - m.xmin[i]);
  __syncthreads();
  if (threadIdx.x == 0) {
    double acc = 0.0;
    for (int i = 0; i < m.nx; i++) acc = acc + w[i];
    double vm = acc / (m.XMAX - m.XMIN);
    vxmed[l2] = vm;
    sh[8] = vm;
  }
  __syncthreads();
  double mn = 1e20;
  if (j < m.ny + m.nghy) {
    for (int i = threadIdx.x; i < m.nx; i += blockDim.x) {
      double d = cell::cfl_cell<G, DZ>(m, i, j, k, base + i, cs, rho, vx, vy, vz, sh[8], cp);
      mn = (d < mn) ? d : mn;
    }
  }
  mn = block_min(mn, sh);
  if (threadIdx.x == 0) ringmin[blockIdx.x] = mn;
  // last workgroup: global MIN + shear limit (src/cfl_b.c:6-37)
  __shared__ int last;
  __threadfence();
  if (threadIdx.x == 0) last = (atomicAdd(counter, 1u) == gridDim.x - 1);
  __syncthreads();
  if (!last) return;
  __threadfence();
  double r = 1e20;
  for (int b = threadIdx.x; b < (int)gridDim.x; b += blockDim.x) {
    double a = __hip_atomic_load(&ringmin[b], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    r = (a < r) ? a : r;
  }
  if (cp.fargo) {
    const int nk = DZ ? m.nz : 1;
    for (int t = threadIdx.x; t < m.ny * nk; t += blockDim.x) {
      int jj = m.nghy + t % m.ny, kk = DZ ? m.nghz + t / m.ny : 0;
      int q2 = jj + kk * m.sy;
      double a = __hip_atomic_load(&vxmed[q2], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      double b = __hip_atomic_load(&vxmed[q2 + 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      double shearlimit = cp.cfl * m.mindx / fabs(a / m.ymed[jj] - b / m.ymed[jj + 1]);
      r = (shearlimit < r) ? shearlimit : r;
    }
  }
  r = block_min(r, sh);
  if (threadIdx.x == 0) {
    result[0] = r;
    *counter = 0u;  // ready for the next call, no memset needed
  }
}

// ------------------------------------------------------------------ sources
// Loader of the pressure field computed on the fly (src/CompPresIso.c:56,
// src/CompPresAd.c:58).
struct PresLoader {
  cell::GL rho, e;
  double gm1;
  int iso;
  __device__ __forceinline__ double operator()(int di, int dj, int dk) const {
    if (iso) {
      double c = e(di, dj, dk);
      return rho(di, dj, dk) * c * c;
    }
    return gm1 * e(di, dj, dk);
  }
};

// Loader of Vx after CorrectVtheta (src/correct_vtheta.c:51-56) applied on the fly.
template <int G>
struct VxcLoader {
  cell::GL v;
  const double* ymed;
  const double* sinzmed;
  int j, k;
  double domega;
  __device__ __forceinline__ double operator()(int di, int dj, int dk) const {
    if (G == SPH) return v(di, dj, dk) - domega * ymed[j + dj] * sinzmed[k + dk];
    return v(di, dj, dk) - domega * ymed[j + dj];
  }
};

template <int G>
__device__ __forceinline__ VxcLoader<G> vxc(const double* vx, const dmesh& m, int l, int i, int j, int k, double dom) {
  return VxcLoader<G>{cell::gl(vx, m, l, i), m.ymed, m.sinzmed, j, k, dom};
}

// kf_src_a: stresses (visctensor ranges j,k in [1, s-2]), artificial pressure
// (SubStep2_a) and the potential, at every cell the active update reads.
template <int G, int DZ>
__global__ void __launch_bounds__(WG) kf_src_a(dmesh m, const double* rho, const double* energy, const double* vx,
                                               const double* vy, const double* vz, double* txx, double* tyy,
                                               double* tzz, double* txz, double* tyx, double* tzy, double* qx,
                                               double* qy, double* qz, double* pot, cell::ViscPar vp, int visc,
                                               cell::PlanetSet ps, int potential, double domega) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m.nx) return;
  const int j = 1 + (int)blockIdx.y;       // 1 .. sy-2
  const int k = DZ ? 1 + (int)blockIdx.z : 0;
  const int l = i + j * m.pitch + k * m.stride;
  const auto VX = vxc<G>(vx, m, l, i, j, k, domega);
  const cell::GL VY = cell::gl(vy, m, l, i);
  const cell::GL VZ = cell::gl(DZ ? vz : vy, m, l, i);
  const cell::GL R = cell::gl(rho, m, l, i);
  if (visc) {
    cell::Taus t = cell::visctensor<G, DZ>(m, i, j, k, R, cell::gl(energy, m, l, i), VX, VY, VZ, vp);
    txx[l] = t.xx;
    tyy[l] = t.yy;
    tyx[l] = t.yx;
    if (DZ) {
      tzz[l] = t.zz;
      txz[l] = t.xz;
      tzy[l] = t.zy;
    }
  }
  // SubStep2_a (src/substep2_a.c:103-134) on j in [1, sy-2] (a subset of its
  // range [0, sy-2] that covers every read of the active update)
  const double r0 = R(0, 0, 0);
  qx[l] = cell::art_q(r0, VX(1, 0, 0) - VX(0, 0, 0));
  qy[l] = cell::art_q(r0, VY(0, 1, 0) - VY(0, 0, 0));
  if (DZ) qz[l] = cell::art_q(r0, VZ(0, 0, 1) - VZ(0, 0, 0));
  if (potential) pot[l] = cell::potential<G>(m, i, j, k, ps);
}

// kf_src_b: the velocity update of SubStep1 + addviscosity + SubStep2_b on
// active cells, in the order the reference applies them.
template <int G, int DZ>
__global__ void __launch_bounds__(WG) kf_src_b(dmesh m, double dt, const double* rho, const double* energy,
                                               const double* vx, const double* vy, const double* vz,
                                               const double* pot, const double* txx, const double* tyy,
                                               const double* tzz, const double* txz, const double* tyx,
                                               const double* tzy, const double* qx, const double* qy,
                                               const double* qz, double* vxt, double* vyt, double* vzt, int visc,
                                               int potential, int iso, double gm1, double omegaframe,
                                               double domega) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m.nx) return;
  const int j = m.nghy + (int)blockIdx.y;
  const int k = DZ ? m.nghz + (int)blockIdx.z : 0;
  const int l = i + j * m.pitch + k * m.stride;
  const cell::GL R = cell::gl(rho, m, l, i);
  const PresLoader P{R, cell::gl(energy, m, l, i), gm1, iso};
  const cell::GL PT = cell::gl(potential ? pot : rho, m, l, i);
  const auto VX = vxc<G>(vx, m, l, i, j, k, domega);
  const cell::GL VY = cell::gl(vy, m, l, i);
  const cell::GL VZ = cell::gl(DZ ? vz : vy, m, l, i);
  double ax = cell::substep1_x<G>(m, i, j, k, dt, R, P, PT, VX, 0, potential);
  double ay = cell::substep1_y<G, DZ>(m, i, j, k, dt, R, P, PT, VX, VY, VZ, 0, potential, omegaframe);
  double az = DZ ? cell::substep1_z<G>(m, i, j, k, dt, R, P, PT, VX, VZ, 0, potential, omegaframe) : 0.0;
  if (visc) {
    const cell::GL TZZ = cell::gl(DZ ? tzz : txx, m, l, i), TXZ = cell::gl(DZ ? txz : txx, m, l, i),
                   TZY = cell::gl(DZ ? tzy : txx, m, l, i);
    cell::addviscosity<G, DZ>(m, i, j, k, dt, R, cell::gl(txx, m, l, i), cell::gl(tyy, m, l, i), TZZ, TXZ,
                              cell::gl(tyx, m, l, i), TZY, ax, ay, az, ax, ay, az);
  }
  double e_unused;
  cell::substep2_b<G, DZ>(m, i, j, k, dt, R, VX, VY, VZ, cell::gl(qx, m, l, i), cell::gl(qy, m, l, i),
                          cell::gl(DZ ? qz : qy, m, l, i), 0, ax, ay, az, 0.0, ax, ay, az, e_unused);
  vxt[l] = ax;
  vyt[l] = ay;
  if (DZ) vzt[l] = az;
}

// --------------------------------------------------------------- transport
// Y sweep (2-D, with momenta): thread per column i, rows [jb, je) active.
// Q order: Mpx, Mmx, Mpy, Mmy; the density is updated last with DensStar.
template <int G, int DZ>
__global__ void __launch_bounds__(WG) kf_sweep_y(dmesh m, double dt, const double* rho_in, const double* vxt,
                                                 const double* vyt, const double* qin0, const double* qin1,
                                                 const double* qin2, const double* qin3, const double* qin4,
                                                 const double* qin5, double* rho_out, double* q0, double* q1,
                                                 double* q2, double* q3, double* q4, double* q5, int rows_per,
                                                 double omegaframe, int with_momenta) {
  constexpr int NQ = DZ ? 6 : 4;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m.nx) return;
  const int k = DZ ? m.nghz + (int)blockIdx.z : 0;
  const int jb = m.nghy + (int)blockIdx.y * rows_per;
  const int je = min(jb + rows_per, m.ny + m.nghy);
  const int ixp = ixp_of(i, m.nx);
  const int kofs = k * m.stride;
  const double* qin[6] = {qin0, qin1, qin2, qin3, qin4, qin5};
  double* qout[6] = {q0, q1, q2, q3, q4, q5};
  // value of quantity c (or the density, c = -1) at row jj
  auto load = [&](int c, int jj) -> double {
    const int l = i + jj * m.pitch + kofs;
    if (c < 0) return rho_in[l];
    if (!with_momenta) return qin[c][l];
    const double r = rho_in[l];
    const double ym = m.ymed[jj];
    switch (c) {
      case 0: {  // Mpx (src/momenta_x.c:77-78)
        double mm, mp;
        cell::momenta_x<G>(m, i, jj, k, r, vxt[l], vxt[i + (ixp - i) + jj * m.pitch + kofs], omegaframe, mm, mp);
        (void)ym;
        return mp;
      }
      case 1: {  // Mmx
        double mm, mp;
        cell::momenta_x<G>(m, i, jj, k, r, vxt[l], vxt[ixp + jj * m.pitch + kofs], omegaframe, mm, mp);
        return mm;
      }
      case 2:  // Mpy (src/momenta_y.c:62)
        return vyt[l + m.pitch] * r;
      default:  // Mmy
        return vyt[l] * r;
    }
  };
  double Qv[NQ + 1][5];  // [NQ] is the density
  for (int o = 0; o < 4; o++) {
    const int jj = jb - 2 + o;
    Qv[NQ][o] = load(-1, jj);
    for (int c = 0; c < NQ; c++) Qv[c][o] = load(c, jj);
  }
  // face values at the face of row jb (needs rows jb-2..jb+1): density first
  auto face = [&](int c, int jj, const double* qq /*q at jj-2..jj+1*/, const double* rr) -> double {
    // q = Q/rho for momenta, q = rho for the density (DensStar)
    double a0, a1, a2, a3;
    if (c < 0) {
      a0 = qq[0]; a1 = qq[1]; a2 = qq[2]; a3 = qq[3];
    } else {
      a0 = qq[0] / rr[0]; a1 = qq[1] / rr[1]; a2 = qq[2] / rr[2]; a3 = qq[3] / rr[3];
    }
    // slopes at rows jj-1 and jj (src/vanleer_y_a.c:63-68)
    double sm = vanleer_slope((a1 - a0) / zone_size_y(m, jj - 1), (a2 - a1) / zone_size_y(m, jj));
    double s0 = vanleer_slope((a2 - a1) / zone_size_y(m, jj), (a3 - a2) / zone_size_y(m, jj + 1));
    const double v = vyt[i + jj * m.pitch + kofs];
    return vanleer_face(v, dt, a1, sm, zone_size_y(m, jj - 1), a2, s0, zone_size_y(m, jj));
  };
  double fprev[NQ + 1], dsprev;
  {
    double rr[4] = {Qv[NQ][0], Qv[NQ][1], Qv[NQ][2], Qv[NQ][3]};
    dsprev = face(-1, jb, rr, rr);
    for (int c = 0; c < NQ; c++) {
      double qq[4] = {Qv[c][0], Qv[c][1], Qv[c][2], Qv[c][3]};
      fprev[c] = face(c, jb, qq, rr);
    }
  }
  for (int jj = jb; jj < je; jj++) {
    // shift in row jj+2
    const int jn = jj + 2;
    Qv[NQ][4] = load(-1, jn);
    for (int c = 0; c < NQ; c++) Qv[c][4] = load(c, jn);
    double rr[4] = {Qv[NQ][1], Qv[NQ][2], Qv[NQ][3], Qv[NQ][4]};
    const double dsn = face(-1, jj + 1, rr, rr);
    const double vy0 = vyt[i + jj * m.pitch + kofs], vy1 = vyt[i + (jj + 1) * m.pitch + kofs];
    const int l = i + jj * m.pitch + kofs;
    for (int c = 0; c < NQ; c++) {
      double qq[4] = {Qv[c][1], Qv[c][2], Qv[c][3], Qv[c][4]};
      const double fn = face(c, jj + 1, qq, rr);
      // UpdateY (src/update_y.c:68-70)
      qout[c][l] = Qv[c][2] + cell::flux_y(m, i, jj, k, dt, vy0 * fprev[c] * dsprev, vy1 * fn * dsn);
      fprev[c] = fn;
    }
    // UpdateDensityY (src/update_density_y.c:66-68)
    rho_out[l] = Qv[NQ][2] + cell::flux_y(m, i, jj, k, dt, vy0 * dsprev, vy1 * dsn);
    dsprev = dsn;
    for (int o = 0; o < 4; o++) {
      Qv[NQ][o] = Qv[NQ][o + 1];
      for (int c = 0; c < NQ; c++) Qv[c][o] = Qv[c][o + 1];
    }
  }
}

// Z sweep (3-D) with momenta: thread per (i, j) column, j in [1, sy-2], k
// active in strips; writes the Z-updated density and six momenta.
template <int G, int DZ>
__global__ void __launch_bounds__(WG) kf_sweep_z(dmesh m, double dt, const double* rho_in, const double* vxt,
                                                 const double* vyt, const double* vzt, double* rho_out, double* q0,
                                                 double* q1, double* q2, double* q3, double* q4, double* q5,
                                                 int planes_per, double omegaframe) {
  constexpr int NQ = 6;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m.nx) return;
  const int j = 1 + (int)blockIdx.y;
  const int kb = m.nghz + (int)blockIdx.z * planes_per;
  const int ke = min(kb + planes_per, m.nz + m.nghz);
  const int ixp = ixp_of(i, m.nx);
  double* qout[6] = {q0, q1, q2, q3, q4, q5};
  auto load = [&](int c, int kk) -> double {
    const int l = i + j * m.pitch + kk * m.stride;
    const double r = rho_in[l];
    if (c < 0) return r;
    double mm, mp;
    switch (c) {
      case 0:
        cell::momenta_x<G>(m, i, j, kk, r, vxt[l], vxt[ixp + j * m.pitch + kk * m.stride], omegaframe, mm, mp);
        return mp;
      case 1:
        cell::momenta_x<G>(m, i, j, kk, r, vxt[l], vxt[ixp + j * m.pitch + kk * m.stride], omegaframe, mm, mp);
        return mm;
      case 2:
        return vyt[l + m.pitch] * r;
      case 3:
        return vyt[l] * r;
      case 4: {  // Mpz (src/momenta_z.c:66-71)
        double p = vzt[l + m.stride] * r;
        if (G == SPH) p *= m.ymed[j];
        return p;
      }
      default: {  // Mmz
        double q = vzt[l] * r;
        if (G == SPH) q *= m.ymed[j];
        return q;
      }
    }
  };
  double Qv[NQ + 1][5];
  for (int o = 0; o < 4; o++) {
    const int kk = kb - 2 + o;
    Qv[NQ][o] = load(-1, kk);
    for (int c = 0; c < NQ; c++) Qv[c][o] = load(c, kk);
  }
  auto face = [&](int c, int kk, const double* qq, const double* rr) -> double {
    double a0, a1, a2, a3;
    if (c < 0) {
      a0 = qq[0]; a1 = qq[1]; a2 = qq[2]; a3 = qq[3];
    } else {
      a0 = qq[0] / rr[0]; a1 = qq[1] / rr[1]; a2 = qq[2] / rr[2]; a3 = qq[3] / rr[3];
    }
    // src/vanleer_z_a.c:63-68 and src/vanleer_z_b.c:60-65
    double sm = vanleer_slope((a1 - a0) / zone_size_z<G>(m, j, kk - 1), (a2 - a1) / zone_size_z<G>(m, j, kk));
    double s0 = vanleer_slope((a2 - a1) / zone_size_z<G>(m, j, kk), (a3 - a2) / zone_size_z<G>(m, j, kk + 1));
    const double v = vzt[i + j * m.pitch + kk * m.stride];
    return vanleer_face(v, dt, a1, sm, zone_size_z<G>(m, j, kk - 1), a2, s0, zone_size_z<G>(m, j, kk));
  };
  double fprev[NQ], dsprev;
  {
    double rr[4] = {Qv[NQ][0], Qv[NQ][1], Qv[NQ][2], Qv[NQ][3]};
    dsprev = face(-1, kb, rr, rr);
    for (int c = 0; c < NQ; c++) {
      double qq[4] = {Qv[c][0], Qv[c][1], Qv[c][2], Qv[c][3]};
      fprev[c] = face(c, kb, qq, rr);
    }
  }
  for (int kk = kb; kk < ke; kk++) {
    const int kn = kk + 2;
    Qv[NQ][4] = load(-1, kn);
    for (int c = 0; c < NQ; c++) Qv[c][4] = load(c, kn);
    double rr[4] = {Qv[NQ][1], Qv[NQ][2], Qv[NQ][3], Qv[NQ][4]};
    const double dsn = face(-1, kk + 1, rr, rr);
    const double vz0 = vzt[i + j * m.pitch + kk * m.stride], vz1 = vzt[i + j * m.pitch + (kk + 1) * m.stride];
    const int l = i + j * m.pitch + kk * m.stride;
    for (int c = 0; c < NQ; c++) {
      double qq[4] = {Qv[c][1], Qv[c][2], Qv[c][3], Qv[c][4]};
      const double fn = face(c, kk + 1, qq, rr);
      qout[c][l] = Qv[c][2] + cell::flux_z(m, i, j, kk, dt, vz0 * fprev[c] * dsprev, vz1 * fn * dsn);
      fprev[c] = fn;
    }
    rho_out[l] = Qv[NQ][2] + cell::flux_z(m, i, j, kk, dt, vz0 * dsprev, vz1 * dsn);
    dsprev = dsn;
    for (int o = 0; o < 4; o++) {
      Qv[NQ][o] = Qv[NQ][o + 1];
      for (int c = 0; c < NQ; c++) Qv[c][o] = Qv[c][o + 1];
    }
  }
}

// X sweep, ring-owning: VxMed of Vx_temp, residual split, van Leer pass with
// the cell residual, PPA pass with the uniform residual, integer shift and
// NewVelocity_x (src/transport.c:129-192).  LDS holds the ring.
template <int G, int DZ>
__global__ void __launch_bounds__(WG) kf_sweep_x(dmesh m, double dt, const double* rho_in, const double* vxt,
                                                 double* mpx, double* mmx, double* mpy, double* mmy, double* mpz,
                                                 double* mmz, double* rho_out, double* vx_out, double* vxmed_out,
                                                 int* nshift_out, double omegaframe, int fargo) {
  constexpr int NQ = DZ ? 6 : 4;
  extern __shared__ double lds[];
  const int nx = m.nx;
  double* R = lds;                 // density
  double* Qa = R + nx;             // NQ quantities
  double* VR = Qa + NQ * nx;       // cell residual / total velocity
  double* VU = VR + nx;            // uniform residual
  double* DS = VU + nx;            // DensStar
  double* QQ = DS + nx;            // q = Q/rho
  double* SL = QQ + nx;            // slope / qL
  double* QR = SL + nx;            // qR
  double* QS = QR + nx;            // interface value
  double* sh = QS + nx;            // 8 doubles
  const int j = m.nghy + (int)(blockIdx.x % m.ny);
  const int k = DZ ? m.nghz + (int)(blockIdx.x / m.ny) : 0;
  const int base = j * m.pitch + k * m.stride;
  double* qg[6] = {mpx, mmx, mpy, mmy, mpz, mmz};
  const int t = threadIdx.x, nt = blockDim.x;
  for (int i = t; i < nx; i += nt) {
    R[i] = rho_in[base + i];
    for (int c = 0; c < NQ; c++) Qa[c * nx + i] = qg[c][base + i];
    QS[i] = vxt[base + i] * (m.xmin[i + 1] - m.xmin[i]);
  }
  __syncthreads();
  // ComputeVmed(Vx_temp) (src/compute_vmed.c:14-30): sequential ring sum
  if (t == 0) {
    double acc = 0.0;
    for (int i = 0; i < nx; i++) acc = acc + QS[i];
    sh[0] = acc / (m.XMAX - m.XMIN);
  }
  __syncthreads();
  const double vmed = sh[0];
  if (t == 0) vxmed_out[j + k * m.sy] = vmed;
  int ns = 0;
  // ComputeResidual (src/compute_residual.c:68-73)
  for (int i = t; i < nx; i += nt) {
    const double v = vxt[base + i];
    if (fargo) {
      const double zs = zone_size_x<G>(m, i, j, k);
      double ntilde = vmed * dt / zs;
      double nround = floor(ntilde + 0.5);
      if (i == 0) sh[1] = nround;
      VR[i] = v - vmed;
      VU[i] = (ntilde - nround) * zs / dt;
    } else {
      VR[i] = v;
    }
  }
  __syncthreads();
  if (fargo) {
    ns = (int)sh[1];
    if (t == 0) nshift_out[j + k * m.sy] = ns;
  }
  // one X advection pass (X_advection, src/transport.c:208-226)
  for (int pass = 0; pass < (fargo ? 2 : 1); pass++) {
    const double* V = pass ? VU : VR;
    const bool ppa = (pass == 1);
    // quantities in order: density (DensStar), NQ momenta, density update
    for (int c = -1; c <= NQ; c++) {
      if (c == NQ) {
        // UpdateDensityX (src/update_density_x.c:66-68)
        for (int i = t; i < nx; i += nt) {
          const int ip = ixp_of(i, nx);
          QQ[i] = R[i] + cell::flux_x(m, i, j, k, dt, V[i] * DS[i], V[ip] * DS[ip]);
        }
        __syncthreads();
        for (int i = t; i < nx; i += nt) R[i] = QQ[i];
        __syncthreads();
        break;
      }
      // q: the density itself for DensStar, Q/rho otherwise (DivideByRho)
      for (int i = t; i < nx; i += nt) QQ[i] = (c < 0) ? R[i] : Qa[c * nx + i] / R[i];
      __syncthreads();
      if (!ppa) {
        // VanLeerX_a (src/vanleer_x_a.c:60-68)
        for (int i = t; i < nx; i += nt) {
          const int im = ixm_of(i, nx), ip = ixp_of(i, nx);
          double dqm = (QQ[i] - QQ[im]) / zone_size_x<G>(m, i, j, k);
          double dqp = (QQ[ip] - QQ[i]) / zone_size_x<G>(m, ip, j, k);
          SL[i] = vanleer_slope(dqm, dqp);
        }
        __syncthreads();
        // VanLeerX_b (src/vanleer_x_b.c:60-65)
        double* out = (c < 0) ? DS : QS;
        for (int i = t; i < nx; i += nt) {
          const int im = ixm_of(i, nx);
          out[i] = vanleer_face(V[i], dt, QQ[im], SL[im], zone_size_x<G>(m, im, j, k), QQ[i], SL[i],
                                zone_size_x<G>(m, i, j, k));
        }
        __syncthreads();
      } else {
        // VanLeerX_PPA_a/b/c/d (src/fargo_ppa_a.c .. src/fargo_ppa_d.c)
        for (int i = t; i < nx; i += nt) SL[i] = cell::ppa_slope(QQ[ixm_of(i, nx)], QQ[i], QQ[ixp_of(i, nx)]);
        __syncthreads();
        // QR[i] = face i+1/2; qL of cell i is face i-1/2 = QR[i-1]
        for (int i = t; i < nx; i += nt) {
          const int ip = ixp_of(i, nx);
          QR[i] = cell::ppa_face(QQ[i], QQ[ip], SL[i], SL[ip]);
        }
        __syncthreads();
        double* out = (c < 0) ? DS : QS;
        // limited (qL, qR) of each cell: store qL in SL, qR in a register pass
        for (int i = t; i < nx; i += nt) {
          double qL = QR[ixm_of(i, nx)], qR = QR[i];
          cell::ppa_limit(QQ[i], qL, qR);
          SL[i] = qL;
          out[i] = qR;  // temporarily the limited qR
        }
        __syncthreads();
        // copy limited qR out of the face buffer before computing faces
        for (int i = t; i < nx; i += nt) QR[i] = out[i];
        __syncthreads();
        for (int i = t; i < nx; i += nt) {
          const int im = ixm_of(i, nx);
          out[i] = cell::ppa_flux_value(V[i], dt, zone_size_x<G>(m, i, j, k), QQ[im], SL[im], QR[im], QQ[i],
                                        SL[i], QR[i]);
        }
        __syncthreads();
      }
      if (c >= 0) {
        // UpdateX (src/update_x.c:69-71)
        double* Qc = Qa + c * nx;
        for (int i = t; i < nx; i += nt) {
          const int ip = ixp_of(i, nx);
          QQ[i] = Qc[i] + cell::flux_x(m, i, j, k, dt, V[i] * QS[i] * DS[i], V[ip] * QS[ip] * DS[ip]);
        }
        __syncthreads();
        for (int i = t; i < nx; i += nt) Qc[i] = QQ[i];
        __syncthreads();
      }
    }
  }
  // AdvectSHIFT (src/advect_shift.c:45-58) and write-back
  if (fargo) {
    for (int c = -1; c < NQ; c++) {
      double* A = (c < 0) ? R : Qa + c * nx;
      for (int i = t; i < nx; i += nt) {
        int it = (i - ns) % nx;
        if (it < 0) it += nx;
        QQ[i] = A[it];
      }
      __syncthreads();
      for (int i = t; i < nx; i += nt) A[i] = QQ[i];
      __syncthreads();
    }
  }
  // NewVelocity_x (src/newvel_x.c:78-86) and outputs
  for (int i = t; i < nx; i += nt) {
    const int im = ixm_of(i, nx);
    vx_out[base + i] = cell::newvel_x<G>(m, i, j, k, R[i], R[im], Qa[1 * nx + i], Qa[0 * nx + im], omegaframe);
    rho_out[base + i] = R[i];
    for (int c = 2; c < NQ; c++) qg[c][base + i] = Qa[c * nx + i];
  }
}

// NewVelocity_y/z on active faces and Stockholm damping on active cells.
template <int G, int DZ>
__global__ void __launch_bounds__(WG) kf_finish(dmesh m, double dt, const double* rho_x, const double* mmy,
                                                const double* mpy, const double* mmz, const double* mpz,
                                                double* rho, double* vx, double* vy, double* vz, const double* rho0,
                                                const double* vx0, const double* vy0, const double* vz0,
                                                cell::DampPar dp, int stockholm) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m.nx) return;
  const int j = m.nghy + (int)blockIdx.y;
  const int k = DZ ? m.nghz + (int)blockIdx.z : 0;
  const int l = i + j * m.pitch + k * m.stride;
  const double r = rho_x[l];
  // the faces j = nghy and k = nghz lie on the mesh boundary; the ghost fill
  // sets them to zero, so they are not computed here
  if (j > m.nghy) vy[l] = cell::newvel_y(m, i, j, k, r, rho_x[l - m.pitch], mmy[l], mpy[l - m.pitch]);
  if (DZ && k > m.nghz) {
    double v = (mul_vol(m, mmz[l], i, j, k) + mul_vol(m, mpz[l - m.stride], i, j, k - 1)) /
               (mul_vol(m, r, i, j, k) + mul_vol(m, rho_x[l - m.stride], i, j, k - 1));
    if (G == SPH) v /= m.ymed[j];
    vz[l] = v;
  }
  rho[l] = r;
  if (stockholm) cell::stockholm<G, DZ>(m, j, k, l, dt, rho, vx, vy, vz, rho0, vx0, vy0, vz0, dp);
}

// ------------------------------------------------------------------ C ABI

#define FDISPATCH(ctx, KERN, grid, block, shmem, ...)                                                  \
  do {                                                                                                  \
    if ((ctx)->m.geom == SPH) {                                                                         \
      if ((ctx)->m.dimz)                                                                                \
        hipLaunchKernelGGL((KERN<SPH, 1>), grid, block, shmem, (ctx)->stream, (ctx)->m, __VA_ARGS__);   \
      else                                                                                              \
        hipLaunchKernelGGL((KERN<SPH, 0>), grid, block, shmem, (ctx)->stream, (ctx)->m, __VA_ARGS__);   \
    } else {                                                                                            \
      if ((ctx)->m.dimz)                                                                                \
        hipLaunchKernelGGL((KERN<CYL, 1>), grid, block, shmem, (ctx)->stream, (ctx)->m, __VA_ARGS__);   \
      else                                                                                              \
        hipLaunchKernelGGL((KERN<CYL, 0>), grid, block, shmem, (ctx)->stream, (ctx)->m, __VA_ARGS__);   \
    }                                                                                                   \
    F3D_LAUNCH_CHECK(ctx);                                                                              \
  } while (0)

static cell::CflPar cfl_par_f(const f3d_hip_ctx* ctx) {
  cell::CflPar c;
  c.cfl = ctx->ph.cfl;
  c.nu = ctx->ph.nu;
  c.alpha = ctx->ph.alpha;
  c.gamma = ctx->ph.gamma;
  c.iso = (ctx->ph.eos != 2);
  c.alpha_visc = (ctx->ph.visc == 2);
  c.fargo = ctx->ph.fargo;
  return c;
}

extern "C" int f3d_boundaries_on(f3d_hip_ctx* ctx, double* dens, double* vx, double* vy, double* vz);

extern "C" {

int f3d_fused_cfl(f3d_hip_ctx* ctx, double* fluid_min) {
  Scratch* S;
  F3D_REQUIRE(ctx, scratch_of(ctx, &S) == 0, "fused scratch allocation failed");
  F3D_REQUIRE(ctx, ctx->ph.eos == 1, "the fused path covers the isothermal equation of state");
  const dmesh& m = ctx->m;
  const int nrings = (m.ny + 1) * (m.dimz ? m.nz : 1);
  const size_t shmem = (m.nx + 16) * sizeof(double);
  f3d_stage_begin(ctx, F3D_ST_CFL);
  FDISPATCH(ctx, kf_cfl, dim3(nrings), dim3(WG), shmem, ctx->f.Energy, ctx->f.Density, ctx->f.Vx, ctx->f.Vy,
            ctx->f.Vz, ctx->f.VxMed, S->ringmin, S->counter, S->result, cfl_par_f(ctx));
  f3d_stage_end(ctx, F3D_ST_CFL);
  F3D_CHECK(ctx, hipMemcpyAsync(ctx->hres, S->result, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  F3D_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  *fluid_min = ctx->hres[0];
  return 0;
}

int f3d_fused_sources(f3d_hip_ctx* ctx, double dt, double domega) {
  Scratch* S;
  F3D_REQUIRE(ctx, scratch_of(ctx, &S) == 0, "fused scratch allocation failed");
  F3D_REQUIRE(ctx, ctx->ph.eos == 1, "the fused path covers the isothermal equation of state");
  const dmesh& m = ctx->m;
  const f3d_fields& f = ctx->f;
  cell::ViscPar vp;
  vp.model = ctx->ph.visc;
  vp.nu = ctx->ph.nu;
  vp.alpha = ctx->ph.alpha;
  vp.gamma = ctx->ph.gamma;
  vp.iso = 1;
  cell::PlanetSet ps;
  ps.n = ctx->ph.nplanets < cell::kMaxPlanets ? ctx->ph.nplanets : cell::kMaxPlanets;
  ps.indirect = ctx->ph.indirectterm;
  for (int n = 0; n < ps.n; n++) {
    ps.x[n] = ctx->ph.px[n];
    ps.y[n] = ctx->ph.py[n];
    ps.z[n] = ctx->ph.pz[n];
    ps.mp[n] = ctx->ph.pmass[n];
    ps.smooth2[n] = ctx->ph.psmooth2[n];
  }
  const int visc = ctx->ph.visc != 0 && ctx->ph.fluidtype == 1;
  const int pot = ctx->ph.potential;
  dim3 ga((m.nx + BXF - 1) / BXF, m.sy - 2, m.dimz ? m.sz - 2 : 1);
  f3d_stage_begin(ctx, F3D_ST_SRC_A);
  FDISPATCH(ctx, kf_src_a, ga, dim3(BXF), 0, f.Density, f.Energy, f.Vx, f.Vy, f.Vz, f.Mmx, f.Mmy, f.Mmz, f.Mpx,
            f.Mpy, f.Mpz, S->f[FS_QX], S->f[FS_QY], S->f[FS_QZ], f.Pot, vp, visc, ps, 0, domega);
  f3d_stage_end(ctx, F3D_ST_SRC_A);
  dim3 gb((m.nx + BXF - 1) / BXF, m.ny, m.dimz ? m.nz : 1);
  f3d_stage_begin(ctx, F3D_ST_SRC_B);
  FDISPATCH(ctx, kf_src_b, gb, dim3(BXF), 0, dt, f.Density, f.Energy, f.Vx, f.Vy, f.Vz, f.Pot, f.Mmx, f.Mmy, f.Mmz,
            f.Mpx, f.Mpy, f.Mpz, S->f[FS_QX], S->f[FS_QY], S->f[FS_QZ], f.Vx_temp, f.Vy_temp, f.Vz_temp, visc, pot,
            1, ctx->ph.gamma - 1.0, ctx->ph.OMEGAFRAME, domega);
  f3d_stage_end(ctx, F3D_ST_SRC_B);
  // mid-step ghost fill (src/algogas.c:116-118) applied to V_temp directly
  f3d_stage_begin(ctx, F3D_ST_BC_MID);
  if (f3d_boundaries_on(ctx, f.Density, f.Vx_temp, f.Vy_temp, f.Vz_temp)) return -1;
  f3d_stage_end(ctx, F3D_ST_BC_MID);
  return 0;
}

int f3d_fused_transport(f3d_hip_ctx* ctx, double dt, int end_of_step) {
  Scratch* S;
  F3D_REQUIRE(ctx, scratch_of(ctx, &S) == 0, "fused scratch allocation failed");
  F3D_REQUIRE(ctx, ctx->ph.eos == 1, "the fused path covers the isothermal equation of state");
  const dmesh& m = ctx->m;
  const f3d_fields& f = ctx->f;
  const double of = ctx->ph.OMEGAFRAME;
  const int rows_per = 16;
  if (m.dimz) {
    // seven separate buffers: six Z-swept momenta and the Z-swept density
    double* const* Z = &S->f[FS_Z0];
    dim3 gz((m.nx + BXF - 1) / BXF, m.sy - 2, (m.nz + rows_per - 1) / rows_per);
    f3d_stage_begin(ctx, F3D_ST_SWEEP_Z);
    FDISPATCH(ctx, kf_sweep_z, gz, dim3(BXF), 0, dt, f.Density, f.Vx_temp, f.Vy_temp, f.Vz_temp, Z[6], Z[0], Z[1],
              Z[2], Z[3], Z[4], Z[5], rows_per, of);
    f3d_stage_end(ctx, F3D_ST_SWEEP_Z);
    dim3 gy((m.nx + BXF - 1) / BXF, (m.ny + rows_per - 1) / rows_per, m.nz);
    f3d_stage_begin(ctx, F3D_ST_SWEEP_Y);
    FDISPATCH(ctx, kf_sweep_y, gy, dim3(BXF), 0, dt, Z[6], f.Vx_temp, f.Vy_temp, Z[0], Z[1], Z[2], Z[3], Z[4], Z[5],
              S->f[FS_RHOY], f.Mpx, f.Mmx, f.Mpy, f.Mmy, f.Mpz, f.Mmz, rows_per, of, 0);
    f3d_stage_end(ctx, F3D_ST_SWEEP_Y);
  } else {
    dim3 gy((m.nx + BXF - 1) / BXF, (m.ny + rows_per - 1) / rows_per, 1);
    f3d_stage_begin(ctx, F3D_ST_SWEEP_Y);
    FDISPATCH(ctx, kf_sweep_y, gy, dim3(BXF), 0, dt, f.Density, f.Vx_temp, f.Vy_temp, nullptr, nullptr, nullptr,
              nullptr, nullptr, nullptr, S->f[FS_RHOY], f.Mpx, f.Mmx, f.Mpy, f.Mmy, nullptr, nullptr, rows_per, of, 1);
    f3d_stage_end(ctx, F3D_ST_SWEEP_Y);
  }
  const int nq = m.dimz ? 6 : 4;
  const size_t shx = ((size_t)(nq + 8) * m.nx + 8) * sizeof(double);
  F3D_REQUIRE(ctx, shx <= 160 * 1024, "ring too long for the LDS X sweep");
  dim3 gx(m.ny * (m.dimz ? m.nz : 1));
  f3d_stage_begin(ctx, F3D_ST_SWEEP_X);
  FDISPATCH(ctx, kf_sweep_x, gx, dim3(WG), shx, dt, S->f[FS_RHOY], f.Vx_temp, f.Mpx, f.Mmx, f.Mpy, f.Mmy, f.Mpz,
            f.Mmz, S->f[FS_RHOX], f.Vx, f.VxMed, f.Nshift, of, ctx->ph.fargo);
  f3d_stage_end(ctx, F3D_ST_SWEEP_X);
  cell::DampPar dp;
  dp.Y_inf = ctx->ph.Y_inf; dp.Y_sup = ctx->ph.Y_sup; dp.Z_inf = ctx->ph.Z_inf; dp.Z_sup = ctx->ph.Z_sup;
  dp.y_min = ctx->ph.ymin_par; dp.y_max = ctx->ph.ymax_par; dp.z_min = ctx->ph.zmin_par; dp.z_max = ctx->ph.zmax_par;
  dp.ds = ctx->ph.taudamp; dp.of = of; dp.of0 = ctx->ph.OMEGAFRAME0; dp.periodic_z = ctx->ph.periodicz;
  dim3 gf((m.nx + BXF - 1) / BXF, m.ny, m.dimz ? m.nz : 1);
  f3d_stage_begin(ctx, F3D_ST_FINISH);
  FDISPATCH(ctx, kf_finish, gf, dim3(BXF), 0, dt, S->f[FS_RHOX], f.Mmy, f.Mpy, f.Mmz, f.Mpz, f.Density, f.Vx, f.Vy,
            f.Vz, f.Density0, f.Vx0, f.Vy0, f.Vz0, dp, ctx->ph.stockholm);
  f3d_stage_end(ctx, F3D_ST_FINISH);
  if (end_of_step) {
    // end-of-step ghost fill (src/main.c:431)
    f3d_stage_begin(ctx, F3D_ST_BC_END);
    if (f3d_boundaries_on(ctx, f.Density, f.Vx, f.Vy, f.Vz)) return -1;
    f3d_stage_end(ctx, F3D_ST_BC_END);
  }
  return 0;
}

}  // extern "C"
