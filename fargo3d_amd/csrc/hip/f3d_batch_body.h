// f3d_batch_body.h -- the batched form: many independent 2-D isothermal disks
// (uneven grids, per-instance parameters and time step) advanced together.
// Every stage of the step is one launch over all instances: a workgroup finds
// its instance in a prefix table of per-instance workgroup counts and runs
// the same device function the single-disk kernels run.  The time step of
// each instance is computed and clamped on the device (src/main.c:383-395),
// so a step needs no host round trip and is captured once in a hipGraph.
// The planets of each instance live on the device too, in FP64: the
// potential of the step from their positions, then the RK5 N-body step and
// the rotation of the frame that follows the first planet (src/potential.c
// :10-49), inside the same graph.
// Included by f3d_fused_body.h in the FP32 build only.
#include "../host/f3d_driver.h"

namespace F3D_NS {
namespace fused {

constexpr int kBP = cell::kMaxPlanets;

// The planetary system of one instance, FP64, in device memory (the host's
// f3d_psys and the frame's scalars).  pm/ps2: tapered mass and squared
// smoothing length of each planet for the next Potential.
struct BPsys {
  int n, indirect, corotating, guiding;
  double x[kBP], y[kBP], z[kBP], vx[kBP], vy[kBP], vz[kBP], m[kBP];
  int feelothers[kBP];
  double pm[kBP], ps2[kBP];
  double masstaper, rochesmoothing, aspectratio, flaringindex, thicknesssmoothing;
  double of;  // OMEGAFRAME
};

// one instance as the batched kernels see it (device memory)
struct BInst {
  dmesh m;
  real *rho, *cs, *vx, *vy, *vxt, *vyt, *mpx, *mmx, *mpy, *mmy, *rho0, *vx0, *vy0;
  double* pot;  // FP64 potential
  real *rhoy, *rhox, *e2, *ey, *xm0, *xm1, *xm2, *xm3;
  double* vxmed;
  int* nshift;
  const real *ydexp0, *ydexp1;
  cell::CflPar cp;
  cell::ViscPar vp;
  cell::DampPar dp;
  cell::BcPar b0, b1;
  real of;                  // OMEGAFRAME of the step (written by kb_dt with the planets)
  double of64;              // the same in FP64 (the centrifugal terms)
  real domega;              // CorrectVtheta's change of the frame rate this step
  int visc, potential, stockholm, fargo;
  double* clk;              // [0] dt, [1] dtemp, [2] PhysicalTime, [3] DT, [4] null-dt flag
  unsigned long long* res;  // the CFL atomic MIN (two slots; kd_cfl resets the other)
  BPsys* ps;                // planets (nullptr: no potential)
  const double *ymed64, *cosx64, *sinx64;  // FP64 mesh of the potential
  double zc;                // Zmed of the single plane
};

// the stages, each with its own per-instance workgroup counts
enum { BK_CFL = 0, BK_SRCA, BK_SRCB, BK_BND, BK_SWY, BK_VMED, BK_XT, BK_FIN, BK_POT, BK_N };

__device__ __forceinline__ int find_inst(const int* off, int n, int b) {
  int lo = 0, hi = n;  // off[lo] <= b < off[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (off[mid] <= b) lo = mid;
    else hi = mid;
  }
  return lo;
}

// local 2-D workgroup index of a grid gx wide
__device__ __forceinline__ dim3 local2(int lb, int gx) { return dim3(lb % gx, lb / gx, 0); }

__global__ void __launch_bounds__(WG) kb_cfl(const BInst* I, const int* off, int n) {
  const int ii = find_inst(off, n, blockIdx.x);
  const BInst& in = I[ii];
  kd_cfl<CYL, 0, 0>((int)blockIdx.x - off[ii], in.m, in.cs, in.rho, in.vx, in.vy, in.vy, in.vx, in.vxmed,
                    in.res, 0, in.cp, nullptr);
}

// Potential (src/potential.c:51-208) of every instance with planets, FP64:
// the star's term, each planet's softened term and its indirect term, at the
// planets' positions at the start of the step (cylindrical 2-D disks).
__global__ void __launch_bounds__(BXF) kb_potential(const BInst* I, const int* off, int n) {
  const int ii = find_inst(off, n, blockIdx.x);
  const BInst& in = I[ii];
  const int gx = (in.m.nx + BXF - 1) / BXF;
  const dim3 b = local2(blockIdx.x - off[ii], gx);
  const int i = b.x * BXF + threadIdx.x;
  if (i >= in.m.nx) return;
  const int j = b.y;
  const BPsys& P = *in.ps;
  const double r = in.ymed64[j], XC = r * in.cosx64[i], YC = r * in.sinx64[i], ZC = in.zc;
  double pot = -1.0 / sqrt(r * r + ZC * ZC);
  for (int p = 0; p < P.n; p++) {
    const double xp = P.x[p], yp = P.y[p], zp = P.z[p], mp = P.pm[p];
    const double d = sqrt(xp * xp + yp * yp + zp * zp);
    const double dist = (XC - xp) * (XC - xp) + (YC - yp) * (YC - yp) + (ZC - zp) * (ZC - zp);
    if (P.indirect) pot += mp * (XC * xp + YC * yp + ZC * zp) / (d * d * d);
    pot += -mp / sqrt(dist + P.ps2[p]);
  }
  in.pot[(size_t)j * in.m.pitch + i] = pot;
}

// Tapered mass and squared smoothing length of each planet at time t
// (src/potential.c:56-59, 115-131).
__device__ void bp_prepare(BPsys& P, double t) {
  const double taper = (P.masstaper == 0.0) ? 1.0
                       : (t >= P.masstaper ? 1.0 : .5 * (1.0 - cos(M_PI * t / P.masstaper)));
  for (int p = 0; p < P.n; p++) {
    const double mp = P.m[p] * taper;
    const double d = sqrt(P.x[p] * P.x[p] + P.y[p] * P.y[p] + P.z[p] * P.z[p]);
    double h = (P.rochesmoothing != 0) ? d * pow(1.0 / 3.0 * mp, 1.0 / 3.0) * P.rochesmoothing
                                       : P.aspectratio * pow(d, P.flaringindex) * d * P.thicknesssmoothing;
    P.pm[p] = mp;
    P.ps2[p] = h * h;
  }
}

// Time derivative of the state q (x, y, z, vx, vy, vz blocks of NP) times
// dt: the star's pull, the indirect acceleration of the star by every
// planet, and the pull of the other planets on those that feel them
// (src/runge_kutta.c:10-58).  NP, an upper bound of the planet count P.n,
// is a template parameter and every loop is unrolled to it, so the state
// stays in registers; slots at or beyond P.n are skipped (and hold 1.0).
template <int NP>
__device__ __forceinline__ void bp_deriv(const BPsys& P, const double (&q)[6 * NP], double (&d)[6 * NP], double dt) {
  double rad[NP];
#pragma unroll
  for (int i = 0; i < NP; i++) rad[i] = sqrt(q[i] * q[i] + q[NP + i] * q[NP + i] + q[2 * NP + i] * q[2 * NP + i]);
#pragma unroll
  for (int i = 0; i < NP; i++) {
    double a[3];
#pragma unroll
    for (int c = 0; c < 3; c++) a[c] = -q[c * NP + i] / rad[i] / rad[i] / rad[i];
#pragma unroll
    for (int k = 0; k < NP; k++) {
      if (k >= P.n) continue;
      if (P.indirect) {
        const double cf = P.m[k] / rad[k] / rad[k] / rad[k];
#pragma unroll
        for (int c = 0; c < 3; c++) a[c] -= cf * q[c * NP + k];
      }
      if (k != i && P.feelothers[i]) {
        const double dx = q[k] - q[i], dy = q[NP + k] - q[NP + i], dz = q[2 * NP + k] - q[2 * NP + i];
        const double sd = sqrt(dx * dx + dy * dy + dz * dz);
        const double cf = P.m[k] / sd / sd / sd;
        a[0] += cf * dx;
        a[1] += cf * dy;
        a[2] += cf * dz;
      }
    }
#pragma unroll
    for (int c = 0; c < 3; c++) {
      d[c * NP + i] = q[(3 + c) * NP + i] * dt;
      d[(3 + c) * NP + i] = a[c] * dt;
    }
  }
}

// One Cash-Karp step of the planets over dt (src/runge_kutta.c:120-157):
// five stages from the tableau's rows, the fifth-order combination.
template <int NP>
__device__ __forceinline__ void bp_rk5(BPsys& P, double dt) {
  constexpr int S = 6 * NP;
  double q0[S], q[S], k[6][S];
#pragma unroll
  for (int i = 0; i < NP; i++) {
    const bool on = i < P.n;
    q0[i] = on ? P.x[i] : 1.0; q0[NP + i] = on ? P.y[i] : 1.0; q0[2 * NP + i] = on ? P.z[i] : 1.0;
    q0[3 * NP + i] = on ? P.vx[i] : 1.0; q0[4 * NP + i] = on ? P.vy[i] : 1.0; q0[5 * NP + i] = on ? P.vz[i] : 1.0;
  }
  constexpr double A[5][5] = {{0.2, 0.0, 0.0, 0.0, 0.0},
                              {0.075, 0.225, 0.0, 0.0, 0.0},
                              {0.3, -0.9, 1.2, 0.0, 0.0},
                              {-11.0 / 54.0, 2.5, -70.0 / 27.0, 35.0 / 27.0, 0.0},
                              {1631.0 / 55296.0, 175.0 / 512.0, 575.0 / 13824.0, 44275.0 / 110592.0,
                               253.0 / 4096.0}};
  bp_deriv<NP>(P, q0, k[0], dt);
#pragma unroll
  for (int s = 1; s < 6; s++) {
#pragma unroll
    for (int a = 0; a < S; a++) {
      double v = q0[a];
#pragma unroll
      for (int r = 0; r < s; r++) v += A[s - 1][r] * k[r][a];
      q[a] = v;
    }
    bp_deriv<NP>(P, q, k[s], dt);
  }
#pragma unroll
  for (int a = 0; a < S; a++)
    q[a] = q0[a] + 37.0 / 378.0 * k[0][a] + 250.0 / 621.0 * k[2][a] + 125.0 / 594.0 * k[3][a] +
           512.0 / 1771.0 * k[5][a];
#pragma unroll
  for (int i = 0; i < NP; i++) {
    if (i >= P.n) continue;
    P.x[i] = q[i]; P.y[i] = q[NP + i]; P.z[i] = q[2 * NP + i];
    P.vx[i] = q[3 * NP + i]; P.vy[i] = q[4 * NP + i]; P.vz[i] = q[5 * NP + i];
  }
}

// The point the frame follows (GetPsysInfo, src/psys.c:362-415): the first
// planet itself, or (Frame G) its guiding centre, the point at mean anomaly M
// on the circle of radius a of its two-body orbit about star + planet,
// azimuth from the perihelion, tilted by the inclination.  Elements from the
// angular momentum h = r x v and the Laplace-Runge-Lenz vector
// A = v x h - G m r/|r| (|A| = G m e), a = h^2 / (G m (1 - e^2)), the
// eccentric anomaly E from r = a (1 - e cos E) (negative while r.v < 0;
// 0 or pi where rounding leaves acos's domain), M = E - e sin E; a circular
// orbit (e < 1e-8) follows the planet's position.
__device__ void bp_follow(const BPsys& P, double& xc, double& yc) {
  const double x = P.x[0], y = P.y[0], z = P.z[0], vx = P.vx[0], vy = P.vy[0], vz = P.vz[0];
  xc = x;
  yc = y;
  if (!P.guiding) return;
  const double gm = 1.0 + P.m[0];
  const double d = sqrt(x * x + y * y + z * z);
  const double hx = y * vz - z * vy, hy = z * vx - x * vz, hz = x * vy - y * vx;
  const double hh = sqrt(hx * hx + hy * hy), h = sqrt(hx * hx + hy * hy + hz * hz);
  const double inc = asin(hh / h);
  const double Ax = vy * hz - vz * hy - gm * x / d, Ay = vz * hx - vx * hz - gm * y / d,
               Az = vx * hy - vy * hx - gm * z / d;
  const double e = sqrt(Ax * Ax + Ay * Ay + Az * Az) / gm;
  if (e < 1e-8) return;
  const double a = h * h / (gm * (1. - e * e));
  double E = acos((1.0 - d / a) / e);
  if (x * vx + y * vy + z * vz < 0) E = -E;
  if (isnan(E)) E = (d < a) ? 0.0 : M_PI;
  const double M = E - e * sin(E);
  const double phi = (Ax * Ax + Ay * Ay > 0.0) ? atan2(Ay, Ax) : atan2(y, x);
  xc = a * cos(M + phi) * cos(inc);
  yc = a * sin(M + phi) * cos(inc);
}

// compute_potential's planet part after Potential (src/potential.c:32-48):
// five RK5 substeps over dt; with a frame that follows the first planet, the
// new frame rate from the angle the followed point swept (GetPsysInfo
// MARK/GET) and its change for CorrectVtheta; then the rotation of the
// system into the new frame (src/psys.c:417-440).
template <int NP>
__device__ void bp_advance(BPsys& P, double dt, double& domega) {
  double x0, y0;
  bp_follow(P, x0, y0);
  for (int s = 0; s < 5; s++) bp_rk5<NP>(P, 1.0 / 5.0 * dt);
  domega = 0.0;
  if (P.corotating) {
    double x1, y1;
    bp_follow(P, x1, y1);
    const double d1 = sqrt(x0 * x0 + y0 * y0), d2 = sqrt(x1 * x1 + y1 * y1);
    const double omeganew = asin((x0 * y1 - x1 * y0) / (d1 * d2)) / dt;
    domega = omeganew - P.of;
    P.of = omeganew;
  }
  const double ang = P.of * dt, sn = sin(ang), cs = cos(ang);
  for (int i = 0; i < P.n; i++) {
    const double x = P.x[i], y = P.y[i], vx = P.vx[i], vy = P.vy[i];
    P.x[i] = x * cs + y * sn;
    P.y[i] = -x * sn + y * cs;
    P.vx[i] = vx * cs + vy * sn;
    P.vy[i] = -vx * sn + vy * cs;
  }
}

// CflFluidsMin and the output-time clamp (src/cfl_fluids_min.c:3-22,
// src/main.c:383-395, 423), one thread per instance; clears the MIN slot.
// With planets, then their step over dt and the frame's new rate, which the
// rest of the step's kernels read from the instance table.  NP: the largest
// planet count of the batch (1, 2, 4 or 8).
template <int NP>
__global__ void __launch_bounds__(64) kb_dt(BInst* I, int n) {
  const int ii = blockIdx.x * blockDim.x + threadIdx.x;
  if (ii >= n) return;
  BInst& in = I[ii];
  double* c = in.clk;
  const double DT = c[3];
  if (c[1] >= DT) c[1] = 0.0;
  const double fluid_min = __longlong_as_double((long long)in.res[0]);
  in.res[0] = kInfBits;
  double step = 1e30;
  if (step > fluid_min) step = fluid_min;
  if (step <= 1e-10) c[4] = 1.0;
  double dt = step;
  c[1] += dt;
  if (c[1] > DT) dt = DT - (c[1] - dt);
  c[0] = dt;
  if (in.ps) {
    BPsys& P = *in.ps;
    double domega = 0.0;
    bp_advance<NP>(P, dt, domega);
    in.of = (real)P.of;
    in.of64 = P.of;
    in.dp.of = P.of;
    in.b0.of = P.of;
    in.b1.of = P.of;
    in.domega = (real)domega;
    bp_prepare(P, c[2] + dt);  // the next step's Potential
  }
  c[2] += dt;
}

__global__ void __launch_bounds__(BXF) kb_src_a(const BInst* I, const int* off, int n) {
  const int ii = find_inst(off, n, blockIdx.x);
  const BInst& in = I[ii];
  const int gx = (in.m.nx + BXF - 1) / BXF;
  kd_src_a<CYL, 0>(local2(blockIdx.x - off[ii], gx), in.m, in.rho, in.cs, in.vx, in.vy, nullptr, in.mmx, in.mmy,
                   nullptr, in.mpx, in.mpy, nullptr, in.vp, in.visc, in.domega);
}

__global__ void __launch_bounds__(BXF) kb_src_b(const BInst* I, const int* off, int n) {
  const int ii = find_inst(off, n, blockIdx.x);
  const BInst& in = I[ii];
  const int gx = (in.m.nx + BXF - 1) / BXF;
  kd_src_b<CYL, 0>(local2(blockIdx.x - off[ii], gx), in.m, (real)in.clk[0], in.rho, in.cs, in.vx, in.vy, nullptr,
                   in.pot, in.mmx, in.mmy, nullptr, in.mpx, in.mpy, nullptr, in.vxt, in.vyt,
                   nullptr, in.e2, in.visc, in.potential, 0, 1, (real)0, in.of64, in.domega);
}

// the ghost fill in Y (both sides in one launch): mid = 1 fills V_temp
__global__ void __launch_bounds__(BXF) kb_bound(const BInst* I, const int* off, int n, int mid) {
  const int ii = find_inst(off, n, blockIdx.x);
  const BInst& in = I[ii];
  const int gx = (in.m.nx + BXF - 1) / BXF;
  kd_bound_y(local2(blockIdx.x - off[ii], gx), in.m, 0, in.rho, mid ? in.vxt : in.vx, mid ? in.vyt : in.vy, nullptr,
             in.cs, in.b0, in.b1, in.ydexp0, in.ydexp1);
}

__global__ void __launch_bounds__(BXF) kb_sweep_y(const BInst* I, const int* off, int n, int rows_per) {
  const int ii = find_inst(off, n, blockIdx.x);
  const BInst& in = I[ii];
  const int gx = (in.m.nx + BXF - 1) / BXF;
  kd_sweep_y2<CYL, 0, 0>(local2(blockIdx.x - off[ii], gx), in.m, (real)in.clk[0], in.rho, in.vxt, in.vyt, in.cs,
                         in.rhoy, in.mpx, in.mmx, in.mpy, in.mmy, in.ey, rows_per, in.of);
}

// one wave per ring, 8 terms per lane (the instances' rings hold 256 to a few
// thousand cells)
__global__ void __launch_bounds__(64) kb_vmed(const BInst* I, const int* off, int n) {
  const int ii = find_inst(off, n, blockIdx.x);
  const BInst& in = I[ii];
  kd_vmed<CYL, 0, 8>(dim3(blockIdx.x - off[ii]), in.m, in.vxt, in.vxmed);
}

template <int T>
__global__ void __launch_bounds__(T) kb_sweep_x(const BInst* I, const int* off, int n) {
  const int ii = find_inst(off, n, blockIdx.x);
  const BInst& in = I[ii];
  const int tx = (in.m.nx + x_tile_cells(T) - 1) / x_tile_cells(T);
  kd_sweep_x_tile<CYL, 0, 0, T>(local2(blockIdx.x - off[ii], tx), in.m, in.clk[0], in.rhoy, in.vxt, in.mpx, in.mmx,
                                in.mpy, in.mmy, nullptr, nullptr, in.ey, in.e2, in.rhox, in.vx, in.xm0, in.xm1,
                                in.xm2, in.xm3, in.vxmed, in.nshift, in.of, in.fargo);
}

__global__ void __launch_bounds__(BXF) kb_finish(const BInst* I, const int* off, int n) {
  const int ii = find_inst(off, n, blockIdx.x);
  const BInst& in = I[ii];
  const int gx = (in.m.nx + BXF - 1) / BXF;
  kd_finish<CYL, 0>(local2(blockIdx.x - off[ii], gx), in.m, (real)in.clk[0], in.rhox, in.xm1, in.xm0, in.xm3, in.xm2,
                    in.rho, in.vx, in.vy, nullptr, in.rho0, in.vx0, in.vy0, nullptr, in.dp, in.stockholm);
}

}  // namespace fused
}  // namespace F3D_NS

// ------------------------------------------------------------------ host side
struct f3d_batch {
  int n = 0;
  f3d_sim** sims = nullptr;
  hipStream_t stream = nullptr;
  F3D_NS::fused::BInst* dI = nullptr;
  double* dclk = nullptr;
  unsigned long long* dres = nullptr;
  F3D_NS::fused::BPsys* dps = nullptr;  // planetary systems, one per instance
  int* doff[F3D_NS::fused::BK_N] = {};
  int total[F3D_NS::fused::BK_N] = {};
  int* dswy = nullptr;  // workgroup offsets of the Y sweep for the chosen rows per thread
  int swy_total = 0;
  int* dxt = nullptr;   // and of the X tiles for the chosen tile
  int xt_total = 0;
  int nxmax = 0;
  int rows_y = 4, tile = 128;
  std::vector<int> has_ps;  // instance has a planetary system on the device
  int npmax = 0;            // the largest planet count
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::vector<double*> save;  // state copies used by the tuner
  char err[512] = {0};
};

namespace F3D_NS {
namespace fused {

#define BCHECK(B, expr)                                                                           \
  do {                                                                                            \
    hipError_t e_ = (expr);                                                                       \
    if (e_ != hipSuccess) {                                                                       \
      snprintf((B)->err, sizeof((B)->err), "%s: %s", #expr, hipGetErrorString(e_));               \
      return -1;                                                                                  \
    }                                                                                             \
  } while (0)

static int batch_offsets(f3d_batch* B, const std::vector<int>& counts, int** dst, int* total) {
  std::vector<int> off(B->n + 1, 0);
  for (int i = 0; i < B->n; i++) off[i + 1] = off[i] + counts[i];
  if (*dst) (void)hipFree(*dst);
  *dst = nullptr;
  BCHECK(B, hipMalloc(dst, off.size() * sizeof(int)));
  BCHECK(B, hipMemcpy(*dst, off.data(), off.size() * sizeof(int), hipMemcpyHostToDevice));
  *total = off[B->n];
  return 0;
}

static int batch_layout(f3d_batch* B) {
  std::vector<std::vector<int>> c(BK_N, std::vector<int>(B->n));
  std::vector<int> swy(B->n), xt(B->n);
  for (int i = 0; i < B->n; i++) {
    const f3d::dmesh& m = f3d_sim_ctx(B->sims[i])->m;
    const int gx = (m.nx + BXF - 1) / BXF;
    c[BK_CFL][i] = m.ny;
    c[BK_SRCA][i] = gx * (m.sy - 2);
    c[BK_SRCB][i] = gx * m.ny;
    c[BK_BND][i] = gx * 2;  // one block row per side, all ghost layers per thread
    c[BK_VMED][i] = m.ny;
    c[BK_FIN][i] = gx * m.ny;
    c[BK_POT][i] = B->has_ps[i] ? gx * m.sy : 0;
    swy[i] = gx * ((m.ny + B->rows_y - 1) / B->rows_y);
    xt[i] = ((m.nx + x_tile_cells(B->tile) - 1) / x_tile_cells(B->tile)) * m.ny;
  }
  for (int k = 0; k < BK_N; k++)
    if (k != BK_SWY && k != BK_XT && batch_offsets(B, c[k], &B->doff[k], &B->total[k])) return -1;
  if (batch_offsets(B, swy, &B->dswy, &B->swy_total)) return -1;
  if (batch_offsets(B, xt, &B->dxt, &B->xt_total)) return -1;
  return 0;
}

// one step: every stage one launch over all instances
static void batch_launch(f3d_batch* B) {
  const int n = B->n;
  hipStream_t s = B->stream;
  const size_t cfl_lds = 16 * sizeof(double);
  if (B->total[BK_POT])
    hipLaunchKernelGGL(kb_potential, dim3(B->total[BK_POT]), dim3(BXF), 0, s, B->dI, B->doff[BK_POT], n);
  hipLaunchKernelGGL(kb_cfl, dim3(B->total[BK_CFL]), dim3(WG), cfl_lds, s, B->dI, B->doff[BK_CFL], n);
  if (B->npmax <= 1) hipLaunchKernelGGL(kb_dt<1>, dim3((n + 63) / 64), dim3(64), 0, s, B->dI, n);
  else if (B->npmax <= 2) hipLaunchKernelGGL(kb_dt<2>, dim3((n + 63) / 64), dim3(64), 0, s, B->dI, n);
  else if (B->npmax <= 4) hipLaunchKernelGGL(kb_dt<4>, dim3((n + 63) / 64), dim3(64), 0, s, B->dI, n);
  else hipLaunchKernelGGL(kb_dt<kBP>, dim3((n + 63) / 64), dim3(64), 0, s, B->dI, n);
  hipLaunchKernelGGL(kb_src_a, dim3(B->total[BK_SRCA]), dim3(BXF), 0, s, B->dI, B->doff[BK_SRCA], n);
  hipLaunchKernelGGL(kb_src_b, dim3(B->total[BK_SRCB]), dim3(BXF), 0, s, B->dI, B->doff[BK_SRCB], n);
  hipLaunchKernelGGL(kb_bound, dim3(B->total[BK_BND]), dim3(BXF), 0, s, B->dI, B->doff[BK_BND], n, 1);
  hipLaunchKernelGGL(kb_sweep_y, dim3(B->swy_total), dim3(BXF), 0, s, B->dI, B->dswy, n, B->rows_y);
  hipLaunchKernelGGL(kb_vmed, dim3(B->total[BK_VMED]), dim3(64), 0, s, B->dI,
                     B->doff[BK_VMED], n);
  if (B->tile == 64) hipLaunchKernelGGL(kb_sweep_x<64>, dim3(B->xt_total), dim3(64), 0, s, B->dI, B->dxt, n);
  else if (B->tile == 256) hipLaunchKernelGGL(kb_sweep_x<256>, dim3(B->xt_total), dim3(256), 0, s, B->dI, B->dxt, n);
  else hipLaunchKernelGGL(kb_sweep_x<128>, dim3(B->xt_total), dim3(128), 0, s, B->dI, B->dxt, n);
  hipLaunchKernelGGL(kb_finish, dim3(B->total[BK_FIN]), dim3(BXF), 0, s, B->dI, B->doff[BK_FIN], n);
  hipLaunchKernelGGL(kb_bound, dim3(B->total[BK_BND]), dim3(BXF), 0, s, B->dI, B->doff[BK_BND], n, 0);
}

static int batch_capture(f3d_batch* B) {
  if (B->exec) (void)hipGraphExecDestroy(B->exec);
  if (B->graph) (void)hipGraphDestroy(B->graph);
  B->exec = nullptr;
  B->graph = nullptr;
  BCHECK(B, hipStreamBeginCapture(B->stream, hipStreamCaptureModeThreadLocal));
  batch_launch(B);
  BCHECK(B, hipStreamEndCapture(B->stream, &B->graph));
  BCHECK(B, hipGraphInstantiate(&B->exec, B->graph, nullptr, nullptr, 0));
  return 0;
}

// copies of the evolving state for the tuner: rho, vx, vy of every
// instance, the clocks, the planetary systems and the instance table (its
// frame rates change with the planets)
static int batch_state(f3d_batch* B, int restore) {
  std::vector<std::pair<void*, size_t>> reg;
  for (int i = 0; i < B->n; i++) {
    f3d_hip_ctx* c = f3d_sim_ctx(B->sims[i]);
    const size_t bytes = c->n3 * sizeof(real);
    reg.push_back({c->f.Density, bytes});
    reg.push_back({c->f.Vx, bytes});
    reg.push_back({c->f.Vy, bytes});
  }
  reg.push_back({B->dclk, (size_t)B->n * 5 * sizeof(double)});
  reg.push_back({B->dps, (size_t)B->n * sizeof(BPsys)});
  reg.push_back({B->dI, (size_t)B->n * sizeof(BInst)});
  for (size_t k = 0; k < reg.size(); k++) {
    if (!restore && k >= B->save.size()) {
      double* p = nullptr;
      BCHECK(B, hipMalloc(&p, reg[k].second));
      B->save.push_back(p);
    }
    if (restore)
      BCHECK(B, hipMemcpyAsync(reg[k].first, B->save[k], reg[k].second, hipMemcpyDeviceToDevice, B->stream));
    else
      BCHECK(B, hipMemcpyAsync(B->save[k], reg[k].first, reg[k].second, hipMemcpyDeviceToDevice, B->stream));
  }
  BCHECK(B, hipStreamSynchronize(B->stream));
  return 0;
}

}  // namespace fused
}  // namespace F3D_NS

extern "C" {

void f3d_batch_destroy(f3d_batch* B) {
  if (!B) return;
  if (B->stream) (void)hipStreamSynchronize(B->stream);
  if (B->exec) (void)hipGraphExecDestroy(B->exec);
  if (B->graph) (void)hipGraphDestroy(B->graph);
  for (int k = 0; k < F3D_NS::fused::BK_N; k++)
    if (B->doff[k]) (void)hipFree(B->doff[k]);
  if (B->dswy) (void)hipFree(B->dswy);
  if (B->dxt) (void)hipFree(B->dxt);
  for (double* p : B->save) (void)hipFree(p);
  if (B->dI) (void)hipFree(B->dI);
  if (B->dclk) (void)hipFree(B->dclk);
  if (B->dres) (void)hipFree(B->dres);
  if (B->dps) (void)hipFree(B->dps);
  if (B->sims)
    for (int i = 0; i < B->n; i++)
      if (B->sims[i]) f3d_sim_destroy(B->sims[i]);
  free(B->sims);
  if (B->stream) (void)hipStreamDestroy(B->stream);
  delete B;
}

f3d_batch* f3d_batch_create(int n, const char* const* parfiles, const char* const* overrides, char* err,
                            int errlen) {
  using namespace F3D_NS;
  using namespace F3D_NS::fused;
  f3d_batch* B = new f3d_batch();
  auto fail = [&](const char* msg) -> f3d_batch* {
    if (err) snprintf(err, errlen, "%s", msg);
    f3d_batch_destroy(B);
    return nullptr;
  };
  if (n < 1) return fail("f3d_batch_create: no instances");
  B->n = n;
  B->sims = (f3d_sim**)calloc(n, sizeof(f3d_sim*));
  std::vector<BInst> h(n);
  std::vector<double> clk(5 * (size_t)n, 0.0);
  if (hipStreamCreateWithFlags(&B->stream, hipStreamNonBlocking) != hipSuccess) return fail("stream creation failed");
  std::vector<BPsys> hps(n);
  B->has_ps.assign(n, 0);
  if (hipMalloc(&B->dclk, clk.size() * sizeof(double)) != hipSuccess ||
      hipMalloc(&B->dres, 2 * (size_t)n * sizeof(unsigned long long)) != hipSuccess ||
      hipMalloc(&B->dps, (size_t)n * sizeof(BPsys)) != hipSuccess)
    return fail("device allocation failed");
  for (int i = 0; i < n; i++) {
    char ov[1024];
    snprintf(ov, sizeof(ov), "%s Precision=32", overrides && overrides[i] ? overrides[i] : "");
    char e2[512] = {0};
    B->sims[i] = f3d_sim_create(parfiles[i], ov, e2, sizeof(e2));
    if (!B->sims[i]) return fail(e2);
    f3d_hip_ctx* c = f3d_sim_ctx(B->sims[i]);
    if (c->m.dimz || c->m.geom != CYL || c->ph.eos != 1 || c->ph.nplanets > cell::kMaxPlanets ||
        c->m.ny <= c->m.nghy || c->m.nx < 12)
      return fail("the batched form runs 2-D cylindrical isothermal disks (ny > 3, nx >= 12)");
    Scratch* S;
    if (scratch_of(c, &S) || F3D_API(fused_potential)(c) || hipStreamSynchronize(c->stream) != hipSuccess)
      return fail("instance set-up failed");
    const RF f = rfields(c);
    BInst& b = h[i];
    b.m = S->m;
    b.rho = f.Density; b.cs = f.Energy; b.vx = f.Vx; b.vy = f.Vy; b.vxt = f.Vx_temp; b.vyt = f.Vy_temp;
    b.mpx = f.Mpx; b.mmx = f.Mmx; b.mpy = f.Mpy; b.mmy = f.Mmy; b.pot = c->f.Pot;
    b.rho0 = f.Density0; b.vx0 = f.Vx0; b.vy0 = f.Vy0;
    b.rhoy = S->f[FS_RHOY]; b.rhox = S->f[FS_RHOX];
    b.e2 = S->f[FS_E2]; b.ey = S->f[FS_EY];
    b.xm0 = S->f[FS_XM0]; b.xm1 = S->f[FS_XM0 + 1]; b.xm2 = S->f[FS_XM0 + 2]; b.xm3 = S->f[FS_XM0 + 3];
    b.vxmed = f.VxMed; b.nshift = f.Nshift;
    b.ydexp0 = S->ymin_dexp; b.ydexp1 = S->ymax_dexp;
    b.cp = cfl_par_f(c);
    b.vp.model = c->ph.visc; b.vp.nu = (real)c->ph.nu; b.vp.alpha = (real)c->ph.alpha;
    b.vp.gamma = (real)c->ph.gamma; b.vp.iso = 1;
    b.dp.Y_inf = (real)c->ph.Y_inf; b.dp.Y_sup = (real)c->ph.Y_sup; b.dp.Z_inf = (real)c->ph.Z_inf;
    b.dp.Z_sup = (real)c->ph.Z_sup; b.dp.y_min = (real)c->ph.ymin_par; b.dp.y_max = (real)c->ph.ymax_par;
    b.dp.z_min = (real)c->ph.zmin_par; b.dp.z_max = (real)c->ph.zmax_par; b.dp.ds = (real)c->ph.taudamp;
    b.dp.of = c->ph.OMEGAFRAME; b.dp.of0 = c->ph.OMEGAFRAME0; b.dp.periodic_z = c->ph.periodicz;
    b.b0 = bc_par_f(c, 0);
    b.b1 = bc_par_f(c, 1);
    b.of = (real)c->ph.OMEGAFRAME;
    b.of64 = c->ph.OMEGAFRAME;
    b.visc = c->ph.visc != 0 && c->ph.fluidtype == 1;
    b.potential = c->ph.potential;
    b.stockholm = c->ph.stockholm;
    b.fargo = c->ph.fargo;
    b.domega = (real)0;
    b.ps = nullptr;
    b.ymed64 = c->m.ymed;
    b.cosx64 = c->m.cosxmed;
    b.sinx64 = c->m.sinxmed;
    b.zc = 0.0;
    {
      const f3d_params* p = f3d_sim_params(B->sims[i]);
      const f3d_psys* sys = f3d_sim_psys(B->sims[i]);
      const f3d_clock* ck = f3d_sim_clock(B->sims[i]);
      BPsys& P = hps[i];
      memset(&P, 0, sizeof(P));
      P.of = ck->OMEGAFRAME;
      if (c->ph.potential && sys->nb > 0) {
        if (p->releasedate > 0.0) return fail("the batched form's planets have no release date");
        for (int k = 0; k < sys->nb; k++)
          if (sys->feeldisk[k]) return fail("the batched form's planets do not feel the disk");
        P.n = sys->nb;
        P.indirect = p->indirectterm;
        P.corotating = p->corotating;
        P.guiding = p->guidingcenter;
        P.masstaper = p->masstaper;
        P.rochesmoothing = p->rochesmoothing;
        P.aspectratio = p->aspectratio;
        P.flaringindex = p->flaringindex;
        P.thicknesssmoothing = p->thicknesssmoothing;
        const double taper = f3d_planet_taper(p, ck->PhysicalTime);
        for (int k = 0; k < sys->nb; k++) {
          P.x[k] = sys->x[k]; P.y[k] = sys->y[k]; P.z[k] = sys->z[k];
          P.vx[k] = sys->vx[k]; P.vy[k] = sys->vy[k]; P.vz[k] = sys->vz[k];
          P.m[k] = sys->mass[k];
          P.feelothers[k] = sys->feelothers[k];
          P.pm[k] = sys->mass[k] * taper;
          P.ps2[k] = f3d_planet_smoothing2(sys, p, k, taper);
        }
        b.ps = B->dps + i;
        B->has_ps[i] = 1;
        if (P.n > B->npmax) B->npmax = P.n;
      }
    }
    b.clk = B->dclk + 5 * (size_t)i;
    b.res = B->dres + 2 * (size_t)i;
    clk[5 * (size_t)i + 3] = f3d_sim_dt_out(B->sims[i]);
    if (c->m.nx > B->nxmax) B->nxmax = c->m.nx;
  }
  std::vector<unsigned long long> res(2 * (size_t)n, kInfBits);
  if (hipMalloc(&B->dI, n * sizeof(BInst)) != hipSuccess ||
      hipMemcpy(B->dI, h.data(), n * sizeof(BInst), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(B->dclk, clk.data(), clk.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(B->dps, hps.data(), hps.size() * sizeof(BPsys), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(B->dres, res.data(), res.size() * sizeof(unsigned long long), hipMemcpyHostToDevice) != hipSuccess)
    return fail("device upload failed");
  if (batch_layout(B) || batch_capture(B)) return fail(B->err);
  return B;
}

const char* f3d_batch_error(f3d_batch* B) { return B ? B->err : "no batch"; }

int f3d_batch_size(f3d_batch* B) { return B->n; }

f3d_sim* f3d_batch_instance(f3d_batch* B, int i) { return (i >= 0 && i < B->n) ? B->sims[i] : nullptr; }

int f3d_batch_step(f3d_batch* B, int nsteps) {
  for (int s = 0; s < nsteps; s++) BCHECK(B, hipGraphLaunch(B->exec, B->stream));
  return 0;
}

int f3d_batch_sync(f3d_batch* B) {
  BCHECK(B, hipStreamSynchronize(B->stream));
  std::vector<double> clk(5 * (size_t)B->n);
  BCHECK(B, hipMemcpy(clk.data(), B->dclk, clk.size() * sizeof(double), hipMemcpyDeviceToHost));
  for (int i = 0; i < B->n; i++)
    if (clk[5 * (size_t)i + 4] != 0.0) {
      snprintf(B->err, sizeof(B->err), "instance %d: null time step", i);
      return -1;
    }
  return 0;
}

int f3d_batch_clock(f3d_batch* B, int i, double out[3]) {
  if (i < 0 || i >= B->n) return -2;
  BCHECK(B, hipStreamSynchronize(B->stream));
  BCHECK(B, hipMemcpy(out, B->dclk + 5 * (size_t)i, 3 * sizeof(double), hipMemcpyDeviceToHost));
  return 0;
}

int f3d_batch_planets(f3d_batch* B, int i, double* out, int maxn, double* omegaframe) {
  using namespace F3D_NS::fused;
  if (i < 0 || i >= B->n) return -2;
  BPsys P;
  BCHECK(B, hipStreamSynchronize(B->stream));
  BCHECK(B, hipMemcpy(&P, B->dps + i, sizeof(BPsys), hipMemcpyDeviceToHost));
  if (omegaframe) *omegaframe = P.of;
  for (int k = 0; k < P.n && k < maxn; k++) {
    double* o = out + 8 * (size_t)k;
    o[0] = P.x[k]; o[1] = P.y[k]; o[2] = P.z[k]; o[3] = P.vx[k]; o[4] = P.vy[k]; o[5] = P.vz[k];
    o[6] = P.m[k]; o[7] = (double)P.n;
  }
  return P.n;
}

int f3d_batch_set_tiling(f3d_batch* B, int rows_y, int tile) {
  using namespace F3D_NS;
  using namespace F3D_NS::fused;
  if (rows_y < 1 || (tile != 64 && tile != 128 && tile != 256)) return -2;
  B->rows_y = rows_y;
  B->tile = tile;
  if (batch_layout(B) || batch_capture(B)) return -1;
  return 0;
}

int f3d_batch_tune(f3d_batch* B, int reps, int* best_rows_y, int* best_tile) {
  using namespace F3D_NS;
  using namespace F3D_NS::fused;
  if (reps < 1) reps = 5;
  if (batch_state(B, 0)) return -1;
  hipEvent_t e0, e1;
  BCHECK(B, hipEventCreate(&e0));
  BCHECK(B, hipEventCreate(&e1));
  float best = 1e30f;
  int br = B->rows_y, bt = B->tile;
  const int rows[] = {1, 2, 4, 8}, tiles[] = {64, 128, 256};
  for (int r : rows)
    for (int t : tiles) {
      if (f3d_batch_set_tiling(B, r, t)) return -1;
      if (f3d_batch_step(B, 2)) return -1;
      BCHECK(B, hipEventRecord(e0, B->stream));
      if (f3d_batch_step(B, reps)) return -1;
      BCHECK(B, hipEventRecord(e1, B->stream));
      BCHECK(B, hipEventSynchronize(e1));
      float ms = 0.f;
      BCHECK(B, hipEventElapsedTime(&ms, e0, e1));
      if (ms < best) {
        best = ms;
        br = r;
        bt = t;
      }
      if (batch_state(B, 1)) return -1;
    }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (f3d_batch_set_tiling(B, br, bt)) return -1;
  if (best_rows_y) *best_rows_y = br;
  if (best_tile) *best_tile = bt;
  return 0;
}

}  // extern "C"
