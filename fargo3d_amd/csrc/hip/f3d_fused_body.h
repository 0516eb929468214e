// f3d_fused_body.h -- the fused step, written once for the type `real` of a
// build: included by f3d_fused.hip (real = double, namespace f3d, entry points
// f3d_fused_*) and f3d_fused32.hip (real = float, namespace f3d32, entry points
// f3d32_fused_*).  The parts that stay FP64 in every build are written with
// `double`: ring sums (VxMed), the residual/shift split, the CFL time step.
// Requires F3D_NS, F3D_PREC (64 or 32) and F3D_API(name) from the includer.
//
// The per-routine entry points (f3d_routines.hip) move every intermediate
// array through HBM, as the reference's one-launch-per-routine GPU build
// does (about 370 array passes per step).  The kernels here compute the same
// values with the same per-cell arithmetic (f3d_cell.h) but keep
// intermediates on chip:
//
//   kf_cfl         ring-owning: VxMed of the ring and of the next ring
//                  (sequential sums, in parallel, or from kf_vmed_rings),
//                  per-cell CFL, shear limit, one MIN per ring (atomic, or
//                  into an array that kf_cfl_fin reduces) -> dt, 8 bytes to host
//   kf_src_a       viscous stress and artificial pressure at every cell the
//                  next kernel reads
//   kf_src_b       SubStep1 + addviscosity + SubStep2_b on active cells
//   (boundaries)   mid-step ghost fill on V_temp
//   kf_sweep_z     3-D: momenta + Z sweep, one thread per (i,j) column
//                  marching in k over a register window
//   kf_sweep_y     momenta (2-D) + Y sweep, columns marching in j
//   kf_sweep_xr_fluids ring-owning (every fluid in one launch): VxMed,
//                  residual, van Leer and PPA X sweeps, integer shift and
//                  NewVelocity_x, cell state in registers, stencils in LDS
//   kf_sweep_x_tile the same X sweep on tiles with halos, for long rings
//   kf_finish      NewVelocity_y/z and Stockholm damping
//   (boundaries)   end-of-step ghost fill
//
// Only the active cells of the primitive fields carry information from one
// stage to the next: every ghost value is rewritten by the ghost fill from
// active values (std/boundaries.txt formulas read mirror rows only), so the
// fused kernels evaluate ghost rows only where an active cell's stencil
// reads them, and there with the reference's formulas.
namespace F3D_NS {
namespace fused {

namespace {

constexpr int WG = 256;
constexpr int BXF = 64;
// halo cells on each side of an X-sweep tile: the van Leer pass reaches 2
// cells, the PPA pass 3 more, NewVelocity_x the left neighbour
constexpr int kXHalo = 6;
constexpr int kNgh = 3;  // ghost layers in Y and Z (src/define.h:57-63)

// Scratch of the fused stages, n3 doubles each, allocated on a handle's
// first fused call and released with the handle: Y- and X-swept density,
// three work arrays (dust diffusion), energy after SubStep2_b and after the Y
// sweep, the X-swept Y/Z momenta (4) and (3-D) the Z-swept density, six
// momenta and energy.
enum { FS_RHOY = 0, FS_RHOX, FS_W0, FS_W1, FS_W2, FS_E2, FS_EY, FS_XM0, FS_Z0 = FS_XM0 + 4, FS_COUNT = FS_Z0 + 8 };

struct Scratch {
  real* f[FS_COUNT];
  unsigned long long* result;  // two slots, alternated per step
  int slot;
  double* ringmin;             // per-ring CFL minima of up to 8 fluids (two-pass MIN)
  real* fset[8][13];           // transport arrays of fluids 1..7 (multifluid transport)
  double* vxmed_f[8];          // VxMed and Nshift of the fluids whose values the
  int* nshift_f[8];            // reference overwrites (all but the last)
  real* ddtemp[8];             // dust diffusion: the new density of fluid f before its copy back
  dmesh m;                     // the build's mesh view
  mreal* mesh_r;               // FP32 build: converted mesh arrays
  real* bctab_r;               // FP32 build: converted boundary tables
  const real *ymin_dexp, *ymax_dexp, *zmin_dens, *zmax_dens;
};

constexpr unsigned long long kInfBits = 0x7FF0000000000000ull;

void scratch_free(void* p) {
  Scratch* S = static_cast<Scratch*>(p);
  for (int a = 0; a < FS_COUNT; a++)
    if (S->f[a]) (void)hipFree(S->f[a]);
  if (S->result) (void)hipFree(S->result);
  if (S->ringmin) (void)hipFree(S->ringmin);
  for (int f = 0; f < 8; f++) {
    for (int a = 0; a < 13; a++)
      if (S->fset[f][a]) (void)hipFree(S->fset[f][a]);
    if (S->vxmed_f[f]) (void)hipFree(S->vxmed_f[f]);
    if (S->nshift_f[f]) (void)hipFree(S->nshift_f[f]);
    if (S->ddtemp[f]) (void)hipFree(S->ddtemp[f]);
  }
  if (S->mesh_r) (void)hipFree(S->mesh_r);
  if (S->bctab_r) (void)hipFree(S->bctab_r);
  delete S;
}

template <class T>
__global__ void k_convert(const double* a, T* b, size_t n) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) b[t] = (T)a[t];
}

// The build's view of the handle's mesh: the FP64 build uses the handle's
// arrays; the FP32 build converts them once (the FP64 copies stay for the
// parts that remain FP64).
int make_mesh(f3d_hip_ctx* ctx, Scratch* S) {
#if F3D_PREC == 64
  S->m = ctx->m;
  S->ymin_dexp = ctx->d_ymin_dexp;
  S->ymax_dexp = ctx->d_ymax_dexp;
  S->zmin_dens = ctx->d_zmin_dens;
  S->zmax_dens = ctx->d_zmax_dens;
#else
  const f3d::dmesh& d = ctx->m;
  const size_t n = ctx->meshlen;
  if (hipMalloc(&S->mesh_r, n * sizeof(mreal)) != hipSuccess) return -1;
  hipLaunchKernelGGL(k_convert<mreal>, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, ctx->meshbuf, S->mesh_r, n);
  dmesh& m = S->m;
  m.nx = d.nx; m.ny = d.ny; m.nz = d.nz; m.nghy = d.nghy; m.nghz = d.nghz; m.dimz = d.dimz; m.geom = d.geom;
  m.sx = d.sx; m.sy = d.sy; m.sz = d.sz; m.pitch = d.pitch; m.stride = d.stride;
  m.XMIN = d.XMIN; m.XMAX = d.XMAX; m.mindx = d.mindx;
  auto conv = [&](const double* p) -> const mreal* { return p ? S->mesh_r + (p - ctx->meshbuf) : nullptr; };
  m.xmin = conv(d.xmin); m.ymin = conv(d.ymin); m.zmin = conv(d.zmin);
  m.xmed = conv(d.xmed); m.ymed = conv(d.ymed); m.zmed = conv(d.zmed);
  m.sxi = conv(d.sxi); m.sxj = conv(d.sxj); m.syj = conv(d.syj); m.szj = conv(d.szj);
  m.sxk = conv(d.sxk); m.syk = conv(d.syk); m.szk = conv(d.szk); m.invvj = conv(d.invvj);
  m.invdiffxmed = conv(d.invdiffxmed); m.cosxmed = conv(d.cosxmed); m.sinxmed = conv(d.sinxmed);
  m.sinzmed = conv(d.sinzmed); m.coszmed = conv(d.coszmed); m.sinzmin = conv(d.sinzmin); m.coszmin = conv(d.coszmin);
  m.xmin64 = d.xmin; m.ymin64 = d.ymin; m.zmin64 = d.zmin; m.ymed64 = d.ymed; m.sinzmed64 = d.sinzmed;
  m.sxi64 = d.sxi; m.zmed64 = d.zmed; m.sinzmin64 = d.sinzmin; m.coszmin64 = d.coszmin;
  if (!ctx->vbase) return -1;  // set with f3d_hip_set_vbase before the first FP32 stage
  m.vb = ctx->vbase;
  if (ctx->bctab) {
    const size_t ny3 = 3 * (size_t)d.sz, nz3 = 3 * (size_t)d.sy, nb = 2 * ny3 + 2 * nz3;
    if (hipMalloc(&S->bctab_r, nb * sizeof(real)) != hipSuccess) return -1;
    hipLaunchKernelGGL(k_convert<real>, dim3((nb + 255) / 256), dim3(256), 0, ctx->stream, ctx->bctab, S->bctab_r, nb);
    S->ymin_dexp = S->bctab_r;
    S->ymax_dexp = S->bctab_r + ny3;
    S->zmin_dens = S->bctab_r + 2 * ny3;
    S->zmax_dens = S->bctab_r + 2 * ny3 + nz3;
  }
  if (hipStreamSynchronize(ctx->stream) != hipSuccess) return -1;
#endif
  return 0;
}

constexpr int kPrec = (F3D_PREC == 32) ? 1 : 0;

int scratch_of(f3d_hip_ctx* ctx, Scratch** out) {
  if (!ctx->fused[kPrec]) {
    Scratch* S = new Scratch();
    memset(S, 0, sizeof(*S));
    ctx->fused[kPrec] = S;
    ctx->fused_free[kPrec] = scratch_free;
    if (make_mesh(ctx, S)) return -1;
    const int nbuf = ctx->m.dimz ? FS_COUNT : FS_Z0;
    for (int a = 0; a < nbuf; a++) {
      if (hipMalloc(&S->f[a], ctx->n3 * sizeof(real)) != hipSuccess) return -1;
      if (hipMemset(S->f[a], 0, ctx->n3 * sizeof(real)) != hipSuccess) return -1;
    }
    if (hipMalloc(&S->result, 2 * sizeof(unsigned long long)) != hipSuccess) return -1;
    const size_t nrings = (size_t)ctx->m.ny * (ctx->m.dimz ? ctx->m.nz : 1);
    if (hipMalloc(&S->ringmin, 8 * nrings * sizeof(double)) != hipSuccess) return -1;
    unsigned long long init[2] = {kInfBits, kInfBits};
    if (hipMemcpy(S->result, init, sizeof(init), hipMemcpyHostToDevice) != hipSuccess) return -1;
  }
  *out = static_cast<Scratch*>(ctx->fused[kPrec]);
  return 0;
}

// Whole-wave rotations by one lane (DPP wave_ror:1 and wave_rol:1 of the
// gfx9 family): from_left gives lane L the value of lane L-1 and lane 0 that
// of lane 63; from_right gives lane L the value of lane L+1 and lane 63 that
// of lane 0.  All 64 lanes must be active.
__device__ __forceinline__ float from_left(float x) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x13C, 0xF, 0xF, false));
}
__device__ __forceinline__ float from_right(float x) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x134, 0xF, 0xF, false));
}
__device__ __forceinline__ double from_left(double x) {
  return __hiloint2double(__builtin_amdgcn_update_dpp(0, __double2hiint(x), 0x13C, 0xF, 0xF, false),
                          __builtin_amdgcn_update_dpp(0, __double2loint(x), 0x13C, 0xF, 0xF, false));
}
__device__ __forceinline__ double from_right(double x) {
  return __hiloint2double(__builtin_amdgcn_update_dpp(0, __double2hiint(x), 0x134, 0xF, 0xF, false),
                          __builtin_amdgcn_update_dpp(0, __double2loint(x), 0x134, 0xF, 0xF, false));
}

__device__ __forceinline__ double wave_max(double v) {
  for (int o = 32; o > 0; o >>= 1) {
    double w = __shfl_down(v, o, 64);
    v = (w > v) ? w : v;
  }
  return v;
}

// Block-wide max; every thread gets the result.
__device__ double block_max(double v, double* sh) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
  v = wave_max(v);
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0) {
    for (int w = 0; w < nw; w++) r = (sh[w] > r) ? sh[w] : r;
    sh[0] = r;
  }
  __syncthreads();
  r = sh[0];
  __syncthreads();
  return r;
}

// Workgroups go to the chip's 8 XCDs round-robin by linear id, and every XCD
// has an L2 of its own.  xcd_linear() renumbers a launch so that each XCD
// takes one contiguous range of the kernel's logical workgroups: neighbours
// that read each other's halo cells (or the same ring) then share one L2.
// A bijection on [0, N): results do not depend on it.  Used where it
// measured faster on MI355X (the sources, Y sweeps, X sweeps and CFL; the Z
// sweep, SubStep3 and the finishing kernel run faster in launch order).
__device__ __forceinline__ int xcd_linear() {
  const int N = (int)(gridDim.x * gridDim.y * gridDim.z);
  const int L = (int)(blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z));
  const int x = L & 7, q = N >> 3, r = N & 7;
  return x * q + min(x, r) + (L >> 3);
}
__device__ __forceinline__ dim3 xcd_bid() {
  const int p = xcd_linear();
  return dim3(p % gridDim.x, (p / gridDim.x) % gridDim.y, p / (gridDim.x * gridDim.y));
}

// The sequential sum in index order (the reduction order of
// src/reduction_generic.c:18-20), in FP64, of a ring's weights
// w_i = v_i (x_{i+1} - x_i) (src/compute_vmed.c:14-21), i in [0, n), by one
// wave (all 64 lanes taking part) with the terms in registers.  The ring goes
// in chunks of 64*S terms, lane L holding terms [c + L*S, c + (L+1)*S) of the
// chunk at c.  The running sum
// passes from lane to lane: every lane adds its own S terms to the sum so far
// (one chain of dependent register additions), and lane L's result -- the
// sequential sum through its last term -- is read out (v_readlane) as the
// start of lane L+1's chain; the other lanes' results are discarded.  Terms
// past the ring's end are +0.0, which leaves a sum started at +0.0 unchanged
// (in round-to-nearest such a sum is never -0.0).  The next chunk's loads are
// issued before the current chunk's chain.  Every lane returns the sum.
// Measured on MI355X: a dependent FP64 addition in registers takes about
// 4.5 ns, the same sum read from LDS about 12 ns per term.
// The chain of wave_ring_sum over one chunk: lane L's S terms w follow the
// terms of lanes 0 .. L-1, after acc; lanes 0 .. lanes-1 hold terms.
template <int S>
__device__ __forceinline__ double wave_chain(const double (&w)[S], double acc, int lanes) {
#pragma clang fp contract(off)
  for (int L = 0; L < lanes; L++) {
    double t = acc;
#pragma unroll
    for (int s = 0; s < S; s++) t = t + w[s];
    acc = __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(t), L),
                           __builtin_amdgcn_readlane(__double2loint(t), L));
  }
  return acc;
}
template <int S>
__device__ __forceinline__ double wave_ring_sum(const real* __restrict__ v, const double* __restrict__ x, int n) {
#pragma clang fp contract(off)
  const int lane = (int)threadIdx.x & 63;
  double a[S], x0[S + 1];
  auto load = [&](int c) {
#pragma unroll
    for (int s = 0; s < S; s++) a[s] = (double)v[min(c + lane * S + s, n - 1)];
#pragma unroll
    for (int s = 0; s <= S; s++) x0[s] = x[min(c + lane * S + s, n)];
  };
  load(0);
  double acc = 0.0;
  for (int c = 0; c < n; c += 64 * S) {
    double w[S];
#pragma unroll
    for (int s = 0; s < S; s++) w[s] = (c + lane * S + s < n) ? a[s] * (x0[s + 1] - x0[s]) : 0.0;
    if (c + 64 * S < n) load(c + 64 * S);
    acc = wave_chain<S>(w, acc, min(64, (n - c + S - 1) / S));
  }
  return acc;
}

// VxMed of ring (j,k) from the ring mean of the stored velocity: the
// precision form stores deviations from V0(j,k), so the full mean adds V0
__device__ __forceinline__ double vmed_full(double stored_mean, const dmesh& m, int j, int k) {
  if constexpr (cell::kFix64) return stored_mean + cell::vbase(m, j, k);
  return stored_mean;
}

}  // namespace

// ----------------------------------------------------------------- kf_cfl
// One workgroup per active ring (j,k), in FP64 in every build.  The cell time
// step CFL/sqrt(s) is monotone non-increasing in s and IEEE sqrt/division
// round monotonically, so the MIN over a ring's cells is CFL/sqrt(max s)
// exactly; the MIN over rings is order independent, so one atomic MIN per
// ring on the bit pattern of the positive double gives the value of
// reduction_full_MIN + cfl_b (src/cfl_b.c:6-37) exactly.
// PRE: VxMed of the ring and the next from kf_vmed_rings; else summed here
// (the in-kernel sum's registers would halve the occupancy of the PRE form,
// hence a template parameter).
template <int G, int DZ, int PRE, int FX = 0>
__device__ __forceinline__ void kd_cfl(int ring, const dmesh& m, const real* cs, const real* rho, const real* vx, const real* vy, const real* vz, const real* vxm, double* vxmed, unsigned long long* result, int slot, cell::CflPar cp, double* ringmin) {
  extern __shared__ double lds[];  // 16
  double* sh = lds;
  const int j = m.nghy + ring % m.ny;
  const int k = DZ ? m.nghz + ring / m.ny : 0;
  const int base = j * m.pitch + k * m.stride;
  const int l2 = j + k * m.sy;
  if (ring == 0 && threadIdx.x == 0) result[slot ^ 1] = kInfBits;  // the next step's slot
  if constexpr (PRE) {
    // VxMed of rings j and j+1 from the pass before (kf_vmed_rings)
    if (threadIdx.x == 0) {
      sh[8] = vxmed[l2];
      sh[9] = vxmed[l2 + 1];
    }
  } else {
    // ComputeVweight + reduction_SUM (src/compute_vmed.c:14-30) of rings j and j+1
    // of vxm (the velocity whose VxMed the reference's cfl reads: with several
    // fluids, the last fluid's, src/main.c:366 then :381): the first wave sums
    // ring j, the second ring j+1, each with its terms in registers
    // (wave_ring_sum; this form runs with workgroups of four waves)
    const int wv = (int)threadIdx.x >> 6;
    if (wv == 0) {
      const double s0 = wave_ring_sum<4>(vxm + base, m.xmin64, m.nx);
      if (threadIdx.x == 0) sh[8] = vmed_full(s0 / (m.XMAX - m.XMIN), m, j, k);
    } else if (wv == 1 && cp.fargo) {
      const double s1 = wave_ring_sum<4>(vxm + base + m.pitch, m.xmin64, m.nx);
      if (threadIdx.x == 64) sh[9] = vmed_full(s1 / (m.XMAX - m.XMIN), m, j + 1, k);
    }
  }
  __syncthreads();
  const double vm0 = sh[8];
  // MAX over the ring's cells; four cells per thread at a time, their loads
  // issued together (indexes clamped into the ring, the extra cells masked)
  double mx = 0.0;
  const int t = (int)threadIdx.x, nt = (int)blockDim.x;
  for (int i0 = 0; i0 < m.nx; i0 += 4 * nt) {
    double s4[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int i = min(i0 + t + u * nt, m.nx - 1);
      s4[u] = cell::cfl_sumsq<G, DZ, FX>(m, i, j, k, base + i, cs, rho, vx, vy, vz, vm0, cp);
    }
#pragma unroll
    for (int u = 0; u < 4; u++)
      if (i0 + t + u * nt < m.nx) mx = (s4[u] > mx) ? s4[u] : mx;
  }
  mx = block_max(mx, sh);
  if (threadIdx.x == 0) {
    if (!PRE) vxmed[l2] = vm0;
    double mn = cp.cfl / sqrt(mx);
    if (cp.fargo) {
      const double shearlimit = cp.cfl * m.mindx / fabs(vm0 / m.ymed64[j] - sh[9] / m.ymed64[j + 1]);
      if (shearlimit < mn) mn = shearlimit;
    }
    if (ringmin) ringmin[ring] = mn;
    else atomicMin(&result[slot], (unsigned long long)__double_as_longlong(mn));
  }
}
template <int G, int DZ, int PRE>
__global__ void __launch_bounds__(WG) kf_cfl(dmesh m, const real* __restrict__ cs, const real* __restrict__ rho, const real* __restrict__ vx,
                                             const real* __restrict__ vy, const real* __restrict__ vz, const real* __restrict__ vxm, double* __restrict__ vxmed,
                                             unsigned long long* __restrict__ result, int slot, cell::CflPar cp,
                                             double* __restrict__ ringmin) {
  kd_cfl<G, DZ, PRE>(xcd_linear(), m, cs, rho, vx, vy, vz, vxm, vxmed, result, slot, cp, ringmin);
}

// ComputeVmed (src/compute_vmed.c:14-30) of rings j = NGHY .. NGHY+nry-1 of
// every active plane: one wave per ring, one sequential sum in i order with
// S terms per lane (wave_ring_sum).
template <int G, int DZ, int S>
__global__ void __launch_bounds__(64) kf_vmed_rings(dmesh m, const real* __restrict__ v, double* __restrict__ vxmed, int nry) {
  const int j = m.nghy + (int)(blockIdx.x % nry);
  const int k = DZ ? m.nghz + (int)(blockIdx.x / nry) : 0;
  const int base = j * m.pitch + k * m.stride;
  const double sum = wave_ring_sum<S>(v + base, m.xmin64, m.nx);
  if (threadIdx.x == 0) vxmed[j + k * m.sy] = vmed_full(sum / (m.XMAX - m.XMIN), m, j, k);
}

// The cfl of nf fluids in one launch: nf consecutive blocks of nrings
// workgroups, each block reading its fluid's arrays from the table.  Every
// fluid's workgroup of a ring computes the same VxMed of vxm and stores the
// same value.
struct CflTab {
  const real* cs[8];
  const real* rho[8];
  const real* vx[8];
  const real* vy[8];
  const real* vz[8];
  double* ringmin[8];
};
// PF bit 0: PRE; PF >> 1: FX of cell::cfl_sumsq (the flags at compile time)
template <int G, int DZ, int PF>
__global__ void __launch_bounds__(WG) kf_cfl_fluids(dmesh m, CflTab T, int nrings, const real* vxm, double* vxmed,
                                                    unsigned long long* result, int slot, cell::CflPar cp) {
  const int b = xcd_linear();
  const int fl = b / nrings, r = b % nrings;
  kd_cfl<G, DZ, PF & 1, (PF >> 1)>(r, m, T.cs[fl], T.rho[fl], T.vx[fl], T.vy[fl], T.vz[fl], vxm, vxmed, result,
                                   slot, cp, T.ringmin[fl]);
}

// Second pass of the two-pass MIN: one workgroup takes the MIN of the n ring
// minima (order independent, so equal to the atomic form) into result[slot].
__global__ void __launch_bounds__(1024) kf_cfl_fin(const double* ringmin, int n, unsigned long long* result,
                                                   int slot, double* host_out, unsigned long long seq) {
  __shared__ double sh[16];
  cfl_min_publish(ringmin, n, result, slot, host_out, seq, sh);
}

// ------------------------------------------------------------------ sources
// Loader of the pressure field computed on the fly (src/CompPresIso.c:56,
// src/CompPresAd.c:58).
struct PresLoader {
  cell::GL rho, e;
  real gm1;
  int iso;
  __device__ __forceinline__ real operator()(int di, int dj, int dk) const {
    if (iso) {
      real c = e(di, dj, dk);
      return rho(di, dj, dk) * c * c;
    }
    return gm1 * e(di, dj, dk);
  }
};

// Loader of Vx after CorrectVtheta (src/correct_vtheta.c:51-56) applied on the fly.
template <int G>
struct VxcLoader {
  cell::GL v;
  const mreal* ymed;
  const mreal* sinzmed;
  int j, k;
  real domega;
  __device__ __forceinline__ real operator()(int di, int dj, int dk) const {
    if (G == SPH) return v(di, dj, dk) - domega * ymed[j + dj] * sinzmed[k + dk];
    return v(di, dj, dk) - domega * ymed[j + dj];
  }
};

template <int G>
__device__ __forceinline__ VxcLoader<G> vxc(const real* vx, const dmesh& m, int l, int i, int j, int k, real dom) {
  return VxcLoader<G>{cell::gl(vx, m, l, i), m.ymed, m.sinzmed, j, k, dom};
}

// The same for the viscous stress in the precision form, which stores the
// deviation vx' from V0(j,k): the stress differentiates vx across rings and
// planes, where V0 varies, so the loader returns the full V0 + vx'.
template <int G>
struct VxfLoader {
  cell::GL v;
  const dmesh* m;
  int j, k;
  real domega;
  __device__ __forceinline__ real operator()(int di, int dj, int dk) const {
    const real full = (real)((double)v(di, dj, dk) + cell::vbase(*m, j + dj, k + dk));
    if (G == SPH) return full - domega * m->ymed[j + dj] * m->sinzmed[k + dk];
    return full - domega * m->ymed[j + dj];
  }
};
template <int G>
__device__ __forceinline__ auto vx_stress(const real* vx, const dmesh& m, int l, int i, int j, int k, real dom) {
  if constexpr (cell::kFix64) return VxfLoader<G>{cell::gl(vx, m, l, i), &m, j, k, dom};
  else return vxc<G>(vx, m, l, i, j, k, dom);
}

// The artificial pressure of SubStep2_a (src/substep2_a.c:103-113) at a
// cell, computed where SubStep2_b reads it, from the same loaders and with
// the same arithmetic as the array the reference stores: q = art_q(rho,
// v(+1 along AX) - v).
template <class LR, class LV, int AX>
struct ArtQ {
  LR r;
  LV v;
  __device__ __forceinline__ real operator()(int di, int dj, int dk) const {
    return cell::art_q(r(di, dj, dk), v(di + (AX == 0), dj + (AX == 1), dk + (AX == 2)) - v(di, dj, dk));
  }
};

// kf_src_a: stresses (visctensor ranges j,k in [1, s-2]) at every cell the
// active update reads (launched only with viscosity).  VM fixes the
// viscosity model at compile time (1 constant nu, 2 alpha isothermal, 3 alpha
// adiabatic; 0: from vp), so that no branch separates the stencil's loads.
template <int G, int DZ, int VM = 0>
__device__ __forceinline__ void kd_src_a(const dim3 bid, const dmesh& m, const real* rho, const real* energy, const real* vx, const real* vy, const real* vz, real* txx, real* tyy, real* tzz, real* txz, real* tyx, real* tzy, cell::ViscPar vp, int visc, real domega) {
  if constexpr (VM != 0) {
    visc = 1;
    vp.model = VM == 1 ? 1 : 2;
    vp.iso = VM != 3;
  }
  const int i = bid.x * blockDim.x + threadIdx.x;
  if (i >= m.nx) return;
  const int j = 1 + (int)bid.y;       // 1 .. sy-2
  const int k = DZ ? 1 + (int)bid.z : 0;
  const int l = i + j * m.pitch + k * m.stride;
  const auto VX = vx_stress<G>(vx, m, l, i, j, k, domega);
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
}
template <int G, int DZ, int VM>
__global__ void __launch_bounds__(WG) kf_src_a(dmesh m, const real* __restrict__ rho, const real* __restrict__ energy, const real* __restrict__ vx,
                                               const real* __restrict__ vy, const real* __restrict__ vz, real* __restrict__ txx, real* __restrict__ tyy,
                                               real* __restrict__ tzz, real* __restrict__ txz, real* __restrict__ tyx, real* __restrict__ tzy,
                                               cell::ViscPar vp, int visc, real domega) {
  kd_src_a<G, DZ, VM>(xcd_bid(), m, rho, energy, vx, vy, vz, txx, tyy, tzz, txz, tyx, tzy, vp, visc, domega);
}

// kf_src_b: the velocity update of SubStep1 + addviscosity + SubStep2_b on
// active cells, in the order the reference applies them.  MODE fixes the
// flags of a viscous gas with a potential at compile time (1 isothermal, 2
// adiabatic; 0: the flags as passed), so that no branch separates the
// stencil's loads.
template <int G, int DZ, int MODE = 0>
__device__ __forceinline__ void kd_src_b(const dim3 bid, const dmesh& m, real dt, const real* __restrict__ rho, const real* __restrict__ energy, const real* __restrict__ vx, const real* __restrict__ vy, const real* __restrict__ vz, const double* __restrict__ pot, const real* __restrict__ txx, const real* __restrict__ tyy, const real* __restrict__ tzz, const real* __restrict__ txz, const real* __restrict__ tyx, const real* __restrict__ tzy, real* __restrict__ vxt, real* __restrict__ vyt, real* __restrict__ vzt, real* __restrict__ e2, int visc, int potential, int dust, int iso, real gm1, double omegaframe, real domega) {
  if constexpr (MODE != 0) {
    visc = 1;
    potential = 1;
    dust = 0;
    iso = MODE == 1;
  }
  const int i = bid.x * blockDim.x + threadIdx.x;
  if (i >= m.nx) return;
  const int j = m.nghy + (int)bid.y;
  const int k = DZ ? m.nghz + (int)bid.z : 0;
  const int l = i + j * m.pitch + k * m.stride;
  const cell::GL R = cell::gl(rho, m, l, i);
  const PresLoader P{R, cell::gl(energy, m, l, i), gm1, iso};
  const cell::GLT<double> PT = cell::gld(pot, m, l, i);  // read only when potential != 0
  const auto VX = vxc<G>(vx, m, l, i, j, k, domega);
  const cell::GL VY = cell::gl(vy, m, l, i);
  const cell::GL VZ = cell::gl(DZ ? vz : vy, m, l, i);
  real ax = cell::substep1_x<G>(m, i, j, k, dt, R, P, PT, VX, dust, potential);
  real ay = cell::substep1_y<G, DZ>(m, i, j, k, dt, R, P, PT, VX, VY, VZ, dust, potential, omegaframe);
  real az = DZ ? cell::substep1_z<G>(m, i, j, k, dt, R, P, PT, VX, VZ, dust, potential, omegaframe) : real(0.0);
  if (visc) {
    const cell::GL TZZ = cell::gl(DZ ? tzz : txx, m, l, i), TXZ = cell::gl(DZ ? txz : txx, m, l, i),
                   TZY = cell::gl(DZ ? tzy : txx, m, l, i);
    cell::addviscosity<G, DZ>(m, i, j, k, dt, R, cell::gl(txx, m, l, i), cell::gl(tyy, m, l, i), TZZ, TXZ,
                              cell::gl(tyx, m, l, i), TZY, ax, ay, az, ax, ay, az);
  }
  real e = iso ? real(0.0) : energy[l];
  // SubStep2_a's q (src/substep2_a.c:103-134) on the fly
  const ArtQ<cell::GL, decltype(VX), 0> QX{R, VX};
  const ArtQ<cell::GL, cell::GL, 1> QY{R, VY};
  const ArtQ<cell::GL, cell::GL, 2> QZ{R, VZ};
  cell::substep2_b<G, DZ>(m, i, j, k, dt, R, VX, VY, VZ, QX, QY, QZ, !iso, ax, ay, az, e, ax, ay, az, e);
  vxt[l] = ax;
  vyt[l] = ay;
  if (DZ) vzt[l] = az;
  if (!iso) e2[l] = e;
}
template <int G, int DZ, int MODE>
__global__ void __launch_bounds__(WG, DZ ? 5 : 8) kf_src_b(dmesh m, real dt, const real* __restrict__ rho, const real* __restrict__ energy,
                                               const real* __restrict__ vx, const real* __restrict__ vy, const real* __restrict__ vz,
                                               const double* __restrict__ pot, const real* __restrict__ txx, const real* __restrict__ tyy,
                                               const real* __restrict__ tzz, const real* __restrict__ txz, const real* __restrict__ tyx,
                                               const real* __restrict__ tzy, real* __restrict__ vxt, real* __restrict__ vyt, real* __restrict__ vzt, real* __restrict__ e2,
                                               int visc, int potential, int dust, int iso, real gm1, double omegaframe,
                                               real domega) {
  kd_src_b<G, DZ, MODE>(xcd_bid(), m, dt, rho, energy, vx, vy, vz, pot, txx, tyy, tzz, txz, tyx, tzy, vxt, vyt, vzt, e2, visc, potential, dust, iso, gm1, omegaframe, domega);
}

// 2-D sources of a viscous gas with a potential in one pass, the viscous
// stress computed where addviscosity reads it instead of going through the
// stress arrays (kf_src_a + kf_src_b, without their 3 arrays written and read
// back).  Each lane marches a column through rows_per rows with the stresses
// of rows j-1, j, j+1 in registers (visctensor once per cell and row, plus two
// rows per segment); the x neighbours' stresses come from the next lanes by
// wave rotation.  A wave holds 64 stress columns, i0-1 .. i0+62, and updates
// the 62 cells between (waves overlap by two columns), so every neighbour a
// cell reads is in the wave.  MODE as kd_src_b (1 isothermal, 2 adiabatic),
// VM as kd_src_a (1 constant nu, 2 alpha isothermal, 3 alpha adiabatic).
struct StressAt {
  real c, xm, xp, ym, yp;  // at the cell, its x-1, x+1, y-1 and y+1 neighbours
  __device__ __forceinline__ real operator()(int di, int dj, int) const {
    return di < 0 ? xm : di > 0 ? xp : dj < 0 ? ym : dj > 0 ? yp : c;
  }
};
constexpr int kSrcCols = 62;  // cells a wave of kf_src_fused2 updates
template <int G, int MODE, int VM>
__global__ void __launch_bounds__(64) kf_src_fused2(dmesh m, real dt, const real* __restrict__ rho,
                                                    const real* __restrict__ energy, const real* __restrict__ vx,
                                                    const real* __restrict__ vy, const double* __restrict__ pot,
                                                    real* __restrict__ vxt, real* __restrict__ vyt,
                                                    real* __restrict__ e2, cell::ViscPar vp, real gm1,
                                                    double omegaframe, real domega, int rows_per, int jend) {
  constexpr int iso = MODE == 1;
  vp.model = VM == 1 ? 1 : 2;
  vp.iso = VM != 3;
  const int lane = (int)threadIdx.x;
  const dim3 b = xcd_bid();
  const int is = kSrcCols * (int)b.x - 1 + lane;  // this lane's stress column
  const int i = is < 0 ? is + m.nx : (is >= m.nx ? is - m.nx : is);
  const bool upd = lane >= 1 && lane <= kSrcCols && is < m.nx;
  const int r0 = m.nghy + (int)b.y * rows_per;
  const int r1 = min(r0 + rows_per, jend);
  // visctensor (src/visctensor_cyl.c:140-186, src/visctensor_sph.c:126-206)
  // at (i, row): the same loaders as kd_src_a
  auto stress = [&](int row) -> cell::Taus {
    const int l = i + row * m.pitch;
    const auto VX = vx_stress<G>(vx, m, l, i, row, 0, domega);
    const cell::GL VY = cell::gl(vy, m, l, i);
    return cell::visctensor<G, 0>(m, i, row, 0, cell::gl(rho, m, l, i), cell::gl(energy, m, l, i), VX, VY, VY, vp);
  };
  cell::Taus Tm = stress(r0 - 1), T0 = stress(r0);
  for (int j = r0; j < r1; j++) {
    const cell::Taus Tp = stress(j + 1);
    // the x neighbours' stresses (every lane takes part in the rotations)
    const real txx_xm = from_left(T0.xx), tyx_xp = from_right(T0.yx);
    if (upd) {
      // kd_src_b at (i, j)
      const int l = i + j * m.pitch;
      const cell::GL R = cell::gl(rho, m, l, i);
      const PresLoader P{R, cell::gl(energy, m, l, i), gm1, iso};
      const cell::GLT<double> PT = cell::gld(pot, m, l, i);
      const auto VX = vxc<G>(vx, m, l, i, j, 0, domega);
      const cell::GL VY = cell::gl(vy, m, l, i);
      real ax = cell::substep1_x<G>(m, i, j, 0, dt, R, P, PT, VX, 0, 1);
      real ay = cell::substep1_y<G, 0>(m, i, j, 0, dt, R, P, PT, VX, VY, VY, 0, 1, omegaframe);
      real az = real(0.0);
      const StressAt TXX{T0.xx, txx_xm, real(0.0), Tm.xx, real(0.0)};
      const StressAt TYY{T0.yy, real(0.0), real(0.0), Tm.yy, real(0.0)};
      const StressAt TYX{T0.yx, real(0.0), tyx_xp, real(0.0), Tp.yx};
      cell::addviscosity<G, 0>(m, i, j, 0, dt, R, TXX, TYY, TXX, TXX, TYX, TXX, ax, ay, az, ax, ay, az);
      real e = iso ? real(0.0) : energy[l];
      const ArtQ<cell::GL, decltype(VX), 0> QX{R, VX};
      const ArtQ<cell::GL, cell::GL, 1> QY{R, VY};
      cell::substep2_b<G, 0>(m, i, j, 0, dt, R, VX, VY, VY, QX, QY, QY, !iso, ax, ay, az, e, ax, ay, az, e);
      vxt[l] = ax;
      vyt[l] = ay;
      if (!iso) e2[l] = e;
    }
    Tm = T0;
    T0 = Tp;
  }
}

// kf_src_b of every fluid in one launch (isothermal runs): blockIdx.z =
// fluid * (planes) + plane; each fluid has its own viscosity and dust flags
struct SrcTab {
  const real* rho[8];
  const real* cs[8];
  const real* vx[8];
  const real* vy[8];
  const real* vz[8];
  real* vxt[8];
  real* vyt[8];
  real* vzt[8];
  int visc[8];
  int dust[8];
};
template <int G, int DZ>
__global__ void __launch_bounds__(WG, DZ ? 5 : 8) kf_src_b_fluids(dmesh m, real dt, SrcTab T, int nzb, const double* __restrict__ pot,
                                                      const real* __restrict__ txx, const real* __restrict__ tyy,
                                                      const real* __restrict__ tzz, const real* __restrict__ txz,
                                                      const real* __restrict__ tyx, const real* __restrict__ tzy,
                                                      int potential, real gm1, double omegaframe, real domega) {
  const dim3 b0 = xcd_bid();
  const int fl = (int)b0.z / nzb;
  const dim3 b(b0.x, b0.y, b0.z % nzb);
  kd_src_b<G, DZ>(b, m, dt, T.rho[fl], T.cs[fl], T.vx[fl], T.vy[fl], T.vz[fl], pot, txx, tyy, tzz, txz, tyx, tzy,
                  T.vxt[fl], T.vyt[fl], T.vzt[fl], nullptr, T.visc[fl], potential, T.dust[fl], 1, gm1, omegaframe,
                  domega);
}

// SubStep3 (src/substep3.c:98-110) on active cells: compression work with the
// divergence of V_temp; reads the energy after SubStep2_b.
template <int G, int DZ>
__global__ void __launch_bounds__(WG) kf_substep3(dmesh m, real dt, const real* __restrict__ e2, const real* __restrict__ vxt,
                                                  const real* __restrict__ vyt, const real* __restrict__ vzt, real* __restrict__ energy,
                                                  real gamma) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m.nx) return;
  const int j = m.nghy + (int)blockIdx.y;
  const int k = DZ ? m.nghz + (int)blockIdx.z : 0;
  const int l = i + j * m.pitch + k * m.stride;
  energy[l] = cell::substep3<DZ>(m, i, j, k, dt, e2[l], cell::gl(vxt, m, l, i), cell::gl(vyt, m, l, i),
                                 cell::gl(DZ ? vzt : vyt, m, l, i), gamma);
}

// --------------------------------------------------------------- transport
// One directional van Leer sweep of one column (the Y or Z transport of
// src/transport.c:81-127): NQ conserved quantities plus the density
// (index NQ), rows r in [r0, r1).  Per row the quantity q = Q/rho
// (src/divrho.c:56), the gradient at the lower face (dqm of
// src/vanleer_y_a.c:63, equal to dqp of the row below), the slope and the
// upwind face value (src/vanleer_y_b.c:60-65) are each computed once and
// kept in a window of registers.
//   col.rho(r)      density at row r (not modified by the sweep)
//   col.Q(c, r)     conserved quantity c at row r
//   col.zs(r)       zone size of row r in the sweep direction
//   col.v(r)        velocity at the lower face of row r
//   col.vol(r)      the divisor of row r's InvVol (rdiv(), shared by its quantities)
//   col.fmesh(r)    the mesh factors of row r's flux difference (cell::FluxMesh)
//   col.flux(F, a, b, V) conservative increment of a row from face fluxes a, b
//   col.out(c, r, value) store of the updated quantity (c == NQ: density)
// In the precision form the first Col::NXQ quantities (the azimuthal momenta)
// are deviations Q' = Q - rho l0(r) from the profile's angular momentum
// (col.base(r) = l0 at row r, FP64); their q' = Q'/rho is carried, and
//   - a gradient takes the profile's jump across the face, db(r) =
//     l0(r) - l0(r-1), so the limiter sees the full gradient;
//   - a face value is held relative to l0 of the row below the face (the
//     upper row's q' enters with + db(r));
//   - the update of row r uses its lower face relative to l0(r) (- db(r)):
//     Q'(r) + dt/V (F_rho(r) (q*(r) - l0(r)) - F_rho(r+1) (q*(r+1) - l0(r))),
//     which is Q(r) + dt/V (F_rho q* - F_rho q*) - l0(r) (rho_new - rho):
// the reference's update of the full momentum, rearranged so that every
// FP32 operand is of the order of the deviation and of l0's jump across one
// row, not of l0.  In the FP64 build NXQ is 0 and the arithmetic is the
// reference's.
// x + d for a deviation quantity; x itself (no addition) otherwise
__device__ __forceinline__ real dsh(real x, real d, bool dev) { return dev ? x + d : x; }
// Holds a value in its registers at this point of the program (an empty asm
// statement that reads it), so that the load producing it stays above.
__device__ __forceinline__ void pin(float x) { asm volatile("" ::"v"(x)); }
__device__ __forceinline__ void pin(double x) { asm volatile("" ::"v"(x)); }

template <int NQ, class Col>
__device__ __forceinline__ void column_sweep(Col& col, int r0, int r1, real dt) {
  constexpr int N = NQ + 1;
  constexpr int NXQ = cell::kFix64 ? Col::NXQ : 0;
  auto db = [&](int r) -> real {
    if constexpr (NXQ > 0) return (real)(col.base(r) - col.base(r - 1));
    return real(0.0);
  };
  real A1[N], A2[N], G2[N], S1[N], F0[N], F1[N];
  // db of the rows r, r+1, r+2 of the loop below (deviation quantities only)
  real D0 = real(0.0), D1 = real(0.0), D2 = real(0.0);
  {
    real a[N][5];
#pragma unroll
    for (int t = 0; t < 5; t++) {
      const int r = r0 - 2 + t;
      const real rr = col.rho(r);
      const Recip R = recip(rr);
#pragma unroll
      for (int c = 0; c < NQ; c++) a[c][t] = rdiv(col.Q(c, r), R);
      a[NQ][t] = rr;
    }
    real dm = real(0.0);
    if constexpr (NXQ > 0) {
      dm = db(r0 - 1);
      D0 = db(r0);
      D1 = db(r0 + 1);
      D2 = db(r0 + 2);
    }
    const Recip Z0 = recip(col.zs(r0 - 1)), Z1 = recip(col.zs(r0)), Z2 = recip(col.zs(r0 + 1)),
                Z3 = recip(col.zs(r0 + 2));
#pragma unroll
    for (int c = 0; c < N; c++) {
      const bool dev = c < NXQ;
      const real g0 = rdiv(dsh(a[c][1] - a[c][0], dm, dev), Z0);
      const real g1 = rdiv(dsh(a[c][2] - a[c][1], D0, dev), Z1);
      const real g2 = rdiv(dsh(a[c][3] - a[c][2], D1, dev), Z2);
      const real g3 = rdiv(dsh(a[c][4] - a[c][3], D2, dev), Z3);
      const real sm = vanleer_slope(g0, g1);
      const real s0 = vanleer_slope(g1, g2);
      const real s1 = vanleer_slope(g2, g3);
      F0[c] = vanleer_face(col.v(r0), dt, a[c][1], sm, col.zs(r0 - 1), dsh(a[c][2], D0, dev), s0,
                           col.zs(r0));
      F1[c] = vanleer_face(col.v(r0 + 1), dt, a[c][2], s0, col.zs(r0), dsh(a[c][3], D1, dev), s1,
                           col.zs(r0 + 1));
      A1[c] = a[c][3];
      A2[c] = a[c][4];
      G2[c] = g3;
      S1[c] = s1;
    }
  }
  for (int r = r0; r < r1; r++) {
    // Every load of the iteration before its first store: the quantities of
    // row r (its update) and of row r+3 (entering the window).  The vector
    // memory counter retires loads and stores in issue order, so a load
    // issued after a store would wait for that store as well.
    // On the last row no row enters, and row r is loaded in its place (no
    // branch around the loads; the values are not used).
    const bool more = r + 1 < r1;
    const int rn = r + 3, rl = more ? rn : r;
    real Qr[NQ], Qn[NQ];
#pragma unroll
    for (int c = 0; c < NQ; c++) Qr[c] = col.Q(c, r);
    const real rhor = col.rho(r);
    const real rr = col.rho(rl);
#pragma unroll
    for (int c = 0; c < NQ; c++) Qn[c] = col.Q(c, rl);
    const real zsn = col.zs(rl);
    const real v0 = col.v(r), v1 = col.v(r + 1), v2 = col.v(r + 2);
    const real zs1 = col.zs(r + 1), zs2 = col.zs(r + 2);
    const Recip V = col.vol(r);
    const cell::FluxMesh FM = col.fmesh(r);
    // the entering row's values are used only after the stores; held here, so
    // that their loads are not moved below them
#pragma unroll
    for (int c = 0; c < NQ; c++) pin(Qn[c]);
    pin(rr);
    pin(zsn);
    pin(v2);
    // UpdateY/Z (src/update_y.c:68-70), then the density last (src/update_density_y.c:66-68)
#pragma unroll
    for (int c = 0; c < NQ; c++) {
      const real f0 = (c < NXQ) ? F0[c] - D0 : F0[c];
      col.out(c, r, Qr[c] + col.flux(FM, v0 * f0 * F0[NQ], v1 * F1[c] * F1[NQ], V));
    }
    col.out(NQ, r, rhor + col.flux(FM, v0 * F0[NQ], v1 * F1[NQ], V));
    if (!more) break;
    // advance the window by one row: row r+3 enters
    const Recip R = recip(rr);
    const Recip Zn = recip(zsn);
    real D3 = real(0.0);
    if constexpr (NXQ > 0) D3 = db(rn);
#pragma unroll
    for (int c = 0; c < N; c++) {
      const bool dev = c < NXQ;
      const real an = (c < NQ) ? rdiv(Qn[c], R) : rr;
      const real g3 = rdiv(dsh(an - A2[c], D3, dev), Zn);
      const real s2 = vanleer_slope(G2[c], g3);
      const real f2 = vanleer_face(v2, dt, A1[c], S1[c], zs1, dsh(A2[c], D2, dev), s2, zs2);
      F0[c] = F1[c];
      F1[c] = f2;
      A1[c] = A2[c];
      A2[c] = an;
      G2[c] = g3;
      S1[c] = s2;
    }
    D0 = D1;
    D1 = D2;
    D2 = D3;
  }
}

// A column carrying part of the quantities: COUNT of them from FIRST on,
// plus the density, which every part needs for q = Q/rho and the mass flux;
// only the part with WRITE_RHO stores the density.  Splitting a column's
// quantities over two threads doubles the threads of a sweep (more waves to
// hide the FP64 latency) for one extra density sweep per column.
template <class Col, int NQ, int FIRST, int COUNT, bool WRITE_RHO>
struct ColPart {
  // the deviation quantities of the column among those carried here
  static constexpr int NXQ = FIRST < Col::NXQ ? (Col::NXQ - FIRST < COUNT ? Col::NXQ - FIRST : COUNT) : 0;
  static_assert(FIRST == 0 || FIRST >= Col::NXQ, "the deviation quantities lead each part");
  const Col& c;
  __device__ __forceinline__ double base(int r) const { return c.base(r); }
  __device__ __forceinline__ real rho(int r) const { return c.rho(r); }
  __device__ __forceinline__ real Q(int q, int r) const { return c.Q(FIRST + q, r); }
  __device__ __forceinline__ real zs(int r) const { return c.zs(r); }
  __device__ __forceinline__ real v(int r) const { return c.v(r); }
  __device__ __forceinline__ Recip vol(int r) const { return c.vol(r); }
  __device__ __forceinline__ cell::FluxMesh fmesh(int r) const { return c.fmesh(r); }
  __device__ __forceinline__ real flux(const cell::FluxMesh& F, real a, real b, const Recip& V) const {
    return c.flux(F, a, b, V);
  }
  __device__ __forceinline__ void out(int q, int r, real val) const {
    if (q < COUNT) c.out(FIRST + q, r, val);
    else if (WRITE_RHO) c.out(NQ, r, val);
  }
};

// the sweep of part grp of a column: 0 every quantity, 1 the first half and
// the density, 2 the second half
template <int NQ, class Col>
__device__ __forceinline__ void column_sweep_part(const Col& col, int grp, int r0, int r1, real dt) {
  constexpr int H = NQ / 2;
  if (grp == 1) {
    ColPart<Col, NQ, 0, H, true> p{col};
    column_sweep<H>(p, r0, r1, dt);
  } else {
    ColPart<Col, NQ, H, NQ - H, false> p{col};
    column_sweep<NQ - H>(p, r0, r1, dt);
  }
}

// Y column of the 2-D step: momenta computed from rho and V_temp on the fly
// (src/momenta_x.c:76-79, src/momenta_y.c:62-63); quantity 4 is the energy
// (adiabatic).  out_[c] receives quantity c, out_[NQ] the density.
template <int G>
struct ColY2 {
  static constexpr int NXQ = 2;  // Mpx, Mmx: deviations in the precision form
  const dmesh* m;
  const real *rho_in, *vxt, *vyt, *e_in;
  real* out_[6];
  int i, ixp;
  real of, dt;
  __device__ __forceinline__ double base(int r) const { return cell::lbase<G>(*m, r, 0, (double)of); }
  __device__ __forceinline__ int L(int r) const { return i + r * m->pitch; }
  __device__ __forceinline__ real rho(int r) const { return rho_in[L(r)]; }
  __device__ __forceinline__ real Q(int c, int r) const {
    const int l = L(r);
    const real rr = rho_in[l];
    if (c < 2) {
      real mm, mp;
      cell::momenta_x<G>(*m, i, r, 0, rr, vxt[l], vxt[ixp + r * m->pitch], of, mm, mp);
      return c == 0 ? mp : mm;
    }
    if (c == 2) return vyt[l + m->pitch] * rr;
    if (c == 3) return vyt[l] * rr;
    return e_in[l];
  }
  __device__ __forceinline__ real zs(int r) const { return zone_size_y(*m, r); }
  __device__ __forceinline__ real v(int r) const { return vyt[L(r)]; }
  __device__ __forceinline__ Recip vol(int r) const { return invvol_recip(*m, i, 0); }
  __device__ __forceinline__ cell::FluxMesh fmesh(int r) const { return cell::flux_y_mesh(*m, i, r, 0); }
  __device__ __forceinline__ real flux(const cell::FluxMesh& F, real a, real b, const Recip& V) const {
    return cell::flux_yz(F, dt, a, b, V);
  }
  __device__ __forceinline__ void out(int c, int r, real val) const { out_[c][L(r)] = val; }
};

// Z column of the 3-D step: momenta from rho and V_temp on the fly
// (src/momenta_x.c:80-84, src/momenta_y.c:62-63, src/momenta_z.c:66-71);
// quantity 6 is the energy (adiabatic).
template <int G>
struct ColZ {
  static constexpr int NXQ = 2;  // Mpx, Mmx: deviations in the precision form
  const dmesh* m;
  const real *rho_in, *vxt, *vyt, *vzt, *e_in;
  real* out_[8];
  int i, j, ixp;
  real of, dt;
  __device__ __forceinline__ double base(int r) const { return cell::lbase<G>(*m, j, r, (double)of); }
  __device__ __forceinline__ int L(int r) const { return i + j * m->pitch + r * m->stride; }
  __device__ __forceinline__ real rho(int r) const { return rho_in[L(r)]; }
  __device__ __forceinline__ real Q(int c, int r) const {
    const int l = L(r);
    const real rr = rho_in[l];
    switch (c) {
      case 0:
      case 1: {
        real mm, mp;
        cell::momenta_x<G>(*m, i, j, r, rr, vxt[l], vxt[ixp + j * m->pitch + r * m->stride], of, mm, mp);
        return c == 0 ? mp : mm;
      }
      case 2:
        return vyt[l + m->pitch] * rr;
      case 3:
        return vyt[l] * rr;
      case 4: {
        real p = vzt[l + m->stride] * rr;
        if (G == SPH) p *= m->ymed[j];
        return p;
      }
      case 5: {
        real q = vzt[l] * rr;
        if (G == SPH) q *= m->ymed[j];
        return q;
      }
      default:
        return e_in[l];
    }
  }
  __device__ __forceinline__ real zs(int r) const { return zone_size_z<G>(*m, j, r); }
  __device__ __forceinline__ real v(int r) const { return vzt[L(r)]; }
  __device__ __forceinline__ Recip vol(int r) const { return invvol_recip(*m, i, r); }
  __device__ __forceinline__ cell::FluxMesh fmesh(int r) const { return cell::flux_z_mesh(*m, i, j, r); }
  __device__ __forceinline__ real flux(const cell::FluxMesh& F, real a, real b, const Recip& V) const {
    return cell::flux_yz(F, dt, a, b, V);
  }
  __device__ __forceinline__ void out(int c, int r, real val) const { out_[c][L(r)] = val; }
};

// Y column of the 3-D step: reads the Z-swept density, momenta (and energy).
template <int G>
struct ColY3 {
  static constexpr int NXQ = 2;  // Mpx, Mmx: deviations in the precision form
  const dmesh* m;
  const real* rho_in;
  const real* q_in[7];
  const real* vyt;
  real* out_[8];
  int i, k;
  real dt, of;
  __device__ __forceinline__ double base(int r) const { return cell::lbase<G>(*m, r, k, (double)of); }
  __device__ __forceinline__ int L(int r) const { return i + r * m->pitch + k * m->stride; }
  __device__ __forceinline__ real rho(int r) const { return rho_in[L(r)]; }
  __device__ __forceinline__ real Q(int c, int r) const { return q_in[c][L(r)]; }
  __device__ __forceinline__ real zs(int r) const { return zone_size_y(*m, r); }
  __device__ __forceinline__ real v(int r) const { return vyt[L(r)]; }
  __device__ __forceinline__ Recip vol(int r) const { return invvol_recip(*m, i, k); }
  __device__ __forceinline__ cell::FluxMesh fmesh(int r) const { return cell::flux_y_mesh(*m, i, r, k); }
  __device__ __forceinline__ real flux(const cell::FluxMesh& F, real a, real b, const Recip& V) const {
    return cell::flux_yz(F, dt, a, b, V);
  }
  __device__ __forceinline__ void out(int c, int r, real val) const { out_[c][L(r)] = val; }
};

// 2-D: momenta + Y sweep; rows [nghy + s*rows_per, ...) of the active range.
// AD = 1 also transports the energy (e_in -> e_out).
template <int G, int DZ, int AD, int SPLIT = 0>
__device__ __forceinline__ void kd_sweep_y2(const dim3 bid, const dmesh& m, real dt, const real* __restrict__ rho_in, const real* __restrict__ vxt, const real* __restrict__ vyt, const real* __restrict__ e_in, real* __restrict__ rho_out, real* __restrict__ q0, real* __restrict__ q1, real* __restrict__ q2, real* __restrict__ q3, real* __restrict__ e_out, int rows_per, real of) {
  const int i = bid.x * blockDim.x + threadIdx.x;
  if (i >= m.nx) return;
  const int seg = SPLIT ? (int)bid.y >> 1 : (int)bid.y;
  const int r0 = m.nghy + seg * rows_per;
  const int r1 = min(r0 + rows_per, m.ny + m.nghy);
  ColY2<G> col{&m, rho_in, vxt, vyt, e_in, {q0, q1, q2, q3, AD ? e_out : rho_out, rho_out}, i, ixp_of(i, m.nx), of,
               dt};
  if (SPLIT) column_sweep_part<4 + AD>(col, 1 + ((int)bid.y & 1), r0, r1, dt);
  else column_sweep<4 + AD>(col, r0, r1, dt);
}
// 2-D Y sweep of every fluid in one launch: blockIdx.z is the fluid
struct Y2Tab {
  const real* rho_in[8];
  const real* vxt[8];
  const real* vyt[8];
  const real* e_in[8];
  real* rho_out[8];
  real* q[8][4];
  real* e_out[8];
};
template <int G, int DZ, int AD, int SPLIT>
__global__ void __launch_bounds__(BXF) kf_sweep_y2(dmesh m, real dt, Y2Tab T, int rows_per, real of) {
  const dim3 b = xcd_bid();
  const int fl = (int)b.z;
  kd_sweep_y2<G, DZ, AD, SPLIT>(b, m, dt, T.rho_in[fl], T.vxt[fl], T.vyt[fl], T.e_in[fl], T.rho_out[fl], T.q[fl][0],
                                T.q[fl][1], T.q[fl][2], T.q[fl][3], T.e_out[fl], rows_per, of);
}

// 3-D: Y sweep of the Z-swept fields
template <int G, int DZ, int AD, int SPLIT>
__global__ void __launch_bounds__(BXF) kf_sweep_y3(dmesh m, real dt, const real* __restrict__ rho_in, const real* __restrict__ z0,
                                                   const real* __restrict__ z1, const real* __restrict__ z2, const real* __restrict__ z3,
                                                   const real* __restrict__ z4, const real* __restrict__ z5, const real* __restrict__ ze,
                                                   const real* __restrict__ vyt, real* __restrict__ rho_out, real* __restrict__ q0, real* __restrict__ q1,
                                                   real* __restrict__ q2, real* __restrict__ q3, real* __restrict__ q4, real* __restrict__ q5, real* __restrict__ e_out,
                                                   int rows_per, real of) {
  const dim3 b = xcd_bid();
  const int i = b.x * blockDim.x + threadIdx.x;
  if (i >= m.nx) return;
  const int k = m.nghz + (int)b.z;
  const int seg = SPLIT ? (int)b.y >> 1 : (int)b.y;
  const int r0 = m.nghy + seg * rows_per;
  const int r1 = min(r0 + rows_per, m.ny + m.nghy);
  ColY3<G> col{&m, rho_in, {z0, z1, z2, z3, z4, z5, ze}, vyt,
               {q0, q1, q2, q3, q4, q5, AD ? e_out : rho_out, rho_out}, i, k, dt, of};
  if (SPLIT) column_sweep_part<6 + AD>(col, 1 + ((int)b.y & 1), r0, r1, dt);
  else column_sweep<6 + AD>(col, r0, r1, dt);
}

// 3-D: momenta + Z sweep for rows j in [1, sy-2] (the Y sweep's stencil)
template <int G, int DZ, int AD, int SPLIT>
__global__ void __launch_bounds__(BXF) kf_sweep_z(dmesh m, real dt, const real* __restrict__ rho_in, const real* __restrict__ vxt,
                                                  const real* __restrict__ vyt, const real* __restrict__ vzt, const real* __restrict__ e_in,
                                                  real* __restrict__ rho_out, real* __restrict__ q0, real* __restrict__ q1, real* __restrict__ q2, real* __restrict__ q3,
                                                  real* __restrict__ q4, real* __restrict__ q5, real* __restrict__ e_out, int planes_per, real of) {
  const dim3 b = blockIdx;
  const int i = b.x * blockDim.x + threadIdx.x;
  if (i >= m.nx) return;
  const int j = 1 + (SPLIT ? (int)b.y >> 1 : (int)b.y);
  const int r0 = m.nghz + (int)b.z * planes_per;
  const int r1 = min(r0 + planes_per, m.nz + m.nghz);
  ColZ<G> col{&m, rho_in, vxt, vyt, vzt, e_in, {q0, q1, q2, q3, q4, q5, AD ? e_out : rho_out, rho_out}, i, j,
              ixp_of(i, m.nx), of, dt};
  if (SPLIT) column_sweep_part<6 + AD>(col, 1 + ((int)b.y & 1), r0, r1, dt);
  else column_sweep<6 + AD>(col, r0, r1, dt);
}

// The ring X sweep with the cell state in registers.  Thread t owns cells
// i = t + s*blockDim.x (s < CPT) of the ring; LDS holds q = Q/rho, the slopes
// (then the limited qL) and the faces (then qR, then the face fluxes) of the
// quantities this workgroup carries, plus the advection velocity.  GRP 0: one
// workgroup per ring carries every quantity; GRP 1/2: two workgroups per ring
// (blockIdx.x = 2*ring + grp - 1), each with the density and half of the
// quantities, so that each needs about half the LDS and more workgroups fit
// on a CU.  Group 1 carries Mpx, Mmx (and Mpy in 3-D) and writes the density,
// Vx, VxMed and Nshift; group 2 carries the other momenta and the energy.
// Every quantity goes through the same arithmetic as in the reference's
// routine sequence, so any split gives the same bits.
template <int DZ, int AD, int GRP>
struct XQuant {
  static constexpr int NM = DZ ? 6 : 4;
  static constexpr int NQ = NM + AD;  // quantities of the whole ring
  static constexpr int NA = DZ ? 3 : 2;  // group 1: Mpx, Mmx (, Mpy)
  static constexpr int NL = GRP == 0 ? NQ : GRP == 1 ? NA : NQ - NA;  // carried here
  static constexpr int first = GRP == 2 ? NA : 0;  // global index of local quantity 0
};

// An empty asm that "changes" the LDS row stride after each barrier: the
// compiler then computes each phase's LDS addresses in that phase instead of
// holding every phase's addresses in registers from the start (measured:
// 118 -> 82 registers with the VPRE form, 4 -> 5 waves per SIMD).
#define XR_REFRESH() asm volatile("" : "+s"(nxv))
// VPRE: VxMed of the ring from kf_vmed (the kernel then carries no sum).
template <int G, int DZ, int AD, int CPT, int GRP, int VPRE>
__device__ __forceinline__ void xsweep_ring(const dmesh& m, int ring, double dt64, const real* __restrict__ rho_in,
                                            const real* __restrict__ vxt, const real* const* qin,
                                            const real* __restrict__ e_in, real* __restrict__ e_out,
                                            real* __restrict__ rho_out, real* __restrict__ vx_out,
                                            real* const* qout, double* __restrict__ vxmed_out,
                                            int* __restrict__ nshift_out, real omegaframe, int fargo) {
  using XQ = XQuant<DZ, AD, GRP>;
  constexpr int NL = XQ::NL;
  constexpr int N = NL + 1;  // + the density (local index NL)
  constexpr int NM = XQ::NM;
  extern __shared__ double lds_d[];
  const int nx = m.nx;
  const real dt = (real)dt64;
  const int P = (N * nx + 1) & ~1;  // N * nx rounded up to 8-byte multiples
  real* AQ = (real*)lds_d;   // q = Q/rho (or rho): N * nx
  real* SL = AQ + P;         // slopes, then limited qL: N * nx
  real* FC = SL + P;         // faces, then qR, then face fluxes: N * nx
  real* V = FC + P;          // advection velocity: nx
  double* sh = (double*)(((size_t)(V + nx) + 15) & ~(size_t)15);  // 4
  const int j = m.nghy + ring % m.ny;
  const int k = DZ ? m.nghz + ring / m.ny : 0;
  const int base = j * m.pitch + k * m.stride;
  const int t = threadIdx.x, nt = blockDim.x;
  real st[CPT][N];
  double vt[CPT];
  real vu[CPT];
#pragma unroll
  for (int s = 0; s < CPT; s++) {
    const int i = t + s * nt;
    if (i < nx) {
      st[s][NL] = rho_in[base + i];
#pragma unroll
      for (int c = 0; c < NL; c++) {
        const int q = XQ::first + c;
        st[s][c] = (q < NM) ? qin[q][base + i] : e_in[base + i];
      }
      vt[s] = vxt[base + i];
    }
  }
  double vmed;
  if constexpr (VPRE) {
    // VxMed of this ring from kf_vmed, run before this launch
    vmed = vxmed_out[j + k * m.sy];
  } else {
    // ComputeVmed(Vx_temp) (src/compute_vmed.c:14-30), FP64, by the first
    // wave with the terms in registers (wave_ring_sum)
    if (t < 64) {
      const double sum = wave_ring_sum<4>(vxt + base, m.xmin64, nx);
      if (t == 0) sh[0] = vmed_full(sum / (m.XMAX - m.XMIN), m, j, k);
    }
    __syncthreads();
    vmed = sh[0];
  }
  // ComputeResidual (src/compute_residual.c:68-73), FP64: cell residual into V
  // (the stored velocity is the deviation from V0 in the precision form)
  int ns = 0;
  const double vb = cell::vbase(m, j, k);
#pragma unroll
  for (int s = 0; s < CPT; s++) {
    const int i = t + s * nt;
    if (i < nx) {
      if (fargo) {
        const double zs = zone_size_x64<G>(m, i, j, k);
        const double ntilde = vmed * dt64 / zs;
        const double nround = floor(ntilde + 0.5);
        if (i == 0) sh[1] = nround;
        if constexpr (cell::kFix64) V[i] = (real)(vt[s] - (vmed - vb));
        else V[i] = (real)(vt[s] - vmed);
        vu[s] = (real)((ntilde - nround) * zs / dt64);
      } else {
        if constexpr (cell::kFix64) V[i] = (real)(vt[s] + vb);
        else V[i] = (real)vt[s];
      }
    }
  }
  __syncthreads();  // V complete
  if (fargo) {
    ns = (int)sh[1];
    if (GRP != 2 && t == 0) {
      nshift_out[j + k * m.sy] = ns;
      if (!VPRE) vxmed_out[j + k * m.sy] = vmed;
    }
  }
  int nxv = nx;  // LDS row stride, re-read after each barrier (keeps the
  // addresses of each phase out of the registers of the others)
  // X_advection (src/transport.c:208-226), twice with FARGO
  for (int pass = 0; pass < (fargo ? 2 : 1); pass++) {
    const bool ppa = (pass == 1);
    if (ppa) {
      __syncthreads();
      XR_REFRESH();  // the last update read V of the neighbours
#pragma unroll
      for (int s = 0; s < CPT; s++) {
        const int i = t + s * nt;
        if (i < nx) V[i] = vu[s];
      }
    }
    // q = Q/rho (DivideByRho) and, for the density, q = rho
#pragma unroll
    for (int s = 0; s < CPT; s++) {
      const int i = t + s * nt;
      if (i < nx) {
        const real r = st[s][NL];
        const Recip R = recip(r);
#pragma unroll
        for (int c = 0; c < NL; c++) AQ[c * nxv + i] = rdiv(st[s][c], R);
        AQ[NL * nxv + i] = r;
      }
    }
    __syncthreads();
      XR_REFRESH();
    if (!ppa) {
      // VanLeerX_a (src/vanleer_x_a.c:60-68).  The gradient at face i,
      // (q_i - q_{i-1}) / zone_i, is cell i's left and cell i-1's right
      // gradient: computed once per face into FC (free until VanLeerX_b).
#pragma unroll
      for (int s = 0; s < CPT; s++) {
        const int i = t + s * nt;
        if (i < nx) {
          const int im = ixm_of(i, nx);
          const Recip Zi = recip(zone_size_x<G>(m, i, j, k));
#pragma unroll
          for (int c = 0; c < N; c++) {
            const real* a = AQ + c * nxv;
            FC[c * nxv + i] = rdiv(a[i] - a[im], Zi);
          }
        }
      }
      __syncthreads();
      XR_REFRESH();
#pragma unroll
      for (int s = 0; s < CPT; s++) {
        const int i = t + s * nt;
        if (i < nx) {
          const int ip = ixp_of(i, nx);
#pragma unroll
          for (int c = 0; c < N; c++) SL[c * nxv + i] = vanleer_slope(FC[c * nxv + i], FC[c * nxv + ip]);
        }
      }
      __syncthreads();
      XR_REFRESH();
      // VanLeerX_b (src/vanleer_x_b.c:60-65)
#pragma unroll
      for (int s = 0; s < CPT; s++) {
        const int i = t + s * nt;
        if (i < nx) {
          const int im = ixm_of(i, nx);
          const real zi = zone_size_x<G>(m, i, j, k), zm = zone_size_x<G>(m, im, j, k);
          const real v = V[i];
#pragma unroll
          for (int c = 0; c < N; c++) {
            const real* a = AQ + c * nxv;
            const real* sl = SL + c * nxv;
            FC[c * nxv + i] = vanleer_face(v, dt, a[im], sl[im], zm, a[i], sl[i], zi);
          }
        }
      }
      __syncthreads();
      XR_REFRESH();
    } else {
      // VanLeerX_PPA_a: monotonised-centred slope (src/fargo_ppa_a.c:58-69)
#pragma unroll
      for (int s = 0; s < CPT; s++) {
        const int i = t + s * nt;
        if (i < nx) {
          const int im = ixm_of(i, nx), ip = ixp_of(i, nx);
#pragma unroll
          for (int c = 0; c < N; c++) {
            const real* a = AQ + c * nxv;
            SL[c * nxv + i] = cell::ppa_slope(a[im], a[i], a[ip]);
          }
        }
      }
      __syncthreads();
      XR_REFRESH();
      // VanLeerX_PPA_b: face value q_{i+1/2} = qR of cell i = qL of cell i+1
#pragma unroll
      for (int s = 0; s < CPT; s++) {
        const int i = t + s * nt;
        if (i < nx) {
          const int ip = ixp_of(i, nx);
#pragma unroll
          for (int c = 0; c < N; c++) {
            const real* a = AQ + c * nxv;
            const real* sl = SL + c * nxv;
            FC[c * nxv + i] = cell::ppa_face(a[i], a[ip], sl[i], sl[ip]);
          }
        }
      }
      __syncthreads();
      XR_REFRESH();
      // VanLeerX_PPA_c (src/fargo_ppa_c.c): the limited (qL, qR) of each cell;
      // qL goes to SL (slopes are no longer read), qR to FC once every
      // neighbour has read its face
      real qr[CPT][N];
#pragma unroll
      for (int s = 0; s < CPT; s++) {
        const int i = t + s * nt;
        if (i < nx) {
          const int im = ixm_of(i, nx);
#pragma unroll
          for (int c = 0; c < N; c++) {
            real qL = FC[c * nxv + im], qR = FC[c * nxv + i];
            cell::ppa_limit(AQ[c * nxv + i], qL, qR);
            SL[c * nxv + i] = qL;
            qr[s][c] = qR;
          }
        }
      }
      __syncthreads();
      XR_REFRESH();
#pragma unroll
      for (int s = 0; s < CPT; s++) {
        const int i = t + s * nt;
        if (i < nx) {
#pragma unroll
          for (int c = 0; c < N; c++) FC[c * nxv + i] = qr[s][c];
        }
      }
      __syncthreads();
      XR_REFRESH();
      // VanLeerX_PPA_d: parabolic average over the swept zone (src/fargo_ppa_d.c:62-70);
      // the fluxes replace qR in FC once every neighbour has read it
#pragma unroll
      for (int s = 0; s < CPT; s++) {
        const int i = t + s * nt;
        if (i < nx) {
          const int im = ixm_of(i, nx);
          const real v = V[i];
          const real ksi = cell::ppa_ksi(v, dt, zone_size_x<G>(m, i, j, k));
#pragma unroll
          for (int c = 0; c < N; c++) {
            const real* a = AQ + c * nxv;
            const real* qL = SL + c * nxv;
            const real* qR = FC + c * nxv;
            qr[s][c] = cell::ppa_flux_ksi(v, ksi, a[im], qL[im], qR[im], a[i], qL[i], qR[i]);
          }
        }
      }
      __syncthreads();
      XR_REFRESH();
#pragma unroll
      for (int s = 0; s < CPT; s++) {
        const int i = t + s * nt;
        if (i < nx) {
#pragma unroll
          for (int c = 0; c < N; c++) FC[c * nxv + i] = qr[s][c];
        }
      }
      __syncthreads();
      XR_REFRESH();
    }
    // UpdateX for every quantity and UpdateDensityX last (src/update_x.c:69-71,
    // src/update_density_x.c:66-68), on the registers
#pragma unroll
    for (int s = 0; s < CPT; s++) {
      const int i = t + s * nt;
      if (i < nx) {
        const int ip = ixp_of(i, nx);
        const real v0 = V[i], v1 = V[ip];
        const real* ds = FC + NL * nxv;
        const Recip Vi = invvol_recip(m, i, k);
#pragma unroll
        for (int c = 0; c < NL; c++) {
          const real* f = FC + c * nxv;
          st[s][c] += cell::flux_x(m, j, k, dt, v0 * f[i] * ds[i], v1 * f[ip] * ds[ip], Vi);
        }
        st[s][NL] += cell::flux_x(m, j, k, dt, v0 * ds[i], v1 * ds[ip], Vi);
      }
    }
  }
  // AdvectSHIFT (src/advect_shift.c:45-58): cell i moves to i + ns (mod nx).
  // Quantities other than the density, Mpx and Mmx go to their shifted
  // positions straight from the registers.
#pragma unroll
  for (int s = 0; s < CPT; s++) {
    const int i = t + s * nt;
    if (i < nx) {
      int it = (i + ns) % nx;
      if (it < 0) it += nx;
#pragma unroll
      for (int c = 0; c < NL; c++) {
        const int q = XQ::first + c;
        if (q >= 2) {
          if (q < NM) qout[q - 2][base + it] = st[s][c];
          else e_out[base + it] = st[s][c];
        }
      }
      if (GRP != 2) {
        // density, Mpx, Mmx of the shifted ring into AQ for NewVelocity_x
        // (the AQ reads of the last pass ended before its final barrier)
        AQ[0 * nx + it] = st[s][0];
        AQ[1 * nx + it] = st[s][1];
        AQ[2 * nx + it] = st[s][NL];
        rho_out[base + it] = st[s][NL];
      }
    }
  }
  if (GRP != 2) {
    __syncthreads();
      XR_REFRESH();
    // NewVelocity_x (src/newvel_x.c:78-86)
#pragma unroll
    for (int s = 0; s < CPT; s++) {
      const int i = t + s * nt;
      if (i < nx) {
        const int im = ixm_of(i, nx);
        const real* R = AQ + 2 * nx;
        vx_out[base + i] = cell::newvel_x<G>(m, i, j, k, R[i], R[im], AQ[1 * nx + i], AQ[0 * nx + im], omegaframe);
      }
    }
  }
}

// A ring of 64*CPT cells held by one wave, CPT consecutive cells per lane:
// lane L holds cells L*CPT .. L*CPT+CPT-1.  The ring is periodic, so the left
// neighbour of a lane's first cell is the last cell of the lane before (lane
// 63 for lane 0) and the right neighbour of its last cell the first cell of
// the lane after: one rotation per array, no LDS and no barrier.
template <int CPT>
struct WRing {
  template <class T>
  __device__ __forceinline__ static void left(const T (&a)[CPT], T (&out)[CPT]) {
    out[0] = from_left(a[CPT - 1]);
#pragma unroll
    for (int s = 1; s < CPT; s++) out[s] = a[s - 1];
  }
  template <class T>
  __device__ __forceinline__ static void right(const T (&a)[CPT], T (&out)[CPT]) {
#pragma unroll
    for (int s = 0; s < CPT - 1; s++) out[s] = a[s + 1];
    out[CPT - 1] = from_right(a[0]);
  }
};

// The X sweep of one ring by one wave (rings of exactly 64*CPT cells), the
// cell state in registers: the same arithmetic as xsweep_ring, cell for cell
// (ComputeResidual, VanLeerX_a/b, VanLeerX_PPA_a-d, UpdateX and
// UpdateDensityX for both passes, AdvectSHIFT, NewVelocity_x), with every
// neighbour read from a register or from the next lane.  Each quantity goes
// through a pass on its own after the density's faces (which every update
// reads), so only one quantity's stencil arrays are live at a time.  VxMed
// is summed in the wave (wave_chain) from the velocities it holds.
// TILE = 1: the wave holds a tile of a longer ring instead, the 64*CPT cells
// from tile*(64*CPT - 2 kXHalo) - kXHalo on (indexes wrapped around the
// ring), and updates the 64*CPT - 2 kXHalo cells inside the halos, the same
// reach as kf_sweep_x_tile's: the rotations wrap the tile's two ends into
// each other, and what that puts into the halo cells does not reach the
// updated ones.  VxMed then comes from kf_vmed.
template <int G, int DZ, int AD, int CPT, int TILE = 0>
__device__ __forceinline__ void xsweep_wave(const dmesh& m, int ring, double dt64, const real* __restrict__ rho_in,
                                            const real* __restrict__ vxt, const real* const* qin,
                                            const real* __restrict__ e_in, real* __restrict__ e_out,
                                            real* __restrict__ rho_out, real* __restrict__ vx_out,
                                            real* const* qout, double* __restrict__ vxmed,
                                            int* __restrict__ nshift_out, real omegaframe, int fargo,
                                            int tile = 0) {
  using W = WRing<CPT>;
  constexpr int NM = DZ ? 6 : 4;
  constexpr int NQ = NM + AD;  // transported quantities besides the density
  constexpr int TOUT = 64 * CPT - 2 * kXHalo;  // cells a tile updates
  const int nx = m.nx;         // == 64 * CPT for a whole ring
  const real dt = (real)dt64;
  const int lane = (int)threadIdx.x;
  const int j = m.nghy + ring % m.ny;
  const int k = DZ ? m.nghz + ring / m.ny : 0;
  const int base = j * m.pitch + k * m.stride;
  // this lane's first cell on the ring (a tile's may lie outside [0, nx))
  const int t0 = tile * TOUT;
  const int u0 = TILE ? t0 - kXHalo + lane * CPT : lane * CPT;
  int ix[CPT];  // the cells' indexes on the ring
#pragma unroll
  for (int s = 0; s < CPT; s++) {
    int g = u0 + s;
    if (TILE) {
      g %= nx;
      if (g < 0) g += nx;
    }
    ix[s] = g;
  }
  real rho[CPT], st[NQ][CPT];
  double vt[CPT];
#pragma unroll
  for (int s = 0; s < CPT; s++) {
    rho[s] = rho_in[base + ix[s]];
#pragma unroll
    for (int c = 0; c < NQ; c++) st[c][s] = (c < NM) ? qin[c][base + ix[s]] : e_in[base + ix[s]];
    vt[s] = vxt[base + ix[s]];
  }
  // ComputeResidual (src/compute_residual.c:68-73), FP64: the cell residual
  // velocity for the van Leer pass, the uniform one for the PPA pass
  real vel[CPT], vu[CPT];
  int ns = 0;
  if (fargo) {
    double vmed;
    if constexpr (TILE) {
      vmed = vxmed[j + k * m.sy];  // from kf_vmed
    } else {
      // ComputeVmed(Vx_temp) (src/compute_vmed.c:14-30): the ring's weights
      // are in the lanes already, in the layout of wave_ring_sum's chunks
      double w[CPT];
#pragma unroll
      for (int s = 0; s < CPT; s++) w[s] = vt[s] * (m.xmin64[ix[s] + 1] - m.xmin64[ix[s]]);
      vmed = vmed_full(wave_chain<CPT>(w, 0.0, 64) / (m.XMAX - m.XMIN), m, j, k);
    }
    const double vb = cell::vbase(m, j, k);
    double nr0 = 0.0;
#pragma unroll
    for (int s = 0; s < CPT; s++) {
      const double zs = zone_size_x64<G>(m, ix[s], j, k);
      const double ntilde = vmed * dt64 / zs;
      const double nround = floor(ntilde + 0.5);
      if (s == 0) nr0 = nround;
      if constexpr (cell::kFix64) vel[s] = (real)(vt[s] - (vmed - vb));
      else vel[s] = (real)(vt[s] - vmed);
      vu[s] = (real)((ntilde - nround) * zs / dt64);
    }
    // the shift of the ring is that of its cell 0: lane 0's, or for a tile
    // the same expression at i = 0
    if constexpr (TILE) ns = (int)floor(vmed * dt64 / zone_size_x64<G>(m, 0, j, k) + 0.5);
    else
      ns = (int)__hiloint2double(__builtin_amdgcn_readlane(__double2hiint(nr0), 0),
                                 __builtin_amdgcn_readlane(__double2loint(nr0), 0));
    if (lane == 0 && (!TILE || tile == 0)) {
      nshift_out[j + k * m.sy] = ns;
      if (!TILE) vxmed[j + k * m.sy] = vmed;
    }
  } else {
    const double vb = cell::vbase(m, j, k);
#pragma unroll
    for (int s = 0; s < CPT; s++) {
      if constexpr (cell::kFix64) vel[s] = (real)(vt[s] + vb);
      else vel[s] = (real)vt[s];
    }
  }
  // per-cell mesh factors shared by the quantities
  real zs[CPT], zsl[CPT];
  Recip Vi[CPT];
#pragma unroll
  for (int s = 0; s < CPT; s++) {
    zs[s] = zone_size_x<G>(m, ix[s], j, k);
    Vi[s] = invvol_recip(m, ix[s], k);
  }
  W::left(zs, zsl);

  // X_advection (src/transport.c:208-226), twice with FARGO
  for (int pass = 0; pass < (fargo ? 2 : 1); pass++) {
    const bool ppa = (pass == 1);
    if (ppa) {
#pragma unroll
      for (int s = 0; s < CPT; s++) vel[s] = vu[s];
    }
    real vr[CPT];
    W::right(vel, vr);
    Recip Rr[CPT];
    real ksi[CPT];
#pragma unroll
    for (int s = 0; s < CPT; s++) {
      Rr[s] = recip(rho[s]);  // DivideByRho of every quantity (the density before this pass)
      ksi[s] = ppa ? cell::ppa_ksi(vel[s], dt, zs[s]) : real(0.0);
    }
    // the face value (van Leer) or swept-zone average (PPA) of q at each cell's
    // left face, for one quantity
    auto faces = [&](const real (&q)[CPT], real (&fc)[CPT]) {
      real ql[CPT], qr[CPT];
      W::left(q, ql);
      W::right(q, qr);
      if (!ppa) {
        // VanLeerX_a (src/vanleer_x_a.c:60-68), gradient once per face
        real g[CPT], gr[CPT], sl[CPT], sll[CPT];
#pragma unroll
        for (int s = 0; s < CPT; s++) g[s] = rdiv(q[s] - ql[s], recip(zs[s]));
        W::right(g, gr);
#pragma unroll
        for (int s = 0; s < CPT; s++) sl[s] = vanleer_slope(g[s], gr[s]);
        W::left(sl, sll);
        // VanLeerX_b (src/vanleer_x_b.c:60-65)
#pragma unroll
        for (int s = 0; s < CPT; s++) fc[s] = vanleer_face(vel[s], dt, ql[s], sll[s], zsl[s], q[s], sl[s], zs[s]);
      } else {
        // VanLeerX_PPA_a/b/c/d (src/fargo_ppa_a.c:58-69, src/fargo_ppa_b.c:52-55,
        // src/fargo_ppa_c.c:52-63, src/fargo_ppa_d.c:62-70)
        real sl[CPT], slr[CPT], fr[CPT], qL[CPT], qR[CPT];
#pragma unroll
        for (int s = 0; s < CPT; s++) sl[s] = cell::ppa_slope(ql[s], q[s], qr[s]);
        W::right(sl, slr);
#pragma unroll
        for (int s = 0; s < CPT; s++) fr[s] = cell::ppa_face(q[s], qr[s], sl[s], slr[s]);  // q at face i+1/2
        W::left(fr, qL);
#pragma unroll
        for (int s = 0; s < CPT; s++) {
          qR[s] = fr[s];
          cell::ppa_limit(q[s], qL[s], qR[s]);
        }
        real qLl[CPT], qRl[CPT];
        W::left(qL, qLl);
        W::left(qR, qRl);
#pragma unroll
        for (int s = 0; s < CPT; s++)
          fc[s] = cell::ppa_flux_ksi(vel[s], ksi[s], ql[s], qLl[s], qRl[s], q[s], qL[s], qR[s]);
      }
    };
    // the density's faces first: every quantity's update reads them
    real fd[CPT], fdr[CPT];
    faces(rho, fd);
    W::right(fd, fdr);
    // UpdateX of every quantity (src/update_x.c:69-71)
#pragma unroll
    for (int c = 0; c < NQ; c++) {
      real q[CPT], fc[CPT], fcr[CPT];
#pragma unroll
      for (int s = 0; s < CPT; s++) q[s] = rdiv(st[c][s], Rr[s]);
      faces(q, fc);
      W::right(fc, fcr);
#pragma unroll
      for (int s = 0; s < CPT; s++)
        st[c][s] += cell::flux_x(m, j, k, dt, vel[s] * fc[s] * fd[s], vr[s] * fcr[s] * fdr[s], Vi[s]);
    }
    // UpdateDensityX last (src/update_density_x.c:66-68)
#pragma unroll
    for (int s = 0; s < CPT; s++) rho[s] += cell::flux_x(m, j, k, dt, vel[s] * fd[s], vr[s] * fdr[s], Vi[s]);
  }
  // AdvectSHIFT (src/advect_shift.c:45-58): cell i moves to i + ns (mod nx);
  // NewVelocity_x (src/newvel_x.c:78-86) at the destination, from the moved
  // cell and its left neighbour
  real rl[CPT], mpl[CPT];
  W::left(rho, rl);
  W::left(st[0], mpl);
#pragma unroll
  for (int s = 0; s < CPT; s++) {
    // a tile writes the cells inside its halos that lie on the ring
    if (TILE && (u0 + s < t0 || u0 + s >= t0 + TOUT || u0 + s >= nx)) continue;
    int it = (ix[s] + ns) % nx;
    if (it < 0) it += nx;
#pragma unroll
    for (int c = 2; c < NQ; c++) {
      if (c < NM) qout[c - 2][base + it] = st[c][s];
      else e_out[base + it] = st[c][s];
    }
    rho_out[base + it] = rho[s];
    vx_out[base + it] = cell::newvel_x<G>(m, it, j, k, rho[s], rl[s], st[1][s], mpl[s], omegaframe);
  }
}

// LDS bytes of xsweep_ring for n carried quantities + the density
inline size_t xsweep_ring_lds(int nx, int ncarried) {
  const size_t P = ((size_t)(ncarried + 1) * nx + 1) & ~(size_t)1;
  return (3 * P + nx) * sizeof(real) + 64;
}

// ComputeVmed(Vx_temp) (src/compute_vmed.c:14-30) for the X sweeps: one wave
// per active ring and fluid, one sequential sum with S terms per lane
// (wave_ring_sum); blocks [f*nrings, (f+1)*nrings) take fluid f.
struct VmTab {
  const real* vxt[8];
  double* vxmed[8];
};
template <int G, int DZ, int S>
__device__ __forceinline__ void kd_vmed(const dim3 bid, const dmesh& m, const real* vxt, double* vxmed) {
  const int j = m.nghy + (int)(bid.x % m.ny);
  const int k = DZ ? m.nghz + (int)(bid.x / m.ny) : 0;
  const int base = j * m.pitch + k * m.stride;
  const double sum = wave_ring_sum<S>(vxt + base, m.xmin64, m.nx);
  if (threadIdx.x == 0) vxmed[j + k * m.sy] = vmed_full(sum / (m.XMAX - m.XMIN), m, j, k);
}
template <int G, int DZ, int S>
__global__ void __launch_bounds__(64) kf_vmed(dmesh m, VmTab T, int nrings) {
  const int fl = (int)blockIdx.x / nrings;
  kd_vmed<G, DZ, S>(dim3(blockIdx.x % nrings), m, T.vxt[fl], T.vxmed[fl]);
}

// The 2-D Y sweep with the X sweep's VxMed pass in the same launch: the first
// nvm workgroups (one wave each) sum one ring of one fluid's Vx_temp
// (kd_vmed), the others sweep (workgroup L of the sweep's grid gy, in
// XCD-contiguous order).  The Y sweep neither reads what the sums write nor
// writes the Vx_temp they read.  The ring sums are one dependent chain each,
// latency bound, so their waves run beside the sweep's without the second
// stream and its events.
template <int G, int AD, int SPLIT>
__global__ void __launch_bounds__(BXF) kf_sweep_y2v(dmesh m, real dt, Y2Tab T, int rows_per, real of, VmTab V,
                                                    int nrings, int nvm, dim3 gy) {
  const int lin = (int)blockIdx.x;
  if (lin < nvm) {
    const int fl = lin / nrings;
    kd_vmed<G, 0, 8>(dim3(lin % nrings), m, V.vxt[fl], V.vxmed[fl]);
    return;
  }
  const int N = (int)(gy.x * gy.y * gy.z), L = lin - nvm;
  const int x = L & 7, q = N >> 3, r = N & 7;
  const int p = x * q + min(x, r) + (L >> 3);
  const dim3 b(p % gy.x, (p / gy.x) % gy.y, p / (gy.x * gy.y));
  const int fl = (int)b.z;
  kd_sweep_y2<G, 0, AD, SPLIT>(b, m, dt, T.rho_in[fl], T.vxt[fl], T.vyt[fl], T.e_in[fl], T.rho_out[fl], T.q[fl][0],
                               T.q[fl][1], T.q[fl][2], T.q[fl][3], T.e_out[fl], rows_per, of);
}

// X sweep of rings too long for LDS: a workgroup of T threads owns a tile of
// T - 2H cells and loads H = 6 halo cells on each side (the van Leer pass
// needs 2, the PPA pass 3 more of the van Leer result, NewVelocity_x the left
// neighbour), so no tile waits for another; thread t owns tile position t.
// VxMed comes from kf_vmed.  The outputs go to their shifted positions
// directly (AdvectSHIFT, src/advect_shift.c:45-58) with NewVelocity_x computed
// at the destination cell from the source cell and its left neighbour.
__host__ __device__ constexpr int x_tile_cells(int threads) { return threads - 2 * kXHalo; }
template <int G, int DZ, int AD, int T>
__device__ __forceinline__ void kd_sweep_x_tile(const dim3 bid, const dmesh& m, double dt64, const real* rho_in, const real* vxt, const real* mpx, const real* mmx, const real* mpy, const real* mmy, const real* mpz, const real* mmz, const real* e_in, real* e_out, real* rho_out, real* vx_out, real* mpy_o, real* mmy_o, real* mpz_o, real* mmz_o, const double* vxmed, int* nshift_out, real omegaframe, int fargo) {
  constexpr int H = kXHalo, W = T, TW = x_tile_cells(T);
  static_assert(2 * H <= TW, "tile narrower than its halos");
  const real dt = (real)dt64;
  constexpr int NM = DZ ? 6 : 4;
  constexpr int NQ = NM + AD;
  constexpr int N = NQ + 1;
  // q = Q/rho, slopes (then limited qL), faces (then qR, then face fluxes),
  // the advection velocity of the pass (the cell residual, then the uniform
  // one); the cell state stays in registers
  __shared__ real AQ[N * W], SL[N * W], FC[N * W];
  __shared__ real V[W];
  const int nx = m.nx;
  const int ring = bid.y;
  const int j = m.nghy + ring % m.ny;
  const int k = DZ ? m.nghz + ring / m.ny : 0;
  const int base = j * m.pitch + k * m.stride;
  const int i0 = bid.x * TW;
  const int p = threadIdx.x;
  const double vmed = vxmed[j + k * m.sy];
  const real* qin[7] = {mpx, mmx, mpy, mmy, DZ ? mpz : e_in, DZ ? mmz : e_in, e_in};
  auto gidx = [&](int q) {
    int g = (i0 - H + q) % nx;
    return g < 0 ? g + nx : g;
  };
  // the integer shift of the ring: ComputeResidual at i = 0 (src/compute_residual.c:68-71)
  const int ns = fargo ? (int)floor(vmed * dt64 / zone_size_x64<G>(m, 0, j, k) + 0.5) : 0;
  if (fargo && bid.x == 0 && p == 0) nshift_out[j + k * m.sy] = ns;
  const int g = gidx(p);
  real st[N];
  st[NQ] = rho_in[base + g];
#pragma unroll
  for (int c = 0; c < NQ; c++) st[c] = (c < NM) ? qin[c][base + g] : e_in[base + g];
  const real zs = zone_size_x<G>(m, g, j, k);
  // the left neighbour's zone (the upwind face reads it)
  const real zsm = zone_size_x<G>(m, gidx(p - 1), j, k);
  real vp, vu = real(0.0);  // this cell's advection velocity in the current pass; the uniform residual
  {
    const double v = vxt[base + g];
    if (fargo) {
      // the residual split in FP64 (src/compute_residual.c:68-73)
      const double zs64 = zone_size_x64<G>(m, g, j, k);
      const double ntilde = vmed * dt64 / zs64;
      const double nround = floor(ntilde + 0.5);
      if constexpr (cell::kFix64) vp = (real)(v - (vmed - cell::vbase(m, j, k)));
      else vp = (real)(v - vmed);
      vu = (real)((ntilde - nround) * zs64 / dt64);
    } else {
      if constexpr (cell::kFix64) vp = (real)(v + cell::vbase(m, j, k));
      else vp = (real)v;
    }
    V[p] = vp;
  }
  for (int pass = 0; pass < (fargo ? 2 : 1); pass++) {
    // valid range of the state entering this pass: [lo, hi)
    const int lo = pass ? 2 : 0, hi = pass ? W - 2 : W;
    if (p >= lo && p < hi) {
      const real r = st[NQ];
      const Recip R = recip(r);
#pragma unroll
      for (int c = 0; c < NQ; c++) AQ[c * W + p] = rdiv(st[c], R);
      AQ[NQ * W + p] = r;
    }
    __syncthreads();  // also orders the first pass after the stores of V
    if (pass) {
      // the uniform residual: every read of the cell residual ended before
      // the barrier above
      vp = vu;
      V[p] = vu;
    }
    int ulo, uhi;  // positions updated by this pass
    if (!pass) {
      // the gradient at face p, (q_p - q_{p-1}) / zone_p, once per face into
      // FC: the left gradient of cell p and the right one of cell p-1
      if (p >= 1) {
        const Recip Z = recip(zs);
#pragma unroll
        for (int c = 0; c < N; c++) {
          const real* a = AQ + c * W;
          FC[c * W + p] = rdiv(a[p] - a[p - 1], Z);
        }
      }
      __syncthreads();
      if (p >= 1 && p < W - 1) {
#pragma unroll
        for (int c = 0; c < N; c++) SL[c * W + p] = vanleer_slope(FC[c * W + p], FC[c * W + p + 1]);
      }
      __syncthreads();
      if (p >= 2 && p < W - 1) {
#pragma unroll
        for (int c = 0; c < N; c++) {
          const real* a = AQ + c * W;
          const real* sl = SL + c * W;
          FC[c * W + p] = vanleer_face(vp, dt, a[p - 1], sl[p - 1], zsm, a[p], sl[p], zs);
        }
      }
      __syncthreads();
      ulo = 2;
      uhi = W - 2;
    } else {
      if (p >= 3 && p < W - 3) {
#pragma unroll
        for (int c = 0; c < N; c++) {
          const real* a = AQ + c * W;
          SL[c * W + p] = cell::ppa_slope(a[p - 1], a[p], a[p + 1]);
        }
      }
      __syncthreads();
      if (p >= 3 && p < W - 4) {
#pragma unroll
        for (int c = 0; c < N; c++) {
          const real* a = AQ + c * W;
          const real* sl = SL + c * W;
          FC[c * W + p] = cell::ppa_face(a[p], a[p + 1], sl[p], sl[p + 1]);
        }
      }
      __syncthreads();
      // limited (qL, qR): qL into SL (slopes are no longer read), qR into FC
      // once every neighbour has read its face
      real qr[N];
      const bool lim = p >= 4 && p < W - 4;
      if (lim) {
#pragma unroll
        for (int c = 0; c < N; c++) {
          real qL = FC[c * W + p - 1], qR = FC[c * W + p];
          cell::ppa_limit(AQ[c * W + p], qL, qR);
          SL[c * W + p] = qL;
          qr[c] = qR;
        }
      }
      __syncthreads();
      if (lim) {
#pragma unroll
        for (int c = 0; c < N; c++) FC[c * W + p] = qr[c];
      }
      __syncthreads();
      // face fluxes into FC once every neighbour has read qR
      const bool flx = p >= 5 && p < W - 4;
      if (flx) {
        const real ksi = cell::ppa_ksi(vp, dt, zs);
#pragma unroll
        for (int c = 0; c < N; c++) {
          const real* a = AQ + c * W;
          const real* qL = SL + c * W;
          const real* qR = FC + c * W;
          qr[c] = cell::ppa_flux_ksi(vp, ksi, a[p - 1], qL[p - 1], qR[p - 1], a[p], qL[p], qR[p]);
        }
      }
      __syncthreads();
      if (flx) {
#pragma unroll
        for (int c = 0; c < N; c++) FC[c * W + p] = qr[c];
      }
      __syncthreads();
      ulo = 5;
      uhi = W - 5;
    }
    // UpdateX / UpdateDensityX (src/update_x.c:69-71, src/update_density_x.c:66-68), on the registers
    if (p >= ulo && p < uhi) {
      const real v0 = vp, v1 = V[p + 1];
      const real* ds = FC + NQ * W;
      const Recip Vg = invvol_recip(m, g, k);
#pragma unroll
      for (int c = 0; c < NQ; c++) {
        const real* f = FC + c * W;
        st[c] += cell::flux_x(m, j, k, dt, v0 * f[p] * ds[p], v1 * f[p + 1] * ds[p + 1], Vg);
      }
      st[NQ] += cell::flux_x(m, j, k, dt, v0 * ds[p], v1 * ds[p + 1], Vg);
    }
    if (pass == 0 && fargo) __syncthreads();  // the PPA pass rewrites AQ and V
  }
  // the updated density and Mpx of positions [H-1, H+TW) for NewVelocity_x
  // (AQ is free: its last reads were before the final barrier of the pass)
  if (p >= H - 1 && p < H + TW) {
    AQ[0 * W + p] = st[0];
    AQ[1 * W + p] = st[NQ];
  }
  __syncthreads();
  // outputs of the tile's cells [H, H+TW) at their shifted positions
  // (AdvectSHIFT, src/advect_shift.c:45-58), NewVelocity_x computed at the
  // destination from the source cell and its left neighbour
  if (p >= H && p < H + TW) {
    const int i = i0 + p - H;
    if (i < nx) {
      int d = (i + ns) % nx;
      if (d < 0) d += nx;
      vx_out[base + d] = cell::newvel_x<G>(m, d, j, k, st[NQ], AQ[1 * W + p - 1], st[1], AQ[0 * W + p - 1],
                                           omegaframe);
      rho_out[base + d] = st[NQ];
#pragma unroll
      for (int c = 2; c < NQ; c++) {
        if (c < NM) (c == 2 ? mpy_o : c == 3 ? mmy_o : c == 4 ? mpz_o : mmz_o)[base + d] = st[c];
        else e_out[base + d] = st[c];
      }
    }
  }
}
template <int G, int DZ, int AD, int T>
__global__ void __launch_bounds__(T) kf_sweep_x_tile(dmesh m, double dt64, const real* __restrict__ rho_in, const real* __restrict__ vxt,
                                                      const real* __restrict__ mpx, const real* __restrict__ mmx, const real* __restrict__ mpy,
                                                      const real* __restrict__ mmy, const real* __restrict__ mpz, const real* __restrict__ mmz,
                                                      const real* __restrict__ e_in, real* __restrict__ e_out, real* __restrict__ rho_out,
                                                      real* __restrict__ vx_out, real* __restrict__ mpy_o, real* __restrict__ mmy_o, real* __restrict__ mpz_o,
                                                      real* __restrict__ mmz_o, const double* __restrict__ vxmed, int* __restrict__ nshift_out,
                                                      real omegaframe, int fargo) {
  kd_sweep_x_tile<G, DZ, AD, T>(xcd_bid(), m, dt64, rho_in, vxt, mpx, mmx, mpy, mmy, mpz, mmz, e_in, e_out, rho_out, vx_out, mpy_o, mmy_o, mpz_o, mmz_o, vxmed, nshift_out, omegaframe, fargo);
}

// NewVelocity_y/z on active faces and Stockholm damping on active cells.
template <int G, int DZ>
__device__ __forceinline__ void kd_finish(const dim3 bid, const dmesh& m, real dt, const real* __restrict__ rho_x, const real* __restrict__ mmy, const real* __restrict__ mpy, const real* __restrict__ mmz, const real* __restrict__ mpz, real* __restrict__ rho, real* __restrict__ vx, real* __restrict__ vy, real* __restrict__ vz, const real* __restrict__ rho0, const real* __restrict__ vx0, const real* __restrict__ vy0, const real* __restrict__ vz0, cell::DampPar dp, int stockholm) {
  const int i = bid.x * blockDim.x + threadIdx.x;
  if (i >= m.nx) return;
  const int j = m.nghy + (int)bid.y;
  const int k = DZ ? m.nghz + (int)bid.z : 0;
  const int l = i + j * m.pitch + k * m.stride;
  real r = rho_x[l];
  // the faces j = nghy and k = nghz lie on the mesh boundary; the ghost fill
  // sets them to zero, so they are not computed here (the damping, which runs
  // on the values in registers, reads the stored ones there)
  real y = (j > m.nghy) ? cell::newvel_y(m, i, j, k, r, rho_x[l - m.pitch], mmy[l], mpy[l - m.pitch]) : real(0.0);
  real z = real(0.0);
  if (DZ && k > m.nghz) {
    z = (mul_vol(m, mmz[l], i, j, k) + mul_vol(m, mpz[l - m.stride], i, j, k - 1)) /
        (mul_vol(m, r, i, j, k) + mul_vol(m, rho_x[l - m.stride], i, j, k - 1));
    if (G == SPH) z /= m.ymed[j];
  }
  if (stockholm) {
    // StockholmBoundary (src/stockholm.c:129-184) after NewVelocity_y/z; the
    // ramps depend on the ring only, so the branches are uniform
    const cell::Ramps R = cell::stockholm_ramps<DZ>(m, j, k, dp);
    const int l2 = j + k * m.sy;
    if (R.ramp > real(0.0)) {
      if (j == m.nghy) y = vy[l];
      real x = vx[l];
      cell::stockholm_rxy<G>(m, j, k, dt, R, r, x, y, rho0[l2], vx0[l2], vy0[l2], dp);
      vx[l] = x;
      if (j == m.nghy) vy[l] = y;
    }
    if (DZ && R.rampi > real(0.0)) {
      if (k == m.nghz) z = vz[l];
      z = cell::stockholm_z(dt, R, z, vz0[l2]);
      if (k == m.nghz) vz[l] = z;
    }
  }
  if (j > m.nghy) vy[l] = y;
  if (DZ && k > m.nghz) vz[l] = z;
  rho[l] = r;
}
// every fluid in one launch: blockIdx.z = fluid * (planes) + plane
struct FinTab {
  const real* rho_x[8];
  const real* mmy[8];
  const real* mpy[8];
  const real* mmz[8];
  const real* mpz[8];
  real* rho[8];
  real* vx[8];
  real* vy[8];
  real* vz[8];
  const real* rho0[8];
  const real* vx0[8];
  const real* vy0[8];
  const real* vz0[8];
};
// The end-of-step Z ghost fill (bound_z_cell, both sides) inside the
// finishing kernel: the thread that finishes plane k of column (i, j) writes
// the ghost values that plane is the source of, from its own final values.
// Lower side: ghost plane mm takes rho, vx, vy (and e) of plane
// 2 nghz - 1 - mm and vz of plane 2 nghz - mm, negated; upper side: ghost
// plane nz + nghz + mm takes all of plane nz + nghz - 1 - mm (vz one plane
// further out, for mm < nghz - 1); the boundary faces vz(nghz) and
// vz(nz + nghz) become zero.  Every source cell is active and written only by
// its own thread, so no thread reads what another writes.  What the separate
// Z fill also wrote in the Y ghost rows of the Z ghost planes, the Y fill that
// follows rewrites in full (rho, vx, vz, e of every Y ghost row; vy of the
// ghost faces and the boundary face).
// on: 1 the Z sides (3-D), 2 the Y sides (2-D, yfill_from_row); zd0/zd1:
// the density factors of the lower and upper Z ghost layers
struct ZFill {
  int on;
  real* e[8];
  cell::BcPar b0, b1;
  const real* zd0;
  const real* zd1;
};
__device__ __forceinline__ void zfill_from_plane(const dmesh& m, int i, int j, int k, real* dens, real* vx, real* vy,
                                                 real* vz, real* e, const ZFill& z) {
  const int nghz = m.nghz, nz = m.nz, stride = m.stride;
  const int col = i + j * m.pitch, l = col + k * stride;
  const int mlo = 2 * nghz - 1 - k;  // lower ghost plane whose rho, vx, vy, e this plane gives
  if (mlo >= 0 && mlo < nghz) {
    const int lg = col + mlo * stride;
    dens[lg] = dens[l] * z.zd0[mlo * m.sy + j];
    vx[lg] = cell::ghost_z_vx(m, z.b0, mlo, j, k, mlo, vx[l]);
    vy[lg] = vy[l];
    if (z.b0.energy) e[lg] = e[l];
  }
  const int mlz = 2 * nghz - k;  // lower ghost plane whose vz this plane gives
  if (mlz >= 0 && mlz < nghz) vz[col + mlz * stride] = -vz[l];
  if (k == nghz) vz[l] = 0;  // after this thread's own store there
  const int mup = nz + nghz - 1 - k;  // upper ghost layer this plane gives
  if (mup >= 0 && mup < nghz) {
    const int gh = nz + nghz + mup, lg = col + gh * stride;
    dens[lg] = dens[l] * z.zd1[mup * m.sy + j];
    vx[lg] = cell::ghost_z_vx(m, z.b1, mup, j, k, gh, vx[l]);
    vy[lg] = vy[l];
    if (mup < nghz - 1) vz[lg + stride] = -vz[l];
    if (mup == 0) vz[col + (nz + nghz) * stride] = 0;
    if (z.b1.energy) e[lg] = e[l];
  }
}
// The 2-D end-of-step Y ghost fill (bound_y_cell of both sides, the disk
// family's rules without the 3-D factors) inside the finishing kernel, as the
// Z fill above: row j of column i gives the lower ghost row 2 nghy - 1 - j
// its rho, vx (and e) and ghost row 2 nghy - j its vy, negated; the upper
// ghost row ny + nghy + (ny + nghy - 1 - j) all of them (vy one row further
// out); the boundary faces vy(nghy) and vy(ny + nghy) become zero.
__device__ __forceinline__ void yfill_from_row(const dmesh& m, int i, int j, real* dens, real* vx, real* vy, real* e,
                                               const ZFill& z) {
  const int nghy = m.nghy, ny = m.ny, pitch = m.pitch;
  const int l = i + j * pitch;
  const int mlo = 2 * nghy - 1 - j;
  if (mlo >= 0 && mlo < nghy) {
    const int lg = i + mlo * pitch;
    real dg, xg;
    cell::ghost_y_dx(m, z.b0, mlo, 0, j, mlo, dens[l], vx[l], nullptr, dg, xg);
    dens[lg] = dg;
    vx[lg] = xg;
    if (z.b0.energy) e[lg] = e[l];
  }
  const int mly = 2 * nghy - j;
  if (mly >= 0 && mly < nghy) vy[i + mly * pitch] = -vy[l];
  if (j == nghy) vy[l] = 0;  // after this thread's own store there
  const int mup = ny + nghy - 1 - j;
  if (mup >= 0 && mup < nghy) {
    const int gh = ny + nghy + mup, lg = i + gh * pitch;
    real dg, xg;
    cell::ghost_y_dx(m, z.b1, mup, 0, j, gh, dens[l], vx[l], nullptr, dg, xg);
    dens[lg] = dg;
    vx[lg] = xg;
    if (mup < nghy - 1) vy[lg + pitch] = -vy[l];
    if (mup == 0) vy[i + (ny + nghy) * pitch] = 0;
    if (z.b1.energy) e[lg] = e[l];
  }
}
template <int G, int DZ>
__global__ void __launch_bounds__(WG) kf_finish(dmesh m, real dt, FinTab T, int nzb, cell::DampPar dp, int stockholm,
                                                ZFill zf) {
  const int fl = (int)blockIdx.z / nzb;
  const dim3 b(blockIdx.x, blockIdx.y, blockIdx.z % nzb);
  kd_finish<G, DZ>(b, m, dt, T.rho_x[fl], T.mmy[fl], T.mpy[fl], T.mmz[fl], T.mpz[fl], T.rho[fl], T.vx[fl], T.vy[fl],
                   T.vz[fl], T.rho0[fl], T.vx0[fl], T.vy0[fl], T.vz0[fl], dp, stockholm);
  if (zf.on) {
    const int i = (int)b.x * (int)blockDim.x + (int)threadIdx.x;
    if (i < m.nx) {
      if (DZ)
        zfill_from_plane(m, i, m.nghy + (int)b.y, m.nghz + (int)b.z, T.rho[fl], T.vx[fl], T.vy[fl], T.vz[fl],
                         zf.e[fl], zf);
      else
        yfill_from_row(m, i, m.nghy + (int)b.y, T.rho[fl], T.vx[fl], T.vy[fl], zf.e[fl], zf);
    }
  }
}

// ------------------------------------------------------------------ ghost fill, potential, damping references
// The same per-cell formulas as the per-routine entry points (f3d_cell.h),
// in this build's type.  Y: both sides in one launch when ny > nghy (the
// sides then write disjoint rows and read none the other writes), Z likewise.
// One thread per (i, side, plane or row) fills the three ghost layers of its
// side: the loads of every layer first, then the stores.
__device__ __forceinline__ void kd_bound_y(const dim3 bid, const dmesh& m, int side0, real* dens, real* vx, real* vy, real* vz, real* e, cell::BcPar bp0, cell::BcPar bp1, const real* dexp0, const real* dexp1) {
  const int i = bid.x * blockDim.x + threadIdx.x;
  if (i >= m.nx) return;
  const int side = side0 + (int)bid.y, k = bid.z;
  const cell::BcPar& b = side ? bp1 : bp0;
  cell::GhostCell c[kNgh];
#pragma unroll
  for (int mm = 0; mm < kNgh; mm++) c[mm] = cell::ghost_y_load(m, side, i, mm, k, dens, vx, vy, vz, e, b);
#pragma unroll
  for (int mm = 0; mm < kNgh; mm++)
    cell::ghost_y_store(m, mm, k, c[mm], dens, vx, vy, vz, e, b, side ? dexp1 : dexp0);
}
__global__ void kf_bound_y(dmesh m, int side0, real* dens, real* vx, real* vy, real* vz, real* e, cell::BcPar bp0,
                           cell::BcPar bp1, const real* dexp0, const real* dexp1) {
  kd_bound_y(blockIdx, m, side0, dens, vx, vy, vz, e, bp0, bp1, dexp0, dexp1);
}
// 2-D: the Y fill of every fluid in one launch (blockIdx.z = fluid), both sides
struct BTab {
  real* dens[8];
  real* vx[8];
  real* vy[8];
  real* e[8];
};
__global__ void kf_bound_y_fluids(dmesh m, BTab T, cell::BcPar bp0, cell::BcPar bp1, const real* dexp0,
                                  const real* dexp1) {
  const int fl = (int)blockIdx.z;
  kd_bound_y(dim3(blockIdx.x, blockIdx.y, 0), m, 0, T.dens[fl], T.vx[fl], T.vy[fl], T.vy[fl], T.e[fl], bp0, bp1,
             dexp0, dexp1);
}

__global__ void kf_bound_z(dmesh m, int side0, real* dens, real* vx, real* vy, real* vz, real* e, cell::BcPar bp0,
                           cell::BcPar bp1, const real* zdens0, const real* zdens1) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m.nx) return;
  const int side = side0 + (int)blockIdx.z, j = blockIdx.y;
  const cell::BcPar& b = side ? bp1 : bp0;
  cell::GhostCell c[kNgh];
#pragma unroll
  for (int mm = 0; mm < kNgh; mm++) c[mm] = cell::ghost_z_load(m, side, i, mm, j, dens, vx, vy, vz, e, b);
#pragma unroll
  for (int mm = 0; mm < kNgh; mm++)
    cell::ghost_z_store(m, mm, j, c[mm], dens, vx, vy, vz, e, b, side ? zdens1 : zdens0);
}

// Potential (src/potential.c:51-208) on every cell
template <int G, int DZ>
__global__ void __launch_bounds__(WG) kf_potential(dmesh m, real* __restrict__ pot, cell::PlanetSet ps) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m.nx) return;
  const int j = blockIdx.y, k = DZ ? (int)blockIdx.z : 0;
  pot[i + j * m.pitch + k * m.stride] = cell::potential<G>(m, i, j, k, ps);
}

// init_stockholm (src/init_stockholm.c:3-126): the damping references are the
// i = 0 column of the initial fields
__global__ void kf_init_stockholm(dmesh m, const real* rho, const real* vx, const real* vy, const real* vz,
                                  const real* e, real* rho0, real* vx0, real* vy0, real* vz0, real* e0,
                                  int adiabatic) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int nk = m.dimz ? m.sz : 1;
  if (t >= m.sy * nk) return;
  const int j = t % m.sy, k = t / m.sy;
  const int l = j * m.pitch + k * m.stride, l2 = j + k * m.sy;
  if (adiabatic) e0[l2] = e[l];
  vx0[l2] = vx[l];
  vy0[l2] = vy[l];
  if (m.dimz) vz0[l2] = vz[l];
  rho0[l2] = rho[l];
}

// ------------------------------------------------------------------ multifluid
// The fluids of a multifluid run, as the drag and dust-diffusion kernels see
// them (one struct by value; at most F3D_MAX_FLUIDS fluids).
struct FluidSet {
  const real* rho[8];
  real* vx[8];
  real* vy[8];
  real* vz[8];
  double alpha[64];  // collision rates Alpha (src/colrate.c:10-23), column + row*NF
  int nf;
};

// One direction of the implicit drag (src/collision_kernel.h:14-94, src/gauss.h
// :1-45) at one cell; v[] holds that direction's velocity of each fluid.
template <int NF>
__device__ __forceinline__ void drag_solve(const FluidSet& F, real* const* v, int l, int idm, double dt,
                                           double omega) {
  // the precision form solves in FP64: the stiff drag of small grains
  // couples the fluids' velocities so strongly that an FP32 solve rounds
  // every ring's cells alike, a torque that accumulates step after step
  using dreal = std::conditional_t<cell::kFix64, double, real>;
  dreal m[NF * NF], b[NF];
#pragma unroll
  for (int o = 0; o < NF; o++) {
#pragma unroll
    for (int p = 0; p < NF; p++) {
      const dreal rho_p = dreal(0.5) * ((dreal)F.rho[p][l] + (dreal)F.rho[p][idm]);
      if (p != o) {
        const dreal rho_o = dreal(0.5) * ((dreal)F.rho[o][l] + (dreal)F.rho[o][idm]);
        if (p > o) m[p + o * NF] = -(dreal)dt * (dreal)omega * (dreal)F.alpha[p + o * NF] * rho_p / rho_o;
        else m[p + o * NF] = -(dreal)dt * (dreal)omega * (dreal)F.alpha[p + o * NF];
      } else {
        dreal sum = dreal(0.0);
#pragma unroll
        for (int q = 0; q < NF; q++) {
          if (q != p) {
            const dreal rho_q = dreal(0.5) * ((dreal)F.rho[q][l] + (dreal)F.rho[q][idm]);
            if (q > p) sum += (dreal)F.alpha[q + p * NF] * rho_q / rho_p;
            else sum += (dreal)F.alpha[q + p * NF];
          }
        }
        m[p + p * NF] = dreal(1.0) + (dreal)dt * (dreal)omega * sum;
      }
    }
    b[o] = (dreal)v[o][l];
  }
#pragma unroll
  for (int ir = 0; ir < NF - 1; ir++) {
    // pivot as src/gauss.h:8-14: the comparison uses |m|, the stored maximum
    // the signed value
    dreal max_value = dreal(-1);
    int ir_max = ir;
#pragma unroll
    for (int ir2 = ir; ir2 < NF; ir2++) {
      if (fabs(m[ir + ir2 * NF]) >= max_value) {
        max_value = m[ir + ir2 * NF];
        ir_max = ir2;
      }
    }
#pragma unroll
    for (int r = ir + 1; r < NF; r++) {
      if (ir_max == r) {
#pragma unroll
        for (int ic = 0; ic < NF; ic++) {
          const dreal temp = m[ic + ir * NF];
          m[ic + ir * NF] = m[ic + r * NF];
          m[ic + r * NF] = temp;
        }
        const dreal temp = b[ir];
        b[ir] = b[r];
        b[r] = temp;
      }
    }
#pragma unroll
    for (int ir2 = ir + 1; ir2 < NF; ir2++) {
      const dreal factor = m[ir + ir2 * NF] / m[ir + ir * NF];
#pragma unroll
      for (int ic = 0; ic < NF; ic++) m[ic + ir2 * NF] -= m[ic + ir * NF] * factor;
      b[ir2] -= b[ir] * factor;
    }
  }
#pragma unroll
  for (int ir = NF - 1; ir > -1; ir--) {
#pragma unroll
    for (int ic = NF - 1; ic > ir; ic--) b[ir] -= m[ic + ir * NF] * b[ic];
    b[ir] /= m[ir + ir * NF];
  }
#pragma unroll
  for (int o = 0; o < NF; o++) v[o][l] = (real)b[o];
}

// Collisions(dt, 1) (src/collisions.c:156-189) on V_temp in place over the
// reference's range (k, j from 1, ghosts included): the X, Y and Z solves of a
// cell touch different velocity components, so one launch does the
// directions dirs (bit 0: x, 1: y, 2: z) in the reference's order.
template <int G, int DZ, int NF>
__global__ void __launch_bounds__(BXF) kf_collisions(dmesh m, FluidSet F, double dt, int dirs) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m.nx) return;
  const int j = 1 + (int)blockIdx.y;
  const int k = DZ ? 1 + (int)blockIdx.z : 0;
  const int l = i + j * m.pitch + k * m.stride;
  const double om_c = sqrt(1.0 * 1.0 / (m.ymed64[j] * m.ymed64[j] * m.ymed64[j]));
  const double om_f = sqrt(1.0 * 1.0 / (m.ymin64[j] * m.ymin64[j] * m.ymin64[j]));
  if (dirs & 1) drag_solve<NF>(F, F.vx, l, lxm_of(l, i, m.nx), dt, om_c);
  if (dirs & 2) drag_solve<NF>(F, F.vy, l, l - m.pitch, dt, om_f);
  if (DZ && (dirs & 4)) drag_solve<NF>(F, F.vz, l, l - m.stride, dt, om_c);
}

// DustDiffusion_Coefficients (src/dust_diffusion_coefficients.c:10-143),
// 2-D isothermal forms, every cell: yczc at cell centres, yfzc at lower Y faces
template <int G, int DZ>
__global__ void __launch_bounds__(BXF) kf_ddiff_coef(dmesh m, const real* cs, real* yczc, real* yfzc, int alpha_visc,
                                                     real alpha, real nu) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m.nx) return;
  const int j = blockIdx.y;
  const int ll = i + j * m.pitch;
  if (alpha_visc) {
    const real c0 = cs[ll];
    const real soundspeed2 = c0 * c0;
    real soundspeedf2;
    if (j == 0) soundspeedf2 = soundspeed2;
    else soundspeedf2 = real(0.5) * (c0 + cs[ll - m.pitch]) * real(0.5) * (c0 + cs[ll - m.pitch]);
    const real r3yczc = m.ymed[j] * m.ymed[j] * m.ymed[j];
    const real r3yfzc = m.ymin[j] * m.ymin[j] * m.ymin[j];
    yczc[ll] = alpha * soundspeed2 / sqrt(real(1.0) * real(1.0) / r3yczc);
    yfzc[ll] = alpha * soundspeedf2 / sqrt(real(1.0) * real(1.0) / r3yfzc);
  } else {
    yczc[ll] = nu;
    yfzc[ll] = nu;
  }
}

// DustDiffusion_Core (src/dust_diffusion_core.c:10-192), CYLINDRICAL 2-D, on
// rows j in [1, sy-2]: the dust fluid's updated density into temp
template <int G, int DZ>
__device__ __forceinline__ void ddiff_core_cell(const dmesh& m, real dt, int bx, int by, const real* __restrict__ rhod,
                                                const real* __restrict__ rhog, const real* __restrict__ yczc,
                                                const real* __restrict__ yfzc, real* __restrict__ temp) {
  const int i = bx * (int)blockDim.x + (int)threadIdx.x;
  if (i >= m.nx) return;
  const int j = 1 + by;
  const int ll = i + j * m.pitch;
  const int llxm = lxm_of(ll, i, m.nx), llxp = lxp_of(ll, i, m.nx), llyp = ll + m.pitch, llym = ll - m.pitch;
  real update = real(0.0);
  const real c = rhod[ll] / (rhod[ll] + rhog[ll]);
  const real dxmin = m.xmin[i + 1] - m.xmin[i];
  const real dxmedp = real(1.0) / m.invdiffxmed[ixp_of(i, m.nx)];
  const real dxmedm = real(1.0) / m.invdiffxmed[i];
  const real d1 = real(0.25) * (rhod[ll] + rhog[ll] + rhod[llxp] + rhog[llxp]) * (yczc[llxp] + yczc[ll]);
  const real d2 = real(0.25) * (rhod[ll] + rhog[ll] + rhod[llxm] + rhog[llxm]) * (yczc[llxm] + yczc[ll]);
  const real cxp = rhod[llxp] / (rhod[llxp] + rhog[llxp]);
  const real cxm = rhod[llxm] / (rhod[llxm] + rhog[llxm]);
  const real ym = m.ymed[j];
  update += real(1.0) / ym / ym / (dxmin) * (d1 * (cxp - c) / (dxmedp) - (d2 * (c - cxm)) / (dxmedm));
  const real d3 = real(0.5) * (rhod[ll] + rhog[ll] + rhod[llyp] + rhog[llyp]) * yfzc[llyp];
  const real d4 = real(0.5) * (rhod[ll] + rhog[ll] + rhod[llym] + rhog[llym]) * yfzc[ll];
  const real cyp = rhod[llyp] / (rhod[llyp] + rhog[llyp]);
  const real cym = rhod[llym] / (rhod[llym] + rhog[llym]);
  update += real(1.0) / ym / (m.ymin[j + 1] - m.ymin[j]) *
            (m.ymin[j + 1] * d3 * (cyp - c) / (m.ymed[j + 1] - ym) -
             (m.ymin[j] * d4 * (c - cym)) / (ym - m.ymed[j - 1]));
  temp[ll] = rhod[ll] + dt * update;
}

// The dust fluids' cores in one launch (blockIdx.z = dust fluid) and the copy
// back with the density ghost fill in one more: Temp -> Density over the
// whole field, then FillGhosts(DENS) (src/dust_diffusion_main.c:20-25), give
// active rows = Temp and ghost rows = the boundary formula applied to Temp's
// active rows, which kf_ddiff_store writes directly.
struct DdTab {
  real* rhod[8];
  real* temp[8];
};
template <int G, int DZ>
__global__ void __launch_bounds__(BXF) kf_ddiff_core_fluids(dmesh m, real dt, DdTab T, const real* rhog,
                                                            const real* yczc, const real* yfzc) {
  const int fl = (int)blockIdx.z;
  ddiff_core_cell<G, DZ>(m, dt, (int)blockIdx.x, (int)blockIdx.y, T.rhod[fl], rhog, yczc, yfzc, T.temp[fl]);
}
__global__ void kf_ddiff_store(dmesh m, DdTab T, cell::BcPar bp0, cell::BcPar bp1) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m.nx) return;
  const int j = (int)blockIdx.y, fl = (int)blockIdx.z;
  const real* __restrict__ tp = T.temp[fl];
  real* __restrict__ rd = T.rhod[fl];
  const int top = m.ny + m.nghy;
  if (j < m.nghy) rd[i + j * m.pitch] = tp[i + (2 * m.nghy - j - 1) * m.pitch] * bp0.dens[j];
  else if (j >= top) rd[i + j * m.pitch] = tp[i + (2 * top - 1 - j) * m.pitch] * bp1.dens[j - top];
  else rd[i + j * m.pitch] = tp[i + j * m.pitch];
}


// ------------------------------------------------------------------ C ABI

#define FDISPATCH(ctx, KERN, grid, block, shmem, ...)                                                  \
  do {                                                                                                  \
    if (m.geom == SPH) {                                                                         \
      if (m.dimz)                                                                                \
        hipLaunchKernelGGL((KERN<SPH, 1>), grid, block, shmem, (ctx)->stream, m, __VA_ARGS__);   \
      else                                                                                              \
        hipLaunchKernelGGL((KERN<SPH, 0>), grid, block, shmem, (ctx)->stream, m, __VA_ARGS__);   \
    } else {                                                                                            \
      if (m.dimz)                                                                                \
        hipLaunchKernelGGL((KERN<CYL, 1>), grid, block, shmem, (ctx)->stream, m, __VA_ARGS__);   \
      else                                                                                              \
        hipLaunchKernelGGL((KERN<CYL, 0>), grid, block, shmem, (ctx)->stream, m, __VA_ARGS__);   \
    }                                                                                                   \
    F3D_LAUNCH_CHECK(ctx);                                                                              \
  } while (0)

// the same with a third template argument AD (1: adiabatic, the energy is transported)
// KERN<G, DZ, T_>
#define FDISPATCH_T(ctx, KERN, T_, grid, block, shmem, ...)                                            \
  do {                                                                                                  \
    if (m.geom == SPH) {                                                                                \
      if (m.dimz)                                                                                       \
        hipLaunchKernelGGL((KERN<SPH, 1, T_>), grid, block, shmem, (ctx)->stream, m, __VA_ARGS__);      \
      else                                                                                              \
        hipLaunchKernelGGL((KERN<SPH, 0, T_>), grid, block, shmem, (ctx)->stream, m, __VA_ARGS__);      \
    } else {                                                                                            \
      if (m.dimz)                                                                                       \
        hipLaunchKernelGGL((KERN<CYL, 1, T_>), grid, block, shmem, (ctx)->stream, m, __VA_ARGS__);      \
      else                                                                                              \
        hipLaunchKernelGGL((KERN<CYL, 0, T_>), grid, block, shmem, (ctx)->stream, m, __VA_ARGS__);      \
    }                                                                                                   \
    F3D_LAUNCH_CHECK(ctx);                                                                              \
  } while (0)

#define FDISPATCH_AD(ctx, KERN, grid, block, shmem, ...)                                                 \
  do {                                                                                                    \
    const bool ad_ = ((ctx)->ph.eos == 2);                                                                \
    if (m.geom == SPH) {                                                                           \
      if (m.dimz) {                                                                                \
        if (ad_) hipLaunchKernelGGL((KERN<SPH, 1, 1>), grid, block, shmem, (ctx)->stream, m, __VA_ARGS__); \
        else hipLaunchKernelGGL((KERN<SPH, 1, 0>), grid, block, shmem, (ctx)->stream, m, __VA_ARGS__);     \
      } else {                                                                                            \
        if (ad_) hipLaunchKernelGGL((KERN<SPH, 0, 1>), grid, block, shmem, (ctx)->stream, m, __VA_ARGS__); \
        else hipLaunchKernelGGL((KERN<SPH, 0, 0>), grid, block, shmem, (ctx)->stream, m, __VA_ARGS__);     \
      }                                                                                                   \
    } else {                                                                                              \
      if (m.dimz) {                                                                                \
        if (ad_) hipLaunchKernelGGL((KERN<CYL, 1, 1>), grid, block, shmem, (ctx)->stream, m, __VA_ARGS__); \
        else hipLaunchKernelGGL((KERN<CYL, 1, 0>), grid, block, shmem, (ctx)->stream, m, __VA_ARGS__);     \
      } else {                                                                                            \
        if (ad_) hipLaunchKernelGGL((KERN<CYL, 0, 1>), grid, block, shmem, (ctx)->stream, m, __VA_ARGS__); \
        else hipLaunchKernelGGL((KERN<CYL, 0, 0>), grid, block, shmem, (ctx)->stream, m, __VA_ARGS__);     \
      }                                                                                                   \
    }                                                                                                     \
    F3D_LAUNCH_CHECK(ctx);                                                                                \
  } while (0)

// FDISPATCH_AD with one more template argument T_ after the AD flag
#define FDISPATCH_AD_T(ctx, KERN, T_, grid, block, shmem, ...)                                                    \
  do {                                                                                                            \
    const bool ad_ = ((ctx)->ph.eos == 2);                                                                        \
    if (m.geom == SPH) {                                                                                          \
      if (m.dimz) {                                                                                               \
        if (ad_) hipLaunchKernelGGL((KERN<SPH, 1, 1, T_>), grid, block, shmem, (ctx)->stream, m, __VA_ARGS__);   \
        else hipLaunchKernelGGL((KERN<SPH, 1, 0, T_>), grid, block, shmem, (ctx)->stream, m, __VA_ARGS__);       \
      } else {                                                                                                    \
        if (ad_) hipLaunchKernelGGL((KERN<SPH, 0, 1, T_>), grid, block, shmem, (ctx)->stream, m, __VA_ARGS__);   \
        else hipLaunchKernelGGL((KERN<SPH, 0, 0, T_>), grid, block, shmem, (ctx)->stream, m, __VA_ARGS__);       \
      }                                                                                                           \
    } else {                                                                                                      \
      if (m.dimz) {                                                                                               \
        if (ad_) hipLaunchKernelGGL((KERN<CYL, 1, 1, T_>), grid, block, shmem, (ctx)->stream, m, __VA_ARGS__);   \
        else hipLaunchKernelGGL((KERN<CYL, 1, 0, T_>), grid, block, shmem, (ctx)->stream, m, __VA_ARGS__);       \
      } else {                                                                                                    \
        if (ad_) hipLaunchKernelGGL((KERN<CYL, 0, 1, T_>), grid, block, shmem, (ctx)->stream, m, __VA_ARGS__);   \
        else hipLaunchKernelGGL((KERN<CYL, 0, 0, T_>), grid, block, shmem, (ctx)->stream, m, __VA_ARGS__);       \
      }                                                                                                           \
    }                                                                                                             \
    F3D_LAUNCH_CHECK(ctx);                                                                                        \
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

// The handle's fields as this build's type (the FP32 form allocates FP32
// arrays behind the same pointers; VxMed stays FP64, Nshift int).
struct RF {
  real *Density, *Energy, *Vx, *Vy, *Vz, *Vx_temp, *Vy_temp, *Vz_temp;
  real *Mpx, *Mmx, *Mpy, *Mmy, *Mpz, *Mmz, *Pot;
  real *Density0, *Energy0, *Vx0, *Vy0, *Vz0;
  double* VxMed;
  int* Nshift;
};

// one fluid's fields (f) with the handle's shared fields (sh)
static RF rfields_of(const f3d_fields& f, const f3d_fields& sh) {
  auto R = [](double* p) { return reinterpret_cast<real*>(p); };
  RF r;
  r.Density = R(f.Density); r.Energy = R(f.Energy); r.Vx = R(f.Vx); r.Vy = R(f.Vy); r.Vz = R(f.Vz);
  r.Vx_temp = R(f.Vx_temp); r.Vy_temp = R(f.Vy_temp); r.Vz_temp = R(f.Vz_temp);
  r.Mpx = R(sh.Mpx); r.Mmx = R(sh.Mmx); r.Mpy = R(sh.Mpy); r.Mmy = R(sh.Mmy); r.Mpz = R(sh.Mpz); r.Mmz = R(sh.Mmz);
  r.Pot = R(sh.Pot);
  r.Density0 = R(f.Density0); r.Energy0 = R(f.Energy0); r.Vx0 = R(f.Vx0); r.Vy0 = R(f.Vy0); r.Vz0 = R(f.Vz0);
  r.VxMed = sh.VxMed;
  r.Nshift = sh.Nshift;
  return r;
}

static RF rfields(const f3d_hip_ctx* ctx) {
  const f3d_fields& f = ctx->f;
  auto R = [](double* p) { return reinterpret_cast<real*>(p); };
  RF r;
  r.Density = R(f.Density); r.Energy = R(f.Energy); r.Vx = R(f.Vx); r.Vy = R(f.Vy); r.Vz = R(f.Vz);
  r.Vx_temp = R(f.Vx_temp); r.Vy_temp = R(f.Vy_temp); r.Vz_temp = R(f.Vz_temp);
  r.Mpx = R(f.Mpx); r.Mmx = R(f.Mmx); r.Mpy = R(f.Mpy); r.Mmy = R(f.Mmy); r.Mpz = R(f.Mpz); r.Mmz = R(f.Mmz);
  r.Pot = R(f.Pot);
  r.Density0 = R(f.Density0); r.Energy0 = R(f.Energy0); r.Vx0 = R(f.Vx0); r.Vy0 = R(f.Vy0); r.Vz0 = R(f.Vz0);
  r.VxMed = f.VxMed;
  r.Nshift = f.Nshift;
  return r;
}

static cell::BcPar bc_par_f(const f3d_hip_ctx* ctx, int side) {
  cell::BcPar b;
  b.p3d = (ctx->ph.bc == 2);
  b.energy = ctx->ph.energy_bc && ctx->f.Energy != nullptr;
  b.of = ctx->ph.OMEGAFRAME;
  for (int q = 0; q < 3; q++) {
    b.dens[q] = (real)(side ? ctx->ph.ymax_dens[q] : ctx->ph.ymin_dens[q]);
    b.vsq[q] = (real)(side ? ctx->ph.ymax_vsq[q] : ctx->ph.ymin_vsq[q]);
    b.vfac[q] = (real)(side ? ctx->ph.ymax_vfac[q] : ctx->ph.ymin_vfac[q]);
    b.zvfac[q] = (real)(side ? ctx->ph.zmax_vfac[q] : ctx->ph.zmin_vfac[q]);
  }
  return b;
}

// The ghost fill (src/boundary.c:2-23: zmin, zmax, then ymin, ymax) on the
// given density/velocity buffers and the energy.
// skipz: the Z fill was done by the kernel before (kf_finish with ZFill).
static int fused_boundaries(f3d_hip_ctx* ctx, Scratch* S, real* dens, real* vx, real* vy, real* vz, real* e,
                            int skipz = 0) {
  const dmesh& m = S->m;
  constexpr int BX = 64;
  dim3 blk(BX);
  const cell::BcPar b0 = bc_par_f(ctx, 0), b1 = bc_par_f(ctx, 1);
  F3D_REQUIRE(ctx, m.nghy == kNgh && (!m.dimz || m.nghz == kNgh), "ghost fill: three ghost layers");
  if (!ctx->ph.periodicz && m.dimz && !skipz) {
    F3D_REQUIRE(ctx, ctx->ph.bc == 2, "3-D boundaries need the spherical-disk family");
    // both sides in one launch when they read and write disjoint planes
    if (m.nz > m.nghz) {
      hipLaunchKernelGGL(kf_bound_z, dim3((m.nx + BX - 1) / BX, m.sy, 2), blk, 0, ctx->stream, m, 0, dens, vx, vy, vz,
                         e, b0, b1, S->zmin_dens, S->zmax_dens);
      F3D_LAUNCH_CHECK(ctx);
    } else {
      for (int side = 0; side < 2; side++) {
        const cell::BcPar b = side ? b1 : b0;
        const real* t = side ? S->zmax_dens : S->zmin_dens;
        hipLaunchKernelGGL(kf_bound_z, dim3((m.nx + BX - 1) / BX, m.sy, 1), blk, 0, ctx->stream, m, side, dens, vx,
                           vy, vz, e, b, b, t, t);
        F3D_LAUNCH_CHECK(ctx);
      }
    }
  }
  if (!ctx->ph.periodicy) {
    if (m.ny > m.nghy) {
      hipLaunchKernelGGL(kf_bound_y, dim3((m.nx + BX - 1) / BX, 2, m.dimz ? m.sz : 1), blk, 0, ctx->stream, m, 0,
                         dens, vx, vy, vz, e, b0, b1, S->ymin_dexp, S->ymax_dexp);
      F3D_LAUNCH_CHECK(ctx);
    } else {
      for (int side = 0; side < 2; side++) {
        const cell::BcPar b = side ? b1 : b0;
        const real* t = side ? S->ymax_dexp : S->ymin_dexp;
        hipLaunchKernelGGL(kf_bound_y, dim3((m.nx + BX - 1) / BX, 1, m.dimz ? m.sz : 1), blk, 0, ctx->stream, m,
                           side, dens, vx, vy, vz, e, b, b, t, t);
        F3D_LAUNCH_CHECK(ctx);
      }
    }
  }
  return 0;
}

// tuning knobs of the column sweeps (rows per thread, f3d_fused_set_tiling;
// rows_y 0: automatic).  Measured on MI355X: 4 rows in Y on the 256x96x32,
// 4096x1024 and 1024x512 disks (2 rows: 4096x1024 step 0.554 -> 0.579 ms),
// 2 rows on the 384x128 disk (0.0587 -> 0.0571 ms), so 2 rows where 4 would
// leave the launch under 1024 waves; 6 planes in Z (the 32 planes then take
// 6 workgroup rows, one wave of launches at 3 waves per SIMD: 42.6 -> 41.3 us
// against 3 planes)
static int g_rows_y = 0, g_rows_z = 6;
static int rows_y_of(const dmesh& m, int nf) {
  if (g_rows_y > 0) return g_rows_y;
  const long w4 = (long)((m.nx + BXF - 1) / BXF) * (m.dimz ? m.nz : nf) * ((m.ny + 3) / 4);
  return w4 < 1024 ? 2 : 4;
}
// column sweeps: 1 two threads per column (each with the density and half of
// the quantities), 0 one, -1 automatic: measured on MI355X, the split only
// pays when a sweep has fewer waves than the GPU has SIMDs (small 2-D grids)
static int g_colsplit = -1;
// X-sweep kernel choice: 0 automatic, 1 ring in LDS, 2 tiles in LDS, 3 one
// wave per ring with the cells in registers (rings of 256 cells), 4 tiles of
// one wave each with the cells in registers
static int g_xmode = 0;
// ring X sweep: 1 two workgroups per ring (half of the quantities each), 0
// one; -1 automatic
static int g_xsplit = -1;
// CFL kernel form, bit 0: ring minima to an array and a one-workgroup
// second MIN pass (else one atomic MIN per ring); bit 1: the ring sums of
// VxMed in a pass of their own (else two sums inside each ring's workgroup).
// -1: chosen from the mesh shape.
static int g_cflmode = -1;
// 2-D sources of a viscous gas in one pass without the stress arrays
// (kf_src_fused2): -1 automatic (on), 0 off (kf_src_a + kf_src_b)
static int g_srcfuse = -1;
// end-of-step ghost fill inside kf_finish (3-D: the Z sides, 2-D: the Y
// sides): -1 automatic (on), 0 off (launches of its own)
static int g_zfill = -1;
// 2-D: the X sweep's VxMed pass inside the Y sweep's launch where it would
// run on the second stream: -1 automatic (on), 0 off
static int g_y2v = -1;

// CFL form, bits: 1 ring minima to an array and a second pass (else one
// atomic MIN per ring), 2 VxMed from kf_vmed_rings, 4 (with 2) one wave per
// workgroup.  Measured on MI355X: the array and its MIN pass, which hands
// dt to the host itself, pay on every mesh (the atomic form's copy and event
// cost more: 384x128 step 0.0609 -> 0.0561 ms, 4096x1024 alike); the VxMed
// pass for many or long rings, one wave per ring for rings of <= 256 cells.
static int cfl_mode(const dmesh& m, int nf) {
  if (g_cflmode >= 0) return g_cflmode;
  const int nrings = m.ny * (m.dimz ? m.nz : 1) * nf;
  int mode = 1 | (nrings >= 2048 ? 2 : 0) | (m.nx >= 1024 ? 2 : 0);
  if ((mode & 2) && m.nx <= 256) mode |= 4;
  return mode;
}

// Terms per lane of the one-wave ring sums (wave_ring_sum): 4, 8 or 16; -1
// automatic.  Measured on MI355X (whole FP64 step): 8 is fastest on rings of
// 256 (3-D disk, 0.2503/0.2488/0.2546 ms for 4/8/16), 1024 (gas + 4 dust,
// 0.4682/0.4573/0.4616 ms) and 4096 cells (adiabatic disk, 0.6811/0.6639/
// 0.6756 ms), against 0.2515 and 0.680 ms with the sums read from LDS.
static int g_vmed_terms = -1;
static int vmed_terms(int nx) {
  (void)nx;
  return g_vmed_terms > 0 ? g_vmed_terms : 8;
}

// VxMed of every active ring and of the ring above the last (the shear limit
// reads ring j+1), for the CFL kernel's pre-pass form: one wave per ring, four
// for rings of 1024 cells or more (measured on MI355X: their loads, issued
// by one wave, would take longer than the sum)
static int vmed_prepass_m(f3d_hip_ctx* ctx, const dmesh& m, const real* v) {
  const int nry = m.ny + (ctx->ph.fargo ? 1 : 0);
  const int nr = nry * (m.dimz ? m.nz : 1);
  switch (vmed_terms(m.nx)) {
    case 4: FDISPATCH_T(ctx, kf_vmed_rings, 4, dim3(nr), dim3(64), 0, v, ctx->f.VxMed, nry); break;
    case 8: FDISPATCH_T(ctx, kf_vmed_rings, 8, dim3(nr), dim3(64), 0, v, ctx->f.VxMed, nry); break;
    default: FDISPATCH_T(ctx, kf_vmed_rings, 16, dim3(nr), dim3(64), 0, v, ctx->f.VxMed, nry); break;
  }
  return 0;
}
static int vmed_prepass(f3d_hip_ctx* ctx, Scratch* S, const real* v) { return vmed_prepass_m(ctx, S->m, v); }

// The arrays one fluid's transport writes between its kernels: the Y-swept
// density and energy, the swept momenta (Mpx, Mmx, Mpy, Mmy, Mpz, Mmz), the
// X-swept density and Y/Z momenta, and where VxMed and Nshift go.  Fluid 0
// uses the handle's arrays; the other fluids of a multifluid transport get
// arrays of their own, allocated on first use, so that every fluid's X sweep
// can run in one launch.  As in the reference, VxMed and Nshift keep the
// last fluid's values.
struct TSet {
  real *rhoy, *ey, *rhox;
  real* m[6];
  real* xm[4];
  double* vxmed;
  int* nshift;
};

static int tset_of(f3d_hip_ctx* ctx, Scratch* S, const RF& f, int fluid, int nf, TSet* t) {
  const size_t bytes = ctx->n3 * sizeof(real);
  if (fluid == 0) {
    t->rhoy = S->f[FS_RHOY]; t->ey = S->f[FS_EY]; t->rhox = S->f[FS_RHOX];
    real* mm[6] = {f.Mpx, f.Mmx, f.Mpy, f.Mmy, f.Mpz, f.Mmz};
    for (int a = 0; a < 6; a++) t->m[a] = mm[a];
    for (int a = 0; a < 4; a++) t->xm[a] = S->f[FS_XM0 + a];
  } else {
    F3D_REQUIRE(ctx, fluid < 8, "transport: at most 8 fluids");
    real** a = S->fset[fluid];
    const int na = ctx->m.dimz ? 13 : 11;  // 2-D: no Z momenta
    for (int q = 0; q < na; q++)
      if (!a[q]) {
        F3D_CHECK(ctx, hipMalloc(&a[q], bytes));
        F3D_CHECK(ctx, hipMemsetAsync(a[q], 0, bytes, ctx->stream));
      }
    t->rhoy = a[0]; t->ey = a[1]; t->rhox = a[2];
    for (int q = 0; q < 4; q++) t->xm[q] = a[3 + q];
    for (int q = 0; q < 6; q++) t->m[q] = (7 + q < na) ? a[7 + q] : nullptr;
  }
  if (fluid == nf - 1) {
    t->vxmed = f.VxMed;
    t->nshift = f.Nshift;
  } else {
    if (!S->vxmed_f[fluid]) {
      F3D_CHECK(ctx, hipMalloc(&S->vxmed_f[fluid], ctx->n2 * sizeof(double)));
      F3D_CHECK(ctx, hipMalloc(&S->nshift_f[fluid], ctx->n2 * sizeof(int)));
    }
    t->vxmed = S->vxmed_f[fluid];
    t->nshift = S->nshift_f[fluid];
  }
  return 0;
}

// NewVelocity_y/z, Stockholm damping and the end-of-step ghost fill
static int finish_step(f3d_hip_ctx* ctx, Scratch* S, double dt, int end_of_step, int nf, const RF* f,
                       const TSet* t) {
  const dmesh& m = S->m;
  cell::DampPar dp;
  dp.Y_inf = (real)ctx->ph.Y_inf; dp.Y_sup = (real)ctx->ph.Y_sup;
  dp.Z_inf = (real)ctx->ph.Z_inf; dp.Z_sup = (real)ctx->ph.Z_sup;
  dp.y_min = (real)ctx->ph.ymin_par; dp.y_max = (real)ctx->ph.ymax_par;
  dp.z_min = (real)ctx->ph.zmin_par; dp.z_max = (real)ctx->ph.zmax_par;
  dp.ds = (real)ctx->ph.taudamp; dp.of = ctx->ph.OMEGAFRAME; dp.of0 = ctx->ph.OMEGAFRAME0;
  dp.periodic_z = ctx->ph.periodicz;
  FinTab T;
  memset(&T, 0, sizeof(T));
  for (int q = 0; q < nf; q++) {
    real* const* XM = t[q].xm;
    T.rho_x[q] = t[q].rhox; T.mmy[q] = XM[1]; T.mpy[q] = XM[0]; T.mmz[q] = XM[3]; T.mpz[q] = XM[2];
    T.rho[q] = f[q].Density; T.vx[q] = f[q].Vx; T.vy[q] = f[q].Vy; T.vz[q] = f[q].Vz;
    T.rho0[q] = f[q].Density0; T.vx0[q] = f[q].Vx0; T.vy0[q] = f[q].Vy0; T.vz0[q] = f[q].Vz0;
  }
  const int nzb = m.dimz ? m.nz : 1;
  dim3 gf((m.nx + BXF - 1) / BXF, m.ny, nzb * nf);
  // the end-of-step ghost fill inside the finishing kernel, one launch
  // fewer: in 3-D the Z sides (the Y fill follows as its own launch), in 2-D
  // the Y sides (the disk family's rules)
  ZFill zf;
  memset(&zf, 0, sizeof(zf));
  const int zmerge = end_of_step && m.dimz && !ctx->ph.periodicz && ctx->ph.bc == 2 && m.nghz == kNgh &&
                     m.nz >= 2 * m.nghz && g_zfill != 0;
  const int ymerge = end_of_step && !m.dimz && !ctx->ph.periodicy && ctx->ph.bc != 2 && m.nghy == kNgh &&
                     m.ny >= 2 * m.nghy && g_zfill != 0;
  if (zmerge || ymerge) {
    zf.on = zmerge ? 1 : 2;
    for (int q = 0; q < nf; q++) zf.e[q] = f[q].Energy;
    zf.b0 = bc_par_f(ctx, 0);
    zf.b1 = bc_par_f(ctx, 1);
    zf.zd0 = S->zmin_dens;
    zf.zd1 = S->zmax_dens;
  }
  f3d_stage_begin(ctx, F3D_ST_FINISH);
  FDISPATCH(ctx, kf_finish, gf, dim3(BXF), 0, (real)dt, T, nzb, dp, ctx->ph.stockholm, zf);
  f3d_stage_end(ctx, F3D_ST_FINISH);
  if (end_of_step && !ymerge) {
    // end-of-step ghost fill (src/main.c:431)
    f3d_stage_begin(ctx, F3D_ST_BC_END);
    if (nf > 1 && !m.dimz && !ctx->ph.periodicy && m.ny > m.nghy) {
      // 2-D: every fluid's Y fill in one launch (the same boundary factors)
      BTab B;
      memset(&B, 0, sizeof(B));
      for (int q = 0; q < nf; q++) {
        B.dens[q] = f[q].Density; B.vx[q] = f[q].Vx; B.vy[q] = f[q].Vy; B.e[q] = f[q].Energy;
      }
      hipLaunchKernelGGL(kf_bound_y_fluids, dim3((m.nx + 63) / 64, 2, nf), dim3(64), 0, ctx->stream, m, B,
                         bc_par_f(ctx, 0), bc_par_f(ctx, 1), S->ymin_dexp, S->ymax_dexp);
      F3D_LAUNCH_CHECK(ctx);
    } else {
      for (int q = 0; q < nf; q++)
        if (fused_boundaries(ctx, S, f[q].Density, f[q].Vx, f[q].Vy, f[q].Vz, f[q].Energy, zmerge)) return -1;
    }
    f3d_stage_end(ctx, F3D_ST_BC_END);
  }
  return 0;
}

// momenta and the Z (3-D) and Y sweeps of every fluid into its transport set
// (2-D: one Y-sweep launch over the fluids; 3-D: one pair of launches per fluid)
// vm (2-D only): the X sweep's VxMed pass rides in the Y sweep's launch
static int yz_sweeps(f3d_hip_ctx* ctx, Scratch* S, double dt, int nf, const RF* fl, const TSet* tl,
                     const VmTab* vm = nullptr) {
  const dmesh& m = S->m;
  const real of = (real)ctx->ph.OMEGAFRAME;
  const real rdt = (real)dt;
  const int ry = rows_y_of(m, nf), rz = g_rows_z;
  const long waves = (long)((m.nx + BXF - 1) / BXF) * (m.dimz ? (long)m.nz * ((m.ny + ry - 1) / ry)
                                                               : (long)nf * ((m.ny + ry - 1) / ry));
  const int split = g_colsplit >= 0 ? g_colsplit : (waves < 1024 ? 1 : 0);
  const int sp = split ? 2 : 1;  // threads per column
#define FDISPATCH_ADS(ctx_, KERN, grid, ...)                                             \
  do {                                                                                  \
    if (split) FDISPATCH_AD_T(ctx_, KERN, 1, grid, dim3(BXF), 0, __VA_ARGS__);         \
    else FDISPATCH_AD_T(ctx_, KERN, 0, grid, dim3(BXF), 0, __VA_ARGS__);               \
  } while (0)
  if (m.dimz) {
    for (int q = 0; q < nf; q++) {
      const RF& f = fl[q];
      const TSet& t = tl[q];
      // eight separate buffers: six Z-swept momenta, the Z-swept density and energy
      real* const* Z = &S->f[FS_Z0];
      dim3 gz((m.nx + BXF - 1) / BXF, (m.sy - 2) * sp, (m.nz + rz - 1) / rz);
      f3d_stage_begin(ctx, F3D_ST_SWEEP_Z);
      FDISPATCH_ADS(ctx, kf_sweep_z, gz, rdt, f.Density, f.Vx_temp, f.Vy_temp, f.Vz_temp, f.Energy, Z[6], Z[0], Z[1],
                    Z[2], Z[3], Z[4], Z[5], Z[7], rz, of);
      f3d_stage_end(ctx, F3D_ST_SWEEP_Z);
      dim3 gy((m.nx + BXF - 1) / BXF, ((m.ny + ry - 1) / ry) * sp, m.nz);
      f3d_stage_begin(ctx, F3D_ST_SWEEP_Y);
      FDISPATCH_ADS(ctx, kf_sweep_y3, gy, rdt, Z[6], Z[0], Z[1], Z[2], Z[3], Z[4], Z[5], Z[7], f.Vy_temp, t.rhoy,
                    t.m[0], t.m[1], t.m[2], t.m[3], t.m[4], t.m[5], t.ey, ry, of);
      f3d_stage_end(ctx, F3D_ST_SWEEP_Y);
    }
  } else {
    Y2Tab T;
    memset(&T, 0, sizeof(T));
    for (int q = 0; q < nf; q++) {
      T.rho_in[q] = fl[q].Density; T.vxt[q] = fl[q].Vx_temp; T.vyt[q] = fl[q].Vy_temp; T.e_in[q] = fl[q].Energy;
      T.rho_out[q] = tl[q].rhoy;
      for (int c = 0; c < 4; c++) T.q[q][c] = tl[q].m[c];
      T.e_out[q] = tl[q].ey;
    }
    dim3 gy((m.nx + BXF - 1) / BXF, ((m.ny + ry - 1) / ry) * sp, nf);
    f3d_stage_begin(ctx, F3D_ST_SWEEP_Y);
    if (vm) {
      const int nrings = m.ny, nvm = nrings * nf;
      const dim3 g1(nvm + gy.x * gy.y * gy.z);
      const int ad = (ctx->ph.eos == 2);
#define Y2V(G_, AD_, SP_)                                                                                     \
  hipLaunchKernelGGL((kf_sweep_y2v<G_, AD_, SP_>), g1, dim3(BXF), 0, ctx->stream, m, rdt, T, ry, of, *vm, nrings, \
                     nvm, gy)
#define Y2V_G(G_) \
  do { if (ad) { if (split) Y2V(G_, 1, 1); else Y2V(G_, 1, 0); } else { if (split) Y2V(G_, 0, 1); else Y2V(G_, 0, 0); } } while (0)
      if (m.geom == SPH) Y2V_G(SPH); else Y2V_G(CYL);
#undef Y2V_G
#undef Y2V
      F3D_LAUNCH_CHECK(ctx);
    } else {
      FDISPATCH_ADS(ctx, kf_sweep_y2, gy, rdt, T, ry, of);
    }
    f3d_stage_end(ctx, F3D_ST_SWEEP_Y);
  }
#undef FDISPATCH_ADS
  return 0;
}

// VxMed of Vx_temp for the tiled X sweep (one sequential sum per ring), on
// stream st
static int x_vmed(f3d_hip_ctx* ctx, Scratch* S, int nf, const RF* f, const TSet* t, hipStream_t st) {
  const dmesh& m = S->m;
  const int nrings = m.ny * (m.dimz ? m.nz : 1);
  F3D_REQUIRE(ctx, (size_t)m.nx * sizeof(double) <= 64 * 1024, "ring too long for the VxMed pass");
  F3D_REQUIRE(ctx, nf >= 1 && nf <= 8, "VxMed pass: 1 to 8 fluids");
  VmTab T;
  memset(&T, 0, sizeof(T));
  for (int q = 0; q < nf; q++) {
    T.vxt[q] = f[q].Vx_temp;
    T.vxmed[q] = t[q].vxmed;
  }
#define XV1(G_, DZ_, S_) hipLaunchKernelGGL((kf_vmed<G_, DZ_, S_>), dim3(nrings * nf), dim3(64), 0, st, m, T, nrings)
#define XV(G_, DZ_)                                \
  do {                                             \
    const int s_ = vmed_terms(m.nx);               \
    if (s_ == 4) XV1(G_, DZ_, 4);                  \
    else if (s_ == 8) XV1(G_, DZ_, 8);             \
    else XV1(G_, DZ_, 16);                         \
  } while (0)
  if (m.geom == SPH) {
    if (m.dimz) XV(SPH, 1); else XV(SPH, 0);
  } else {
    if (m.dimz) XV(CYL, 1); else XV(CYL, 0);
  }
#undef XV
#undef XV1
  F3D_LAUNCH_CHECK(ctx);
  return 0;
}

// tile width of the tiled X sweep (cells per workgroup): 64, 128 or 256
static int g_xtile = 128;

// Where the X sweep's VxMed comes from: 0 = the ring kernel's own sum,
// 1 = kf_vmed on the side stream, overlapping the Z/Y sweeps, 2 = kf_vmed on
// the stream just before the X sweep; -1 = chosen here.  The tiled sweep
// always takes it from kf_vmed.  Measured on MI355X: below about four ring
// workgroups per CU the ring kernel's own sum costs least; above, the pass
// pays, on the stream for short rings (a sum of <= 512 terms takes under
// 3 us) and on the side stream for long ones, whose sums it hides.
static int g_xvmed = -1;
static int x_vmed_form(const dmesh& m, int nf, bool tiles) {
  const int nrings = m.ny * (m.dimz ? m.nz : 1);
  if (tiles) return (g_xvmed == 2) ? 2 : 1;
  if (g_xvmed >= 0) return g_xvmed;
  if (nrings * nf < 1024) return 0;
  return m.nx <= 512 ? 2 : 1;
}

// X sweep by tiles (rings longer than the LDS ring kernel holds), once
// VxMed of every ring is known.
static int x_sweep_tiled(f3d_hip_ctx* ctx, Scratch* S, double dt, const RF& f, const TSet& t) {
  const dmesh& m = S->m;
  const real of = (real)ctx->ph.OMEGAFRAME;
  real* const* XM = t.xm;
  const int nrings = m.ny * (m.dimz ? m.nz : 1);
  const int ad = (ctx->ph.eos == 2);
  const int T = g_xtile;
  F3D_REQUIRE(ctx, m.nx >= 12, "the tiled X sweep needs at least 12 cells per ring");
  f3d_stage_begin(ctx, F3D_ST_SWEEP_X);
  dim3 g((m.nx + x_tile_cells(T) - 1) / x_tile_cells(T), nrings);
#define XTILE1(G_, DZ_, AD_, T_)                                                                                \
  hipLaunchKernelGGL((kf_sweep_x_tile<G_, DZ_, AD_, T_>), g, dim3(T_), 0, ctx->stream, m, dt, t.rhoy, f.Vx_temp, \
                     t.m[0], t.m[1], t.m[2], t.m[3], t.m[4], t.m[5], t.ey, f.Energy, t.rhox, f.Vx, XM[0], XM[1],  \
                     XM[2], XM[3], t.vxmed, t.nshift, of, ctx->ph.fargo)
#define XTILE(G_, DZ_, AD_) \
  do { if (T == 64) XTILE1(G_, DZ_, AD_, 64); else if (T == 256) XTILE1(G_, DZ_, AD_, 256); \
       else XTILE1(G_, DZ_, AD_, 128); } while (0)
  if (m.geom == SPH) {
    if (m.dimz) { if (ad) XTILE(SPH, 1, 1); else XTILE(SPH, 1, 0); }
    else { if (ad) XTILE(SPH, 0, 1); else XTILE(SPH, 0, 0); }
  } else {
    if (m.dimz) { if (ad) XTILE(CYL, 1, 1); else XTILE(CYL, 1, 0); }
    else { if (ad) XTILE(CYL, 0, 1); else XTILE(CYL, 0, 0); }
  }
#undef XTILE
#undef XTILE1
  F3D_LAUNCH_CHECK(ctx);
  f3d_stage_end(ctx, F3D_ST_SWEEP_X);
  return 0;
}

// the side stream and its events, created on first use
static int side_stream(f3d_hip_ctx* ctx) {
  if (!ctx->side) {
    F3D_CHECK(ctx, hipStreamCreateWithFlags(&ctx->side, hipStreamNonBlocking));
    F3D_CHECK(ctx, hipEventCreateWithFlags(&ctx->side_go, hipEventDisableTiming));
    F3D_CHECK(ctx, hipEventCreateWithFlags(&ctx->side_done, hipEventDisableTiming));
  }
  return 0;
}

// Ring X sweep of nf fluids in one launch: nf consecutive blocks of
// per_fluid workgroups, each block reading its fluid's arrays from the table.
struct XTab {
  const real* rho_in[8];
  const real* vxt[8];
  const real* q[8][6];
  const real* e_in[8];
  real* e_out[8];
  real* rho_out[8];
  real* vx_out[8];
  real* qo[8][4];
  double* vxmed[8];
  int* nshift[8];
};

// One cell per thread, VxMed from the pass, no energy: 6 waves per SIMD fit
// in 78 registers without spills (measured on the 3-D disk: 63.1 -> 61.0 us;
// the energy form would spill)
template <int G, int DZ, int AD, int CPT, int SPLIT, int VPRE>
__global__ void __launch_bounds__(CPT == 2 ? 512 : 256, (CPT == 1 && VPRE && !AD) ? 6 : 1) kf_sweep_xr_fluids(dmesh m, double dt64, XTab T, int per_fluid,
                                                           real omegaframe, int fargo) {
  const int lin = xcd_linear();
  const int fl = lin / per_fluid, b = lin % per_fluid;
  real* qout[4] = {T.qo[fl][0], T.qo[fl][1], T.qo[fl][2], T.qo[fl][3]};
  if (!SPLIT) {
    xsweep_ring<G, DZ, AD, CPT, 0, VPRE>(m, b, dt64, T.rho_in[fl], T.vxt[fl], T.q[fl], T.e_in[fl], T.e_out[fl],
                                   T.rho_out[fl], T.vx_out[fl], qout, T.vxmed[fl], T.nshift[fl], omegaframe, fargo);
  } else if ((b & 1) == 0) {
    xsweep_ring<G, DZ, AD, CPT, 1, VPRE>(m, b >> 1, dt64, T.rho_in[fl], T.vxt[fl], T.q[fl], T.e_in[fl], T.e_out[fl],
                                   T.rho_out[fl], T.vx_out[fl], qout, T.vxmed[fl], T.nshift[fl], omegaframe, fargo);
  } else {
    xsweep_ring<G, DZ, AD, CPT, 2, VPRE>(m, b >> 1, dt64, T.rho_in[fl], T.vxt[fl], T.q[fl], T.e_in[fl], T.e_out[fl],
                                   T.rho_out[fl], T.vx_out[fl], qout, T.vxmed[fl], T.nshift[fl], omegaframe, fargo);
  }
}

// one wave per ring and fluid; blocks [f*nrings, (f+1)*nrings) take fluid f
template <int G, int DZ, int AD, int CPT>
__global__ void __launch_bounds__(64, 2) kf_sweep_xw(dmesh m, double dt64, XTab T, int nrings, real omegaframe,
                                                  int fargo) {
  const int lin = xcd_linear();
  const int fl = lin / nrings, ring = lin % nrings;
  real* qout[4] = {T.qo[fl][0], T.qo[fl][1], T.qo[fl][2], T.qo[fl][3]};
  xsweep_wave<G, DZ, AD, CPT>(m, ring, dt64, T.rho_in[fl], T.vxt[fl], T.q[fl], T.e_in[fl], T.e_out[fl],
                              T.rho_out[fl], T.vx_out[fl], qout, T.vxmed[fl], T.nshift[fl], omegaframe, fargo);
}

// tiles of long rings, one wave each (xsweep_wave with TILE = 1, four cells
// per lane: 244 cells updated per tile); blocks [f*nrings*ntiles, ...) take
// fluid f, a ring's tiles consecutive
template <int G, int DZ, int AD>
__global__ void __launch_bounds__(64, 2) kf_sweep_xtw(dmesh m, double dt64, XTab T, int nrings, int ntiles,
                                                      real omegaframe, int fargo) {
  const int lin = xcd_linear();
  const int per = nrings * ntiles;
  const int fl = lin / per, rt = lin % per;
  real* qout[4] = {T.qo[fl][0], T.qo[fl][1], T.qo[fl][2], T.qo[fl][3]};
  xsweep_wave<G, DZ, AD, 4, 1>(m, rt / ntiles, dt64, T.rho_in[fl], T.vxt[fl], T.q[fl], T.e_in[fl], T.e_out[fl],
                               T.rho_out[fl], T.vx_out[fl], qout, T.vxmed[fl], T.nshift[fl], omegaframe, fargo,
                               rt % ntiles);
}

// ring X sweep form: the split (two workgroups per ring) and its LDS bytes
static void x_ring_form(const f3d_hip_ctx* ctx, const dmesh& m, int* split, size_t* shx) {
  const int ad = (ctx->ph.eos == 2);
  const int nq = (m.dimz ? 6 : 4) + ad, na = m.dimz ? 3 : 2;
  // measured on MI355X: the split pays while a thread owns one cell of the
  // ring or two of a short one (nx <= 512); beyond, its extra density work
  // costs more than the occupancy it gains
  *split = (g_xsplit < 0) ? (m.nx <= 512 ? 1 : 0) : g_xsplit;
  *shx = *split ? xsweep_ring_lds(m.nx, na > nq - na ? na : nq - na) : xsweep_ring_lds(m.nx, nq);
}

// true when the ring X sweep runs (the ring fits in LDS and no tiles are asked for)
static bool x_ring_fits(const f3d_hip_ctx* ctx, Scratch* S) {
  int split;
  size_t shx;
  x_ring_form(ctx, S->m, &split, &shx);
  const bool fits = shx <= 160 * 1024 && S->m.nx <= 1024;
  return g_xmode == 1 || ((g_xmode == 0 || g_xmode == 3) && fits);
}

// the ring X sweeps' table of every fluid's arrays
static void xtab_fill(XTab* T, int nf, const RF* f, const TSet* t) {
  memset(T, 0, sizeof(*T));
  for (int k = 0; k < nf; k++) {
    T->rho_in[k] = t[k].rhoy; T->vxt[k] = f[k].Vx_temp;
    for (int q = 0; q < 6; q++) T->q[k][q] = t[k].m[q];
    T->e_in[k] = t[k].ey; T->e_out[k] = f[k].Energy; T->rho_out[k] = t[k].rhox; T->vx_out[k] = f[k].Vx;
    for (int q = 0; q < 4; q++) T->qo[k][q] = t[k].xm[q];
    T->vxmed[k] = t[k].vxmed; T->nshift[k] = t[k].nshift;
  }
}

// The one-wave ring X sweep (kf_sweep_xw): rings of 256 cells (four per
// lane; longer rings need more registers than two waves per SIMD leave),
// chosen automatically from 2048 rings x fluids on (enough waves to fill the
// GPU), or asked for with f3d_fused_set_xmode(3).
static bool x_wave_form(const dmesh& m, int nf) {
  if (m.nx != 256) return false;
  if (g_xmode == 3) return true;
  if (g_xmode != 0) return false;
  return (long)m.ny * (m.dimz ? m.nz : 1) * nf >= 2048;
}

static int x_sweep_wave(f3d_hip_ctx* ctx, Scratch* S, double dt, int nf, const RF* f, const TSet* t) {
  const dmesh& m = S->m;
  const real of = (real)ctx->ph.OMEGAFRAME;
  const int ad = (ctx->ph.eos == 2);
  F3D_REQUIRE(ctx, m.nx == 256, "one-wave X sweep: rings of 256 cells");
  F3D_REQUIRE(ctx, nf >= 1 && nf <= 8, "X sweep: 1 to 8 fluids");
  const int nrings = m.ny * (m.dimz ? m.nz : 1);
  XTab T;
  xtab_fill(&T, nf, f, t);
  f3d_stage_begin(ctx, F3D_ST_SWEEP_X);
#define XW0(G_, DZ_, AD_, CPT_)                                                                             \
  hipLaunchKernelGGL((kf_sweep_xw<G_, DZ_, AD_, CPT_>), dim3(nrings * nf), dim3(64), 0, ctx->stream, m, dt, T, \
                     nrings, of, ctx->ph.fargo)
#define XW(G_, DZ_) \
  do { if (ad) XW0(G_, DZ_, 1, 4); else XW0(G_, DZ_, 0, 4); } while (0)
  if (m.geom == SPH) {
    if (m.dimz) XW(SPH, 1); else XW(SPH, 0);
  } else {
    if (m.dimz) XW(CYL, 1); else XW(CYL, 0);
  }
#undef XW
#undef XW0
  F3D_LAUNCH_CHECK(ctx);
  f3d_stage_end(ctx, F3D_ST_SWEEP_X);
  return 0;
}

// Ring X sweep of nf fluids in one launch.
// Tiles of one wave each (kf_sweep_xtw) for the rings the LDS ring kernel
// does not take: automatic in 2-D (the 3-D forms need more registers than
// two waves per SIMD leave and spill), or asked for with
// f3d_fused_set_xmode(4).
static bool x_wtile_form(const dmesh& m) { return g_xmode == 4 || (g_xmode == 0 && !m.dimz); }
// ... and in place of the LDS ring kernel where the tiles make enough waves:
// measured on MI355X, gas + 4 dust 1024x512 (12,800 waves) FP64 step 0.442 ->
// 0.383 ms, FP32 0.299 -> 0.284 ms; the 384x128 disk (256 waves) 0.064 ->
// 0.091 ms, so from 2048 waves on, in 2-D
static bool x_wtile_pays(const dmesh& m, int nf) {
  if (g_xmode == 4) return true;
  if (g_xmode != 0 || m.dimz) return false;
  const long ntiles = (m.nx + 243) / 244;
  return (long)m.ny * ntiles * nf >= 2048;
}

static int x_sweep_wtiles(f3d_hip_ctx* ctx, Scratch* S, double dt, int nf, const RF* f, const TSet* t) {
  const dmesh& m = S->m;
  const real of = (real)ctx->ph.OMEGAFRAME;
  const int ad = (ctx->ph.eos == 2);
  F3D_REQUIRE(ctx, nf >= 1 && nf <= 8, "X sweep: 1 to 8 fluids");
  const int nrings = m.ny * (m.dimz ? m.nz : 1);
  const int ntiles = (m.nx + 243) / 244;
  XTab T;
  xtab_fill(&T, nf, f, t);
  f3d_stage_begin(ctx, F3D_ST_SWEEP_X);
#define XTW(G_, DZ_, AD_)                                                                                     \
  hipLaunchKernelGGL((kf_sweep_xtw<G_, DZ_, AD_>), dim3(nrings * ntiles * nf), dim3(64), 0, ctx->stream, m, dt, \
                     T, nrings, ntiles, of, ctx->ph.fargo)
#define XTW_G(G_, DZ_) \
  do { if (ad) XTW(G_, DZ_, 1); else XTW(G_, DZ_, 0); } while (0)
  if (m.geom == SPH) {
    if (m.dimz) XTW_G(SPH, 1); else XTW_G(SPH, 0);
  } else {
    if (m.dimz) XTW_G(CYL, 1); else XTW_G(CYL, 0);
  }
#undef XTW_G
#undef XTW
  F3D_LAUNCH_CHECK(ctx);
  f3d_stage_end(ctx, F3D_ST_SWEEP_X);
  return 0;
}

static int x_sweep_ring(f3d_hip_ctx* ctx, Scratch* S, double dt, int nf, const RF* f, const TSet* t, int vpre) {
  const dmesh& m = S->m;
  const real of = (real)ctx->ph.OMEGAFRAME;
  const int ad = (ctx->ph.eos == 2);
  int split;
  size_t shx;
  x_ring_form(ctx, m, &split, &shx);
  F3D_REQUIRE(ctx, shx <= 160 * 1024 && m.nx <= 1024, "ring too long for the LDS X sweep");
  F3D_REQUIRE(ctx, nf >= 1 && nf <= 8, "X sweep: 1 to 8 fluids");
  // one cell per thread up to 256 cells, else two (up to 512 threads)
  const int cpt = (m.nx <= 256) ? 1 : 2;
  const int nthreads = (((m.nx + cpt - 1) / cpt + 63) / 64) * 64;
  const int per = m.ny * (m.dimz ? m.nz : 1) * (split ? 2 : 1);
  XTab T;
  xtab_fill(&T, nf, f, t);
  dim3 gx(per * nf);
  f3d_stage_begin(ctx, F3D_ST_SWEEP_X);
#define XSWEEP0(G_, DZ_, AD_, CPT_, SP_, VP_)                                                                  \
  hipLaunchKernelGGL((kf_sweep_xr_fluids<G_, DZ_, AD_, CPT_, SP_, VP_>), gx, dim3(nthreads), shx, ctx->stream, m, \
                     dt, T, per, of, ctx->ph.fargo)
#define XSWEEP1(G_, DZ_, AD_, CPT_, SP_) \
  do { if (vpre) XSWEEP0(G_, DZ_, AD_, CPT_, SP_, 1); else XSWEEP0(G_, DZ_, AD_, CPT_, SP_, 0); } while (0)
#define XSWEEP2(G_, DZ_, AD_, CPT_) \
  do { if (split) XSWEEP1(G_, DZ_, AD_, CPT_, 1); else XSWEEP1(G_, DZ_, AD_, CPT_, 0); } while (0)
#define XSWEEP(G_, DZ_, CPT_) \
  do { if (ad) XSWEEP2(G_, DZ_, 1, CPT_); else XSWEEP2(G_, DZ_, 0, CPT_); } while (0)
#define XSWEEP_G(G_, DZ_) \
  do { if (cpt == 1) XSWEEP(G_, DZ_, 1); else XSWEEP(G_, DZ_, 2); } while (0)
  if (m.geom == SPH) {
    if (m.dimz) XSWEEP_G(SPH, 1); else XSWEEP_G(SPH, 0);
  } else {
    if (m.dimz) XSWEEP_G(CYL, 1); else XSWEEP_G(CYL, 0);
  }
#undef XSWEEP_G
#undef XSWEEP0
#undef XSWEEP1
#undef XSWEEP2
#undef XSWEEP
  F3D_LAUNCH_CHECK(ctx);
  f3d_stage_end(ctx, F3D_ST_SWEEP_X);
  return 0;
}

// Transport of nf fluids (fields fl[k], the handle's shared fields): every
// fluid's momenta and Z/Y sweeps, one X-sweep launch over all fluids (or each
// fluid's tiles), then each fluid's NewVelocity_y/z, damping and ghost fill.
// Fluids do not interact in Transport, so this equals transporting them one
// after the other (src/main.c:420).
static int transport_fluids(f3d_hip_ctx* ctx, double dt, int nf, const f3d_fields* fl, int end_of_step) {
  Scratch* S;
  F3D_REQUIRE(ctx, scratch_of(ctx, &S) == 0, "fused scratch allocation failed");
  F3D_REQUIRE(ctx, nf >= 1 && nf <= 8, "transport: 1 to 8 fluids");
  RF f[8];
  TSet t[8];
  for (int k = 0; k < nf; k++) {
    f[k] = fl ? rfields_of(fl[k], ctx->f) : rfields(ctx);
    if (tset_of(ctx, S, f[k], k, nf, &t[k])) return -1;
  }
  const bool wave = x_wave_form(S->m, nf);
  const bool tiles = !wave && !x_ring_fits(ctx, S);
  const bool wtiles = !wave && (tiles ? x_wtile_form(S->m) : x_wtile_pays(S->m, nf));
  // the one-wave sweep sums VxMed itself; the tiled ones take it from the pass
  const int vpre = (ctx->ph.fargo && !wave) ? x_vmed_form(S->m, nf, tiles || wtiles) : 0;
  // 2-D: the pass in the Y sweep's launch instead of on the second stream
  const bool vin = vpre == 1 && !S->m.dimz && g_y2v != 0 && vmed_terms(S->m.nx) == 8 &&
                   (size_t)S->m.nx * sizeof(double) <= 64 * 1024;
  VmTab VT;
  memset(&VT, 0, sizeof(VT));
  for (int q = 0; q < nf; q++) {
    VT.vxt[q] = f[q].Vx_temp;
    VT.vxmed[q] = t[q].vxmed;
  }
  if (vpre == 1 && !vin) {
    // VxMed of each fluid's Vx_temp (one long sequential sum per ring) on the
    // side stream, overlapping the Z/Y sweeps, which do not write Vx_temp
    if (side_stream(ctx)) return -1;
    F3D_CHECK(ctx, hipEventRecord(ctx->side_go, ctx->stream));
    F3D_CHECK(ctx, hipStreamWaitEvent(ctx->side, ctx->side_go, 0));
    if (x_vmed(ctx, S, nf, f, t, ctx->side)) return -1;
    F3D_CHECK(ctx, hipEventRecord(ctx->side_done, ctx->side));
  }
  if (yz_sweeps(ctx, S, dt, nf, f, t, vin ? &VT : nullptr)) return -1;
  if (vpre == 1 && !vin) F3D_CHECK(ctx, hipStreamWaitEvent(ctx->stream, ctx->side_done, 0));
  if (vpre == 2 && x_vmed(ctx, S, nf, f, t, ctx->stream)) return -1;
  if (wave) {
    if (x_sweep_wave(ctx, S, dt, nf, f, t)) return -1;
  } else if (wtiles) {
    if (x_sweep_wtiles(ctx, S, dt, nf, f, t)) return -1;
  } else if (!tiles) {
    if (x_sweep_ring(ctx, S, dt, nf, f, t, vpre != 0)) return -1;
  } else {
    for (int k = 0; k < nf; k++)
      if (x_sweep_tiled(ctx, S, dt, f[k], t[k])) return -1;
  }
  if (finish_step(ctx, S, dt, end_of_step, nf, f, t)) return -1;
  return 0;
}

extern "C" {

int F3D_API(fused_set_cflmode)(int mode) {
  if (mode < -1 || mode > 7 || (mode >= 4 && !(mode & 2))) return -2;
  g_cflmode = mode;
  return 0;
}

int F3D_API(fused_set_xsplit)(int split) {
  if (split < -1 || split > 1) return -2;
  g_xsplit = split;
  return 0;
}

int F3D_API(fused_set_colsplit)(int split) {
  if (split < -1 || split > 1) return -2;
  g_colsplit = split;
  return 0;
}

int F3D_API(fused_set_xvmed)(int form) {
  if (form < -1 || form > 2) return -2;
  g_xvmed = form;
  return 0;
}

int F3D_API(fused_set_xtile)(int cells) {
  if (cells != 64 && cells != 128 && cells != 256) return -2;
  g_xtile = cells;
  return 0;
}

int F3D_API(fused_set_xmode)(int mode) {
  if (mode < 0 || mode > 4) return -2;
  g_xmode = mode;
  return 0;
}

int F3D_API(fused_set_srcfuse)(int on) {
  if (on < -1 || on > 1) return -2;
  g_srcfuse = on;
  return 0;
}

int F3D_API(fused_set_y2v)(int on) {
  if (on < -1 || on > 1) return -2;
  g_y2v = on;
  return 0;
}

int F3D_API(fused_set_zfill)(int on) {
  if (on < -1 || on > 1) return -2;
  g_zfill = on;
  return 0;
}

int F3D_API(fused_set_vmedterms)(int terms) {
  if (terms != -1 && terms != 4 && terms != 8 && terms != 16) return -2;
  g_vmed_terms = terms;
  return 0;
}

int F3D_API(fused_set_tiling)(int rows_y, int rows_z) {
  if (rows_y < 0 || rows_z < 1) return -2;
  g_rows_y = rows_y;
  g_rows_z = rows_z;
  return 0;
}

int F3D_API(fused_cfl_vm)(f3d_hip_ctx* ctx, const double* vxm_in, double* fluid_min);

int F3D_API(fused_cfl)(f3d_hip_ctx* ctx, double* fluid_min) { return F3D_API(fused_cfl_vm)(ctx, nullptr, fluid_min); }

int F3D_API(fused_cfl_vm)(f3d_hip_ctx* ctx, const double* vxm_in, double* fluid_min) {
  const real* vxm = reinterpret_cast<const real*>(vxm_in);
  Scratch* S;
  F3D_REQUIRE(ctx, scratch_of(ctx, &S) == 0, "fused scratch allocation failed");
  const dmesh& m = S->m;
  const RF f = rfields(ctx);
  const int nrings = m.ny * (m.dimz ? m.nz : 1);
  const int slot = S->slot;
  f3d_stage_begin(ctx, F3D_ST_CFL);
  const int mode = cfl_mode(m, 1);
  double* rm = (mode & 1) ? S->ringmin : nullptr;
  const int pre = (mode & 2) ? 1 : 0;
  if (pre && vmed_prepass(ctx, S, vxm ? vxm : f.Vx)) return -1;
  const int wg = (mode & 4) ? 64 : WG;
#define CFL1(P_)                                                                                                 \
  FDISPATCH_T(ctx, kf_cfl, P_, dim3(nrings), dim3(wg), 16 * sizeof(double), f.Energy, f.Density,            \
              f.Vx, f.Vy, f.Vz, vxm ? vxm : f.Vx, f.VxMed, S->result, slot, cfl_par_f(ctx), rm)
  if (pre) CFL1(1);
  else CFL1(0);
#undef CFL1
  if (rm) {
    hipLaunchKernelGGL(kf_cfl_fin, dim3(1), dim3(1024), 0, ctx->stream, rm, nrings, S->result, slot, ctx->dhres,
                       0ull);
    F3D_LAUNCH_CHECK(ctx);
  }
  f3d_stage_end(ctx, F3D_ST_CFL);
  if (!rm)
    F3D_CHECK(ctx, hipMemcpyAsync(ctx->hres, S->result + slot, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  F3D_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  S->slot ^= 1;
  *fluid_min = ctx->hres[0];
  return 0;
}

// Launches the queued CFL's MIN pass if no potential launch has carried it.
static int cfl_fin_flush(f3d_hip_ctx* ctx) {
  if (!ctx->fin_pending) return 0;
  ctx->fin_pending = 0;
  const f3d_hip_ctx::CflFin& F = ctx->fin;
  hipLaunchKernelGGL(kf_cfl_fin, dim3(1), dim3(1024), 0, ctx->stream, F.ringmin, F.n, F.result, F.slot, ctx->dhres,
                     ctx->cfl_seq);
  F3D_LAUNCH_CHECK(ctx);
  return 0;
}

// cfl of nf fluids with one read-back: every launch takes its MIN into the
// same slot (MIN is order independent, so the result is CflFluidsMin's).
// The launches are queued here; fused_cfl_wait reads the step back.
int F3D_API(fused_cfl_fluids_async)(f3d_hip_ctx* ctx, int nf, double* const* cs, double* const* rho,
                                    double* const* vx, double* const* vy, double* const* vz, const double* vxm_in) {
  Scratch* S;
  F3D_REQUIRE(ctx, scratch_of(ctx, &S) == 0, "fused scratch allocation failed");
  F3D_REQUIRE(ctx, nf >= 1 && nf <= 8, "cfl: 1 to 8 fluids");
  if (cfl_fin_flush(ctx)) return -1;  // a CFL queued before and not waited for
  const dmesh& m = S->m;
  const RF f = rfields(ctx);
  const int nrings = m.ny * (m.dimz ? m.nz : 1);
  const int slot = S->slot;
  const real* vxm = reinterpret_cast<const real*>(vxm_in);
  auto R = [](double* p) { return reinterpret_cast<real*>(p); };
  f3d_stage_begin(ctx, F3D_ST_CFL);
  const int mode = cfl_mode(m, nf);
  const int pre = (mode & 2) ? 1 : 0;
  if (pre && vmed_prepass(ctx, S, vxm)) return -1;
  CflTab T;
  memset(&T, 0, sizeof(T));
  for (int k = 0; k < nf; k++) {
    T.cs[k] = R(cs[k]);
    T.rho[k] = R(rho[k]);
    T.vx[k] = R(vx[k]);
    T.vy[k] = R(vy[k]);
    T.vz[k] = m.dimz ? R(vz[k]) : R(vy[k]);
    T.ringmin[k] = (mode & 1) ? S->ringmin + (size_t)k * nrings : nullptr;
  }
  const int wg = (mode & 4) ? 64 : WG;
#define CFLF(P_)                                                                                             \
  FDISPATCH_T(ctx, kf_cfl_fluids, P_, dim3(nrings * nf), dim3(wg), 16 * sizeof(double), T,                  \
              nrings, vxm, f.VxMed, S->result, slot, cfl_par_f(ctx))
  // the flag set of the common configurations at compile time (FARGO:
  // isothermal with constant nu or alpha, adiabatic with constant nu)
  const cell::CflPar cpf = cfl_par_f(ctx);
  const int fx = !cpf.fargo ? 0 : cpf.iso ? (cpf.alpha_visc ? 2 : 1) : (cpf.alpha_visc ? 0 : 3);
  switch (fx * 2 + pre) {
    case 2: CFLF(2); break;
    case 3: CFLF(3); break;
    case 4: CFLF(4); break;
    case 5: CFLF(5); break;
    case 6: CFLF(6); break;
    case 7: CFLF(7); break;
    case 1: CFLF(1); break;
    default: CFLF(0); break;
  }
#undef CFLF
  ctx->cfl_spin = 0;
  if (mode & 1) {
    // the MIN pass rides on the next launch of the potential (which needs
    // no dt; f3d_Potential), else fused_cfl_wait launches it: one launch
    // fewer per step
    ctx->cfl_spin = 1;
    ++ctx->cfl_seq;
    ctx->fin = f3d_hip_ctx::CflFin{S->ringmin, nf * nrings, slot, S->result};
    ctx->fin_pending = 1;
  }
  f3d_stage_end(ctx, F3D_ST_CFL);
  // the atomic form copies the step to the host and records an event for
  // the wait; the two-pass form publishes it itself (an event record would
  // hold the next launch back by a few microseconds)
  if (!(mode & 1)) {
    F3D_CHECK(ctx, hipMemcpyAsync(ctx->hres, S->result + slot, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    if (!ctx->cfl_done) F3D_CHECK(ctx, hipEventCreateWithFlags(&ctx->cfl_done, hipEventDisableTiming));
    F3D_CHECK(ctx, hipEventRecord(ctx->cfl_done, ctx->stream));
  }
  ctx->cfl_queued = 1;
  S->slot ^= 1;
  return 0;
}

// The step queued by fused_cfl_fluids_async, once its launches are done
// (work queued behind them may still run).
int F3D_API(fused_cfl_wait)(f3d_hip_ctx* ctx, double* step_min) {
  F3D_REQUIRE(ctx, ctx->cfl_queued, "cfl_wait: no CFL queued");
  if (cfl_fin_flush(ctx)) return -1;
  ctx->cfl_queued = 0;
  if (ctx->cfl_spin) {
    // poll the sequence number the last CFL kernel writes after dt: the
    // value is there as soon as the kernel's stores land, before the
    // runtime has processed the kernel's completion.  Past two seconds the
    // wait goes to the stream, which also reports a failed launch.
    const volatile unsigned long long* flag = reinterpret_cast<volatile unsigned long long*>(ctx->hres + 1);
    const unsigned long long want = ctx->cfl_seq;
    const auto t0 = std::chrono::steady_clock::now();
    bool seen = true;
    for (long n = 1; *flag != want; n++) {
      if ((n & 0xFFFF) == 0 && std::chrono::steady_clock::now() - t0 > std::chrono::seconds(2)) {
        seen = false;
        break;
      }
    }
    if (!seen) F3D_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    std::atomic_thread_fence(std::memory_order_acquire);
    *step_min = ctx->hres[0];
    return 0;
  }
  F3D_CHECK(ctx, hipEventSynchronize(ctx->cfl_done));
  *step_min = ctx->hres[0];
  return 0;
}

int F3D_API(fused_cfl_fluids)(f3d_hip_ctx* ctx, int nf, double* const* cs, double* const* rho, double* const* vx,
                              double* const* vy, double* const* vz, const double* vxm_in, double* step_min) {
  if (F3D_API(fused_cfl_fluids_async)(ctx, nf, cs, rho, vx, vy, vz, vxm_in)) return -1;
  return F3D_API(fused_cfl_wait)(ctx, step_min);
}

int F3D_API(fused_potential)(f3d_hip_ctx* ctx) {
#if F3D_PREC == 32
  // the potential is an FP64 array in every form: its radial and vertical
  // differences balance the centrifugal terms, so it is computed in FP64
  f3d_stage_begin(ctx, F3D_ST_POTENTIAL);
  const int rc = f3d_Potential(ctx);
  f3d_stage_end(ctx, F3D_ST_POTENTIAL);
  return rc;
#else
  Scratch* S;
  F3D_REQUIRE(ctx, scratch_of(ctx, &S) == 0, "fused scratch allocation failed");
  F3D_REQUIRE(ctx, ctx->ph.nplanets <= cell::kMaxPlanets, "Potential: too many planets");
  const dmesh& m = S->m;
  const RF f = rfields(ctx);
  cell::PlanetSet ps;
  ps.n = ctx->ph.nplanets;
  ps.indirect = ctx->ph.indirectterm;
  for (int n = 0; n < ps.n; n++) {
    ps.x[n] = (real)ctx->ph.px[n];
    ps.y[n] = (real)ctx->ph.py[n];
    ps.z[n] = (real)ctx->ph.pz[n];
    ps.mp[n] = (real)ctx->ph.pmass[n];
    ps.smooth2[n] = (real)ctx->ph.psmooth2[n];
  }
  dim3 g((m.nx + BXF - 1) / BXF, m.sy, m.dimz ? m.sz : 1);
  f3d_stage_begin(ctx, F3D_ST_POTENTIAL);
  FDISPATCH(ctx, kf_potential, g, dim3(BXF), 0, f.Pot, ps);
  f3d_stage_end(ctx, F3D_ST_POTENTIAL);
  return 0;
#endif
}

int F3D_API(fused_init_stockholm)(f3d_hip_ctx* ctx) {
  Scratch* S;
  F3D_REQUIRE(ctx, scratch_of(ctx, &S) == 0, "fused scratch allocation failed");
  const dmesh& m = S->m;
  const RF f = rfields(ctx);
  const int n = m.sy * (m.dimz ? m.sz : 1);
  hipLaunchKernelGGL(kf_init_stockholm, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, m, f.Density, f.Vx, f.Vy,
                     f.Vz, f.Energy, f.Density0, f.Vx0, f.Vy0, f.Vz0, f.Energy0, (int)(ctx->ph.eos == 2));
  F3D_LAUNCH_CHECK(ctx);
  return 0;
}

int F3D_API(fused_boundaries)(f3d_hip_ctx* ctx) {
  Scratch* S;
  F3D_REQUIRE(ctx, scratch_of(ctx, &S) == 0, "fused scratch allocation failed");
  const RF f = rfields(ctx);
  return fused_boundaries(ctx, S, f.Density, f.Vx, f.Vy, f.Vz, f.Energy);
}

// Implicit drag of all fluids (Collisions(dt, 1), src/collisions.c:156-189),
// V_temp in place; the arrays hold each fluid's device pointers.
int F3D_API(fused_collisions)(f3d_hip_ctx* ctx, double dt, int nf, double* const* rho, double* const* vxt,
                              double* const* vyt, double* const* vzt, const double* alpha) {
  Scratch* S;
  F3D_REQUIRE(ctx, scratch_of(ctx, &S) == 0, "fused scratch allocation failed");
  F3D_REQUIRE(ctx, nf >= 2 && nf <= 8, "collisions: 2 to 8 fluids");
  const dmesh& m = S->m;
  FluidSet F;
  memset(&F, 0, sizeof(F));
  F.nf = nf;
  for (int f = 0; f < nf; f++) {
    F.rho[f] = reinterpret_cast<const real*>(rho[f]);
    F.vx[f] = reinterpret_cast<real*>(vxt[f]);
    F.vy[f] = reinterpret_cast<real*>(vyt[f]);
    F.vz[f] = reinterpret_cast<real*>(m.dimz ? vzt[f] : vyt[f]);
  }
  for (int a = 0; a < nf * nf; a++) F.alpha[a] = alpha[a];
  dim3 g((m.nx + BXF - 1) / BXF, m.sy - 1, m.dimz ? m.sz - 1 : 1);
  const int dirs = m.dimz ? 7 : 3;
#define COLL(NF_)                                                                                  \
  case NF_:                                                                                        \
    if (m.geom == SPH) {                                                                           \
      if (m.dimz) hipLaunchKernelGGL((kf_collisions<SPH, 1, NF_>), g, dim3(BXF), 0, ctx->stream, m, F, dt, dirs); \
      else hipLaunchKernelGGL((kf_collisions<SPH, 0, NF_>), g, dim3(BXF), 0, ctx->stream, m, F, dt, dirs);        \
    } else {                                                                                       \
      if (m.dimz) hipLaunchKernelGGL((kf_collisions<CYL, 1, NF_>), g, dim3(BXF), 0, ctx->stream, m, F, dt, dirs); \
      else hipLaunchKernelGGL((kf_collisions<CYL, 0, NF_>), g, dim3(BXF), 0, ctx->stream, m, F, dt, dirs);        \
    }                                                                                              \
    break;
  switch (nf) {
    COLL(2) COLL(3) COLL(4) COLL(5) COLL(6) COLL(7) COLL(8)
  }
#undef COLL
  F3D_LAUNCH_CHECK(ctx);
  return 0;
}

// DustDiffusion_Main (src/dust_diffusion_main.c:3-28): coefficients from the
// gas (fluid 0), then for each dust fluid the core, the copy of the result
// into its density (copy_field, src/copy_field.c:3-12) and its density ghost
// fill.  2-D cylindrical, isothermal gas.
int F3D_API(fused_dust_diffusion)(f3d_hip_ctx* ctx, double dt, int nf, double* const* rho, const int* type,
                                  const double* cs_gas) {
  Scratch* S;
  F3D_REQUIRE(ctx, scratch_of(ctx, &S) == 0, "fused scratch allocation failed");
  const dmesh& m = S->m;
  F3D_REQUIRE(ctx, !m.dimz && m.geom == CYL, "dust diffusion: 2-D cylindrical meshes");
  F3D_REQUIRE(ctx, type[0] == 1, "dust diffusion: fluid 0 must be the gas");
  real* yczc = S->f[FS_W1];
  real* yfzc = S->f[FS_W2];
  hipLaunchKernelGGL((kf_ddiff_coef<CYL, 0>), dim3((m.nx + BXF - 1) / BXF, m.sy), dim3(BXF), 0, ctx->stream, m,
                     reinterpret_cast<const real*>(cs_gas), yczc, yfzc, (int)(ctx->ph.visc == 2),
                     (real)ctx->ph.alpha, (real)ctx->ph.nu);
  F3D_LAUNCH_CHECK(ctx);
  const cell::BcPar b0 = bc_par_f(ctx, 0), b1 = bc_par_f(ctx, 1);
  // every dust fluid at once, each with a Temp of its own
  F3D_REQUIRE(ctx, nf >= 1 && nf <= 8, "dust diffusion: 1 to 8 fluids");
  DdTab T;
  memset(&T, 0, sizeof(T));
  int nd = 0;
  for (int f = 1; f < nf; f++) {
    if (type[f] != 2) continue;
    if (!S->ddtemp[f] && hipMalloc(&S->ddtemp[f], ctx->n3 * sizeof(real)) != hipSuccess) {
      S->ddtemp[f] = nullptr;
      F3D_REQUIRE(ctx, false, "dust diffusion: scratch allocation failed");
    }
    T.rhod[nd] = reinterpret_cast<real*>(rho[f]);
    T.temp[nd] = S->ddtemp[f];
    nd++;
  }
  if (!nd) return 0;
  hipLaunchKernelGGL((kf_ddiff_core_fluids<CYL, 0>), dim3((m.nx + BXF - 1) / BXF, m.sy - 2, nd), dim3(BXF), 0,
                     ctx->stream, m, (real)dt, T, reinterpret_cast<const real*>(rho[0]), yczc, yfzc);
  F3D_LAUNCH_CHECK(ctx);
  if (!ctx->ph.periodicy) {
    // Temp -> Density and FillGhosts(DENS) in one pass
    hipLaunchKernelGGL(kf_ddiff_store, dim3((m.nx + BXF - 1) / BXF, m.sy, nd), dim3(BXF), 0, ctx->stream, m, T, b0,
                       b1);
    F3D_LAUNCH_CHECK(ctx);
  } else {
    for (int q = 0; q < nd; q++)
      F3D_CHECK(ctx, hipMemcpyAsync(T.rhod[q], T.temp[q], ctx->n3 * sizeof(real), hipMemcpyDeviceToDevice,
                                    ctx->stream));
  }
  return 0;
}

int F3D_API(fused_sources)(f3d_hip_ctx* ctx, double dt, double domega) {
  Scratch* S;
  F3D_REQUIRE(ctx, scratch_of(ctx, &S) == 0, "fused scratch allocation failed");
  const dmesh& m = S->m;
  const RF f = rfields(ctx);
  cell::ViscPar vp;
  vp.model = ctx->ph.visc;
  vp.nu = (real)ctx->ph.nu;
  vp.alpha = (real)ctx->ph.alpha;
  vp.gamma = (real)ctx->ph.gamma;
  const int iso = (ctx->ph.eos != 2);
  vp.iso = iso;
  const int visc = ctx->ph.visc != 0 && ctx->ph.fluidtype == 1;
  const int pot = ctx->ph.potential;
  const real rdt = (real)dt, rdom = (real)domega;
  const int vm = vp.model == 1 ? 1 : vp.model == 2 ? (iso ? 2 : 3) : 0;
  const int dust = (ctx->ph.fluidtype == 2);
  // adiabatic: SubStep3 reads V_temp one row (and plane) past the active
  // range, which the reference computes there with the same formulas
  // (src/substep1_y.c:78-80, src/substep2_b.c:84-86), so those are included
  const int ext = iso ? 0 : 1;
  if (!m.dimz && visc && pot && !dust && vm != 0 && g_srcfuse != 0) {
    // one pass without the stress arrays (kf_src_fused2); rows per lane
    // halved from 8 while the launch has fewer than 4096 waves
    const int wx = (m.nx + kSrcCols - 1) / kSrcCols, rows = m.ny + ext;
    int rp = 8;
    while (rp > 1 && (long)wx * ((rows + rp - 1) / rp) < 4096) rp /= 2;
    dim3 g(wx, (rows + rp - 1) / rp);
    f3d_stage_begin(ctx, F3D_ST_SRC_B);
#define SRCF(G_, MODE_, VM_)                                                                                    \
  hipLaunchKernelGGL((kf_src_fused2<G_, MODE_, VM_>), g, dim3(64), 0, ctx->stream, m, rdt, f.Density, f.Energy, \
                     f.Vx, f.Vy, ctx->f.Pot, f.Vx_temp, f.Vy_temp, S->f[FS_E2], vp, (real)(ctx->ph.gamma - 1.0), \
                     ctx->ph.OMEGAFRAME, rdom, rp, m.nghy + rows)
#define SRCF_G(G_)                                                        \
  do {                                                                    \
    if (iso) { if (vm == 1) SRCF(G_, 1, 1); else SRCF(G_, 1, 2); }       \
    else { if (vm == 1) SRCF(G_, 2, 1); else SRCF(G_, 2, 3); }           \
  } while (0)
    if (m.geom == SPH) SRCF_G(SPH); else SRCF_G(CYL);
#undef SRCF_G
#undef SRCF
    F3D_LAUNCH_CHECK(ctx);
    f3d_stage_end(ctx, F3D_ST_SRC_B);
  } else {
  dim3 ga((m.nx + BXF - 1) / BXF, m.sy - 2, m.dimz ? m.sz - 2 : 1);
  if (visc) {
    f3d_stage_begin(ctx, F3D_ST_SRC_A);
#define SRCA(VM_) FDISPATCH_T(ctx, kf_src_a, VM_, ga, dim3(BXF), 0, f.Density, f.Energy, f.Vx, f.Vy, f.Vz, f.Mmx, \
                              f.Mmy, f.Mmz, f.Mpx, f.Mpy, f.Mpz, vp, visc, rdom)
    if (vm == 1) SRCA(1); else if (vm == 2) SRCA(2); else if (vm == 3) SRCA(3); else SRCA(0);
#undef SRCA
    f3d_stage_end(ctx, F3D_ST_SRC_A);
  }
  dim3 gb((m.nx + BXF - 1) / BXF, m.ny + ext, m.dimz ? m.nz + ext : 1);
  f3d_stage_begin(ctx, F3D_ST_SRC_B);
  // measured on MI355X: the fixed flags pay in 2-D (adiabatic 4096x1024 step
  // 0.662 -> 0.645 ms); in 3-D the kernel took longer (34.5 -> 38.8 us)
  const int mode = (visc && pot && !dust && !m.dimz) ? (iso ? 1 : 2) : 0;
#define SRCB(MODE_) FDISPATCH_T(ctx, kf_src_b, MODE_, gb, dim3(BXF), 0, rdt, f.Density, f.Energy, f.Vx, f.Vy, f.Vz, \
                                ctx->f.Pot, f.Mmx, f.Mmy, f.Mmz, f.Mpx, f.Mpy, f.Mpz, f.Vx_temp, f.Vy_temp,        \
                                f.Vz_temp, S->f[FS_E2], visc, pot, dust, iso, (real)(ctx->ph.gamma - 1.0),        \
                                ctx->ph.OMEGAFRAME, rdom)
  if (mode == 1) SRCB(1); else if (mode == 2) SRCB(2); else SRCB(0);
#undef SRCB
  f3d_stage_end(ctx, F3D_ST_SRC_B);
  }
  if (!iso && ctx->ph.fluidtype == 1) {
    // SubStep3 (src/algogas.c:110-112)
    dim3 g3((m.nx + BXF - 1) / BXF, m.ny, m.dimz ? m.nz : 1);
    f3d_stage_begin(ctx, F3D_ST_SUBSTEP3);
    FDISPATCH(ctx, kf_substep3, g3, dim3(BXF), 0, rdt, S->f[FS_E2], f.Vx_temp, f.Vy_temp, f.Vz_temp, f.Energy,
              (real)ctx->ph.gamma);
    f3d_stage_end(ctx, F3D_ST_SUBSTEP3);
  }
  // mid-step ghost fill (src/algogas.c:116-118) applied to V_temp directly
  f3d_stage_begin(ctx, F3D_ST_BC_MID);
  if (fused_boundaries(ctx, S, f.Density, f.Vx_temp, f.Vy_temp, f.Vz_temp, f.Energy)) return -1;
  f3d_stage_end(ctx, F3D_ST_BC_MID);
  return 0;
}

// Sources of every fluid (isothermal runs, at most one viscous fluid): the
// viscous fluid's stresses, one kf_src_b launch over all fluids, the
// mid-step ghost fill (2-D: one launch over all fluids).
int F3D_API(fused_sources_fluids)(f3d_hip_ctx* ctx, double dt, double domega, int nf, const f3d_fields* fl,
                                  const int* type) {
  Scratch* S;
  F3D_REQUIRE(ctx, scratch_of(ctx, &S) == 0, "fused scratch allocation failed");
  F3D_REQUIRE(ctx, nf >= 1 && nf <= 8 && fl && type, "sources: 1 to 8 fluids");
  F3D_REQUIRE(ctx, ctx->ph.eos != 2, "sources of several fluids in one launch: isothermal runs");
  const dmesh& m = S->m;
  RF f[8];
  int nvisc = 0, vq = -1;
  for (int q = 0; q < nf; q++) {
    f[q] = rfields_of(fl[q], ctx->f);
    if (ctx->ph.visc != 0 && type[q] == 1) {
      nvisc++;
      vq = q;
    }
  }
  F3D_REQUIRE(ctx, nvisc <= 1, "sources of several fluids in one launch: one viscous fluid at most");
  cell::ViscPar vp;
  vp.model = ctx->ph.visc;
  vp.nu = (real)ctx->ph.nu;
  vp.alpha = (real)ctx->ph.alpha;
  vp.gamma = (real)ctx->ph.gamma;
  vp.iso = 1;
  const real rdt = (real)dt, rdom = (real)domega;
  // the stresses land in the shared stress arrays, which only the viscous fluid reads
  const RF& sh = f[vq < 0 ? 0 : vq];
  if (vq >= 0) {
    dim3 ga((m.nx + BXF - 1) / BXF, m.sy - 2, m.dimz ? m.sz - 2 : 1);
    f3d_stage_begin(ctx, F3D_ST_SRC_A);
    // isothermal runs: the model alone selects the form
#define SRCA(VM_) FDISPATCH_T(ctx, kf_src_a, VM_, ga, dim3(BXF), 0, sh.Density, sh.Energy, sh.Vx, sh.Vy, sh.Vz, \
                              sh.Mmx, sh.Mmy, sh.Mmz, sh.Mpx, sh.Mpy, sh.Mpz, vp, 1, rdom)
    if (vp.model == 1) SRCA(1); else if (vp.model == 2) SRCA(2); else SRCA(0);
#undef SRCA
    f3d_stage_end(ctx, F3D_ST_SRC_A);
  }
  SrcTab T;
  memset(&T, 0, sizeof(T));
  for (int q = 0; q < nf; q++) {
    T.rho[q] = f[q].Density; T.cs[q] = f[q].Energy; T.vx[q] = f[q].Vx; T.vy[q] = f[q].Vy; T.vz[q] = f[q].Vz;
    T.vxt[q] = f[q].Vx_temp; T.vyt[q] = f[q].Vy_temp; T.vzt[q] = f[q].Vz_temp;
    T.visc[q] = (q == vq);
    T.dust[q] = (type[q] == 2);
  }
  const int nzb = m.dimz ? m.nz : 1;
  dim3 gb((m.nx + BXF - 1) / BXF, m.ny, nzb * nf);
  f3d_stage_begin(ctx, F3D_ST_SRC_B);
  FDISPATCH(ctx, kf_src_b_fluids, gb, dim3(BXF), 0, rdt, T, nzb, ctx->f.Pot, sh.Mmx, sh.Mmy, sh.Mmz, sh.Mpx, sh.Mpy,
            sh.Mpz, ctx->ph.potential, (real)(ctx->ph.gamma - 1.0), ctx->ph.OMEGAFRAME, rdom);
  f3d_stage_end(ctx, F3D_ST_SRC_B);
  // mid-step ghost fill (src/algogas.c:116-118) on every fluid's V_temp
  f3d_stage_begin(ctx, F3D_ST_BC_MID);
  if (nf > 1 && !m.dimz && !ctx->ph.periodicy && m.ny > m.nghy) {
    BTab B;
    memset(&B, 0, sizeof(B));
    for (int q = 0; q < nf; q++) {
      B.dens[q] = f[q].Density; B.vx[q] = f[q].Vx_temp; B.vy[q] = f[q].Vy_temp; B.e[q] = f[q].Energy;
    }
    hipLaunchKernelGGL(kf_bound_y_fluids, dim3((m.nx + 63) / 64, 2, nf), dim3(64), 0, ctx->stream, m, B,
                       bc_par_f(ctx, 0), bc_par_f(ctx, 1), S->ymin_dexp, S->ymax_dexp);
    F3D_LAUNCH_CHECK(ctx);
  } else {
    for (int q = 0; q < nf; q++)
      if (fused_boundaries(ctx, S, f[q].Density, f[q].Vx_temp, f[q].Vy_temp, f[q].Vz_temp, f[q].Energy)) return -1;
  }
  f3d_stage_end(ctx, F3D_ST_BC_MID);
  return 0;
}

int F3D_API(fused_transport)(f3d_hip_ctx* ctx, double dt, int end_of_step) {
  return transport_fluids(ctx, dt, 1, nullptr, end_of_step);
}

int F3D_API(fused_transport_fluids)(f3d_hip_ctx* ctx, double dt, int nf, const f3d_fields* fl) {
  F3D_REQUIRE(ctx, fl != nullptr, "transport: no fluid fields");
  return transport_fluids(ctx, dt, nf, fl, 1);
}

}  // extern "C"

}  // namespace fused
}  // namespace F3D_NS

#if F3D_PREC == 32
#include "f3d_batch_body.h"
#endif
