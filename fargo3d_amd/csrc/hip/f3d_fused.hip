// f3d_fused.hip -- the FP64 build of the fused step (entry points f3d_fused_*):
// bit-for-bit the reference's arithmetic.  The kernels are in f3d_fused_body.h.
#include <cstring>
#include <type_traits>

#include "f3d_internal.h"
#include "f3d_cell.h"

#define F3D_NS f3d
#define F3D_PREC 64
#define F3D_API(name) f3d_##name
#include "f3d_fused_body.h"

// ---------------------------------------------------------------------------
// The CFL time step with no allocation.  f3d_fused_cfl_fluids keeps its
// per-ring minima and result slots in scratch that the handle allocates on
// its first fused call; this variant keeps them in a fixed device array
// instead, for callers that cannot allocate (code under stream capture,
// serving loops, harnesses that forbid hipMalloc).  The array holds the
// per-ring minima of at most kNoAllocRingMins rings x fluids and the two
// result slots.  Every call on one GPU shares it and it keeps the last
// call's contents: the kernels write every ring's minimum before the final
// pass reads it, so nothing needs clearing, but calls that overlap on one
// GPU (two host threads, two streams) race on it.
#include <mutex>

static constexpr size_t kNoAllocRingMins = (size_t)1 << 18;  // 2 MiB of ring minima
static __device__ __align__(256) double scratch[kNoAllocRingMins + 2];

namespace {
std::mutex g_scratch_mu;
double* g_scratch_dev[64] = {};  // the array's device address on each GPU

// The device address of `scratch` on the handle's GPU, looked up on the first
// call on that GPU with that GPU current (each GPU holds its own copy).
int noalloc_scratch(f3d_hip_ctx* ctx, double** out) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  const int dev = ctx->device;
  if (dev < 0 || dev >= 64) return -1;
  if (!g_scratch_dev[dev]) {
    int cur = -1;
    if (hipGetDevice(&cur) != hipSuccess) return -1;
    if (cur != dev && hipSetDevice(dev) != hipSuccess) return -1;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != dev) (void)hipSetDevice(cur);
    if (e != hipSuccess || !p) return -1;
    g_scratch_dev[dev] = static_cast<double*>(p);
  }
  *out = g_scratch_dev[dev];
  return 0;
}
}  // namespace

namespace f3d {
namespace fused {

extern "C" int f3d_fused_cfl_fluids_noalloc(f3d_hip_ctx* ctx, int nf, double* const* cs, double* const* rho,
                                            double* const* vx, double* const* vy, double* const* vz,
                                            const double* vxm, double* step_min) {
  F3D_REQUIRE(ctx, nf >= 1 && nf <= 8, "cfl_noalloc: 1 to 8 fluids");
  F3D_REQUIRE(ctx, vxm != nullptr && step_min != nullptr, "cfl_noalloc: no VxMed velocity or result");
  const dmesh& m = ctx->m;
  const int nrings = m.ny * (m.dimz ? m.nz : 1);
  if ((size_t)nf * (size_t)nrings > kNoAllocRingMins) {
    snprintf(ctx->err, sizeof(ctx->err),
             "cfl_noalloc: %d fluids x %d rings exceed the fixed scratch of %zu ring minima", nf, nrings,
             kNoAllocRingMins);
    return -2;
  }
  double* sp = nullptr;
  F3D_REQUIRE(ctx, noalloc_scratch(ctx, &sp) == 0, "cfl_noalloc: hipGetSymbolAddress of the scratch array failed");
  unsigned long long* result = reinterpret_cast<unsigned long long*>(sp + kNoAllocRingMins);
  // the two-pass MIN: ring minima into the array, one workgroup takes their MIN
  const int mode = cfl_mode(m, nf) | 1;
  const int pre = (mode & 2) ? 1 : 0;
  if (pre && vmed_prepass_m(ctx, m, vxm)) return -1;
  CflTab T;
  memset(&T, 0, sizeof(T));
  for (int k = 0; k < nf; k++) {
    T.cs[k] = cs[k];
    T.rho[k] = rho[k];
    T.vx[k] = vx[k];
    T.vy[k] = vy[k];
    T.vz[k] = m.dimz ? vz[k] : vy[k];
    T.ringmin[k] = sp + (size_t)k * nrings;
  }
  const int wg = (mode & 4) ? 64 : WG;
  const size_t shmem = 16 * sizeof(double);
  if (pre)
    FDISPATCH_T(ctx, kf_cfl_fluids, 1, dim3(nrings * nf), dim3(wg), shmem, T, nrings, vxm, ctx->f.VxMed, result, 0,
                cfl_par_f(ctx));
  else
    FDISPATCH_T(ctx, kf_cfl_fluids, 0, dim3(nrings * nf), dim3(wg), shmem, T, nrings, vxm, ctx->f.VxMed, result, 0,
                cfl_par_f(ctx));
  hipLaunchKernelGGL(kf_cfl_fin, dim3(1), dim3(1024), 0, ctx->stream, sp, nf * nrings, result, 0, ctx->dhres, 0ull);
  F3D_LAUNCH_CHECK(ctx);
  F3D_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  *step_min = ctx->hres[0];
  return 0;
}

}  // namespace fused
}  // namespace f3d
