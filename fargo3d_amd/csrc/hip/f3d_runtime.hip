// f3d_runtime.hip -- handle lifetime, mesh upload, field binding, memory helpers.
#include <cstdlib>
#include <cstring>
#include <vector>

#include "f3d_internal.h"

static thread_local char g_create_err[512];

extern "C" {

int f3d_hip_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

int f3d_hip_create(const f3d_mesh_desc* d, f3d_hip_ctx** out) {
  *out = nullptr;
  if (!d || d->nx < 1 || d->ny < 1 || d->nz < 1) {
    snprintf(g_create_err, sizeof(g_create_err), "f3d_hip_create: bad mesh description");
    return -2;
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
    snprintf(g_create_err, sizeof(g_create_err), "f3d_hip_create: no HIP device is visible");
    return -3;
  }
  // spin while the host waits on the stream: the step's one dt read-back is
  // on the critical path (harmless if the device is already initialised)
  if (hipSetDeviceFlags(hipDeviceScheduleSpin) != hipSuccess) (void)hipGetLastError();
  f3d_hip_ctx* c = new f3d_hip_ctx();
  hipGetDevice(&c->device);
  const int sx = d->nx, sy = d->ny + 2 * d->nghy, sz = d->nz + 2 * d->nghz;
  f3d::dmesh& m = c->m;
  m.nx = d->nx; m.ny = d->ny; m.nz = d->nz; m.nghy = d->nghy; m.nghz = d->nghz;
  m.dimz = d->dimz; m.geom = d->geometry;
  m.sx = sx; m.sy = sy; m.sz = sz; m.pitch = d->pitch; m.stride = d->stride;
  m.XMIN = d->XMIN; m.XMAX = d->XMAX; m.mindx = d->mindx;
  if (m.pitch < sx || m.stride < m.pitch * sy) {
    snprintf(g_create_err, sizeof(g_create_err), "f3d_hip_create: pitch/stride smaller than the mesh");
    delete c;
    return -2;
  }
  c->n3 = (size_t)m.stride * sz;
  c->n2 = (size_t)sy * sz;
  struct Arr { const double* src; size_t n; const double** dst; };
  Arr arrs[] = {
      {d->xmin, (size_t)sx + 1, &m.xmin}, {d->ymin, (size_t)sy + 1, &m.ymin}, {d->zmin, (size_t)sz + 1, &m.zmin},
      {d->xmed, (size_t)sx, &m.xmed}, {d->ymed, (size_t)sy, &m.ymed}, {d->zmed, (size_t)sz, &m.zmed},
      {d->sxi, (size_t)sx, &m.sxi}, {d->sxj, (size_t)sy, &m.sxj}, {d->syj, (size_t)sy, &m.syj},
      {d->szj, (size_t)sy, &m.szj}, {d->sxk, (size_t)sz, &m.sxk}, {d->syk, (size_t)sz, &m.syk},
      {d->szk, (size_t)sz, &m.szk}, {d->invvj, (size_t)sy, &m.invvj}, {d->invdiffxmed, (size_t)sx, &m.invdiffxmed},
      {d->cosxmed, (size_t)sx, &m.cosxmed}, {d->sinxmed, (size_t)sx, &m.sinxmed},
      {d->sinzmed, (size_t)sz, &m.sinzmed}, {d->coszmed, (size_t)sz, &m.coszmed},
      {d->sinzmin, (size_t)sz + 1, &m.sinzmin}, {d->coszmin, (size_t)sz + 1, &m.coszmin}};
  const int na = sizeof(arrs) / sizeof(arrs[0]);
  size_t total = 0;
  std::vector<size_t> off(na);
  for (int a = 0; a < na; a++) {
    off[a] = total;
    total += (arrs[a].n + 7) & ~(size_t)7;  // 64-byte aligned slices
  }
  std::vector<double> host(total, 0.0);
  for (int a = 0; a < na; a++)
    if (arrs[a].src) memcpy(&host[off[a]], arrs[a].src, arrs[a].n * sizeof(double));
  hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipMalloc(&c->meshbuf, total * sizeof(double));
  if (e == hipSuccess) e = hipMemcpy(c->meshbuf, host.data(), total * sizeof(double), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMalloc(&c->dred, (c->n2 + 64) * sizeof(double));
  if (e == hipSuccess) e = hipHostMalloc((void**)&c->hres, 64 * sizeof(double), hipHostMallocDefault);
  if (e == hipSuccess) e = hipHostGetDevicePointer((void**)&c->dhres, c->hres, 0);
  if (e != hipSuccess) {
    snprintf(g_create_err, sizeof(g_create_err), "f3d_hip_create: %s", hipGetErrorString(e));
    f3d_hip_destroy(c);
    return -1;
  }
  for (int a = 0; a < na; a++) *arrs[a].dst = c->meshbuf + off[a];
  // the FP64 parts read these; in the FP64 build they are the same arrays
  m.xmin64 = m.xmin; m.ymin64 = m.ymin; m.zmin64 = m.zmin;
  m.ymed64 = m.ymed; m.sinzmed64 = m.sinzmed; m.sxi64 = m.sxi;
  m.zmed64 = m.zmed; m.sinzmin64 = m.sinzmin; m.coszmin64 = m.coszmin;
  m.vb = nullptr;  // the FP64 form stores full azimuthal velocities
  c->meshlen = total;
  *out = c;
  return 0;
}

static const char* kStageNames[F3D_NSTAGES] = {"cfl", "sources_a", "sources_b", "ghosts_mid", "sweep_z",
                                                "sweep_y", "sweep_x", "finish", "ghosts_end", "potential",
                                                "substep3"};

int f3d_hip_stage_count(void) { return F3D_NSTAGES; }
const char* f3d_hip_stage_name(int s) { return (s >= 0 && s < F3D_NSTAGES) ? kStageNames[s] : ""; }

int f3d_hip_timing_enable(f3d_hip_ctx* c, int on) { return f3d_hip_timing_mask(c, on ? ~0u : 0u); }

int f3d_hip_timing_mask(f3d_hip_ctx* c, unsigned mask) {
  if (mask && !c->ev[0][0]) {
    for (int s = 0; s < F3D_NSTAGES; s++)
      for (int e = 0; e < 2; e++) F3D_CHECK(c, hipEventCreate(&c->ev[s][e]));
  }
  F3D_CHECK(c, hipStreamSynchronize(c->stream));
  for (int s = 0; s < F3D_NSTAGES; s++) {
    c->pending[s] = 0;
    c->stage_ms[s] = 0.0;
    c->stage_n[s] = 0;
    c->stage_calls[s] = 0;
    c->skipped[s] = 0;
  }
  c->timing = mask & ((1u << F3D_NSTAGES) - 1u);
  return 0;
}

int f3d_hip_timing_every(f3d_hip_ctx* c, int n) {
  if (n < 1) return -2;
  c->timing_every = n;
  for (int s = 0; s < F3D_NSTAGES; s++) c->stage_calls[s] = 0;
  return 0;
}

int f3d_hip_timing_read(f3d_hip_ctx* c, double* ms, long* counts) {
  F3D_CHECK(c, hipStreamSynchronize(c->stream));
  for (int s = 0; s < F3D_NSTAGES; s++) {
    if (c->pending[s]) {
      float t = 0.f;
      if (hipEventElapsedTime(&t, c->ev[s][0], c->ev[s][1]) == hipSuccess) {
        c->stage_ms[s] += t;
        c->stage_n[s] += 1;
      }
      c->pending[s] = 0;
    }
    ms[s] = c->stage_ms[s];
    counts[s] = c->stage_n[s];
  }
  return 0;
}

void f3d_hip_destroy(f3d_hip_ctx* c) {
  if (!c) return;
  if (c->stream) hipStreamSynchronize(c->stream);
  for (int s = 0; s < F3D_NSTAGES; s++)
    for (int e = 0; e < 2; e++)
      if (c->ev[s][e]) hipEventDestroy(c->ev[s][e]);
  for (int b = 0; b < 2; b++)
    if (c->fused[b] && c->fused_free[b]) c->fused_free[b](c->fused[b]);
  if (c->meshbuf) hipFree(c->meshbuf);
  if (c->bctab) hipFree(c->bctab);
  if (c->vbase) hipFree(c->vbase);
  if (c->dred) hipFree(c->dred);
  if (c->hres) hipHostFree(c->hres);
  if (c->side) hipStreamSynchronize(c->side);
  if (c->side_go) hipEventDestroy(c->side_go);
  if (c->side_done) hipEventDestroy(c->side_done);
  if (c->cfl_done) hipEventDestroy(c->cfl_done);
  if (c->side) hipStreamDestroy(c->side);
  if (c->stream) hipStreamDestroy(c->stream);
  delete c;
}

int f3d_hip_bind(f3d_hip_ctx* c, const f3d_fields* f) {
  if (!c || !f) return -2;
  c->f = *f;
  return 0;
}

int f3d_hip_set_vbase(f3d_hip_ctx* ctx, const double* host_v0) {
  if (!ctx || !host_v0) return -1;
  const size_t bytes = (size_t)ctx->m.sy * (size_t)ctx->m.sz * sizeof(double);
  if (!ctx->vbase && hipMalloc(&ctx->vbase, bytes) != hipSuccess) {
    snprintf(ctx->err, sizeof(ctx->err), "f3d_hip_set_vbase: device allocation failed");
    return -1;
  }
  if (hipMemcpy(ctx->vbase, host_v0, bytes, hipMemcpyHostToDevice) != hipSuccess) {
    snprintf(ctx->err, sizeof(ctx->err), "f3d_hip_set_vbase: copy failed");
    return -1;
  }
  return 0;
}

int f3d_hip_set_phys(f3d_hip_ctx* c, const f3d_phys* p) {
  if (!c || !p) return -2;
  const bool had_tables = c->bctab != nullptr;
  c->ph = *p;
  c->ph.ymin_dexp = c->ph.ymax_dexp = c->ph.zmin_dens = c->ph.zmax_dens = nullptr;
  // the per-(m,k) / per-(m,j) boundary factors go to the device once
  if (!had_tables && p->ymin_dexp && p->ymax_dexp && p->zmin_dens && p->zmax_dens) {
    const size_t ny3 = 3 * (size_t)c->m.sz, nz3 = 3 * (size_t)c->m.sy;
    std::vector<double> h(2 * ny3 + 2 * nz3);
    memcpy(&h[0], p->ymin_dexp, ny3 * sizeof(double));
    memcpy(&h[ny3], p->ymax_dexp, ny3 * sizeof(double));
    memcpy(&h[2 * ny3], p->zmin_dens, nz3 * sizeof(double));
    memcpy(&h[2 * ny3 + nz3], p->zmax_dens, nz3 * sizeof(double));
    F3D_CHECK(c, hipMalloc(&c->bctab, h.size() * sizeof(double)));
    F3D_CHECK(c, hipMemcpy(c->bctab, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
    c->d_ymin_dexp = c->bctab;
    c->d_ymax_dexp = c->bctab + ny3;
    c->d_zmin_dens = c->bctab + 2 * ny3;
    c->d_zmax_dens = c->bctab + 2 * ny3 + nz3;
  }
  return 0;
}

int f3d_hip_sync(f3d_hip_ctx* c) {
  F3D_CHECK(c, hipStreamSynchronize(c->stream));
  return 0;
}

const char* f3d_hip_last_error(f3d_hip_ctx* c) { return c ? c->err : g_create_err; }

int f3d_hip_malloc(void** ptr, size_t bytes) {
  if (hipMalloc(ptr, bytes ? bytes : 8) != hipSuccess) return -1;
  return 0;
}

int f3d_hip_free(void* ptr) { return hipFree(ptr) == hipSuccess ? 0 : -1; }

int f3d_hip_memcpy_h2d(f3d_hip_ctx* c, void* dst, const void* src, size_t bytes) {
  F3D_CHECK(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
  F3D_CHECK(c, hipStreamSynchronize(c->stream));
  return 0;
}

int f3d_hip_memcpy_d2h(f3d_hip_ctx* c, void* dst, const void* src, size_t bytes) {
  F3D_CHECK(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
  F3D_CHECK(c, hipStreamSynchronize(c->stream));
  return 0;
}

int f3d_hip_memcpy_d2d(f3d_hip_ctx* c, void* dst, const void* src, size_t bytes) {
  F3D_CHECK(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, c->stream));
  return 0;
}

int f3d_hip_memset(f3d_hip_ctx* c, void* dst, int value, size_t bytes) {
  F3D_CHECK(c, hipMemsetAsync(dst, value, bytes, c->stream));
  return 0;
}

}  // extern "C"
