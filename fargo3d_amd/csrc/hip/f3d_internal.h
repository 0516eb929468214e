// f3d_internal.h -- the handle behind the C ABI and launch helpers.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdio>

#include "../../../include/fargo3d_hip.h"
#include "f3d_device.h"

struct f3d_hip_ctx {
  hipStream_t stream = nullptr;
  int device = 0;
  f3d::dmesh m{};
  double* meshbuf = nullptr;        // one device buffer for all 1-D mesh arrays
  f3d_fields f{};                   // bound device buffers (selected fluid + scratch)
  f3d_phys ph{};                    // host copy of the run parameters
  double* bctab = nullptr;          // device copy of the per-(m,k)/(m,j) boundary factors
  const double* d_ymin_dexp = nullptr;
  const double* d_ymax_dexp = nullptr;
  const double* d_zmin_dens = nullptr;
  const double* d_zmax_dens = nullptr;
  double* dred = nullptr;           // device reduction workspace (n2d + 64 doubles)
  double* hres = nullptr;           // pinned host result slots (64 doubles)
  size_t n3 = 0, n2 = 0;
  char err[512] = {0};
};

#define F3D_CHECK(ctx, call)                                                               \
  do {                                                                                     \
    hipError_t e_ = (call);                                                                \
    if (e_ != hipSuccess) {                                                                \
      snprintf((ctx)->err, sizeof((ctx)->err), "%s failed: %s (%s:%d)", #call,             \
               hipGetErrorString(e_), __FILE__, __LINE__);                                 \
      return -1;                                                                           \
    }                                                                                      \
  } while (0)

#define F3D_LAUNCH_CHECK(ctx)                                                              \
  do {                                                                                     \
    hipError_t e_ = hipGetLastError();                                                     \
    if (e_ != hipSuccess) {                                                                \
      snprintf((ctx)->err, sizeof((ctx)->err), "kernel launch failed: %s (%s:%d)",         \
               hipGetErrorString(e_), __FILE__, __LINE__);                                 \
      return -1;                                                                           \
    }                                                                                      \
  } while (0)

#define F3D_REQUIRE(ctx, cond, msg)                                                        \
  do {                                                                                     \
    if (!(cond)) {                                                                         \
      snprintf((ctx)->err, sizeof((ctx)->err), "%s", msg);                                 \
      return -2;                                                                           \
    }                                                                                      \
  } while (0)
