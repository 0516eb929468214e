// f3d_internal.h -- the handle behind the C ABI and launch helpers.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <chrono>
#include <cstdio>

#include "../../../include/fargo3d_hip.h"
#include "f3d_device.h"

// stages of the fused step, in launch order (names in f3d_runtime.hip)
enum {
  F3D_ST_CFL = 0,
  F3D_ST_SRC_A,
  F3D_ST_SRC_B,
  F3D_ST_BC_MID,
  F3D_ST_SWEEP_Z,
  F3D_ST_SWEEP_Y,
  F3D_ST_SWEEP_X,
  F3D_ST_FINISH,
  F3D_ST_BC_END,
  F3D_ST_POTENTIAL,
  F3D_ST_SUBSTEP3,
  F3D_NSTAGES
};

struct f3d_hip_ctx {
  hipStream_t stream = nullptr;
  // a second stream (default priority) for work that overlaps the main
  // stream's, with the events that order it: side_go (main -> side) and
  // side_done (side -> main); created on first use
  hipStream_t side = nullptr;
  hipEvent_t side_go = nullptr, side_done = nullptr;
  // recorded after the fused CFL's last launch: the dt read-back waits for
  // it, not for work queued behind the CFL (created on first use)
  hipEvent_t cfl_done = nullptr;
  // the fused CFL's last kernel writes dt to hres[0] and then this sequence
  // number to hres[1] (pinned host memory); the host polls it (cfl_spin: the
  // queued CFL publishes the number)
  unsigned long long cfl_seq = 0;
  int cfl_spin = 0;
  int cfl_queued = 0;  // a fused CFL is queued and not yet waited for
  // the queued CFL's MIN pass over its ring minima, not yet launched: the
  // next potential launch carries it (f3d_Potential), else the wait launches
  // it; ringmin[0, n) -> result[slot], then hres with sequence number cfl_seq
  struct CflFin {
    const double* ringmin;
    int n, slot;
    unsigned long long* result;
  } fin{};
  int fin_pending = 0;
  int device = 0;
  f3d::dmesh m{};
  double* meshbuf = nullptr;        // one device buffer for all 1-D mesh arrays
  f3d_fields f{};                   // bound device buffers (selected fluid + scratch)
  f3d_phys ph{};                    // host copy of the run parameters
  double* bctab = nullptr;          // device copy of the per-(m,k)/(m,j) boundary factors
  // precision form: the FP64 ring profile V0(j,k) (sy*sz) that its azimuthal
  // velocities are stored as FP32 deviations from (f3d_hip_set_vbase)
  double* vbase = nullptr;
  const double* d_ymin_dexp = nullptr;
  const double* d_ymax_dexp = nullptr;
  const double* d_zmin_dens = nullptr;
  const double* d_zmax_dens = nullptr;
  double* dred = nullptr;           // device reduction workspace (n2d + 64 doubles)
  double* hres = nullptr;           // pinned host result slots (64 doubles)
  double* dhres = nullptr;          // the same slots as the device addresses them
  size_t n3 = 0, n2 = 0;
  size_t meshlen = 0;                // doubles in meshbuf
  char err[512] = {0};
  // scratch of the fused stages of the FP64 [0] and FP32 [1] builds, released
  // with the handle
  void* fused[2] = {nullptr, nullptr};
  void (*fused_free[2])(void*) = {nullptr, nullptr};
  // per-stage event timing of the fused kernels (f3d_hip_timing_*)
  unsigned timing = 0;  // bit s: time stage s
  hipEvent_t ev[F3D_NSTAGES][2] = {};
  int pending[F3D_NSTAGES] = {};
  double stage_ms[F3D_NSTAGES] = {};
  long stage_n[F3D_NSTAGES] = {};
  int timing_every = 1;             // time every n-th launch of a stage only
  long stage_calls[F3D_NSTAGES] = {};
  int skipped[F3D_NSTAGES] = {};
};

// Event pair around one launch of stage s when timing is on; the previous
// pair of the same stage is complete by then (every step synchronises on dt).
inline void f3d_stage_begin(f3d_hip_ctx* c, int s) {
  if (!((c->timing >> s) & 1u)) return;
  c->skipped[s] = (c->stage_calls[s]++ % c->timing_every) != 0;
  if (c->skipped[s]) return;
  if (c->pending[s]) {
    float ms = 0.f;
    if (hipEventSynchronize(c->ev[s][1]) == hipSuccess &&
        hipEventElapsedTime(&ms, c->ev[s][0], c->ev[s][1]) == hipSuccess) {
      c->stage_ms[s] += ms;
      c->stage_n[s] += 1;
    }
    c->pending[s] = 0;
  }
  (void)hipEventRecord(c->ev[s][0], c->stream);
}
inline void f3d_stage_end(f3d_hip_ctx* c, int s) {
  if (!((c->timing >> s) & 1u) || c->skipped[s]) return;
  (void)hipEventRecord(c->ev[s][1], c->stream);
  c->pending[s] = 1;
}

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
