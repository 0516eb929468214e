// f3d_fused.hip -- fused stage kernels of the fast path.
#include "f3d_internal.h"
#include "f3d_cell.h"

extern "C" {
int f3d_fused_cfl(f3d_hip_ctx* ctx, double* fluid_min) {
  (void)fluid_min;
  F3D_REQUIRE(ctx, false, "fused cfl stage not built");
  return 0;
}
int f3d_fused_sources(f3d_hip_ctx* ctx, double dt, double domega) {
  (void)dt; (void)domega;
  F3D_REQUIRE(ctx, false, "fused sources stage not built");
  return 0;
}
int f3d_fused_transport(f3d_hip_ctx* ctx, double dt, int end_of_step) {
  (void)dt; (void)end_of_step;
  F3D_REQUIRE(ctx, false, "fused transport stage not built");
  return 0;
}
}
