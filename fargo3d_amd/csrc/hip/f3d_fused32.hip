// f3d_fused32.hip -- the reduced-precision build of the fused step (entry
// points f3d32_fused_*): FP32 fields and stencil arithmetic, FP64 ring sums,
// residual/shift split and CFL.  The kernels are in f3d_fused_body.h.
#include <cstring>
#include <type_traits>
#include <utility>
#include <vector>

#include "f3d_internal.h"
#include "f3d_cell32.h"

#define F3D_NS f3d32
#define F3D_PREC 32
#define F3D_API(name) f3d32_##name
#include "f3d_fused_body.h"
