// f3d_device32.h -- the FP32 build of the device mesh view (namespace f3d32),
// used by the reduced-precision form of the fused step.
#pragma once

#include <hip/hip_runtime.h>

#define F3D_NS f3d32
#define F3D_REAL float
#include "f3d_device_body.h"
#undef F3D_NS
#undef F3D_REAL
