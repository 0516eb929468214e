// f3d_device.h -- the FP64 build of the device mesh view (namespace f3d).
#pragma once

#include <hip/hip_runtime.h>

#define F3D_NS f3d
#define F3D_REAL double
#include "f3d_device_body.h"
#undef F3D_NS
#undef F3D_REAL
