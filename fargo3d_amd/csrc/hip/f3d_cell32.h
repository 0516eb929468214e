// f3d_cell32.h -- the FP32 build of the per-cell arithmetic (namespace f3d32::cell).
#pragma once

#include "f3d_device32.h"

#define F3D_NS f3d32
#define F3D_REAL float
#include "f3d_cell_body.h"
#undef F3D_NS
#undef F3D_REAL
