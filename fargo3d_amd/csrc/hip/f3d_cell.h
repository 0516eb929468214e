// f3d_cell.h -- the FP64 build of the per-cell arithmetic (namespace f3d::cell).
#pragma once

#include "f3d_device.h"

#define F3D_NS f3d
#define F3D_REAL double
#include "f3d_cell_body.h"
#undef F3D_NS
#undef F3D_REAL
