// f3d_fused.hip -- the FP64 build of the fused step (entry points f3d_fused_*):
// bit-for-bit the reference's arithmetic.  The kernels are in f3d_fused_body.h.
#include <cstring>

#include "f3d_internal.h"
#include "f3d_cell.h"

#define F3D_NS f3d
#define F3D_PREC 64
#define F3D_API(name) f3d_##name
#include "f3d_fused_body.h"
