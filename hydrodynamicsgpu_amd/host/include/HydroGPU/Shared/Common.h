// HydroGPU/Shared/Common.h -- scalar type of the host interface.
// The reference's host `real` is double in its FP64 build
// (res/include/HydroGPU/Shared/Common.h:3-30); the host API keeps FP64 at the
// boundary for both device precisions.
#pragma once

namespace HydroGPU {
typedef double real;
}
