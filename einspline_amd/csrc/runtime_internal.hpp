// runtime_internal.hpp -- helpers shared between runtime.hip and serve.hip.
#pragma once
#include <hip/hip_runtime.h>
#include "multi_bspline_hip.h"
#include "launch.hpp"

// Record an error for einspline_hip_last_error() and return `code`.
int esk_fail(int code, const char *fmt, ...);
int esk_hip_fail(hipError_t e, const char *what);
// Kernel arguments and tile for `num` points of table `t` (outputs unset).
int esk_prepare(const multi_UBspline_3d_hip *t, int num, esk::Launch &L, esk::Params &P);
// Wrap device memory holding a table laid out as `meta` (takes ownership).
multi_UBspline_3d_hip *esk_table_wrap(const multi_UBspline_3d_hip *meta, void *dcoefs);
