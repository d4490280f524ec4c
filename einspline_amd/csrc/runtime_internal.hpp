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

// Locality traversal (order.hip).  esk_order_begin sorts the batch's points
// by cell on stream `s` into the table's workspace (*slot, created on first
// use) and returns the permutation; the workspace stays locked until
// esk_order_end records the stream's last launch that reads it.
struct EskOrderWs;
void esk_order_free(EskOrderWs *w);
int esk_order_begin(EskOrderWs **slot, const esk::Params &P, const void *pos, bool pos_f32, int num,
                    hipStream_t s, const int **perm);
void esk_order_end(EskOrderWs *w, hipStream_t s);
// The same sort into a file-scope device array (no allocation, at most
// EINSPLINE_NOALLOC_MAX_POINTS points); shared by every such call on the GPU.
int esk_order_begin_scratch(const esk::Params &P, const void *pos, bool pos_f32, int num, hipStream_t s,
                            const int **perm);

// Single point of a 3-D table through the resident polling kernel
// (point_service.hip): 1 when served (results copied out), 0 when the call
// is not one it takes, negative on error.
int esk_point_service_eval(const multi_UBspline_3d_hip *t, int mode, double x, double y, double z, void *vals,
                           void *grads, void *lh);
