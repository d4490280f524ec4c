/*
 * multi_bspline_hip.h -- device-resident 3-D multi-spline tables and
 * batched many-point evaluation on AMD Instinct MI355X (gfx950).
 *
 * Replaces the reference's GPU path:
 *   src/multi_bspline_create_cuda.h:60-78  create_multi_UBspline_3d_*_cuda{,_conv}
 *   src/multi_bspline_eval_cuda.h:62-137   eval_multi_multi_UBspline_3d_*{,_vgh,_vgl}_cuda
 *   src/multi_bspline_structs_cuda.h       multi_UBspline_3d_*_cuda
 * The `_hip` functions keep the reference's argument meaning (a device table
 * built from a host spline, `num` points per call, device arrays of
 * per-point output pointers).  The `_batch` functions are this library's
 * main entry points: outputs are strided device arrays, positions are
 * doubles, work is enqueued on a caller-chosen stream and nothing
 * synchronises.
 *
 * All pointers named *_d are device memory.  `stream` is a hipStream_t
 * (NULL = the null stream); it is typed void* so that this header needs no
 * HIP headers.  Functions returning int give 0 on success and a negative
 * EINSPLINE_E* code on failure; einspline_hip_last_error() describes the
 * most recent failure of the calling thread.
 */
#ifndef EINSPLINE_MULTI_BSPLINE_HIP_H
#define EINSPLINE_MULTI_BSPLINE_HIP_H

#include <stddef.h>
#include <stdint.h>
#include "bspline_base.h"
#include "multi_bspline_structs.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- error codes ------------------------------------------------------ */
#define EINSPLINE_OK          0
#define EINSPLINE_EINVAL     -1   /* bad argument (NULL, size, type mismatch) */
#define EINSPLINE_ENOMEM     -2   /* device or host allocation failed */
#define EINSPLINE_EHIP       -3   /* HIP runtime error (launch, copy, no device) */
#define EINSPLINE_ERANGE     -4   /* value not representable in the storage type */
#define EINSPLINE_EBUDGET    -5   /* serving: table larger than the HBM budget */
#define EINSPLINE_EBUSY      -6   /* serving: no free slot / queue full */

/* ---- device tables ---------------------------------------------------- */

/* Element type of a device table's coefficients.  Evaluation accumulates in
 * double for F64 tables and in float otherwise. */
typedef enum {
  EINSPLINE_STORE_F32  = 0,
  EINSPLINE_STORE_F64  = 1,
  EINSPLINE_STORE_BF16 = 2,
  EINSPLINE_STORE_F16  = 3
} einspline_storage;

/* Device copy of a multi_UBspline_3d_* table.  Coefficients are laid out as
 * in the host struct, coefs[ix][iy][iz][r], with the row padded with zeros
 * to `row_reals` reals (r counts reals: a complex orbital n owns reals 2n and
 * 2n+1).  Strides are in reals. */
typedef struct multi_UBspline_3d_hip {
  void     *coefs;                       /* device memory, owned */
  intptr_t  x_stride, y_stride, z_stride;
  Ugrid     x_grid, y_grid, z_grid;
  int       Nx, Ny, Nz;                  /* table extent incl. the 3 (or 2) halo planes */
  int       num_splines;                 /* orbitals; a complex orbital counts once */
  int       row_reals;                   /* == z_stride */
  type_code tcode;                       /* result type of evaluation */
  einspline_storage storage;
  size_t    bytes;                       /* device bytes of coefs */
  int       device;                      /* HIP device ordinal */
  int       dim;                         /* 3, or 1 / 2 for tables from the 1d / 2d creates
                                            (Ny, Nz = 1 past dim) */
} multi_UBspline_3d_hip;

typedef multi_UBspline_3d_hip multi_UBspline_3d_s_hip;
typedef multi_UBspline_3d_hip multi_UBspline_3d_d_hip;
typedef multi_UBspline_3d_hip multi_UBspline_3d_c_hip;
typedef multi_UBspline_3d_hip multi_UBspline_3d_z_hip;

/* Upload a host table to the current device (src/multi_bspline_create_cuda.h:60-78).
 * Unlike the reference, any boundary condition is accepted and grid.start
 * is honoured.  Returns NULL on failure (see einspline_hip_last_error). */
multi_UBspline_3d_s_hip *create_multi_UBspline_3d_s_hip(multi_UBspline_3d_s *spline);
multi_UBspline_3d_s_hip *create_multi_UBspline_3d_s_hip_conv(multi_UBspline_3d_d *spline);
multi_UBspline_3d_d_hip *create_multi_UBspline_3d_d_hip(multi_UBspline_3d_d *spline);
multi_UBspline_3d_c_hip *create_multi_UBspline_3d_c_hip(multi_UBspline_3d_c *spline);
multi_UBspline_3d_c_hip *create_multi_UBspline_3d_c_hip_conv(multi_UBspline_3d_z *spline);
multi_UBspline_3d_z_hip *create_multi_UBspline_3d_z_hip(multi_UBspline_3d_z *spline);

/* Precision form (no reference counterpart): a table solved in double and
 * stored as F32, BF16 or F16 (round to nearest even); evaluation returns
 * float results with float accumulation.  F16 fails with EINSPLINE_ERANGE
 * if a coefficient exceeds the half-precision range. */
multi_UBspline_3d_s_hip *create_multi_UBspline_3d_s_hip_store(const multi_UBspline_3d_d *spline,
                                                             einspline_storage storage);

/* Free a device table (the reference has none; its callers cudaFree by hand). */
void destroy_multi_UBspline_3d_hip(multi_UBspline_3d_hip *table);

/* ---- 1-D and 2-D device tables -----------------------------------------
 * src/multi_bspline_create_cuda.h:10-52 (create_multi_UBspline_{1,2}d_X_cuda
 * and the _conv forms).  The device table has the 3-D descriptor with dim 1
 * or 2; FP32 (s, c) or FP64 (d, z) storage.  They evaluate through the
 * strided einspline_eval_{v,vg,vgl,vgh}_batch calls below, with `dim`
 * coordinates per point in pos_d and per-point outputs vals N, grads dim*N,
 * lapl N, hess 4N (2-D) or N (1-D, where vgh is vgl).  The tile setting
 * applies; the point order is always the given one (the locality order,
 * the autotuner, vghgh and packed Hessians are for 3-D tables). */
typedef multi_UBspline_3d_hip multi_UBspline_1d_s_hip;
typedef multi_UBspline_3d_hip multi_UBspline_1d_d_hip;
typedef multi_UBspline_3d_hip multi_UBspline_1d_c_hip;
typedef multi_UBspline_3d_hip multi_UBspline_1d_z_hip;
typedef multi_UBspline_3d_hip multi_UBspline_2d_s_hip;
typedef multi_UBspline_3d_hip multi_UBspline_2d_d_hip;
typedef multi_UBspline_3d_hip multi_UBspline_2d_c_hip;
typedef multi_UBspline_3d_hip multi_UBspline_2d_z_hip;
multi_UBspline_1d_s_hip *create_multi_UBspline_1d_s_hip(multi_UBspline_1d_s *spline);
multi_UBspline_1d_s_hip *create_multi_UBspline_1d_s_hip_conv(multi_UBspline_1d_d *spline);
multi_UBspline_1d_d_hip *create_multi_UBspline_1d_d_hip(multi_UBspline_1d_d *spline);
multi_UBspline_1d_c_hip *create_multi_UBspline_1d_c_hip(multi_UBspline_1d_c *spline);
multi_UBspline_1d_c_hip *create_multi_UBspline_1d_c_hip_conv(multi_UBspline_1d_z *spline);
multi_UBspline_1d_z_hip *create_multi_UBspline_1d_z_hip(multi_UBspline_1d_z *spline);
multi_UBspline_2d_s_hip *create_multi_UBspline_2d_s_hip(multi_UBspline_2d_s *spline);
multi_UBspline_2d_s_hip *create_multi_UBspline_2d_s_hip_conv(multi_UBspline_2d_d *spline);
multi_UBspline_2d_d_hip *create_multi_UBspline_2d_d_hip(multi_UBspline_2d_d *spline);
multi_UBspline_2d_c_hip *create_multi_UBspline_2d_c_hip(multi_UBspline_2d_c *spline);
multi_UBspline_2d_c_hip *create_multi_UBspline_2d_c_hip_conv(multi_UBspline_2d_z *spline);
multi_UBspline_2d_z_hip *create_multi_UBspline_2d_z_hip(multi_UBspline_2d_z *spline);

/* 1-D CUDA-style batched evaluation (src/multi_bspline_eval_cuda.h:10-51):
 * pos_d[i] is point i (float for s/c, double for d/z); vals_d[i],
 * grads_d[i], lapl_d[i] / grad_lapl_d[i] are device pointers to point i's
 * outputs.  The d/c/z vgl form writes the gradient row (N elements) and the
 * Laplacian row_stride elements after it.  Derivatives are in physical units
 * for every orbital (the reference's s kernel leaves them in grid units and
 * skips the last N % 64 orbitals' derivatives).  Synchronous. */
int eval_multi_multi_UBspline_1d_s_hip(multi_UBspline_1d_s_hip *spline, float *pos_d, float *vals_d[],
                                       int num);
int eval_multi_multi_UBspline_1d_s_vgl_hip(multi_UBspline_1d_s_hip *spline, float *pos_d,
                                           float *vals_d[], float *grads_d[], float *lapl_d[], int num);
int eval_multi_multi_UBspline_1d_d_hip(multi_UBspline_1d_d_hip *spline, double *pos_d, double *vals_d[],
                                       int num);
int eval_multi_multi_UBspline_1d_d_vgl_hip(multi_UBspline_1d_d_hip *spline, double *pos_d,
                                           double *vals_d[], double *grad_lapl_d[], int num,
                                           int row_stride);
int eval_multi_multi_UBspline_1d_c_hip(multi_UBspline_1d_c_hip *spline, float *pos_d,
                                       complex_float *vals_d[], int num);
int eval_multi_multi_UBspline_1d_c_vgl_hip(multi_UBspline_1d_c_hip *spline, float *pos_d,
                                           complex_float *vals_d[], complex_float *grad_lapl_d[],
                                           int num, int row_stride);
int eval_multi_multi_UBspline_1d_z_hip(multi_UBspline_1d_z_hip *spline, double *pos_d,
                                       complex_double *vals_d[], int num);
int eval_multi_multi_UBspline_1d_z_vgl_hip(multi_UBspline_1d_z_hip *spline, double *pos_d,
                                           complex_double *vals_d[], complex_double *grad_lapl_d[],
                                           int num, int row_stride);

/* ---- CUDA-style batched evaluation -------------------------------------
 * src/multi_bspline_eval_cuda.h:62-137.  pos_d[3*i+{0,1,2}] is point i;
 * vals_d[i], grads_d[i], hess_d[i] are device pointers (held in device
 * arrays) to point i's outputs.  Gradients are [n][3]; Hessians are packed
 * as 6 unique entries [xx,xy,xz,yy,yz,zz] per orbital, as the reference's
 * CUDA kernels write them.  Complex tables are written as those kernels
 * write them, as 2N reals: per orbital the gradient is [re x,y,z][im x,y,z]
 * and the Hessian [re 6][im 6] (src/multi_bspline_cuda_z_impl.h:205-228,
 * src/multi_bspline_cuda_c_impl.h:389-412); the complex_* pointer types only
 * give the blocks' sizes.
 * These calls synchronise the device before returning, as the reference's
 * do. */
int eval_multi_multi_UBspline_3d_s_hip(multi_UBspline_3d_s_hip *spline, float *pos_d,
                                       float *vals_d[], int num);
int eval_multi_multi_UBspline_3d_s_vgh_hip(multi_UBspline_3d_s_hip *spline, float *pos_d,
                                           float *vals_d[], float *grads_d[], float *hess_d[],
                                           int num);
int eval_multi_multi_UBspline_3d_d_hip(multi_UBspline_3d_d_hip *spline, double *pos_d,
                                       double *vals_d[], int num);
int eval_multi_multi_UBspline_3d_d_vgh_hip(multi_UBspline_3d_d_hip *spline, double *pos_d,
                                           double *vals_d[], double *grads_d[], double *hess_d[],
                                           int num);
int eval_multi_multi_UBspline_3d_c_hip(multi_UBspline_3d_c_hip *spline, float *pos_d,
                                       complex_float *vals_d[], int num);
int eval_multi_multi_UBspline_3d_c_vgh_hip(multi_UBspline_3d_c_hip *spline, float *pos_d,
                                           complex_float *vals_d[], complex_float *grads_d[],
                                           complex_float *hess_d[], int num);
int eval_multi_multi_UBspline_3d_z_hip(multi_UBspline_3d_z_hip *spline, double *pos_d,
                                       complex_double *vals_d[], int num);
int eval_multi_multi_UBspline_3d_z_vgh_hip(multi_UBspline_3d_z_hip *spline, double *pos_d,
                                           complex_double *vals_d[], complex_double *grads_d[],
                                           complex_double *hess_d[], int num);

/* QMC-shaped outputs (src/multi_bspline_eval_cuda.h:67-79, 96-99, 112-115,
 * 131-134): for point i, vals_d[i][n] is the value (times sign_d[i] for the
 * _sign variants); grad_lapl_d[i] + d*row_stride, d = 0..2, holds the
 * gradient transformed by the row-major 3x3 lattice matrix Linv_d
 * (g'_d = sum_j Linv[3d+j] g_j), and grad_lapl_d[i] + 3*row_stride holds
 * the Laplacian tr((Linv^T Linv) H).  row_stride counts result elements
 * (complex ones for c/z) and must be >= num_splines.  Synchronous. */
int eval_multi_multi_UBspline_3d_s_sign_hip(multi_UBspline_3d_s_hip *spline, float *pos_d,
                                            float *sign_d, float *vals_d[], int num);
int eval_multi_multi_UBspline_3d_s_vgl_hip(multi_UBspline_3d_s_hip *spline, float *pos_d,
                                           float *Linv_d, float *vals_d[], float *grad_lapl_d[],
                                           int num, int row_stride);
int eval_multi_multi_UBspline_3d_s_vgl_sign_hip(multi_UBspline_3d_s_hip *spline, float *pos_d,
                                                float *sign_d, float *Linv_d, float *vals_d[],
                                                float *grad_lapl_d[], int num, int row_stride);
int eval_multi_multi_UBspline_3d_d_vgl_hip(multi_UBspline_3d_d_hip *spline, double *pos_d,
                                           double *Linv_d, double *vals_d[],
                                           double *grad_lapl_d[], int num, int row_stride);
int eval_multi_multi_UBspline_3d_c_vgl_hip(multi_UBspline_3d_c_hip *spline, float *pos_d,
                                           float *Linv_d, complex_float *vals_d[],
                                           complex_float *grad_lapl_d[], int num, int row_stride);
int eval_multi_multi_UBspline_3d_z_vgl_hip(multi_UBspline_3d_z_hip *spline, double *pos_d,
                                           double *Linv_d, complex_double *vals_d[],
                                           complex_double *grad_lapl_d[], int num, int row_stride);

/* ---- strided batched evaluation (main entry points) --------------------
 * Evaluates `num` points pos_d[3*i+{0,1,2}] (doubles, any storage type) in
 * one launch on `stream`.  Point i's results go to
 *   vals_d  + i*N          (N = num_splines values)
 *   grads_d + i*3N         ([n][3])
 *   lapl_d  + i*N          (vgl)
 *   hess_d  + i*9N or i*6N (vgh; full row-major 3x3, or packed 6 if hess_packed)
 * counted in result elements (float for F32/BF16/F16 tables, double for
 * F64, complex for c/z).  Results are bit-for-bit independent of `num` and
 * of the other points in the batch.  Does not synchronise. */
int einspline_eval_v_batch(const multi_UBspline_3d_hip *spline, const double *pos_d, int num,
                           void *vals_d, void *stream);
int einspline_eval_vg_batch(const multi_UBspline_3d_hip *spline, const double *pos_d, int num,
                            void *vals_d, void *grads_d, void *stream);
int einspline_eval_vgl_batch(const multi_UBspline_3d_hip *spline, const double *pos_d, int num,
                             void *vals_d, void *grads_d, void *lapl_d, void *stream);
int einspline_eval_vgh_batch(const multi_UBspline_3d_hip *spline, const double *pos_d, int num,
                             void *vals_d, void *grads_d, void *hess_d, int hess_packed,
                             void *stream);
/* vghgh: as vgh with the full 3x3 Hessian, plus gradhess_d + i*27N holding the
 * 27 third derivatives of each orbital ([n][a][b][c], symmetric).  Batched
 * form of eval_multi_UBspline_3d_X_vghgh (src/multi_bspline_eval_std_d_impl.h:868-1049). */
int einspline_eval_vghgh_batch(const multi_UBspline_3d_hip *spline, const double *pos_d, int num,
                               void *vals_d, void *grads_d, void *hess_d, void *gradhess_d,
                               void *stream);

/* Allocation-free vgh (no counterpart in the reference).  For callers that
 * cannot allocate device memory: code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc.  Same arguments and results as
 * einspline_eval_vgh_batch on a 3-D table, bit for bit, but no call
 * allocates: when the table's point order is EINSPLINE_ORDER_LOCALITY, the
 * sort works in a file-scope device array (12.6 MB per GPU, kept until the
 * process ends) instead of the table's workspace, and it is also done under
 * stream capture.  num must be at most EINSPLINE_NOALLOC_MAX_POINTS
 * (EINSPLINE_EINVAL otherwise).  Every such call on one GPU shares that
 * GPU's array: calls that overlap in time on one GPU (other streams, other
 * threads) race and must be ordered by the caller.  Does not synchronise. */
#define EINSPLINE_NOALLOC_MAX_POINTS (1 << 20)
int einspline_eval_vgh_batch_noalloc(const multi_UBspline_3d_hip *spline, const double *pos_d, int num,
                                     void *vals_d, void *grads_d, void *hess_d, int hess_packed,
                                     void *stream);

/* Tile shape of the batched kernels: bytes of an orbital row one lane loads
 * per stencil row (8, 16 or 32; F64 tables 16 or 32, 16-bit tables 8 or 16)
 * and work-items per block (64..256, a multiple of 64).  0 selects the
 * built-in default.  Applies to subsequent launches on `spline`; returns
 * EINSPLINE_EINVAL for unsupported shapes. */
int einspline_set_tile(multi_UBspline_3d_hip *spline, int lane_bytes, int block_threads);

/* Result store policy (results are identical under all three):
 *   0 (default) gradient/Hessian blocks are transposed through LDS so each
 *     wavefront store writes contiguous memory, with streaming
 *     (non-temporal) stores that keep outputs from evicting the table;
 *   1 the same transposition with ordinary stores;
 *   2 each lane stores its own results with ordinary stores. */
int einspline_set_store_policy(multi_UBspline_3d_hip *spline, int policy);

/* Point traversal order (no reference counterpart; the reference evaluates
 * points in the order given).
 *   EINSPLINE_ORDER_GIVEN     one point after another as passed (default);
 *   EINSPLINE_ORDER_LOCALITY  batches of 1024 points or more are first sorted
 *     by the Morton code of their cell (a key kernel and a radix sort on the
 *     call's stream), and each XCD then walks one contiguous run of the sorted
 *     points, so that points sharing coefficient rows run together and the
 *     rows are served from L2 and the Infinity Cache instead of HBM.
 * Results are identical in both orders: each point's outputs go to its own
 * index and are computed by the same instructions.  The sort uses a workspace
 * owned by the table (allocated on first use, grown with hipMalloc when a
 * larger batch arrives); calls on the same table from different streams are
 * ordered one after another through it.  Under stream capture the given order
 * is used.
 * Applies to the 3-D batched calls: the strided einspline_eval_*_batch
 * functions and the pointer-array eval_multi_multi_UBspline_3d_*_hip ones. */
enum { EINSPLINE_ORDER_GIVEN = 0, EINSPLINE_ORDER_LOCALITY = 1 };
/* Either order may add EINSPLINE_ORDER_SLABS (order | EINSPLINE_ORDER_SLABS):
 * orbital slabs outermost.  When a point's row spans several wavefronts'
 * widths (vector kernel: more than 64 lanes per point; matrix kernels: more
 * than 256 orbitals for 16-bit tables, or 64 for FP32 tables), all points evaluate the first
 * slab of orbitals, then the second, and so on, so the rows of one slab are
 * what the caches hold at a time.  Results are identical.  Applies to the
 * strided einspline_eval_*_batch calls; elsewhere it is ignored. */
enum { EINSPLINE_ORDER_SLABS = 2 };
int einspline_set_point_order(multi_UBspline_3d_hip *spline, int order);

/* Which kernel evaluates vgh on a BF16, FP16 or FP32 table: the vector
 * kernel (default; the tile applies) or the matrix-core kernel, which
 * multiplies each point's weight matrix (one row per output: v, gradient,
 * Hessian; 64 stencil columns) with its 64 stencil rows as MFMA tiles, FP32
 * accumulation (strided einspline_eval_vgh_batch of real tables only; other
 * modes and entry points use the vector kernel; the store policy's
 * streaming/ordinary choice applies).  Its results differ from the vector
 * kernel's in rounding only.  The autotuner tries both.  FP64 and complex
 * tables ignore the setting.  einspline_get_kernel returns the setting or a
 * negative code.  No counterpart in the reference (its kernels are vector
 * code). */
enum { EINSPLINE_KERNEL_VECTOR = 0, EINSPLINE_KERNEL_MATRIX = 1 };
int einspline_set_kernel(multi_UBspline_3d_hip *spline, int kernel);
int einspline_get_kernel(const multi_UBspline_3d_hip *spline);

/* The table's current tile, store policy and point order (defaults filled in). */
int einspline_get_tuning(const multi_UBspline_3d_hip *spline, int *lane_bytes, int *block_threads,
                         int *store_policy, int *point_order);

/* Autotuner (no reference counterpart): times every supported tile shape
 * (lane_bytes x block_threads x store policy x point order) of `mode` (0 v, 1 vg, 2 vgl,
 * 3 vgh) on `sample_points` uniform random points of the table's box, applies the
 * fastest with einspline_set_tile / einspline_set_store_policy /
 * einspline_set_point_order and returns the tile and policy (the order is read
 * back with einspline_get_tuning).  Results are cached per
 * (storage, type, mode, row bytes, sample size, Nx, Ny, Nz) for the process, and in the text file
 * named by $EINSPLINE_TUNE_CACHE when set.  Synchronises the device. */
int einspline_autotune(multi_UBspline_3d_hip *spline, int mode, int sample_points,
                       int *lane_bytes, int *block_threads, int *store_policy,
                       float *best_ms);

/* Autotuner on the caller's own positions (no reference counterpart): the same
 * candidates as einspline_autotune, timed on `npoints` host positions
 * pos[3*i + {0,1,2}] (a representative sample of the points the caller will
 * evaluate; at most the first 65536 are used), the fastest applied to the table.  Not cached: the best point
 * order depends on how the points are distributed (points clustered in a few
 * cells favour the given order, uniform points the locality order).
 * Synchronises the device. */
int einspline_autotune_points(multi_UBspline_3d_hip *spline, int mode, const double *pos, int npoints,
                              int *lane_bytes, int *block_threads, int *store_policy,
                              float *best_ms);

/* ---- single-point path support ----------------------------------------
 * The single-point eval_multi_UBspline_3d_* functions keep a device mirror
 * of the host table and refresh it after set_*.  A caller that writes
 * spline->coefs directly must call this before the next evaluation. */
void einspline_hip_invalidate(void *spline);

/* Calls on 3-D tables (v, vg, vgl, vgh; no reference counterpart: the
 * reference evaluates on the CPU) are served by a resident kernel (one block
 * of four wavefronts) that polls a mailbox in pinned host memory, so a call
 * costs no launch and no HIP call.  The kernel leaves after `idle_us`
 * microseconds without a call (default 2000; 100 .. 1000000) and is started
 * again by the next call.  While it is resident it holds one block slot, a
 * device-wide synchronisation waits for it to go idle, and so does work on
 * any stream that the HIP runtime maps to the same hardware queue (streams
 * share queues once GPU_MAX_HW_QUEUES are in use).  enable = 0 makes
 * every call one launch plus one stream synchronisation; idle_us <= 0
 * keeps the current idle time.  Also set by the environment variables
 * EINSPLINE_POINT_SERVICE=0 and EINSPLINE_POINT_SERVICE_IDLE_US at first
 * use.  Resident kernels leave before the call returns. */
int einspline_set_point_service(int enable, int idle_us);
/* Number of devices on which the resident kernel is running now. */
int einspline_point_service_resident(void);

/* ---- minimal device runtime (so C and ctypes callers need no HIP) ------ */
/* ---- tables solved on the GPU ------------------------------------------ */
/* Build a 3-D device table from samples already in device memory: the solve
 * of set_multi_UBspline_3d_* (src/multi_bspline_create.c:323-1477) runs on
 * the GPU, so neither the host solve nor a coefficient upload is needed.
 *   spline   a 3-D host spline from create_multi_UBspline_3d_{s,d,c,z}: its
 *            grids, boundary conditions and num_splines (its host coefficients
 *            are neither read nor written);
 *   data_d   num_splines blocks of Mx*My*Mz samples, in the order successive
 *            set_multi_UBspline_3d_* calls take them (float for s/c, double
 *            for d/z; complex samples as re,im pairs);
 *   storage  -1 for the spline's own type (s/c F32, d/z F64); a d spline may
 *            also ask for F32, BF16 or F16 (as ..._s_hip_store) and a z spline
 *            for F32 (as ..._c_hip_conv).
 * The solve is done in double, with the host's operations in the host's
 * order, so the table equals the set + upload one bit for bit.  Runs on
 * `stream` and waits for it.  NULL on failure (einspline_hip_last_error). */
multi_UBspline_3d_hip *einspline_build_table_hip(void *spline, const void *data_d, int storage, void *stream);

/* ---- tables on disk (no reference counterpart: einspline has no table file) */
/* Save a device table of any dimension and storage to `path`, as the kernels
 * read it: a 512-byte header (grids, sizes, storage type, tuned tile, point
 * order) and the coefficient rows byte for byte, with 64-bit hashes of both.
 * Written to `path`.tmp, then renamed, so a reader never sees a partial file. */
int einspline_table_save(const multi_UBspline_3d_hip *spline, const char *path);
/* Load a saved table onto the current device: no solve and no conversion,
 * the saved tile and point order restored.  NULL for a missing, damaged or
 * foreign file (see einspline_hip_last_error). */
multi_UBspline_3d_hip *einspline_table_load(const char *path);

int         einspline_hip_device_count(void);
const char *einspline_hip_last_error(void);
void       *einspline_hip_malloc(size_t bytes);             /* NULL on failure */
int         einspline_hip_free(void *ptr_d);
int         einspline_hip_memcpy_htod(void *dst_d, const void *src, size_t bytes);
int         einspline_hip_memcpy_dtoh(void *dst, const void *src_d, size_t bytes);
int         einspline_hip_memset(void *ptr_d, int value, size_t bytes);
int         einspline_hip_synchronize(void);
void       *einspline_hip_stream_create(void);
int         einspline_hip_stream_destroy(void *stream);
int         einspline_hip_stream_synchronize(void *stream);
void       *einspline_hip_event_create(void);
int         einspline_hip_event_destroy(void *event);
int         einspline_hip_event_record(void *event, void *stream);
int         einspline_hip_event_synchronize(void *event);
/* Milliseconds between two recorded events; negative on error. */
float       einspline_hip_event_elapsed_ms(void *start, void *stop);
/* Plain device-to-device copy kernel (bytes a multiple of 16), used to
 * measure the achievable HBM bandwidth and to calibrate traffic counters. */
int         einspline_hip_stream_copy(void *dst_d, const void *src_d, size_t bytes, void *stream);
/* Plain fill kernel (bytes a multiple of 16; streaming stores if nontemporal),
 * used to measure the achievable HBM write bandwidth. */
int         einspline_hip_stream_fill(void *dst_d, unsigned value, size_t bytes, int nontemporal,
                                      void *stream);

#ifdef __cplusplus
}
#endif

#endif /* EINSPLINE_MULTI_BSPLINE_HIP_H */
