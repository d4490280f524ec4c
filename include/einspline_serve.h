/*
 * einspline_serve.h -- serving runtime: resident spline tables under an HBM
 * budget, continuous batching of evaluation requests of uneven size into
 * persistent, hipGraph-captured launches.
 *
 * No reference counterpart: einspline has no serving layer.  Its batched GPU
 * call (src/multi_bspline_eval_cuda.h:62-137) evaluates one table per call
 * and synchronises; this runtime keeps several tables resident, fuses
 * requests for different tables into one batch, and keeps several batches in
 * flight on separate streams.
 *
 * Semantics
 *  - Tables are registered from host splines and uploaded ("admitted") when
 *    a request needs them and the budget allows.  To make room, the set of
 *    idle tables with the fewest bytes is evicted; when the cheapest set
 *    includes a table with work in flight, that table gets no new work until
 *    it is idle (einspline_server_plan_admission).  A request whose table
 *    cannot be admitted yet waits.
 *  - A request is `npts` points (positions copied at submit) of one table;
 *    results (vgh) go to caller-owned device buffers laid out as in
 *    einspline_eval_vgh_batch.  Each result is bitwise identical to
 *    evaluating the request alone: every point is computed by the same
 *    instruction sequence whatever else is in the batch.
 *  - Requests are admitted into batches first-come first-served, skipping
 *    (not reordering past more than `max_batch_requests`) requests that do
 *    not fit the current batch.
 *  - Latency of a request = submit -> completion observed by the runtime.
 *
 * All functions return 0 or a negative EINSPLINE_E* code
 * (see multi_bspline_hip.h); einspline_hip_last_error() has the details.
 * The server's worker thread is the only one that touches the GPU.
 */
#ifndef EINSPLINE_SERVE_H
#define EINSPLINE_SERVE_H

#include <stddef.h>
#include <stdint.h>
#include "multi_bspline_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct einspline_server einspline_server;

typedef struct {
  size_t hbm_budget;          /* bytes of device memory for resident tables (0 = 90% of free) */
  int    slots;               /* batches in flight, one stream each (default 4) */
  int    max_batch_points;    /* points per batch (default 16384) */
  int    max_batch_requests;  /* requests per batch (default 256) */
  int    hess_packed;         /* 1: 6-entry Hessians, 0: full 3x3 */
} einspline_server_config;

typedef struct {
  uint64_t submitted, completed, batches;
  double   evals;             /* points x orbitals completed */
  uint64_t admissions, evictions, deferrals;
  int      resident_tables;
  size_t   resident_bytes;
} einspline_server_stats;

einspline_server *einspline_server_create(const einspline_server_config *cfg);
void einspline_server_destroy(einspline_server *srv);

/* Register a host spline (multi_UBspline_3d_s with EINSPLINE_STORE_F32, or
 * multi_UBspline_3d_d with F32/BF16/F16 storage).  The spline must outlive
 * the server.  Nothing is uploaded until a request needs the table. */
int einspline_server_register_table(einspline_server *srv, void *host_spline,
                                    einspline_storage storage, int *table_id);

/* Queue `npts` (1..max_batch_points) points pos_host[3*i+{0,1,2}] of table
 * `table_id`.  Outputs: vals_d[npts*N], grads_d[npts*3N], hess_d[npts*9N]
 * (or 6N) floats in device memory, owned by the caller and untouched until
 * the request completes. */
int einspline_server_submit(einspline_server *srv, int table_id, const double *pos_host,
                            int npts, void *vals_d, void *grads_d, void *hess_d,
                            uint64_t *ticket);

/* 1 if the request finished (latency in seconds stored), 0 if pending.  A
 * request whose batch or table upload failed returns the failure's code once
 * collected; after any failure the server forms no more batches and every
 * call that submits or waits returns that code. */
int einspline_server_test(einspline_server *srv, uint64_t ticket, double *latency_s);
/* Block until the request finished (EINSPLINE_OK), or a failure code as for
 * einspline_server_test. */
int einspline_server_wait(einspline_server *srv, uint64_t ticket, double *latency_s);
/* Block until every submitted request finished. */
int einspline_server_drain(einspline_server *srv);

int einspline_server_get_stats(einspline_server *srv, einspline_server_stats *out);
/* 1 if the table is resident on the device now. */
int einspline_server_table_resident(einspline_server *srv, int table_id);

/* The server's admission decision on its own: no server, no GPU, no clock
 * (for planning budgets and for testing the policy).  `tables` describes
 * every registered table; the request needs `table_id` resident and `used`
 * bytes of the `budget` are taken (resident tables plus uploads in
 * progress).  Returns
 *   EINSPLINE_ADMIT_RESIDENT  the table is resident;
 *   EINSPLINE_ADMIT_UPLOAD    upload it now, after evicting the *nvictims
 *                             tables listed in victims[] (possibly none):
 *                             the set of idle tables with the fewest bytes
 *                             that frees enough room (ties: fewer with
 *                             queued requests, then least recently used);
 *   EINSPLINE_ADMIT_DEFER     wait: the table is being uploaded, or the
 *                             cheapest set holds tables with work in flight,
 *                             listed in drain[] (*ndrain): they get no new
 *                             work until they are idle and can be evicted,
 *                             instead of a larger idle table being evicted;
 *   EINSPLINE_EBUDGET         the table is larger than the budget.
 * victims and drain must hold ntables entries. */
typedef struct {
  size_t   bytes;       /* device bytes of the table when resident */
  int      state;       /* 0 absent, 1 resident, 2 being uploaded */
  int      busy;        /* nonzero: work in flight */
  int      pending;     /* requests queued for it */
  uint64_t last_use;    /* logical time of its last batch */
} einspline_admission_table;
enum { EINSPLINE_ADMIT_RESIDENT = 0, EINSPLINE_ADMIT_UPLOAD = 1, EINSPLINE_ADMIT_DEFER = 2 };
int einspline_server_plan_admission(const einspline_admission_table *tables, int ntables, int table_id,
                                    size_t used, size_t budget, int *victims, int *nvictims, int *drain,
                                    int *ndrain);

/* Benchmark helper: generate requests on the calling thread and collect
 * their latencies.  Arrivals are Poisson at rate_hz (rate_hz <= 0: closed
 * loop, keeping at most slots*max_batch_requests requests queued) for
 * duration_s seconds; each request picks one of the `ntables` tables
 * uniformly, a log-uniform point count on [1, max_points] and uniform
 * positions in [0, box).  Outputs of all requests of table k go to the
 * same sink buffers (vals_d[k], grads_d[k], hess_d[k], sized for
 * max_points points), so their contents are not meaningful.  Returns after
 * every request completed; latencies (seconds, completion order) are
 * copied to `latencies` (up to cap), their count to *nreq. */
typedef struct {
  double rate_hz;
  double duration_s;
  int max_points;
  uint64_t seed;
  int ntables;
  const int *table_ids;
  void *const *vals_d;
  void *const *grads_d;
  void *const *hess_d;
  double box[3];
} einspline_load;

int einspline_server_run_load(einspline_server *srv, const einspline_load *load,
                              double *latencies, size_t cap, size_t *nreq, double *wall_s);

#ifdef __cplusplus
}
#endif

#endif
