/*
 * loop_tool_serve.h -- C ABI of the serving runtime of loop_tool_amd.
 *
 * No counterpart in the reference, whose only execution entry point is
 * Compiled::run (include/loop_tool/backend.h:18-22) driven by the lazy
 * frontend one program at a time (src/frontends/lazy.cpp:430-438).  The
 * server takes lazy programs (ltc_tensor from loop_tool_c.h) of any shape as
 * they arrive and runs them on one GPU with continuous batching:
 *   - a program is keyed by the structure of its lazy graph with sizes masked
 *     (the lazy hash of src/frontends/lazy.cpp:112-133 without the bound
 *     sizes); programs with the structure of matmul, row softmax, 3x3
 *     pad-1 convolution or 2x2 max-pool run on shape-generic kernels, any
 *     other program on the hip backend, compiled once per exact lazy hash;
 *   - inputs are materialised in HBM from per-input seeds and outputs stay
 *     in HBM (optionally checksummed on the device or copied back);
 *   - requests are admitted in arrival order while a slot (one stream each)
 *     and the device memory budget allow; work of admitted requests overlaps
 *     on the device; completions are polled without blocking.
 */
#ifndef LOOP_TOOL_SERVE_H_
#define LOOP_TOOL_SERVE_H_

#include <stdint.h>

#include "loop_tool_c.h"

#ifdef __cplusplus
extern "C" {
#endif

#pragma GCC visibility push(default)

typedef struct lts_server lts_server;

enum {
  LTS_MATMUL = 0,
  LTS_SOFTMAX = 1,
  LTS_CONV2D = 2,
  LTS_MAXPOOL = 3,
  LTS_GENERIC = 4,
  LTS_FAMILIES = 5
};

typedef struct {
  double budget_bytes; /* device bytes admitted requests may hold at once */
  int slots;           /* requests in flight at once, one stream each */
  int keep_outputs;    /* keep outputs in HBM until lts_release */
  int checksums;       /* checksum every output on the device */
  int dtype;           /* contraction operands: 0 bf16, 1 fp16 */
  int compiled_cap;    /* programs of no known family kept compiled, least
                          recently used dropped first (0: 256) */
} lts_config;

typedef struct {
  int id;
  int family;          /* LTS_* */
  int state;           /* 0 queued, 1 running, 2 done, -1 failed */
  double arrival_s;    /* seconds on the server clock */
  double start_s;
  double done_s;
  int64_t out_numel;
  int64_t flops;       /* algorithmic flops */
  int64_t bytes;       /* algorithmic HBM bytes (inputs read, output written) */
  uint64_t checksum;   /* of the output bits, when checksums are on */
  char plan[96];       /* kernels and tile of the plan */
} lts_request_info;

typedef struct {
  int64_t submitted, completed, failed;
  double wall_s;                 /* first arrival to last completion */
  double p50_ms, p99_ms, mean_ms, max_ms;
  double flops, bytes;           /* sums over completed requests */
  int64_t family_count[LTS_FAMILIES];
  int64_t struct_hits, struct_misses, exact_hits, compiles;
  double peak_bytes;             /* most device bytes admitted at once */
  int64_t admission_waits;       /* admission passes the head request waited for memory */
  double client_lag_ms;          /* mean lateness of submissions (workload driver) */
  double host_plan_ms;           /* host time: structure keys and plans (submit) */
  double host_launch_ms;         /* host time: allocation and kernel launches */
  double host_retire_ms;         /* host time: completion polling and retirement */
  int64_t device_allocs;         /* device allocations made (blocks are reused) */
  double cached_bytes;           /* idle device bytes held for reuse */
  int64_t compiled_evictions;    /* compiled programs dropped from the cache */
} lts_stats;

typedef struct {
  double rate;         /* Poisson arrivals per second */
  double duration_s;   /* arrivals stop after this */
  uint64_t seed;
  double mix[4];       /* fractions of matmul, softmax, conv2d, max-pool */
  int max_requests;    /* 0 = unbounded */
} lts_workload;

const char *lts_last_error(void);
lts_server *lts_create(const lts_config *cfg);
void lts_free(lts_server *s);
/* structural key of a lazy program (sizes masked) and the family it maps to */
uint64_t lts_structural_hash(const ltc_tensor *program);
/* structural key of family f's template program (LTS_MATMUL .. LTS_MAXPOOL) */
uint64_t lts_template_hash(int family);
int lts_family(lts_server *s, const ltc_tensor *program);
/* queue a program: input i is uniform [-1, 1) * scales[i] from seeds[i]
 * (inputs in the order of the lowered IR's inputs).  Returns the request id */
int lts_submit(lts_server *s, const ltc_tensor *program, const uint64_t *seeds,
               const float *scales, int n_inputs);
/* one admission and completion pass; returns requests queued or running */
int lts_poll(lts_server *s);
/* poll until every submitted request has completed */
int lts_drain(lts_server *s);
int lts_info(lts_server *s, int id, lts_request_info *out);
/* copy a kept output to host memory (n floats) */
int lts_output(lts_server *s, int id, float *host, int64_t n);
int lts_release(lts_server *s, int id);
int lts_stats_get(lts_server *s, lts_stats *out);
int lts_stats_reset(lts_server *s);
/* open-loop synthetic workload: a client thread builds the programs and
 * submits them at their Poisson arrival times while this thread serves */
int lts_run_workload(lts_server *s, const lts_workload *w, lts_stats *out);

#pragma GCC visibility pop

#ifdef __cplusplus
}
#endif

#endif /* LOOP_TOOL_SERVE_H_ */
