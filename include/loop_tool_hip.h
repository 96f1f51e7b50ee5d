/*
 * loop_tool_hip.h -- C ABI extensions of the "hip" backend's Compiled objects
 * (handles from loop_tool_c.h compiled with backend "hip" or "hip_bf16").
 *
 * No counterpart in the reference: its Compiled::run(memory, sync)
 * (include/loop_tool/backend.h:18-22) is the only launch entry point.  These
 * calls let callers that keep tensors resident in HBM launch a compiled nest
 * on their own stream with their own scratch (serving, graph capture,
 * benchmarks) and read the plan's per-stage accounting.
 */
#ifndef LOOP_TOOL_HIP_H_
#define LOOP_TOOL_HIP_H_

#include <stdint.h>

#include "loop_tool_c.h"

#ifdef __cplusplus
extern "C" {
#endif

#pragma GCC visibility push(default)

/* 1 if the Compiled was produced by the hip backend */
int lth_is_hip(const ltc_compiled *cc);
/* the calling thread's stream of the Compiled (created on the thread's first
 * use): ltc_compiled_run launches there, with that thread's own scratch, so
 * runs from different threads proceed concurrently */
int lth_compiled_stream(const ltc_compiled *cc, void **stream);
/* number of plan stages (kernel launches per run) */
int lth_num_stages(const ltc_compiled *cc);
/* kernel name, kind, algorithmic HBM bytes and flops of stage i */
int lth_stage_info(const ltc_compiled *cc, int i, char *name, int name_cap, char *kind,
                   int kind_cap, int64_t *bytes, int64_t *flops);
/* number and sizes (float elements) of the scratch buffers a launch needs */
int lth_scratch_sizes(const ltc_compiled *cc, int64_t *out, int cap);
/* launch on `stream` with device buffers io (inputs then outputs) and
 * scratch; no allocation, copy or synchronisation */
int lth_run_on(const ltc_compiled *cc, void *const *io, int n_io, void *const *scratch,
               int n_scratch, void *stream);
/* time `iters` back-to-back launches on the Compiled's stream with HIP
 * events after `warmup` untimed ones; io are device buffers.  Writes the
 * mean milliseconds per launch of the whole plan to *ms */
int lth_time(const ltc_compiled *cc, void *const *io, int n_io, int warmup, int iters,
             float *ms);
/* as lth_time for stage `stage` alone (after one full run, so the stage
 * reads the scratch contents a full run leaves): mean ms per launch */
int lth_time_stage(const ltc_compiled *cc, void *const *io, int n_io, int stage, int warmup,
                   int iters, float *ms);

/* ---- tuned-schedule store.  No counterpart in the reference, whose
 * schedules live only in memory; the file format is the IR text of
 * serialize/deserialize (include/loop_tool/serialization.h).  The hip
 * backend's compile looks every program up here and, when a schedule of the
 * same program is stored, lowers that schedule instead (the Compiled's
 * string property "tuned_schedule" then holds the key, and its dump shows
 * the schedule).  One IR text file per program under <dir>/gfx950/<key>.ir;
 * <dir> is the last lth_tuned_set_dir, else $LT_HIP_TUNED_DIR, else
 * $HOME/.cache/loop_tool_amd/tuned.  LT_HIP_TUNED=0 turns lookups off. */
int lth_tuned_set_dir(const char *dir);
/* the program's key (hex): its IR text with var names replaced by positions,
 * hashed */
int lth_schedule_key(const ltc_looptree *program, char *buf, int cap);
/* store `schedule` (a LoopTree of the same program) as the program's tuned
 * schedule; refused when nodes, vars or var sizes differ */
int lth_tuned_put(const ltc_looptree *program, const ltc_looptree *schedule);
/* the stored schedule of the program, or NULL when there is none */
ltc_looptree *lth_tuned_get(const ltc_looptree *program);

/* ---- allocation-free launch.  No counterpart in the reference: for callers
 * that cannot allocate device memory -- code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc.  As lth_run_on, but the
 * plan's scratch buffers come from one static device array of
 * LTH_STATIC_SCRATCH_BYTES per GPU (held from the first call on that GPU
 * until the process ends); no device memory is allocated or freed and
 * nothing is copied or synchronised.  The Compiled's code must already be
 * loaded (any earlier ltc_compiled_run, or lth_compiled_stream).  Plans
 * needing more scratch are rejected (return -1, lth_last_error()).  Calls
 * on one GPU share its array: calls that overlap on one GPU race. */
#define LTH_STATIC_SCRATCH_BYTES (64u << 20)
int lth_run_static(const ltc_compiled *cc, void *const *io, int n_io, void *stream);

/* ---- packs of compiled nests in one graph.  No counterpart in the
 * reference, which runs one Compiled at a time (backend.h:18-22).  Any hip
 * Compiled plans (pointwise, row reductions, contractions, convolutions,
 * several stages each) with their device buffers (inputs then outputs of
 * each nest, concatenated: n_io[i] pointers for nest i) are captured into
 * one hipGraph: the pack allocates every nest's scratch once, spreads the
 * nests over `lanes` (1..16) streams forked from `stream` and joined back
 * (longest first, each to the lane with the least estimated time, from the
 * plans' algorithmic bytes and flops), so independent nests run
 * concurrently, and a replay is one graph launch.  The nests must not write
 * a buffer another nest of the pack reads or writes.  `stream` must be a
 * created stream; replays of one pack must not overlap.
 * With LTH_PACK_GROUP_MATMULS in `flags`, the nests that are plain 16-bit
 * matmuls (C[M][N] = A[M][K] B[K][N] over row-major float buffers, compiled
 * with bf16 or fp16 operands, K a multiple of 8) run instead as their
 * operands' conversions plus one persistent grouped launch per operand type
 * (the lth_batch kernels, 128x128x64 tiles): the same operand rounding, but
 * the grouped tiles' accumulation order, so their outputs agree with each
 * nest's own run to float32 rounding rather than bit for bit. */
#define LTH_PACK_GROUP_MATMULS 1
typedef struct lth_pack lth_pack;
lth_pack *lth_pack_create(const ltc_compiled *const *cc, void *const *io, const int *n_io,
                          int n, int lanes, int flags, void *stream);
int lth_pack_replay(lth_pack *p, void *stream);
void lth_pack_free(lth_pack *p);

/* ---- batches of run-time-sized matmul nests (C = A B, 16-bit operands,
 * float32 accumulation and output).  Operands live in HBM as A[M][ld] and the
 * transposed Bt[N][ld], ld = LTH_BATCH_PITCH(K) (K rounded up to a multiple
 * of 64, so every row starts on a 128-byte line; columns K..ld-1 are never
 * read), as lth_convert writes them; C is [M][N].  Tile configurations 0..6
 * are 64x64x32, 128x64x32, 64x128x32, 128x128x64, 256x128x32, 128x256x32 and
 * 128x128x32 (rows x cols of C x K tile). */
#define LTH_BATCH_PITCH(k) (((k) + 63) / 64 * 64)
typedef struct lth_batch lth_batch;
const char *lth_batch_last_error(void);
/* dtype: 0 bf16, 1 fp16; K must be a multiple of 8 */
lth_batch *lth_batch_create(int n, const int *M, const int *N, const int *K, int dtype);
int lth_batch_set_operands(lth_batch *b, void *const *A, void *const *Bt, void *const *C);
/* tile configuration per GEMM (as chosen per shape bucket by the autotuner) */
int lth_batch_set_configs(lth_batch *b, const int *cfg);
/* number of output tiles in the grouped work list */
int lth_batch_num_items(lth_batch *b);
/* one launch per GEMM with tile config `cfg` (-1: each GEMM's own config) */
int lth_batch_run_plain(lth_batch *b, int cfg, void *stream);
/* one persistent launch of `grid` workgroups over every tile of the batch
 * (grid <= 0: as many as the GPU holds at once for the batch's kernel).  The
 * tiles are handed out through a counter owned by the batch, so grouped
 * launches (and graph replays) of one batch must not run concurrently. */
int lth_batch_run_grouped(lth_batch *b, int grid, void *stream);
/* resolve the grid of a grouped launch (grid <= 0: the occupancy-sized one)
 * and upload everything that launch reads; returns that grid (negative on
 * error).  A grouped launch with the returned grid then allocates and copies
 * nothing, so a caller can capture it into a graph of its own. */
int lth_batch_prepare(lth_batch *b, int grid);
/* capture the grouped launch into a hipGraph on `stream`, replay it later */
int lth_batch_capture(lth_batch *b, int grid, void *stream);
int lth_batch_replay(lth_batch *b, void *stream);
void lth_batch_free(lth_batch *b);
/* float32 [rows][cols] -> 16-bit operand [rows][LTH_BATCH_PITCH(cols)], or
 * transposed to [cols][LTH_BATCH_PITCH(rows)] if transpose != 0 (the padding
 * is not written), on `stream` */
int lth_convert(const float *src, void *dst, int rows, int cols, int transpose, int dtype,
                void *stream);
int lth_tile_config(int cfg, int *bm, int *bn);

#pragma GCC visibility pop

#ifdef __cplusplus
}
#endif

#endif /* LOOP_TOOL_HIP_H_ */
