/*
 * lt_hip.h -- the thin C ABI between loop_tool's C++ "hip" backend and the
 * HIP runtime / hipRTC on AMD Instinct MI355X (gfx950).
 *
 * Every entry point returns 0 on success and a non-zero code on failure; the
 * failure text is kept per thread and read with lt_hip_last_error().  No C++
 * exception crosses this boundary; the C++ backend turns failures into
 * loop_tool ASSERT errors.
 *
 * Each group names the reference interface it replaces.  The reference's only
 * GPU path is its (deprecated, uncompilable) CUDA backend, which dlopens the
 * CUDA driver, runtime and NVRTC (src/backends/cuda/cuda.cpp, cited below);
 * this layer is its MI355X counterpart and has no CUDA-compatibility shims.
 */
#ifndef LT_HIP_H_
#define LT_HIP_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- errors --------------------------------------------------------------
 * Replaces CUDA_SAFE_CALL / NVRTC_SAFE_CALL error reporting
 * (src/backends/cuda/cuda.cpp:20-31). */
const char *lt_hip_last_error(void);

/* ---- device ----------------------------------------------------------------
 * Replaces cuInit/cuDeviceGet/cuCtxCreate (cuda.cpp:861-873),
 * cuDeviceGetCount (cuda.cpp:796) and cuDeviceGetAttribute
 * (cuda.cpp:912-915, 960-969, 1020-1026). */
typedef struct lt_hip_props {
  char name[256];
  char arch[64];            /* e.g. "gfx950:sramecc+:xnack-" */
  int compute_units;
  int max_threads_per_block;
  int warp_size;            /* 64 on CDNA */
  int clock_khz;
  int memory_clock_khz;
  int memory_bus_width;     /* bits */
  int l2_cache_bytes;
  int lds_bytes_per_block;
  size_t total_global_mem;
} lt_hip_props;

int lt_hip_init(void);
int lt_hip_device_count(int *count);
int lt_hip_device_props(int device, lt_hip_props *out);

/* ---- runtime compilation ---------------------------------------------------
 * Replaces nvrtcCreateProgram/CompileProgram/GetPTX (cuda.cpp:893-950).
 * Compiles HIP C++ source to a gfx950 code object with hipRTC.  `options`
 * are appended after the defaults (--offload-arch=gfx950 -O3
 * -ffp-contract=off, packed FP32 disabled).  On success *code is malloc'd and
 * must be released with lt_hip_free_code(). On failure the compiler log is
 * the last-error text. */
int lt_hip_compile(const char *source, const char *name, const char **options,
                   int num_options, void **code, size_t *code_size);
void lt_hip_free_code(void *code);

/* ---- modules ---------------------------------------------------------------
 * Replaces cuModuleLoadDataEx / cuModuleGetFunction (cuda.cpp:951-958). */
typedef struct lt_hip_module_s *lt_hip_module;
int lt_hip_module_load(const void *code, lt_hip_module *module);
int lt_hip_module_get_function(lt_hip_module module, const char *name, void **function);
int lt_hip_module_unload(lt_hip_module module);

/* ---- launch ----------------------------------------------------------------
 * Replaces cuLaunchKernel + cuCtxSynchronize (cuda.cpp:837-850).  `args` is
 * the usual array of pointers to each kernel argument. stream 0 = null
 * stream. */
int lt_hip_launch(void *function, unsigned grid_x, unsigned grid_y, unsigned grid_z,
                  unsigned block_x, unsigned block_y, unsigned block_z,
                  unsigned shared_bytes, void *stream, void **args);

/* Workgroups of `function` (with `block` threads and `shared_bytes` of
 * dynamic LDS) that the current device holds at once: per-CU occupancy times
 * the CU count.  No counterpart in the reference, whose CUDA backend sized
 * its 1-D grids by the nest alone (cuda.cpp:776-792); persistent launches
 * here size their grids with it. */
int lt_hip_max_resident_blocks(void *function, unsigned block, unsigned shared_bytes,
                               int *blocks);

/* ---- memory ----------------------------------------------------------------
 * Replaces cudaMallocManaged / cudaFree (cuda.cpp:800-816).  Device memory is
 * plain hipMalloc (HBM); host staging is done by the caller with the copies
 * below.  kind: 0 host->device, 1 device->host, 2 device->device,
 * 3 default (runtime infers). */
int lt_hip_malloc(void **ptr, size_t bytes);
int lt_hip_free(void *ptr);
int lt_hip_host_malloc(void **ptr, size_t bytes); /* pinned host memory */
int lt_hip_host_free(void *ptr);
int lt_hip_memcpy(void *dst, const void *src, size_t bytes, int kind);
int lt_hip_memcpy_async(void *dst, const void *src, size_t bytes, int kind, void *stream);
int lt_hip_memset(void *dst, int value, size_t bytes);
/* 1 if `ptr` is device (HBM) memory of this process, 0 otherwise */
int lt_hip_is_device_pointer(const void *ptr);
int lt_hip_mem_info(size_t *free_bytes, size_t *total_bytes);
/* stream-ordered allocation from the device's default memory pool, which
 * keeps freed blocks cached for reuse (no driver call on the steady state) */
int lt_hip_malloc_async(void **ptr, size_t bytes, void *stream);
int lt_hip_free_async(void *ptr, void *stream);

/* ---- streams, events, graphs -----------------------------------------------
 * No counterpart in the reference (it launched on the null stream and
 * synchronized the context after every launch, cuda.cpp:843-849).  Streams
 * are always created at the default priority. */
int lt_hip_stream_create(void **stream);
int lt_hip_stream_destroy(void *stream);
int lt_hip_stream_synchronize(void *stream);
int lt_hip_device_synchronize(void);
int lt_hip_event_create(void **event);
int lt_hip_event_destroy(void *event);
int lt_hip_event_record(void *event, void *stream);
int lt_hip_event_synchronize(void *event);
/* 1 if every work recorded before the event has completed, 0 if not yet */
int lt_hip_event_query(void *event);
int lt_hip_event_elapsed_ms(void *start, void *stop, float *ms);
/* work enqueued on `stream` after this call waits for `event` */
int lt_hip_stream_wait_event(void *stream, void *event);
int lt_hip_graph_begin_capture(void *stream);
int lt_hip_graph_end_capture(void *stream, void **graph_exec);
int lt_hip_graph_launch(void *graph_exec, void *stream);
int lt_hip_graph_destroy(void *graph_exec);

#ifdef __cplusplus
}
#endif

#endif /* LT_HIP_H_ */
