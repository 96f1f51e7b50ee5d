// Thin C ABI over the HIP runtime and hipRTC (see include/lt_hip.h).
#include "lt_hip.h"

#include <hip/hip_runtime_api.h>
#include <hip/hiprtc.h>

#include <mutex>

#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace {

thread_local std::string g_last_error;

int fail(const std::string &msg) {
  g_last_error = msg;
  return 1;
}

int check(hipError_t e, const char *what) {
  if (e == hipSuccess) return 0;
  return fail(std::string(what) + ": " + hipGetErrorString(e));
}

int check_rtc(hiprtcResult r, const char *what) {
  if (r == HIPRTC_SUCCESS) return 0;
  return fail(std::string(what) + ": " + hiprtcGetErrorString(r));
}

}  // namespace

struct lt_hip_module_s {
  hipModule_t mod;
};

extern "C" {

const char *lt_hip_last_error(void) { return g_last_error.c_str(); }

int lt_hip_init(void) {
  int n = 0;
  if (check(hipGetDeviceCount(&n), "hipGetDeviceCount")) return 1;
  if (n < 1) return fail("no HIP device visible");
  return check(hipSetDevice(0), "hipSetDevice");
}

int lt_hip_device_count(int *count) {
  return check(hipGetDeviceCount(count), "hipGetDeviceCount");
}

int lt_hip_device_props(int device, lt_hip_props *out) {
  hipDeviceProp_t p;
  if (check(hipGetDeviceProperties(&p, device), "hipGetDeviceProperties")) return 1;
  std::memset(out, 0, sizeof(*out));
  std::strncpy(out->name, p.name, sizeof(out->name) - 1);
  std::strncpy(out->arch, p.gcnArchName, sizeof(out->arch) - 1);
  out->compute_units = p.multiProcessorCount;
  out->max_threads_per_block = p.maxThreadsPerBlock;
  out->warp_size = p.warpSize;
  out->clock_khz = p.clockRate;
  out->memory_clock_khz = p.memoryClockRate;
  out->memory_bus_width = p.memoryBusWidth;
  out->l2_cache_bytes = p.l2CacheSize;
  out->lds_bytes_per_block = (int)p.sharedMemPerBlock;
  out->total_global_mem = p.totalGlobalMem;
  return 0;
}

int lt_hip_compile(const char *source, const char *name, const char **options,
                   int num_options, void **code, size_t *code_size) {
  hiprtcProgram prog;
  if (check_rtc(hiprtcCreateProgram(&prog, source, name, 0, nullptr, nullptr),
                "hiprtcCreateProgram"))
    return 1;
  std::vector<const char *> opts = {"--offload-arch=gfx950", "-O3", "-ffp-contract=off",
                                    "-Xclang", "-target-feature", "-Xclang",
                                    "-packed-fp32-ops", "-std=c++17"};
  for (int i = 0; i < num_options; ++i) opts.push_back(options[i]);
  auto r = hiprtcCompileProgram(prog, (int)opts.size(), opts.data());
  if (r != HIPRTC_SUCCESS) {
    size_t log_size = 0;
    hiprtcGetProgramLogSize(prog, &log_size);
    std::string log(log_size + 1, '\0');
    hiprtcGetProgramLog(prog, &log[0]);
    hiprtcDestroyProgram(&prog);
    return fail(std::string("hipRTC compile of ") + name + " failed: " + log);
  }
  size_t size = 0;
  if (check_rtc(hiprtcGetCodeSize(prog, &size), "hiprtcGetCodeSize")) {
    hiprtcDestroyProgram(&prog);
    return 1;
  }
  void *buf = std::malloc(size);
  if (!buf) {
    hiprtcDestroyProgram(&prog);
    return fail("out of host memory for code object");
  }
  if (check_rtc(hiprtcGetCode(prog, (char *)buf), "hiprtcGetCode")) {
    std::free(buf);
    hiprtcDestroyProgram(&prog);
    return 1;
  }
  hiprtcDestroyProgram(&prog);
  *code = buf;
  *code_size = size;
  return 0;
}

void lt_hip_free_code(void *code) { std::free(code); }

int lt_hip_module_load(const void *code, lt_hip_module *module) {
  hipModule_t m;
  if (check(hipModuleLoadData(&m, code), "hipModuleLoadData")) return 1;
  *module = new lt_hip_module_s{m};
  return 0;
}

int lt_hip_module_get_function(lt_hip_module module, const char *name, void **function) {
  hipFunction_t f;
  if (check(hipModuleGetFunction(&f, module->mod, name), "hipModuleGetFunction"))
    return 1;
  *function = (void *)f;
  return 0;
}

int lt_hip_module_unload(lt_hip_module module) {
  int rc = check(hipModuleUnload(module->mod), "hipModuleUnload");
  delete module;
  return rc;
}

int lt_hip_launch(void *function, unsigned gx, unsigned gy, unsigned gz, unsigned bx,
                  unsigned by, unsigned bz, unsigned shared_bytes, void *stream,
                  void **args) {
  return check(hipModuleLaunchKernel((hipFunction_t)function, gx, gy, gz, bx, by, bz,
                                     shared_bytes, (hipStream_t)stream, args, nullptr),
               "hipModuleLaunchKernel");
}

int lt_hip_max_resident_blocks(void *function, unsigned block, unsigned shared_bytes,
                               int *blocks) {
  int per_cu = 0, dev = 0, cus = 0;
  if (int rc = check(hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(
                         &per_cu, (hipFunction_t)function, (int)block, shared_bytes),
                     "hipModuleOccupancyMaxActiveBlocksPerMultiprocessor"))
    return rc;
  if (int rc = check(hipGetDevice(&dev), "hipGetDevice")) return rc;
  if (int rc = check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev),
                     "hipDeviceGetAttribute"))
    return rc;
  *blocks = per_cu * cus;
  return 0;
}

int lt_hip_malloc(void **ptr, size_t bytes) {
  return check(hipMalloc(ptr, bytes ? bytes : 1), "hipMalloc");
}
int lt_hip_free(void *ptr) { return check(hipFree(ptr), "hipFree"); }
int lt_hip_host_malloc(void **ptr, size_t bytes) {
  return check(hipHostMalloc(ptr, bytes ? bytes : 1, hipHostMallocDefault),
               "hipHostMalloc");
}
int lt_hip_host_free(void *ptr) { return check(hipHostFree(ptr), "hipHostFree"); }

static hipMemcpyKind kind_of(int kind) {
  switch (kind) {
    case 0: return hipMemcpyHostToDevice;
    case 1: return hipMemcpyDeviceToHost;
    case 2: return hipMemcpyDeviceToDevice;
    default: return hipMemcpyDefault;
  }
}

int lt_hip_memcpy(void *dst, const void *src, size_t bytes, int kind) {
  return check(hipMemcpy(dst, src, bytes, kind_of(kind)), "hipMemcpy");
}
int lt_hip_memcpy_async(void *dst, const void *src, size_t bytes, int kind, void *stream) {
  return check(hipMemcpyAsync(dst, src, bytes, kind_of(kind), (hipStream_t)stream),
               "hipMemcpyAsync");
}
int lt_hip_malloc_async(void **ptr, size_t bytes, void *stream) {
  static std::once_flag once;
  std::call_once(once, [] {
    hipMemPool_t pool;
    int dev = 0;
    if (hipGetDevice(&dev) == hipSuccess && hipDeviceGetDefaultMemPool(&pool, dev) == hipSuccess) {
      uint64_t keep = UINT64_MAX;  // never trim the pool at synchronisation points
      (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
    }
    (void)hipGetLastError();
  });
  return check(hipMallocAsync(ptr, bytes, (hipStream_t)stream), "hipMallocAsync");
}

int lt_hip_free_async(void *ptr, void *stream) {
  return check(hipFreeAsync(ptr, (hipStream_t)stream), "hipFreeAsync");
}

int lt_hip_memset(void *dst, int value, size_t bytes) {
  return check(hipMemset(dst, value, bytes), "hipMemset");
}

int lt_hip_is_device_pointer(const void *ptr) {
  if (!ptr) return 0;
  hipPointerAttribute_t attr;
  if (hipPointerGetAttributes(&attr, ptr) != hipSuccess) {
    (void)hipGetLastError();  // unregistered host memory reports an error
    return 0;
  }
  return attr.type == hipMemoryTypeDevice ? 1 : 0;
}

int lt_hip_mem_info(size_t *free_bytes, size_t *total_bytes) {
  return check(hipMemGetInfo(free_bytes, total_bytes), "hipMemGetInfo");
}

int lt_hip_stream_create(void **stream) {
  hipStream_t s;
  if (check(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate"))
    return 1;
  *stream = s;
  return 0;
}
int lt_hip_stream_destroy(void *stream) {
  return check(hipStreamDestroy((hipStream_t)stream), "hipStreamDestroy");
}
int lt_hip_stream_synchronize(void *stream) {
  return check(hipStreamSynchronize((hipStream_t)stream), "hipStreamSynchronize");
}
int lt_hip_device_synchronize(void) {
  return check(hipDeviceSynchronize(), "hipDeviceSynchronize");
}
int lt_hip_event_create(void **event) {
  hipEvent_t e;
  if (check(hipEventCreate(&e), "hipEventCreate")) return 1;
  *event = e;
  return 0;
}
int lt_hip_event_destroy(void *event) {
  return check(hipEventDestroy((hipEvent_t)event), "hipEventDestroy");
}
int lt_hip_event_record(void *event, void *stream) {
  return check(hipEventRecord((hipEvent_t)event, (hipStream_t)stream), "hipEventRecord");
}
int lt_hip_stream_wait_event(void *stream, void *event) {
  return check(hipStreamWaitEvent((hipStream_t)stream, (hipEvent_t)event, 0),
               "hipStreamWaitEvent");
}
int lt_hip_event_synchronize(void *event) {
  return check(hipEventSynchronize((hipEvent_t)event), "hipEventSynchronize");
}
int lt_hip_event_query(void *event) {
  hipError_t e = hipEventQuery((hipEvent_t)event);
  if (e == hipSuccess) return 1;
  if (e == hipErrorNotReady) {
    (void)hipGetLastError();
    return 0;
  }
  return check(e, "hipEventQuery");
}
int lt_hip_event_elapsed_ms(void *start, void *stop, float *ms) {
  return check(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop),
               "hipEventElapsedTime");
}
int lt_hip_graph_begin_capture(void *stream) {
  return check(hipStreamBeginCapture((hipStream_t)stream, hipStreamCaptureModeThreadLocal),
               "hipStreamBeginCapture");
}
int lt_hip_graph_end_capture(void *stream, void **graph_exec) {
  hipGraph_t g;
  if (check(hipStreamEndCapture((hipStream_t)stream, &g), "hipStreamEndCapture")) return 1;
  hipGraphExec_t ex;
  int rc = check(hipGraphInstantiate(&ex, g, nullptr, nullptr, 0), "hipGraphInstantiate");
  hipGraphDestroy(g);
  if (rc) return rc;
  *graph_exec = ex;
  return 0;
}
int lt_hip_graph_launch(void *graph_exec, void *stream) {
  return check(hipGraphLaunch((hipGraphExec_t)graph_exec, (hipStream_t)stream),
               "hipGraphLaunch");
}
int lt_hip_graph_destroy(void *graph_exec) {
  return check(hipGraphExecDestroy((hipGraphExec_t)graph_exec), "hipGraphExecDestroy");
}

}  // extern "C"
