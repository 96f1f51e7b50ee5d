// Allocation-free launch of a hip Compiled: lth_run_static (declared in
// include/loop_tool_hip.h).
//
// The plan's scratch buffers (staged contraction operands, materialised
// reductions) are carved out of one file-scope device array, `scratch`, that
// every GPU the library runs on holds from the first call on that device
// until the process ends.  Host code never uses the array's name as an
// address: on the first call on a device (that device current) its device
// address is looked up with hipGetSymbolAddress and kept per device; the
// kernels receive offsets into that address.
//
// Every call on one GPU shares that GPU's array, and the array keeps the
// last call's contents: each plan writes every scratch element it reads
// before reading it, so nothing needs clearing between calls, but calls that
// overlap on one GPU (different streams, or host threads) race on it.
#include <hip/hip_runtime.h>

#include <exception>
#include <mutex>
#include <string>
#include <vector>

#include "hip_backend.h"
#include "loop_tool_hip.h"

struct ltc_compiled {
  std::shared_ptr<loop_tool::Compiled> cc;
};

static __device__ __align__(256) unsigned char scratch[LTH_STATIC_SCRATCH_BYTES];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void *g_scratch[kMaxDevices];  // device address of `scratch`, per device

void check(hipError_t e, const char *what) {
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

// The array's address on the current device, looked up once per device.
void *scratch_address() {
  int dev = 0;
  check(hipGetDevice(&dev), "hipGetDevice");
  if (dev < 0 || dev >= kMaxDevices) throw std::runtime_error("device ordinal out of range");
  std::lock_guard<std::mutex> g(g_mu);
  if (!g_scratch[dev]) {
    void *p = nullptr;
    check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
    if (!p) throw std::runtime_error("hipGetSymbolAddress(scratch) returned no address");
    g_scratch[dev] = p;
  }
  return g_scratch[dev];
}

}  // namespace

namespace loop_tool {
namespace hip {
void set_lth_error(const std::string &msg);  // capi_hip.cpp: lth_last_error()
}  // namespace hip
}  // namespace loop_tool

extern "C" {

int lth_run_static(const ltc_compiled *cc, void *const *io, int n_io, void *stream) {
  try {
    auto *h = dynamic_cast<const loop_tool::hip::HipCompiled *>(cc ? cc->cc.get() : nullptr);
    if (!h) throw std::runtime_error("lth_run_static: not a hip Compiled");
    if (n_io < 0 || (size_t)n_io < h->num_io())
      throw std::runtime_error("lth_run_static: expected " + std::to_string(h->num_io()) +
                               " device buffers (inputs then outputs), got " +
                               std::to_string(n_io));
    // the plan's scratch buffers, each at a 256-byte aligned offset
    std::vector<size_t> offsets;
    size_t total = 0;
    for (int64_t n : h->scratch_sizes()) {
      offsets.push_back(total);
      total += ((size_t)n * sizeof(float) + 255) / 256 * 256;
    }
    if (total > (size_t)LTH_STATIC_SCRATCH_BYTES)
      throw std::runtime_error(
          "lth_run_static: the plan needs " + std::to_string(total) +
          " bytes of scratch; the static scratch holds " +
          std::to_string((size_t)LTH_STATIC_SCRATCH_BYTES) +
          " (use lth_run_on with caller-owned scratch)");
    if (!h->loaded())
      throw std::runtime_error(
          "lth_run_static: the plan's code is not loaded on the device (loading allocates): "
          "run it once with ltc_compiled_run or call lth_compiled_stream first");
    std::vector<void *> bufs;
    if (!offsets.empty()) {
      char *base = static_cast<char *>(scratch_address());
      for (size_t o : offsets) bufs.push_back(base + o);
    }
    h->run_on(std::vector<void *>(io, io + n_io), bufs, stream);
    return 0;
  } catch (const std::exception &e) {
    loop_tool::hip::set_lth_error(e.what());
  }
  return -1;
}

}  // extern "C"
