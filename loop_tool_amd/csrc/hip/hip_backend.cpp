// The "hip" backend: loop_tool::Backend / Compiled for AMD Instinct MI355X.
//
// Registered beside the reference's CPU backends through the same plugin
// interface (include/loop_tool/backend.h:52-92; registration pattern of
// src/backends/cpu/cpp.cpp:505-514).  compile_impl() lowers the LoopTree to a
// launch plan (emit.cpp), generates one HIP module for it and compiles it for
// gfx950 with hipRTC through the C ABI of include/lt_hip.h; run() launches the
// plan's kernels on the Compiled's stream.
//
// Buffers handed to run() may be device (HBM) pointers, used in place, or
// host pointers, staged through device buffers owned by the Compiled (copied
// in before and out after the launch); either way the call is synchronous
// when staging is involved.
//
// Knobs are read from LoopTree annotations (loop or node), which the
// reference C backend ignores, so an annotated tree still runs on the oracle:
//   hip:plain      only the exact pointwise lowering
//   mfma_bf16      bf16 operands, fp32 accumulation for contractions
//   mfma_f16       fp16 operands, fp32 accumulation for contractions
//   block=N        threads per workgroup of pointwise stages
//   vec=N          elements per lane for row reductions
//   tile=MxN       workgroup tile of contractions
//   bk=N           K tile (32 or 64) of 16-bit contractions
//   srows=N        rows of a source plane per workgroup when contraction
//                  operands are staged through LDS plane tiles
//   contract=gather  contractions gather their operands in the tile loader
//                  instead of staging them
//   halo=direct|staged  the halo tile reads the activations' float source
//                  itself (rounding to 16 bits as rows go to LDS; only the
//                  weights are staged) or a staged 16-bit copy (default)
//   overlap=0|1    consecutive runs of a plan whose first stage stages
//                  operands into scratch overlap that stage with the
//                  previous run's main kernels (two scratch sets, a second
//                  stream; default off)
//   ksplit=0|1     gathered contractions: 32x32 tiles with each K tile split
//                  over the workgroup's four waves (default: on for nests
//                  under 64 tiles of 64x64)
//   streamk=0|1    staged float contractions: stream-K work split over 512
//                  workgroups off / on (default: on when the tiles fill the
//                  workgroup slots unevenly)
//   contract=staged  float contractions are staged (and run the pipelined
//                  float tile) whatever their size; by default only nests of
//                  2^28 or more multiply-adds are
//   igemm=plain    staged 16-bit contractions use the single-stage tile
//                  (default: the software-pipelined tile where it applies)
//   waves=2|4|8    waves per workgroup of the pipelined tile
//   lds=dl|swz|pad LDS of the pipelined 16-bit tile: filled by direct-to-LDS
//                  loads into XOR-permuted rows (dl, the default outside the
//                  halo form), or through registers into XOR-permuted (swz)
//                  or 16-byte padded (pad) rows
//   stage=blocked|flat  16-bit contraction operands staged k-blocked
//                  ([K/BK][...][BK]) or with the whole k_in dim innermost
//   kb=outer|inner order of the K blocks among the outer-K positions
//   igemm=halo|halo1|pipe  staged 16-bit contractions whose outer-K positions
//                  shift the Q rows (convolution taps) load those rows once
//                  per K block into an LDS halo, double-buffered (halo) or
//                  single (halo1, the default where it applies), or per K
//                  tile (pipe)
#include <atomic>
#include <cstdlib>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <list>
#include <mutex>
#include <set>
#include <sstream>
#include <unordered_map>

#include <unistd.h>

#include "emit.h"
#include "loop_tool/backend.h"
#include "loop_tool/hardware.h"
#include "lt_hip.h"
#include "nest.h"
#include "hip_backend.h"
#include "tuned_store.h"

namespace loop_tool {
namespace hip {

namespace {

void hip_check(int rc, const char *what) {
  ASSERT(rc == 0) << what << ": " << lt_hip_last_error();
}

std::mutex &runtime_mutex() {
  static std::mutex m;
  return m;
}

bool &runtime_ready() {
  static bool ready = false;
  return ready;
}

// Set by an exit handler registered right after the HIP runtime initialises,
// so it runs before the runtime's own teardown: device resources still held
// by static caches at process exit are then left to the driver instead of
// being released into a runtime that is already gone.
std::atomic<bool> g_exiting{false};
void mark_exiting() { g_exiting = true; }

uint64_t fnv1a(const std::string &s) {
  uint64_t h = 1469598103934665603ULL;
  for (unsigned char c : s) {
    h ^= c;
    h *= 1099511628211ULL;
  }
  return h;
}

// In-process and on-disk cache of code objects, keyed by module source.  The
// in-process map keeps the most recently used kMemEntries; a disk entry
// stores its source beside the code and is used only when the source
// matches (the file name is a hash of it).
struct CodeCache {
  static constexpr size_t kMemEntries = 512;
  using Code = std::shared_ptr<std::vector<char>>;
  std::mutex mu;
  std::list<std::string> lru;  // sources, most recently used first
  std::unordered_map<std::string, std::pair<Code, std::list<std::string>::iterator>> mem;
  std::string dir;
  CodeCache() {
    if (const char *d = std::getenv("LT_HIP_CACHE_DIR")) {
      dir = d;
    } else if (const char *h = std::getenv("HOME")) {
      dir = std::string(h) + "/.cache/loop_tool_amd";
    }
    std::error_code ec;
    if (!dir.empty() && !std::filesystem::create_directories(dir, ec) && ec) dir.clear();
  }
  // callers hold mu
  Code find(const std::string &source) {
    auto it = mem.find(source);
    if (it == mem.end()) return nullptr;
    lru.splice(lru.begin(), lru, it->second.second);
    return it->second.first;
  }
  void put(const std::string &source, Code code) {
    if (find(source)) return;
    lru.push_front(source);
    mem[source] = {std::move(code), lru.begin()};
    while (mem.size() > kMemEntries) {
      mem.erase(lru.back());
      lru.pop_back();
    }
  }
};

CodeCache &code_cache() {
  static CodeCache c;
  return c;
}

}  // namespace

void ensure_runtime() {
  std::lock_guard<std::mutex> g(runtime_mutex());
  if (runtime_ready()) return;
  hip_check(lt_hip_init(), "HIP runtime initialisation (is a GPU visible?)");
  std::atexit(mark_exiting);
  runtime_ready() = true;
}

namespace {
std::string disk_path(const CodeCache &cc, const std::string &source) {
  if (cc.dir.empty()) return "";
  std::stringstream p;
  p << cc.dir << "/" << std::hex << fnv1a(source) << "_" << source.size() << ".co";
  return p.str();
}
}  // namespace

void evict_code(const std::string &source) {
  auto &cc = code_cache();
  {
    std::lock_guard<std::mutex> g(cc.mu);
    auto it = cc.mem.find(source);
    if (it != cc.mem.end()) {
      cc.lru.erase(it->second.second);
      cc.mem.erase(it);
    }
  }
  auto path = disk_path(cc, source);
  if (!path.empty()) std::remove(path.c_str());
}

std::shared_ptr<std::vector<char>> compile_source(const std::string &source,
                                                  const std::string &name,
                                                  bool *from_cache) {
  auto &cc = code_cache();
  {
    std::lock_guard<std::mutex> g(cc.mu);
    if (auto code = cc.find(source)) {
      if (from_cache) *from_cache = true;
      return code;
    }
  }
  // disk entry: [source size: 8 bytes][source][code object]
  const std::string path = disk_path(cc, source);
  if (!path.empty()) {
    std::ifstream f(path, std::ios::binary);
    uint64_t n = 0;
    if (f && f.read((char *)&n, sizeof(n)) && n == source.size()) {
      std::string stored(n, '\0');
      if (f.read(&stored[0], (std::streamsize)n) && stored == source) {
        auto buf = std::make_shared<std::vector<char>>((std::istreambuf_iterator<char>(f)),
                                                       std::istreambuf_iterator<char>());
        if (!buf->empty()) {
          std::lock_guard<std::mutex> g(cc.mu);
          cc.put(source, buf);
          if (from_cache) *from_cache = true;
          return buf;
        }
      }
    }
  }
  void *code = nullptr;
  size_t size = 0;
  hip_check(lt_hip_compile(source.c_str(), name.c_str(), nullptr, 0, &code, &size),
            "hipRTC");
  auto buf = std::make_shared<std::vector<char>>((char *)code, (char *)code + size);
  lt_hip_free_code(code);
  if (!path.empty()) {
    // written under a name of this process and call, published by rename
    // only when every byte reached the file
    static std::atomic<uint64_t> serial{0};
    const std::string tmp = path + ".tmp" + std::to_string((long)getpid()) + "_" +
                            std::to_string(serial++);
    bool ok;
    {
      std::ofstream f(tmp, std::ios::binary);
      const uint64_t n = source.size();
      f.write((const char *)&n, sizeof(n));
      f.write(source.data(), (std::streamsize)n);
      f.write(buf->data(), (std::streamsize)buf->size());
      f.close();
      ok = f.good();
    }
    if (!ok || std::rename(tmp.c_str(), path.c_str()) != 0) std::remove(tmp.c_str());
  }
  std::lock_guard<std::mutex> g(cc.mu);
  cc.put(source, buf);
  if (from_cache) *from_cache = false;
  return buf;
}

Knobs knobs_from(const LoopTree &lt, bool bf16_variant) {
  Knobs k;
  k.bf16 = bf16_variant;
  // "key=value" tokens whose key is a knob of this backend must carry a
  // valid value; other tokens are annotations for other consumers
  auto number = [](const std::string &tok, const std::string &v) {
    ASSERT(!v.empty() && v.size() < 6 && v.find_first_not_of("0123456789") == std::string::npos)
        << "knob " << tok << ": expected a number";
    return std::stoi(v);
  };
  auto scan = [&](const std::string &annot) {
    std::stringstream ss(annot);
    std::string tok;
    while (ss >> tok) {
      if (tok == "hip:plain") k.force_plain = true;
      else if (tok == "mfma_bf16") k.bf16 = true;
      else if (tok == "mfma_f16") k.f16 = true;
      auto eq = tok.find('=');
      if (eq == std::string::npos) continue;
      const std::string key = tok.substr(0, eq), v = tok.substr(eq + 1);
      auto bad = [&]() { ASSERT(0) << "knob " << tok << ": unknown value"; };
      if (key == "contract") {
        if (v == "gather") k.gather = true;
        else if (v == "staged") k.staged = true;
        else bad();
      } else if (key == "streamk") {
        if (v == "0" || v == "1") k.streamk = v == "1";
        else bad();
      } else if (key == "halo") {
        if (v == "direct" || v == "staged") k.halo_direct = v == "direct";
        else bad();
      } else if (key == "overlap") {
        if (v == "0" || v == "1") k.overlap = v == "1";
        else bad();
      } else if (key == "ksplit") {
        if (v == "0" || v == "1") k.ksplit = v == "1";
        else bad();
      } else if (key == "igemm") {
        if (v == "plain") k.igemm_plain = true;
        else if (v == "halo") k.igemm_halo = 1;
        else if (v == "halo1") k.igemm_halo = 2;
        else if (v == "pipe") k.igemm_halo = 0;
        else bad();
      } else if (key == "lds") {
        if (v == "swz" || v == "pad") k.lds_swz = v == "swz";
        else if (v == "dl") k.lds_swz = 2;
        else bad();
      } else if (key == "stage") {
        if (v == "blocked" || v == "flat") k.stage_blocked = v == "blocked";
        else bad();
      } else if (key == "kb") {
        if (v == "outer" || v == "inner") k.kb_outer = v == "outer";
        else bad();
      } else if (key == "waves") {
        k.waves = number(tok, v);
      } else if (key == "srows") {
        k.stage_rows = number(tok, v);
        ASSERT(k.stage_rows >= 1 && k.stage_rows <= 64)
            << "knob " << tok << ": srows=N must be in [1, 64]";
      } else if (key == "bk") {
        k.bk = number(tok, v);
        ASSERT(k.bk == 32 || k.bk == 64) << "knob " << tok << ": bk=N must be 32 or 64";
      } else if (key == "block") {
        k.block = number(tok, v);
        k.block_set = true;
      } else if (key == "vec") {
        k.vec = number(tok, v);
        ASSERT(k.vec == 1 || k.vec == 2 || k.vec == 4) << "knob " << tok << ": vec=N must be 1, 2 or 4";
      } else if (key == "tile") {
        auto x = v.find('x');
        ASSERT(x != std::string::npos) << "knob " << tok << ": expected tile=MxN";
        k.tile_m = number(tok, v.substr(0, x));
        k.tile_n = number(tok, v.substr(x + 1));
        for (int d : {k.tile_m, k.tile_n})
          ASSERT(d >= 32 && d <= 512 && d % 32 == 0)
              << "knob " << tok << ": tile sides must be multiples of 32 in [32, 512]";
      }
    }
  };
  for (const auto &a : lt.annotations) scan(a);
  for (auto n : lt.ir.nodes()) {
    scan(lt.ir.annotation(n));
    for (const auto &a : lt.ir.loop_annotations(n)) scan(a);
  }
  ASSERT(k.block >= 64 && k.block <= 1024 && k.block % 64 == 0)
      << "block=N must be a multiple of 64 in [64, 1024]";
  ASSERT(k.waves == 0 || k.waves == 2 || k.waves == 4 || k.waves == 8)
      << "waves=N must be 2, 4 or 8";
  return k;
}

namespace {
// The first stage can run ahead of the rest of the plan on a second stream
// when it reads only inputs and writes only scratch that the later stages
// read and never write (the operand staging of contractions; not a staging
// launch that also zeroes output blocks for stream-K).
bool overlappable(const Plan &plan) {
  if (plan.stages.size() < 2) return false;
  const auto &s0 = plan.stages[0];
  std::set<int> w0(s0.writes.begin(), s0.writes.end());
  for (int b : s0.writes)
    if (plan.buffers.at(b).kind != BufferSpec::SCRATCH) return false;
  for (int b : s0.buffers)
    if (!w0.count(b) && plan.buffers.at(b).kind != BufferSpec::INPUT) return false;
  for (size_t k = 1; k < plan.stages.size(); ++k)
    for (int b : plan.stages[k].writes)
      if (w0.count(b) || plan.buffers.at(b).kind == BufferSpec::INPUT) return false;
  return true;
}
}  // namespace

// ------------------------------------------------------------ HipCompiled ----
HipCompiled::HipCompiled(const LoopTree &lt, bool bf16_variant) {
  auto df = analyze(lt);
  knobs_ = knobs_from(lt, bf16_variant);
  plan_ = build_plan(*df, knobs_);
  bool cached = false;
  code_ = compile_source(plan_.source, "lt_module.hip", &cached);
  int_properties["stages"] = (int)plan_.stages.size();
  for (const auto &p : plan_.config) int_properties[p.first] = p.second;
  int_properties["code_cached"] = cached ? 1 : 0;
  std::string kinds;
  for (const auto &s : plan_.stages) kinds += (kinds.empty() ? "" : ",") + s.kind;
  string_properties["kinds"] = kinds;
  string_properties["code"] = plan_.source;
  std::string grids;  // "<kernel> grid <workgroups> block <threads>" per stage
  for (const auto &s : plan_.stages)
    grids += (grids.empty() ? "" : "; ") + s.kernel + " grid " +
             std::to_string((uint64_t)s.grid[0] * s.grid[1] * s.grid[2]) + " block " +
             std::to_string(s.block[0] * s.block[1] * s.block[2]);
  string_properties["grid"] = grids;
  // opt-in: on the bf16 conv 32x128->128x56^2 the overlapped plan measured
  // 62.0 us per run against 45.9 us serial (the staging launch takes CU
  // slots from the MFMA kernel)
  overlap_ = knobs_.overlap == 1 && overlappable(plan_);
  int_properties["overlap"] = overlap_ ? 1 : 0;
  int64_t bytes = 0, flops = 0;
  for (const auto &s : plan_.stages) {
    bytes += s.bytes;
    flops += s.flops;
  }
  algorithmic_bytes_ = bytes;
  algorithmic_flops_ = flops;
}

bool runtime_exiting() { return g_exiting.load(); }

HipCompiled::~HipCompiled() {
  if (!loaded_ || runtime_exiting()) return;
  for (auto &kv : contexts_) {
    auto &c = *kv.second;
    // work still queued on the context's streams (a run with sync=false)
    // finishes before its buffers go
    if (c.side) lt_hip_stream_synchronize(c.side);
    if (c.stream) lt_hip_stream_synchronize(c.stream);
    for (auto p : c.scratch) lt_hip_free(p);
    for (auto p : c.scratch2) lt_hip_free(p);
    for (auto &p : c.staging) lt_hip_free(p.second.first);
    for (auto *e : {c.ev_staged[0], c.ev_staged[1], c.ev_done[0], c.ev_done[1]})
      if (e) lt_hip_event_destroy(e);
    if (c.side) lt_hip_stream_destroy(c.side);
    if (c.stream) lt_hip_stream_destroy(c.stream);
  }
  if (module_) lt_hip_module_unload(module_);
}

void HipCompiled::load() const {
  if (loaded_) return;
  ensure_runtime();
  if (lt_hip_module_load(code_->data(), &module_) != 0) {
    // a cached code object that does not load (a damaged cache file): drop
    // it and compile the source again
    if (!int_properties.count("code_cached") || !int_properties.at("code_cached"))
      hip_check(1, "module load");
    evict_code(plan_.source);
    code_ = compile_source(plan_.source, "lt_module.hip", nullptr);
    hip_check(lt_hip_module_load(code_->data(), &module_), "module load (recompiled)");
  }
  for (const auto &s : plan_.stages) {
    void *fn = nullptr;
    hip_check(lt_hip_module_get_function(module_, s.kernel.c_str(), &fn), "get function");
    functions_.push_back(fn);
  }
  loaded_ = true;
}

HipCompiled::RunContext &HipCompiled::context() const {
  std::lock_guard<std::mutex> g(mu_);
  load();
  auto &slot = contexts_[std::this_thread::get_id()];
  if (!slot) {
    auto c = std::make_unique<RunContext>();
    hip_check(lt_hip_stream_create(&c->stream), "stream");
    slot = std::move(c);
  }
  return *slot;
}

// The Compiled's own scratch, allocated on first use by run()/run_stage():
// run_on() callers (the serving runtime) bring theirs, so a program compiled
// only for them holds no device memory beyond its code.
void HipCompiled::ensure_scratch(RunContext &c, int set) const {
  auto &sc = set ? c.scratch2 : c.scratch;
  if (!sc.empty()) return;
  for (const auto &b : plan_.buffers) {
    if (b.kind != BufferSpec::SCRATCH) continue;
    void *p = nullptr;
    hip_check(lt_hip_malloc(&p, b.numel * sizeof(float)), "scratch allocation");
    sc.push_back(p);
  }
}



void HipCompiled::launch(const std::vector<void *> &device_buffers, void *stream, int first,
                         int last) const {
  if (last < 0) last = (int)plan_.stages.size() - 1;
  for (int k = first; k <= last; ++k) {
    const auto &s = plan_.stages[k];
    std::vector<void *> ptrs;
    for (auto b : s.buffers) ptrs.push_back(device_buffers.at(b));
    std::vector<void *> args;
    for (auto &p : ptrs) args.push_back(&p);
    hip_check(lt_hip_launch(functions_[k], s.grid[0], s.grid[1], s.grid[2], s.block[0],
                            s.block[1], s.block[2], s.shared_bytes, stream, args.data()),
              "kernel launch");
  }
}

std::vector<void *> HipCompiled::device_buffers(RunContext &c, const std::vector<void *> &io,
                                                int set) const {
  ensure_scratch(c, set);
  const auto &sc = set ? c.scratch2 : c.scratch;
  std::vector<void *> bufs;
  size_t scratch_i = 0;
  for (const auto &b : plan_.buffers) {
    if (b.kind == BufferSpec::SCRATCH) {
      bufs.push_back(sc[scratch_i++]);
    } else {
      size_t pos = b.kind == BufferSpec::INPUT ? b.index : num_inputs() + b.index;
      bufs.push_back(io.at(pos));
    }
  }
  return bufs;
}

size_t HipCompiled::num_inputs() const {
  size_t n = 0;
  for (const auto &b : plan_.buffers) n += b.kind == BufferSpec::INPUT;
  return n;
}

size_t HipCompiled::num_io() const {
  size_t n = 0;
  for (const auto &b : plan_.buffers) n += b.kind != BufferSpec::SCRATCH;
  return n;
}

// Every call runs on the calling thread's context (stream, scratch, staging
// copies): calls from one thread are ordered on its stream, calls from
// different threads run concurrently.  sync=false returns once the work is
// queued on the thread's stream.
void HipCompiled::run(const std::vector<void *> &memory, bool sync) const {
  RunContext &c = context();
  ASSERT(memory.size() >= num_io())
      << "hip backend expects " << num_io() << " buffers (inputs then outputs), got "
      << memory.size();
  std::vector<void *> io(memory.begin(), memory.begin() + num_io());
  std::vector<std::pair<size_t, void *>> copy_out;  // (io position, host ptr)
  bool staged = false;
  for (size_t i = 0; i < io.size(); ++i) {
    if (!io[i] || lt_hip_is_device_pointer(io[i])) continue;
    const auto &spec = plan_.buffers[i];
    size_t bytes = spec.numel * sizeof(float);
    auto &st = c.staging[i];
    if (st.second < bytes) {
      // the stream may still read the old copy (an earlier sync=false run)
      if (st.first) {
        hip_check(lt_hip_stream_synchronize(c.stream), "synchronize");
        lt_hip_free(st.first);
      }
      hip_check(lt_hip_malloc(&st.first, bytes), "staging allocation");
      st.second = bytes;
    }
    if (spec.kind == BufferSpec::INPUT)
      hip_check(lt_hip_memcpy_async(st.first, io[i], bytes, 0, c.stream), "copy in");
    else
      copy_out.emplace_back(i, io[i]);
    io[i] = st.first;
    staged = true;
  }
  if (overlap_ && !staged) {
    // Runs alternate between two scratch sets.  The first stage (operand
    // staging) runs on the side stream as soon as the run that last used
    // its set has finished; the later stages run on the context's stream
    // after it.  Run i+1's staging thus overlaps run i's main kernels.
    if (!c.side) {
      hip_check(lt_hip_stream_create(&c.side), "stream");
      for (int b = 0; b < 2; ++b) {
        hip_check(lt_hip_event_create(&c.ev_staged[b]), "event");
        hip_check(lt_hip_event_create(&c.ev_done[b]), "event");
        hip_check(lt_hip_event_record(c.ev_done[b], c.stream), "event record");
      }
    }
    const int set = (int)(c.runs++ & 1);
    auto bufs = device_buffers(c, io, set);
    hip_check(lt_hip_stream_wait_event(c.side, c.ev_done[set]), "stream wait");
    launch(bufs, c.side, 0, 0);
    hip_check(lt_hip_event_record(c.ev_staged[set], c.side), "event record");
    hip_check(lt_hip_stream_wait_event(c.stream, c.ev_staged[set]), "stream wait");
    launch(bufs, c.stream, 1);
    hip_check(lt_hip_event_record(c.ev_done[set], c.stream), "event record");
  } else {
    launch(device_buffers(c, io), c.stream);
    // a later run-ahead staging into set 0 must wait for this run
    if (c.side) hip_check(lt_hip_event_record(c.ev_done[0], c.stream), "event record");
  }
  for (const auto &p : copy_out)
    hip_check(lt_hip_memcpy_async(p.second, io[p.first],
                                  plan_.buffers[p.first].numel * sizeof(float), 1, c.stream),
              "copy out");
  if (sync || staged) hip_check(lt_hip_stream_synchronize(c.stream), "synchronize");
}

void HipCompiled::run_on(const std::vector<void *> &device_io,
                         const std::vector<void *> &scratch, void *stream) const {
  {
    std::lock_guard<std::mutex> g(mu_);
    load();
  }
  std::vector<void *> bufs;
  size_t si = 0;
  for (const auto &b : plan_.buffers) {
    if (b.kind == BufferSpec::SCRATCH) {
      ASSERT(si < scratch.size()) << "missing scratch buffer";
      bufs.push_back(scratch[si++]);
    } else {
      size_t pos = b.kind == BufferSpec::INPUT ? b.index : num_inputs() + b.index;
      bufs.push_back(device_io.at(pos));
    }
  }
  launch(bufs, stream);
}

void HipCompiled::run_stage(const std::vector<void *> &device_io, int k, void *stream) const {
  RunContext &c = context();
  ASSERT(k >= 0 && k < (int)plan_.stages.size()) << "no stage " << k;
  ASSERT(device_io.size() >= num_io()) << "expected " << num_io() << " device buffers";
  launch(device_buffers(c, device_io), stream, k, k);
  if (c.side) hip_check(lt_hip_event_record(c.ev_done[0], stream), "event record");
}

std::vector<int64_t> HipCompiled::scratch_sizes() const {
  std::vector<int64_t> out;
  for (const auto &b : plan_.buffers)
    if (b.kind == BufferSpec::SCRATCH) out.push_back(b.numel);
  return out;
}

std::string HipCompiled::dump() const {
  return plan_.description + (schedule_note.empty() ? "" : schedule_note) + "\n" + plan_.source;
}

// ------------------------------------------------------------ registration ----
struct HipBackend : public Backend {
  bool bf16;
  HipBackend(std::string name, bool bf16_) : Backend(std::move(name)), bf16(bf16_) {}
  // A tuned schedule stored for this program (tuned_store.h) is lowered in
  // place of the tree as given; the Compiled names it in its properties and
  // shows it in dump().
  std::unique_ptr<Compiled> compile_impl(const LoopTree &lt) const override {
    std::string key;
    auto tuned = tuned_get(lt, &key);
    if (!tuned) return std::make_unique<HipCompiled>(lt, bf16);
    auto cc = std::make_unique<HipCompiled>(*tuned, bf16);
    cc->string_properties["tuned_schedule"] = key;
    cc->schedule_note = "[hip] tuned schedule " + key + " from the store:\n" + tuned->dump();
    return cc;
  }
  // accepts CPU-addressable memory (staged) as well as device pointers
  int hardware_requirement() const override { return 1 << 0; }
};

// Pinned host memory that kernels can also address; the lazy frontend can
// hand its pointers straight to users (include/loop_tool/hardware.h:24-33).
struct HipHardware : public Hardware {
  HipHardware() : Hardware("hip", 1) {}
  Memory alloc(size_t size) override {
    ensure_runtime();
    void *p = nullptr;
    hip_check(lt_hip_host_malloc(&p, size), "pinned allocation");
    return Memory{(1 << 0) | (1 << id()), p};
  }
  void free(Memory &data) override {
    if (data.address && !runtime_exiting()) lt_hip_host_free(data.address);
    data.address = nullptr;
    data.compatible = 0;
  }
};

static RegisterBackend register_hip(std::make_shared<HipBackend>("hip", false));
static RegisterBackend register_hip_bf16(std::make_shared<HipBackend>("hip_bf16", true));
static RegisterHardware register_hip_hw(std::make_shared<HipHardware>());

}  // namespace hip
}  // namespace loop_tool
