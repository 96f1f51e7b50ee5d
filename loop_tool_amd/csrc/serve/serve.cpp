// Serving runtime: lazy programs of ever-changing shapes on one GPU with
// continuous batching (C ABI in include/loop_tool_serve.h).
//
// Request path:
//   submit   structural key of the lazy graph (sizes masked) -> family, from
//            a cache filled by comparing against template programs built at
//            start-up; the input sizes -> a launch plan on the shape-generic
//            kernels of csrc/kernels/serve.hip (no compilation).  Programs of
//            no known family are lowered and compiled by the hip backend once
//            per exact lazy hash (the reference's compile cache key,
//            src/frontends/lazy.cpp:430-438) and run through run_on().
//   admit    in arrival order, while a slot is free and the request's device
//            bytes fit the budget: stream-ordered allocation from the device
//            memory pool, inputs filled on the device from their seeds, the
//            plan's kernels, an optional device checksum, a completion event
//            -- all on the slot's stream, so admitted requests overlap.
//   retire   completion events are polled (never waited on); buffers are
//            returned to the pool on the slot's stream.
// Latency is measured on the host clock from the arrival time to the poll
// that sees the completion.
#include <dlfcn.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <fstream>
#include <list>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../hip/hip_backend.h"
#include "loop_tool/lazy.h"
#include "loop_tool/nn.h"
#include "loop_tool_c.h"
#include "loop_tool_serve.h"
#include "lt_hip.h"

struct ltc_tensor {
  loop_tool::lazy::Tensor t;
};

namespace lz = loop_tool::lazy;
namespace sym = loop_tool::symbolic;

namespace loop_tool {
namespace serve {
namespace {

using Clock = std::chrono::steady_clock;

void chk(int rc, const char *what) { ASSERT(rc == 0) << what << ": " << lt_hip_last_error(); }

std::string library_dir() {
  Dl_info info;
  ASSERT(dladdr((void *)&library_dir, &info) && info.dli_fname) << "cannot locate library";
  std::string p = info.dli_fname;
  return p.substr(0, p.find_last_of('/'));
}

// ---------------------------------------------------------------- kernels --
struct Kernels {
  lt_hip_module mod = nullptr;
  std::vector<char> code;
  std::unordered_map<std::string, void *> fns;
  void *fn(const std::string &name) {
    auto it = fns.find(name);
    if (it != fns.end()) return it->second;
    void *f = nullptr;
    chk(lt_hip_module_get_function(mod, name.c_str(), &f), name.c_str());
    fns[name] = f;
    return f;
  }
};

Kernels &kernels() {
  static std::mutex mu;
  static Kernels *k = nullptr;
  std::lock_guard<std::mutex> g(mu);
  if (!k) {
    hip::ensure_runtime();
    std::string path = library_dir() + "/kernels/serve.hsaco";
    std::ifstream f(path, std::ios::binary);
    ASSERT(f.is_open()) << "kernel library missing: " << path << " (python loop_tool_amd/build.py)";
    auto *kk = new Kernels();
    kk->code.assign(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>());
    chk(lt_hip_module_load(kk->code.data(), &kk->mod), "load serve.hsaco");
    k = kk;  // kept for the process lifetime
  }
  return *k;
}

// device-side argument blocks (layouts of LtSvStage / LtSvGeo in serve.hip)
struct StageDesc {
  int64_t ssize[4], src_stride[4], src_size[4], padlo[4];
  int64_t kin;
  int32_t ld, pad_;
};
struct Geo {
  int32_t KO, KI;
  int64_t MD, ND, p_rs, qd1, qd2, qc0, qc1, qc2, qcst, ko2, pk0, pk1, qk0, qk1, o_rs, oc0,
      oc1, oc2;
  uint32_t pbytes, qbytes;  // staged operand sizes (the tile's buffer range)
};

void launch(void *fn, int64_t blocks, unsigned block, void *stream, std::vector<void *> args,
            unsigned gy = 1, unsigned gz = 1) {
  ASSERT(blocks >= 1 && blocks < (1LL << 31)) << "grid out of range";
  chk(lt_hip_launch(fn, (unsigned)blocks, gy, gz, block, 1, 1, 0, stream, args.data()),
      "launch");
}

int64_t grid_for(int64_t n, int64_t per_block = 256, int64_t cap = 8192) {
  return std::max<int64_t>(1, std::min<int64_t>((n + per_block - 1) / per_block, cap));
}

// --------------------------------------------------------- program shape --
struct Structure {
  uint64_t key = 0;
  std::vector<std::vector<int64_t>> leaves;  // bound sizes of the inputs, DFS order
};

bool is_bound_size(const sym::Constraint &c) {
  return c.first.type() == sym::Expr::Type::function && c.first.op() == sym::Op::size &&
         c.second.type() == sym::Expr::Type::value;
}

uint64_t structural(const lz::TensorImpl *t,
                    std::unordered_map<const lz::TensorImpl *, uint64_t> &memo,
                    std::vector<std::vector<int64_t>> &leaves) {
  auto it = memo.find(t);
  if (it != memo.end()) return it->second;
  uint64_t h = sym::hash((uint64_t)t->op_ + 0x5151);
  h = sym::hash_combine(h, sym::hash(t->shape_.size()));
  for (const auto &c : t->constraints_) {
    if (is_bound_size(c)) continue;  // concrete sizes are masked
    h = sym::hash_combine(h, c.first.hash());
    h = sym::hash_combine(h, c.second.hash());
  }
  std::unordered_map<int, uint64_t> numbering;
  auto hash_sym = [&](const sym::Symbol &s) {
    auto f = numbering.find(s.id());
    if (f == numbering.end()) f = numbering.emplace(s.id(), numbering.size()).first;
    return sym::hash(f->second);
  };
  for (const auto &d : t->deps_) {
    h = sym::hash_combine(h, structural(d.get(), memo, leaves));
    for (const auto &s : d->shape_) h = sym::hash_combine(h, hash_sym(s));
  }
  for (const auto &s : t->shape_) h = sym::hash_combine(h, hash_sym(s));
  if (t->deps_.empty()) {
    auto sc = t->size_constraints();
    std::vector<int64_t> sizes;
    for (const auto &s : t->shape_) {
      auto f = sc.find(s.id());
      sizes.push_back(f != sc.end() && f->second.type() == sym::Expr::Type::value
                          ? f->second.value()
                          : -1);
    }
    leaves.push_back(sizes);
  }
  memo[t] = h;
  return h;
}

Structure structure_of(const lz::Tensor &t) {
  Structure s;
  std::unordered_map<const lz::TensorImpl *, uint64_t> memo;
  s.key = structural(t.impl_.get(), memo, s.leaves);
  return s;
}

// ------------------------------------------------------ family templates --
// The same lazy programs as loop_tool_amd/programs.py.
lz::Tensor prog_matmul(int64_t M, int64_t N, int64_t K) {
  sym::Symbol m("M"), n("N"), k("K");
  lz::Tensor A(std::vector<int64_t>{M, K}), B(std::vector<int64_t>{K, N});
  return (A.as(m, k) * B.as(k, n)).sum(k);
}
lz::Tensor prog_softmax(int64_t R, int64_t C) {
  sym::Symbol r("R"), c("C");
  auto X = lz::Tensor(std::vector<int64_t>{R, C}).as(r, c);
  auto mx = X.max(std::vector<sym::Symbol>{c});
  auto e = (X - mx).exp();
  return e / e.sum(c);
}
lz::Tensor prog_conv2d(int64_t N, int64_t Ci, int64_t Co, int64_t H, int64_t W) {
  sym::Symbol n("N"), ci("Ci"), h("H"), w("W"), co("Co"), kh("Kh"), kw("Kw");
  auto X = lz::Tensor(std::vector<int64_t>{N, Ci, H, W}).as(n, ci, h, w);
  auto Wt = lz::Tensor(std::vector<int64_t>{Co, Ci, 3, 3}).as(co, ci, kh, kw);
  X = X.pad(h, 1).pad(w, 1);
  auto sh = X.shape()[2], sw = X.shape()[3];
  auto Y = nn::convolve(X, Wt, {sh, sw}, {kh, kw}, 1);
  auto s = Y.shape();
  return Y.transpose(std::vector<sym::Symbol>{s[0], s[3], s[1], s[2]});
}
lz::Tensor prog_maxpool(int64_t N, int64_t C, int64_t H, int64_t W) {
  sym::Symbol n("N"), c("C"), h("H"), w("W");
  auto X = lz::Tensor(std::vector<int64_t>{N, C, H, W}).as(n, c, h, w);
  return nn::maxpool(X, {h, w}, 2, 2);
}

// structural keys of the family templates (sizes chosen distinct; masked)
const uint64_t *template_keys() {
  static const uint64_t keys[4] = {
      structure_of(prog_matmul(3, 5, 7)).key,
      structure_of(prog_softmax(3, 5)).key,
      structure_of(prog_conv2d(2, 3, 5, 7, 11)).key,
      structure_of(prog_maxpool(2, 3, 8, 10)).key,
  };
  return keys;
}

// ------------------------------------------------------------------ plans --
struct Plan {
  int family = LTS_GENERIC;
  std::vector<int64_t> in_numel;
  int64_t out_numel = 0;
  std::vector<int64_t> scratch_bytes;
  int64_t flops = 0, bytes = 0;
  std::string desc;
  // matmul / conv2d: input i is staged by st[i] into scratch i; the GEMM's row
  // operand P is scratch p_of, the column operand Q the other one
  StageDesc st[2];
  int p_of = 0;
  Geo geo;
  std::string stage_fn, gemm_fn;
  unsigned gx = 1, gy = 1;
  // softmax
  std::string sm_fn;
  unsigned sm_block = 256;
  int64_t R = 0, C = 0;
  // max-pool
  int64_t Q = 0, H = 0, W = 0;
  // any other program
  std::shared_ptr<Compiled> compiled;

  int64_t device_bytes() const {
    int64_t b = out_numel * 4 + 8;
    for (auto n : in_numel) b += n * 4;
    for (auto n : scratch_bytes) b += n;
    return b;
  }
};

int64_t roundup(int64_t x, int64_t m) { return (x + m - 1) / m * m; }

struct Tile {
  int bm, bn, bk;
};

Tile pick_tile(int64_t MD, int64_t ND) {
  const Tile cands[] = {{128, 128, 64}, {128, 64, 32}, {64, 128, 32}, {64, 64, 32}};
  for (const auto &t : cands) {
    if (t.bm > roundup(MD, 64) || t.bn > roundup(ND, 64)) continue;
    if (((MD + t.bm - 1) / t.bm) * ((ND + t.bn - 1) / t.bn) >= 512) return t;
  }
  return {64, 64, 32};
}

void finish_gemm(Plan &p, const Tile &t, const char *tn) {
  p.gemm_fn = std::string("lt_sv_igemm_") + tn + "_" + std::to_string(t.bm) + "x" +
              std::to_string(t.bn) + "_" + std::to_string(t.bk);
  p.stage_fn = std::string("lt_sv_stage_") + tn;
  p.gx = (unsigned)((p.geo.ND + t.bn - 1) / t.bn);
  p.gy = (unsigned)((p.geo.MD + t.bm - 1) / t.bm);
  p.desc += " tile " + std::to_string(t.bm) + "x" + std::to_string(t.bn) + "x" +
            std::to_string(t.bk);
}

StageDesc stage(std::initializer_list<int64_t> ssize, std::initializer_list<int64_t> stride,
                std::initializer_list<int64_t> size, std::initializer_list<int64_t> padlo,
                int64_t kin, int ld) {
  StageDesc d{};
  std::copy(ssize.begin(), ssize.end(), d.ssize);
  std::copy(stride.begin(), stride.end(), d.src_stride);
  std::copy(size.begin(), size.end(), d.src_size);
  std::copy(padlo.begin(), padlo.end(), d.padlo);
  d.kin = kin;
  d.ld = ld;
  return d;
}

bool plan_matmul(const Structure &s, int dtype, Plan &p) {
  if (s.leaves.size() != 2 || s.leaves[0].size() != 2 || s.leaves[1].size() != 2) return false;
  const int64_t M = s.leaves[0][0], K = s.leaves[0][1], N = s.leaves[1][1];
  if (s.leaves[1][0] != K || M < 1 || N < 1 || K < 1) return false;
  Tile t = pick_tile(M, N);
  const int64_t KP = roundup(K, t.bk);
  p.in_numel = {M * K, K * N};
  p.out_numel = M * N;
  p.scratch_bytes = {M * KP * 2, N * KP * 2};
  // the tile addresses staged operands with 32-bit byte offsets
  if (p.scratch_bytes[0] >= (1LL << 31) || p.scratch_bytes[1] >= (1LL << 31)) return false;
  // A[M][K] -> A'[M][KP]; B[K][N] -> Bt'[N][KP] (lanes along n, contiguous in B)
  p.st[0] = stage({1, 1, M, KP}, {0, 0, K, 1}, {1, 1, M, K}, {0, 0, 0, 0}, K, -1);
  p.st[1] = stage({1, 1, N, KP}, {0, 0, 1, N}, {1, 1, N, K}, {0, 0, 0, 0}, K, 2);
  Geo &g = p.geo;
  g = Geo{};
  g.KO = 1;
  g.KI = (int32_t)KP;
  g.MD = M;
  g.ND = N;
  g.p_rs = KP;
  g.qd1 = g.qd2 = 1;
  g.qc0 = KP;
  g.ko2 = 1;
  g.o_rs = N;
  g.oc0 = 1;
  p.flops = 2 * M * N * K;
  p.bytes = 4 * (M * K + K * N + M * N);
  p.desc = "matmul " + std::to_string(M) + "x" + std::to_string(N) + "x" + std::to_string(K);
  finish_gemm(p, t, dtype == 0 ? "bf16" : "f16");
  return true;
}

bool plan_conv(const Structure &s, int dtype, Plan &p) {
  if (s.leaves.size() != 2 || s.leaves[0].size() != 4 || s.leaves[1].size() != 4) return false;
  const auto &x = s.leaves[0], &w = s.leaves[1];
  const int64_t N = x[0], C = x[1], H = x[2], W = x[3], Co = w[0];
  if (w[1] != C || w[2] != 3 || w[3] != 3 || N < 1 || C < 1 || H < 1 || W < 1 || Co < 1)
    return false;
  Tile t = pick_tile(Co, N * H * W);
  const int64_t CP = roundup(C, t.bk), Hp = H + 2, Wp = W + 2;
  p.in_numel = {N * C * H * W, Co * C * 9};
  p.out_numel = N * Co * H * W;
  p.scratch_bytes = {N * Hp * Wp * CP * 2, Co * 9 * CP * 2};
  // the tile addresses staged operands with 32-bit byte offsets
  if (p.scratch_bytes[0] >= (1LL << 31) || p.scratch_bytes[1] >= (1LL << 31)) return false;
  // X[N][C][H][W] -> X'[N][H+2][W+2][CP] (zero halo); W[Co][C][3][3] -> W'[Co][3][3][CP]
  p.st[0] = stage({N, Hp, Wp, CP}, {C * H * W, W, 1, H * W}, {N, H, W, C}, {0, 1, 1, 0}, C, 2);
  p.st[1] = stage({Co, 3, 3, CP}, {C * 9, 3, 1, 9}, {Co, 3, 3, C}, {0, 0, 0, 0}, C, 2);
  Geo &g = p.geo;
  g = Geo{};
  g.KO = 9;
  g.KI = (int32_t)CP;
  g.MD = Co;
  g.ND = N * H * W;
  g.p_rs = 9 * CP;
  g.qd1 = H;
  g.qd2 = W;
  g.qc0 = Hp * Wp * CP;
  g.qc1 = Wp * CP;
  g.qc2 = CP;
  g.qcst = 0;
  g.ko2 = 3;
  g.pk0 = 3 * CP;
  g.pk1 = CP;
  g.qk0 = Wp * CP;
  g.qk1 = CP;
  g.o_rs = H * W;
  g.oc0 = Co * H * W;
  g.oc1 = W;
  g.oc2 = 1;
  p.flops = 2 * N * Co * H * W * C * 9;
  p.bytes = 4 * (N * C * H * W + Co * C * 9 + N * Co * H * W);
  p.desc = "conv2d N" + std::to_string(N) + " C" + std::to_string(C) + " Co" +
           std::to_string(Co) + " H" + std::to_string(H) + " W" + std::to_string(W);
  p.p_of = 1;  // P = W' (rows co), Q = X' (rows n, ho, wo)
  finish_gemm(p, t, dtype == 0 ? "bf16" : "f16");
  return true;
}

bool plan_softmax(const Structure &s, Plan &p) {
  if (s.leaves.size() != 1 || s.leaves[0].size() != 2) return false;
  const int64_t R = s.leaves[0][0], C = s.leaves[0][1];
  if (R < 1 || C < 1 || C > 32768 || R >= (1LL << 31)) return false;
  p.R = R;
  p.C = C;
  const int variants[][2] = {{64, 4}, {64, 16}, {256, 8}, {256, 16}, {256, 32}, {1024, 32}};
  for (const auto &v : variants)
    if (C <= (int64_t)v[0] * v[1]) {
      p.sm_fn = "lt_sv_softmax_" + std::to_string(v[0]) + "_" + std::to_string(v[1]);
      p.sm_block = v[0];
      break;
    }
  p.in_numel = {R * C};
  p.out_numel = R * C;
  p.flops = 5 * R * C;
  p.bytes = 8 * R * C;
  p.desc = "softmax " + std::to_string(R) + "x" + std::to_string(C) + " " + p.sm_fn;
  return true;
}

bool plan_maxpool(const Structure &s, Plan &p) {
  if (s.leaves.size() != 1 || s.leaves[0].size() != 4) return false;
  const auto &x = s.leaves[0];
  if (x[2] < 2 || x[3] < 2) return false;
  p.Q = x[0] * x[1];
  p.H = x[2];
  p.W = x[3];
  p.in_numel = {p.Q * p.H * p.W};
  p.out_numel = p.Q * (p.H / 2) * (p.W / 2);
  p.flops = 3 * p.out_numel;
  p.bytes = 4 * (p.in_numel[0] + p.out_numel);
  p.desc = "maxpool2d " + std::to_string(p.Q) + "x" + std::to_string(p.H) + "x" +
           std::to_string(p.W);
  return true;
}

// ---------------------------------------------------------------- server --
struct Request {
  int id = 0;
  std::shared_ptr<lz::TensorImpl> program;
  std::vector<uint64_t> seeds;
  std::vector<float> scales;
  double arrival = 0, start = 0, done = 0;
  int state = 0;
  std::string error;
  std::shared_ptr<Plan> plan;
  std::vector<void *> bufs;  // inputs, output, scratch
  void *out = nullptr;
  int64_t bytes = 0;
  int slot = -1;
  uint64_t checksum = 0;
};

struct Slot {
  void *stream = nullptr;
  void *event = nullptr;
  uint64_t *d_sum = nullptr;   // device checksum word
  uint64_t *h_sum = nullptr;   // pinned host copy
  int req = -1;
};

// Device memory of requests: blocks of geometric size classes (2^(k/4)
// bytes, 4 KiB granules, <= 19% slack), reused once the request holding them
// has been seen complete, so steady-state serving makes no allocation call.
// Cached blocks are returned to the device only when an allocation fails.
struct BlockCache {
  std::map<size_t, std::vector<void *>> idle;
  std::unordered_map<void *, size_t> class_of;
  size_t cached = 0;
  int64_t device_allocs = 0;

  static size_t size_class(size_t n) {
    if (n <= 4096) return 4096;
    double q = std::ceil(std::log2((double)n) * 4.0) / 4.0;
    size_t c = (size_t)std::ceil(std::exp2(q));
    return (std::max(c, n) + 4095) / 4096 * 4096;
  }
  void *get(size_t n) {
    const size_t c = size_class(n);
    auto it = idle.find(c);
    if (it != idle.end() && !it->second.empty()) {
      void *p = it->second.back();
      it->second.pop_back();
      cached -= c;
      return p;
    }
    void *p = nullptr;
    if (lt_hip_malloc(&p, c) != 0) {
      release_idle();
      chk(lt_hip_malloc(&p, c), "device allocation");
    }
    ++device_allocs;
    class_of[p] = c;
    return p;
  }
  void put(void *p) {
    const size_t c = class_of.at(p);
    idle[c].push_back(p);
    cached += c;
  }
  void release_idle() {
    lt_hip_device_synchronize();
    for (auto &kv : idle)
      for (auto p : kv.second) {
        lt_hip_free(p);
        class_of.erase(p);
      }
    idle.clear();
    cached = 0;
  }
};

double seconds_since(Clock::time_point t) {
  return std::chrono::duration<double>(Clock::now() - t).count();
}

}  // namespace

struct Server {
  lts_config cfg;
  Clock::time_point t0 = Clock::now();
  std::vector<Slot> slots;
  std::map<int, std::unique_ptr<Request>> reqs;
  std::deque<int> queue;
  int next_id = 0;
  int64_t in_use = 0;
  BlockCache blocks;
  int64_t allocs_base = 0;
  std::vector<int> retired;  // completed since the last take_retired()
  // caches
  std::unordered_map<uint64_t, int> family_of;  // structural key -> family
  // exact lazy hash -> compiled program, least recently used last (the
  // eviction order of the reference's JS CompilationCache LRU,
  // javascript/loop_tool.mjs:68-130); plans in flight keep their own
  // reference, so an evicted program lives until its requests retire
  std::list<uint64_t> lru;
  std::unordered_map<uint64_t, std::pair<std::shared_ptr<Compiled>, std::list<uint64_t>::iterator>>
      compiled;
  // stats
  lts_stats st{};
  std::vector<double> lat;
  double first_arrival = -1, last_done = 0;
  double lag_sum = 0;
  int64_t lag_n = 0;

  double now() const {
    return std::chrono::duration<double>(Clock::now() - t0).count();
  }

  explicit Server(const lts_config &c) : cfg(c) {
    ASSERT(cfg.slots >= 1 && cfg.slots <= 64) << "slots must be in [1, 64]";
    ASSERT(cfg.dtype == 0 || cfg.dtype == 1) << "dtype must be 0 (bf16) or 1 (fp16)";
    ASSERT(cfg.compiled_cap >= 0) << "compiled_cap must be >= 0";
    kernels();
    template_keys();
    slots.resize(cfg.slots);
    for (auto &s : slots) {
      chk(lt_hip_stream_create(&s.stream), "stream");
      chk(lt_hip_event_create(&s.event), "event");
      chk(lt_hip_malloc((void **)&s.d_sum, 8), "checksum word");
      chk(lt_hip_host_malloc((void **)&s.h_sum, 8), "pinned checksum word");
      *s.h_sum = 0;
    }
  }

  ~Server() {
    if (hip::runtime_exiting()) return;
    for (auto &s : slots)
      if (s.stream) lt_hip_stream_synchronize(s.stream);
    for (auto &p : reqs)
      for (auto b : p.second->bufs)
        if (b) blocks.put(b);
    blocks.release_idle();
    for (auto &s : slots) {
      lt_hip_stream_synchronize(s.stream);
      lt_hip_free(s.d_sum);
      lt_hip_host_free(s.h_sum);
      lt_hip_event_destroy(s.event);
      lt_hip_stream_destroy(s.stream);
    }
  }

  int family(const Structure &s) {
    auto it = family_of.find(s.key);
    if (it != family_of.end()) {
      ++st.struct_hits;
      return it->second;
    }
    ++st.struct_misses;
    int f = LTS_GENERIC;
    for (int i = 0; i < 4; ++i)
      if (template_keys()[i] == s.key) f = i;
    family_of[s.key] = f;
    return f;
  }

  std::shared_ptr<Plan> make_plan(const lz::Tensor &t) {
    auto p = std::make_shared<Plan>();
    Structure s = structure_of(t);
    int f = family(s);
    bool ok = false;
    switch (f) {
      case LTS_MATMUL: ok = plan_matmul(s, cfg.dtype, *p); break;
      case LTS_SOFTMAX: ok = plan_softmax(s, *p); break;
      case LTS_CONV2D: ok = plan_conv(s, cfg.dtype, *p); break;
      case LTS_MAXPOOL: ok = plan_maxpool(s, *p); break;
      default: break;
    }
    if (ok) {
      p->family = f;
      return p;
    }
    // any other program (or sizes a family kernel does not take): the hip
    // backend, compiled once per exact lazy hash; a family plan that gave up
    // may have filled fields
    *p = Plan();
    p->family = LTS_GENERIC;
    uint64_t h = t.hash();
    auto it = compiled.find(h);
    if (it != compiled.end()) {
      ++st.exact_hits;
      p->compiled = it->second.first;
      lru.splice(lru.begin(), lru, it->second.second);
    } else {
      ++st.compiles;
      auto lt = t.loop_tree();
      p->compiled = getBackends().at("hip")->compile(lt);
      lru.push_front(h);
      compiled[h] = {p->compiled, lru.begin()};
      const size_t cap = cfg.compiled_cap > 0 ? (size_t)cfg.compiled_cap : 256;
      while (compiled.size() > cap) {
        compiled.erase(lru.back());
        lru.pop_back();
        ++st.compiled_evictions;
      }
    }
    auto *hc = dynamic_cast<const hip::HipCompiled *>(p->compiled.get());
    ASSERT(hc) << "hip backend expected";
    for (const auto &b : hc->plan().buffers) {
      if (b.kind == hip::BufferSpec::INPUT) p->in_numel.push_back(b.numel);
      else if (b.kind == hip::BufferSpec::OUTPUT) p->out_numel += b.numel;
      else p->scratch_bytes.push_back(b.numel * 4);
    }
    p->flops = hc->algorithmic_flops();
    p->bytes = hc->algorithmic_bytes();
    p->desc = "hip backend: " + p->compiled->string_properties["kinds"];
    return p;
  }

  int submit(const lz::Tensor &t, const uint64_t *seeds, const float *scales, int n,
             double arrival) {
    const auto t_in = Clock::now();
    auto r = std::make_unique<Request>();
    r->id = next_id++;
    r->program = t.impl_;
    r->arrival = arrival;
    r->plan = make_plan(t);
    ASSERT((int)r->plan->in_numel.size() == n)
        << "program has " << r->plan->in_numel.size() << " inputs, " << n << " seeds given";
    r->seeds.assign(seeds, seeds + n);
    r->scales.assign(scales, scales + n);
    r->bytes = r->plan->device_bytes();
    if (first_arrival < 0 || arrival < first_arrival) first_arrival = arrival;
    ++st.submitted;
    int id = r->id;
    queue.push_back(id);
    reqs[id] = std::move(r);
    st.host_plan_ms += seconds_since(t_in) * 1e3;
    return id;
  }

  void start(Request &r, int si) {
    const auto t_in = Clock::now();
    Slot &sl = slots[si];
    void *s = sl.stream;
    const Plan &p = *r.plan;
    auto &K = kernels();
    r.start = now();
    r.slot = si;
    sl.req = r.id;
    in_use += r.bytes;
    st.peak_bytes = std::max(st.peak_bytes, (double)in_use);
    auto alloc = [&](int64_t bytes) {
      void *ptr = blocks.get((size_t)std::max<int64_t>(bytes, 16));
      r.bufs.push_back(ptr);
      return ptr;
    };
    std::vector<void *> in;
    for (auto n : p.in_numel) in.push_back(alloc(n * 4));
    r.out = alloc(p.out_numel * 4);
    std::vector<void *> scratch;
    for (auto b : p.scratch_bytes) scratch.push_back(alloc(b));
    // inputs in HBM from their seeds
    for (size_t i = 0; i < in.size(); ++i) {
      void *ptr = in[i];
      long n = (long)p.in_numel[i];
      unsigned long long seed = r.seeds[i];
      float scale = r.scales[i];
      launch(K.fn("lt_sv_fill"), grid_for(n, 256, 4096), 256, s, {&ptr, &n, &seed, &scale});
    }
    if (p.family == LTS_MATMUL || p.family == LTS_CONV2D) {
      for (int o = 0; o < 2; ++o) {
        StageDesc d = p.st[o];
        const void *src = in[o];
        void *dst = scratch[o];
        int64_t nvec = d.ssize[0] * d.ssize[1] * d.ssize[2] * d.ssize[3] / 8;
        launch(K.fn(p.stage_fn), grid_for(nvec, 256, 16384), 256, s, {&d, &src, &dst});
      }
      Geo g = p.geo;
      g.pbytes = (uint32_t)p.scratch_bytes[p.p_of];
      g.qbytes = (uint32_t)p.scratch_bytes[1 - p.p_of];
      const void *P = scratch[p.p_of], *Q = scratch[1 - p.p_of];
      void *out = r.out;
      launch(K.fn(p.gemm_fn), p.gx, 256, s, {&g, &P, &Q, &out}, p.gy, 1);
    } else if (p.family == LTS_SOFTMAX) {
      const void *x = in[0];
      void *y = r.out;
      long R = (long)p.R;
      int C = (int)p.C;
      launch(K.fn(p.sm_fn), p.R, p.sm_block, s, {&x, &y, &R, &C});
    } else if (p.family == LTS_MAXPOOL) {
      const void *x = in[0];
      void *y = r.out;
      long Q = (long)p.Q;
      int H = (int)p.H, W = (int)p.W;
      launch(K.fn("lt_sv_maxpool2"), grid_for(p.out_numel, 256, 8192), 256, s,
             {&x, &y, &Q, &H, &W});
    } else {
      auto *hc = dynamic_cast<const hip::HipCompiled *>(p.compiled.get());
      std::vector<void *> io = in;
      io.push_back(r.out);
      hc->run_on(io, scratch, s);
    }
    if (cfg.checksums) {
      void *d = sl.d_sum;
      launch(K.fn("lt_sv_zero64"), 1, 64, s, {&d});
      const void *o = r.out;
      long n = (long)p.out_numel;
      launch(K.fn("lt_sv_checksum"), grid_for(n, 256, 1024), 256, s, {&o, &n, &d});
      chk(lt_hip_memcpy_async(sl.h_sum, sl.d_sum, 8, 1, s), "checksum copy");
    }
    chk(lt_hip_event_record(sl.event, s), "completion event");
    r.state = 1;
    st.host_launch_ms += seconds_since(t_in) * 1e3;
  }

  void complete(Request &r) {
    Slot &sl = slots[r.slot];
    r.done = now();
    r.state = 2;
    if (cfg.checksums) r.checksum = *sl.h_sum;
    for (auto b : r.bufs)
      if (b != r.out || !cfg.keep_outputs) blocks.put(b);
    int64_t kept = cfg.keep_outputs ? r.plan->out_numel * 4 : 0;
    r.bufs.clear();
    if (cfg.keep_outputs) {
      r.bufs.push_back(r.out);
    } else {
      r.out = nullptr;
    }
    in_use -= r.bytes - kept;
    r.bytes = kept;
    sl.req = -1;
    ++st.completed;
    ++st.family_count[r.plan->family];
    st.flops += (double)r.plan->flops;
    st.bytes += (double)r.plan->bytes;
    lat.push_back(r.done - r.arrival);
    last_done = std::max(last_done, r.done);
    r.program.reset();  // the graph is no longer needed
    retired.push_back(r.id);
  }

  int poll() {
    const auto t_in = Clock::now();
    for (auto &sl : slots) {
      if (sl.req < 0) continue;
      int q = lt_hip_event_query(sl.event);
      ASSERT(q >= 0) << "event query: " << lt_hip_last_error();
      if (q == 1) complete(*reqs.at(sl.req));
    }
    st.host_retire_ms += seconds_since(t_in) * 1e3;
    while (!queue.empty()) {
      int si = -1;
      for (size_t i = 0; i < slots.size(); ++i)
        if (slots[i].req < 0) {
          si = (int)i;
          break;
        }
      if (si < 0) break;
      Request &r = *reqs.at(queue.front());
      if (in_use + r.bytes > cfg.budget_bytes) {
        bool busy = false;
        for (auto &sl : slots) busy |= sl.req >= 0;
        if (busy) {
          ++st.admission_waits;
          break;
        }
        if (r.bytes > cfg.budget_bytes) {  // can never fit
          r.state = -1;
          r.error = "request needs more device memory than the budget";
          ++st.failed;
          queue.pop_front();
          continue;
        }
      }
      queue.pop_front();
      start(r, si);
    }
    int active = (int)queue.size();
    for (auto &sl : slots) active += sl.req >= 0;
    return active;
  }

  int drain() {
    while (poll() > 0) std::this_thread::yield();
    return 0;
  }

  void fill_stats(lts_stats *out) {
    lts_stats s = st;
    s.wall_s = first_arrival >= 0 ? std::max(0.0, last_done - first_arrival) : 0.0;
    std::vector<double> l = lat;
    std::sort(l.begin(), l.end());
    auto pct = [&](double q) {
      if (l.empty()) return 0.0;
      size_t i = (size_t)std::min<double>(l.size() - 1, std::ceil(q * l.size()) - 1);
      return l[i] * 1e3;
    };
    s.p50_ms = pct(0.50);
    s.p99_ms = pct(0.99);
    double sum = 0;
    for (double x : l) sum += x;
    s.mean_ms = l.empty() ? 0 : sum / l.size() * 1e3;
    s.max_ms = l.empty() ? 0 : l.back() * 1e3;
    s.client_lag_ms = lag_n ? lag_sum / lag_n * 1e3 : 0;
    s.device_allocs = blocks.device_allocs - allocs_base;
    s.cached_bytes = (double)blocks.cached;
    *out = s;
  }

  void reset_stats() {
    st = lts_stats{};
    allocs_base = blocks.device_allocs;
    lat.clear();
    first_arrival = -1;
    last_done = 0;
    lag_sum = 0;
    lag_n = 0;
  }
};

// ------------------------------------------------------- workload driver --
namespace {

struct Rng {
  uint64_t s;
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
  }
  double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
  int64_t range(int64_t lo, int64_t hi) {  // inclusive
    return lo + (int64_t)(uniform() * (double)(hi - lo + 1));
  }
  int64_t log_uniform16(int64_t lo, int64_t hi) {
    double v = std::exp(std::log((double)lo) + uniform() * (std::log((double)hi) - std::log((double)lo)));
    int64_t r = (int64_t)std::llround(v / 16.0) * 16;
    return std::max(lo, std::min(hi, r));
  }
};

struct Incoming {
  std::shared_ptr<lz::TensorImpl> program;
  std::vector<uint64_t> seeds;
  std::vector<float> scales;
  double arrival;
  double submitted;
};

}  // namespace

int run_workload(Server &srv, const lts_workload &w, lts_stats *out) {
  ASSERT(w.rate > 0 && w.duration_s > 0) << "rate and duration must be positive";
  std::mutex mu;
  std::deque<Incoming> inbox;
  std::atomic<bool> client_done{false};
  std::string client_error;
  const double t_start = srv.now() + 0.05;
  std::thread client([&] {
    try {
      Rng rng{w.seed};
      double t = 0;
      double cum[4];
      double tot = w.mix[0] + w.mix[1] + w.mix[2] + w.mix[3];
      double acc = 0;
      for (int i = 0; i < 4; ++i) cum[i] = (acc += w.mix[i] / tot);
      for (int64_t i = 0; w.max_requests <= 0 || i < w.max_requests; ++i) {
        t += -std::log(1.0 - rng.uniform()) / w.rate;
        if (t >= w.duration_s) break;
        double u = rng.uniform();
        int f = u < cum[0] ? 0 : u < cum[1] ? 1 : u < cum[2] ? 2 : 3;
        Incoming in;
        uint64_t base = rng.next();
        if (f == LTS_MATMUL) {
          int64_t M = rng.log_uniform16(32, 2048), N = rng.log_uniform16(32, 2048),
                  K = rng.log_uniform16(32, 2048);
          in.program = prog_matmul(M, N, K).impl_;
          in.scales = {1.0f, 1.0f};
        } else if (f == LTS_SOFTMAX) {
          int64_t R = rng.range(256, 8192), C = rng.range(256, 8192);
          in.program = prog_softmax(R, C).impl_;
          in.scales = {3.0f};
        } else if (f == LTS_CONV2D) {
          int64_t N = rng.range(1, 16), C = rng.range(32, 256), H = rng.range(14, 112);
          in.program = prog_conv2d(N, C, C, H, H).impl_;
          in.scales = {1.0f, 1.0f / std::sqrt(9.0f * (float)C)};
        } else {
          int64_t N = rng.range(1, 16), C = rng.range(32, 256), H = rng.range(14, 112);
          in.program = prog_maxpool(N, C, H, H).impl_;
          in.scales = {1.0f};
        }
        for (size_t k = 0; k < in.scales.size(); ++k) in.seeds.push_back(base + 7919 * k);
        in.arrival = t_start + t;
        // open loop: hand the request over at its arrival time
        while (srv.now() < in.arrival) {
          double left = in.arrival - srv.now();
          if (left > 200e-6) std::this_thread::sleep_for(std::chrono::microseconds(100));
        }
        in.submitted = srv.now();
        std::lock_guard<std::mutex> g(mu);
        inbox.push_back(std::move(in));
      }
    } catch (const std::exception &e) {
      client_error = e.what();
    }
    client_done = true;
  });
  try {
    while (true) {
      std::deque<Incoming> got;
      {
        std::lock_guard<std::mutex> g(mu);
        got.swap(inbox);
      }
      for (auto &in : got) {
        srv.lag_sum += in.submitted - in.arrival;
        ++srv.lag_n;
        srv.submit(lz::Tensor(in.program), in.seeds.data(), in.scales.data(), (int)in.seeds.size(),
                   in.arrival);
        // free the per-request bookkeeping of finished requests as we go
      }
      int active = srv.poll();
      if (!srv.cfg.keep_outputs)
        for (int id : srv.retired) srv.reqs.erase(id);
      srv.retired.clear();
      if (client_done && active == 0) {
        std::lock_guard<std::mutex> g(mu);
        if (inbox.empty()) break;
      }
      if (got.empty() && active == 0) std::this_thread::sleep_for(std::chrono::microseconds(20));
    }
  } catch (...) {
    client.join();
    throw;
  }
  client.join();
  ASSERT(client_error.empty()) << "workload client: " << client_error;
  srv.fill_stats(out);
  return 0;
}

}  // namespace serve
}  // namespace loop_tool

// ------------------------------------------------------------------ C ABI --
struct lts_server {
  std::unique_ptr<loop_tool::serve::Server> s;
};

namespace {
thread_local std::string g_serve_err;
template <typename F>
int guarded(F f) {
  try {
    return f();
  } catch (const std::exception &e) {
    g_serve_err = e.what();
  }
  return -1;
}
}  // namespace

extern "C" {

const char *lts_last_error(void) { return g_serve_err.c_str(); }

lts_server *lts_create(const lts_config *cfg) {
  try {
    auto *s = new lts_server();
    s->s = std::make_unique<loop_tool::serve::Server>(*cfg);
    return s;
  } catch (const std::exception &e) {
    g_serve_err = e.what();
    return nullptr;
  }
}

void lts_free(lts_server *s) { delete s; }

uint64_t lts_structural_hash(const ltc_tensor *program) {
  try {
    return loop_tool::serve::structure_of(program->t).key;
  } catch (const std::exception &e) {
    g_serve_err = e.what();
    return 0;
  }
}

uint64_t lts_template_hash(int family) {
  if (family < 0 || family >= 4) return 0;
  try {
    return loop_tool::serve::template_keys()[family];
  } catch (const std::exception &e) {
    g_serve_err = e.what();
    return 0;
  }
}

int lts_family(lts_server *s, const ltc_tensor *program) {
  return guarded([&] { return s->s->family(loop_tool::serve::structure_of(program->t)); });
}

int lts_submit(lts_server *s, const ltc_tensor *program, const uint64_t *seeds,
               const float *scales, int n_inputs) {
  return guarded([&] {
    return s->s->submit(program->t, seeds, scales, n_inputs, s->s->now());
  });
}

int lts_poll(lts_server *s) {
  return guarded([&] { return s->s->poll(); });
}

int lts_drain(lts_server *s) {
  return guarded([&] { return s->s->drain(); });
}

int lts_info(lts_server *s, int id, lts_request_info *out) {
  return guarded([&] {
    const auto &r = *s->s->reqs.at(id);
    lts_request_info i{};
    i.id = r.id;
    i.family = r.plan->family;
    i.state = r.state;
    i.arrival_s = r.arrival;
    i.start_s = r.start;
    i.done_s = r.done;
    i.out_numel = r.plan->out_numel;
    i.flops = r.plan->flops;
    i.bytes = r.plan->bytes;
    i.checksum = r.checksum;
    std::snprintf(i.plan, sizeof(i.plan), "%s", r.plan->desc.c_str());
    *out = i;
    return 0;
  });
}

int lts_output(lts_server *s, int id, float *host, int64_t n) {
  return guarded([&] {
    const auto &r = *s->s->reqs.at(id);
    ASSERT(r.state == 2 && r.out) << "request " << id << " has no kept output";
    ASSERT(n == r.plan->out_numel) << "output has " << r.plan->out_numel << " elements";
    loop_tool::serve::chk(lt_hip_memcpy(host, r.out, (size_t)n * 4, 1), "output copy");
    return 0;
  });
}

int lts_release(lts_server *s, int id) {
  return guarded([&] {
    auto &srv = *s->s;
    auto it = srv.reqs.find(id);
    ASSERT(it != srv.reqs.end()) << "no request " << id;
    auto &r = *it->second;
    ASSERT(r.state == 2 || r.state == -1) << "request " << id << " still in flight";
    if (r.out) {
      srv.blocks.put(r.out);
      srv.in_use -= r.bytes;
    }
    srv.reqs.erase(it);
    return 0;
  });
}

int lts_stats_get(lts_server *s, lts_stats *out) {
  return guarded([&] {
    s->s->fill_stats(out);
    return 0;
  });
}

int lts_stats_reset(lts_server *s) {
  return guarded([&] {
    s->s->reset_stats();
    return 0;
  });
}

int lts_run_workload(lts_server *s, const lts_workload *w, lts_stats *out) {
  return guarded([&] { return loop_tool::serve::run_workload(*s->s, *w, out); });
}

}  // extern "C"
