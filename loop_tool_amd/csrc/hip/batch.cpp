// Batches of independent, run-time-sized matmul nests on MI355X
// (see include/loop_tool_hip.h, lth_batch_*).
//
// Operands are 16-bit (bf16 or fp16) and resident in HBM: A[M][K] and B
// stored transposed as Bt[N][K]; outputs are float32 C[M][N].  Two ways to
// run a batch:
//   plain    one launch of a fixed-tile kernel per GEMM (the hip backend's
//            per-nest behaviour, one Compiled::run per matmul);
//   grouped  one persistent launch over a work list holding every output
//            tile of every GEMM, each with the tile configuration chosen for
//            its shape bucket, ordered by decreasing cost and taken by the
//            workgroups as a queue, so that the last tiles to start are the
//            cheap ones; optionally captured once into a hipGraph and
//            replayed.
// The kernels are hand-written HIP (csrc/kernels/gemm.hip, on the tile of
// csrc/kernels/igemm_tile.h) built by hipcc into kernels/gemm.hsaco next to
// this library.
#include <dlfcn.h>

#include <algorithm>
#include <cstring>
#include <fstream>
#include <functional>
#include <map>
#include <mutex>
#include <numeric>
#include <queue>
#include <string>
#include <vector>

#include "hip_backend.h"
#include "loop_tool/error.h"
#include "loop_tool_hip.h"
#include "lt_hip.h"

namespace {

struct Desc {
  const void *A;
  const void *Bt;
  float *C;
  int M, N, K, pad;
};
// a work item as the grouped kernels read it (LtWorkItem of gemm.hip): the
// GEMM's operands and sizes, the tile's first row and column, its tile
// configuration
struct Item {
  const void *A;
  const void *Bt;
  float *C;
  int M, N, K, m0, n0, cfg;
};

// tile configurations of csrc/kernels/gemm.hip (BM x BN x BK)
constexpr int kNumCfg = 7;
constexpr int kXcd = 8;  // XCDs of an MI355X (workgroup w runs on XCD w % 8)
const int kTileM[kNumCfg] = {64, 128, 64, 128, 256, 128, 128};
const int kTileN[kNumCfg] = {64, 64, 128, 128, 128, 256, 128};
const int kTileK[kNumCfg] = {32, 32, 32, 64, 32, 32, 32};
// The fixed cost of a work item in K tiles, for balancing the per-workgroup
// lists: uniform batches of 128x128x64 items measured 8.4 us per item with
// one K tile and 22.7 us with eight, i.e. about 2 us per K tile plus 6.4 us
// (three K tiles) per item.
constexpr double kItemFixed = 3.0;

// LDS of the direct-to-LDS tile of configuration c (igemm_tile.h,
// lt_igemm_dl_lds_bytes: two stages of both operand tiles, unpadded rows)
int lds_bytes(int c) { return 2 * (kTileM[c] + kTileN[c]) * kTileK[c] * 2; }

void chk(int rc, const char *what) {
  ASSERT(rc == 0) << what << ": " << lt_hip_last_error();
}

std::string library_dir() {
  Dl_info info;
  ASSERT(dladdr((void *)&library_dir, &info) && info.dli_fname) << "cannot locate library";
  std::string p = info.dli_fname;
  return p.substr(0, p.find_last_of('/'));
}

struct GemmModule {
  lt_hip_module mod = nullptr;
  std::vector<char> code;
  void *fn(const std::string &name) {
    void *f = nullptr;
    chk(lt_hip_module_get_function(mod, name.c_str(), &f), name.c_str());
    return f;
  }
};

GemmModule &gemm_module() {
  static std::mutex mu;
  static GemmModule *m = nullptr;
  std::lock_guard<std::mutex> g(mu);
  if (!m) {
    loop_tool::hip::ensure_runtime();
    std::string path = library_dir() + "/kernels/gemm.hsaco";
    std::ifstream f(path, std::ios::binary);
    ASSERT(f.is_open()) << "kernel library missing: " << path << " (python loop_tool_amd/build.py)";
    auto *mm = new GemmModule();
    mm->code.assign(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>());
    chk(lt_hip_module_load(mm->code.data(), &mm->mod), "load gemm.hsaco");
    m = mm;  // kept for the process lifetime
  }
  return *m;
}

}  // namespace

struct lth_batch {
  int dtype = 0;  // 0 bf16, 1 f16
  std::vector<Desc> descs;
  std::vector<int> cfg;
  void *d_items = nullptr;
  void *d_queue = nullptr;  // grouped kernel's segment offsets and queues (see upload)
  int n_items = 0;
  // the work list in segment order on the host, and its segment offsets
  std::vector<Item> h_items;
  std::vector<int> h_xoff;
  // per-workgroup lists of the one-configuration kernels, one set per grid
  // size launched since the work list was built (a graph captured with one
  // grid keeps reading its set while launches with other grids run)
  struct Lists {
    void *items = nullptr, *offsets = nullptr;
  };
  std::map<int, Lists> lists;
  void free_lists() {
    for (auto &kv : lists) {
      if (kv.second.items) lt_hip_free(kv.second.items);
      if (kv.second.offsets) lt_hip_free(kv.second.offsets);
    }
    lists.clear();
  }
  bool dirty = true;
  void *graph = nullptr;
  void *graph_stream = nullptr;

  const char *tname() const { return dtype == 0 ? "bf16" : "f16"; }

  // The work list: every output tile of every GEMM, cut into kXcd segments,
  // one per XCD (workgroup w of the grouped launch runs on XCD w % kXcd and
  // takes items of segment w % kXcd only).  A GEMM's tiles go to one XCD as
  // long as that XCD stays within the average load, so the tiles that share
  // an A or B panel fetch it into one L2; row strips (the tiles sharing an A
  // panel) of GEMMs larger than that spill to the least-loaded XCDs.  Within
  // a segment, items are ordered by decreasing cost, a GEMM's tiles adjacent.
  void upload() {
    if (!dirty) return;
    struct Strip {
      int g, m0;
      int64_t cost;
    };
    std::vector<Strip> strips;
    std::vector<int64_t> gcost(descs.size(), 0);
    for (size_t g = 0; g < descs.size(); ++g) {
      const int c = cfg[g], bm = kTileM[c], bn = kTileN[c];
      const int64_t tile = (int64_t)bm * bn * descs[g].K;
      const int64_t ntn = (descs[g].N + bn - 1) / bn;
      for (int m0 = 0; m0 < descs[g].M; m0 += bm) {
        strips.push_back({(int)g, m0, ntn * tile});
        gcost[g] += ntn * tile;
      }
    }
    int64_t total = 0;
    for (auto c : gcost) total += c;
    const double target = (double)total / kXcd;
    // GEMMs by decreasing cost, strips of one GEMM in row order
    std::stable_sort(strips.begin(), strips.end(), [&](const Strip &a, const Strip &b) {
      return gcost[a.g] != gcost[b.g] ? gcost[a.g] > gcost[b.g] : a.g < b.g;
    });
    std::vector<double> load(kXcd, 0.0);
    std::vector<std::vector<Item>> seg(kXcd);
    std::vector<std::vector<int64_t>> segcost(kXcd);
    int prev_g = -1, prev_x = 0;
    for (const auto &st : strips) {
      int x = (int)(std::min_element(load.begin(), load.end()) - load.begin());
      if (st.g == prev_g && load[prev_x] + st.cost <= target) x = prev_x;
      load[x] += (double)st.cost;
      prev_g = st.g;
      prev_x = x;
      const int c = cfg[st.g], bm = kTileM[c], bn = kTileN[c];
      const Desc &d = descs[st.g];
      for (int n0 = 0; n0 < d.N; n0 += bn) {
        seg[x].push_back({d.A, d.Bt, d.C, d.M, d.N, d.K, st.m0, n0, c});
        segcost[x].push_back((int64_t)bm * bn * descs[st.g].K);
      }
    }
    std::vector<Item> sorted;
    std::vector<int> xoff(kXcd + 1, 0);
    for (int x = 0; x < kXcd; ++x) {
      std::vector<size_t> order(seg[x].size());
      std::iota(order.begin(), order.end(), 0);
      std::stable_sort(order.begin(), order.end(),
                       [&](size_t a, size_t b) { return segcost[x][a] > segcost[x][b]; });
      for (auto i : order) sorted.push_back(seg[x][i]);
      xoff[x + 1] = (int)sorted.size();
    }
    if (d_items) lt_hip_free(d_items);
    chk(lt_hip_malloc(&d_items, std::max<size_t>(sorted.size(), 1) * sizeof(Item)), "items");
    if (!sorted.empty())
      chk(lt_hip_memcpy(d_items, sorted.data(), sorted.size() * sizeof(Item), 0), "items copy");
    // queue: [kXcd + 1] segment offsets, then kXcd segment counters and the
    // count of workgroups done
    if (!d_queue) chk(lt_hip_malloc(&d_queue, (2 * kXcd + 2) * sizeof(int)), "queue");
    std::vector<int> q(2 * kXcd + 2, 0);
    std::copy(xoff.begin(), xoff.end(), q.begin());
    chk(lt_hip_memcpy(d_queue, q.data(), q.size() * sizeof(int), 0), "queue reset");
    n_items = (int)sorted.size();
    h_items = sorted;
    h_xoff = xoff;
    free_lists();  // (the graph that read them is dropped below)
    dirty = false;
    if (graph) {
      lt_hip_graph_destroy(graph);
      graph = nullptr;
    }
  }

  // the grouped kernel for the batch's configurations: the one-configuration
  // kernel when every item uses the same one, else the smallest register
  // class covering them (gemm.hip: small 0-2, mid adds 6, the full kernel
  // all); its dynamic LDS is sized for the largest configuration present
  bool one_config() const {
    for (int c : cfg)
      if (c != cfg[0]) return false;
    return !cfg.empty();
  }

  void *grouped_kernel(unsigned *lds) {
    int cls = 0;
    *lds = 0;
    for (int c : cfg) {
      cls = std::max(cls, c <= 2 ? 0 : c == 6 ? 1 : 2);
      *lds = std::max(*lds, (unsigned)lds_bytes(c));
    }
    if (one_config())
      return gemm_module().fn("lt_ggemm_c" + std::to_string(cfg[0]) + "_" + tname());
    const char *kname[3] = {"lt_ggemm_small_", "lt_ggemm_mid_", "lt_ggemm_"};
    return gemm_module().fn(std::string(kname[cls]) + tname());
  }

  // The per-workgroup lists of the one-configuration kernels for a grid of
  // G workgroups.  The items of XCD segment x go to the workgroups on that
  // XCD (w % 8 == x; with fewer than 8 workgroups, workgroup x % G takes the
  // whole segment), longest first, each to the least-loaded of them at a
  // cost of kItemFixed plus its K tiles; a list keeps the segment's order
  // (decreasing cost).  Uploaded with blocking copies: built before a graph
  // capture, never during one.
  const Lists &static_lists(int G) {
    auto found = lists.find(G);
    if (found != lists.end()) return found->second;
    std::vector<std::vector<Item>> per(G);
    typedef std::pair<double, int> Load;
    for (int x = 0; x < kXcd; ++x) {
      std::priority_queue<Load, std::vector<Load>, std::greater<Load>> least;
      if (G >= kXcd)
        for (int w = x; w < G; w += kXcd) least.push({0.0, w});
      else
        least.push({0.0, x % G});
      for (int i = h_xoff[x]; i < h_xoff[x + 1]; ++i) {
        const Item &it = h_items[i];
        const Load l = least.top();
        least.pop();
        per[l.second].push_back(it);
        const int kt = (it.K + kTileK[it.cfg] - 1) / kTileK[it.cfg];
        least.push({l.first + kItemFixed + kt, l.second});
      }
    }
    std::vector<Item> flat;
    std::vector<int> off(G + 1, 0);
    for (int w = 0; w < G; ++w) {
      flat.insert(flat.end(), per[w].begin(), per[w].end());
      off[w + 1] = (int)flat.size();
    }
    Lists l;
    chk(lt_hip_malloc(&l.items, std::max<size_t>(flat.size(), 1) * sizeof(Item)), "lists");
    lists[G] = l;  // freed with the others if a copy below fails
    chk(lt_hip_malloc(&lists[G].offsets, off.size() * sizeof(int)), "list offsets");
    if (!flat.empty())
      chk(lt_hip_memcpy(lists[G].items, flat.data(), flat.size() * sizeof(Item), 0),
          "lists copy");
    chk(lt_hip_memcpy(lists[G].offsets, off.data(), off.size() * sizeof(int), 0),
        "offsets copy");
    return lists[G];
  }

  // the workgroups of a grouped launch: grid <= 0 means as many as the GPU
  // holds at once (workgroups beyond that would find the queue empty or
  // nearly so), and never more than there are items; prepares everything
  // the launch reads (work list, per-workgroup lists)
  int launch_grid(int grid) {
    upload();
    if (grid <= 0) {
      unsigned lds = 0;
      void *fn = grouped_kernel(&lds);
      chk(lt_hip_max_resident_blocks(fn, 256, lds, &grid), "occupancy");
    }
    grid = std::max(1, std::min(grid, std::max(n_items, 1)));
    if (n_items && one_config()) static_lists(grid);
    return grid;
  }

  void grouped(int grid, void *stream) {
    grid = launch_grid(grid);
    if (!n_items) return;
    unsigned lds = 0;
    void *fn = grouped_kernel(&lds);
    const bool one = one_config();
    void *a0 = d_items, *a1 = d_queue;
    if (one) {
      const Lists &l = static_lists(grid);
      a0 = l.items;
      a1 = l.offsets;
    }
    void *args[] = {&a0, &a1};
    chk(lt_hip_launch(fn, grid, 1, 1, 256, 1, 1, lds, stream, args), "grouped launch");
  }

  void plain(int cfg_all, void *stream) {
    for (size_t i = 0; i < descs.size(); ++i) {
      int c = cfg_all >= 0 ? cfg_all : cfg[i];
      auto &d = descs[i];
      std::string name = std::string("lt_gemm_") + tname() + "_" + std::to_string(kTileM[c]) +
                         "x" + std::to_string(kTileN[c]) + "x" + std::to_string(kTileK[c]);
      void *fn = gemm_module().fn(name);
      const void *A = d.A, *Bt = d.Bt;
      float *C = d.C;
      int M = d.M, N = d.N, K = d.K;
      void *args[] = {&A, &Bt, &C, &M, &N, &K};
      unsigned gx = (N + kTileN[c] - 1) / kTileN[c], gy = (M + kTileM[c] - 1) / kTileM[c];
      chk(lt_hip_launch(fn, gx, gy, 1, 256, 1, 1, 0, stream, args), "plain launch");
    }
  }

  ~lth_batch() {
    if (loop_tool::hip::runtime_exiting()) return;
    if (graph) lt_hip_graph_destroy(graph);
    if (d_items) lt_hip_free(d_items);
    if (d_queue) lt_hip_free(d_queue);
    free_lists();
  }
};

namespace {
thread_local std::string g_batch_err;
template <typename F>
int guarded(F f) {
  try {
    return f();
  } catch (const std::exception &e) {
    g_batch_err = e.what();
  }
  return -1;
}
}  // namespace

extern "C" {

const char *lth_batch_last_error(void) { return g_batch_err.c_str(); }

lth_batch *lth_batch_create(int n, const int *M, const int *N, const int *K, int dtype) {
  try {
    ASSERT(n > 0) << "empty batch";
    ASSERT(dtype == 0 || dtype == 1) << "dtype must be 0 (bf16) or 1 (f16)";
    auto *b = new lth_batch();
    b->dtype = dtype;
    for (int i = 0; i < n; ++i) {
      ASSERT(M[i] > 0 && N[i] > 0 && K[i] > 0 && K[i] % 8 == 0)
          << "GEMM " << i << ": sizes must be positive and K a multiple of 8";
      // the tile addresses operands through 32-bit buffer ranges and row
      // offsets: each 16-bit operand must stay under 2^31 bytes
      const int64_t ld = LTH_BATCH_PITCH((int64_t)K[i]);
      ASSERT((int64_t)M[i] * ld * 2 < (1LL << 31) && (int64_t)N[i] * ld * 2 < (1LL << 31))
          << "GEMM " << i << ": a 16-bit operand of " << M[i] << "x" << ld << " or "
          << N[i] << "x" << ld << " reaches 2^31 bytes";
      b->descs.push_back({nullptr, nullptr, nullptr, M[i], N[i], K[i], 0});
    }
    b->cfg.assign(n, 3);
    return b;
  } catch (const std::exception &e) {
    g_batch_err = e.what();
    return nullptr;
  }
}

int lth_batch_set_operands(lth_batch *b, void *const *A, void *const *Bt, void *const *C) {
  return guarded([&] {
    for (size_t i = 0; i < b->descs.size(); ++i) {
      b->descs[i].A = A[i];
      b->descs[i].Bt = Bt[i];
      b->descs[i].C = (float *)C[i];
    }
    b->dirty = true;
    return 0;
  });
}

int lth_batch_set_configs(lth_batch *b, const int *cfg) {
  return guarded([&] {
    for (size_t i = 0; i < b->descs.size(); ++i) {
      ASSERT(cfg[i] >= 0 && cfg[i] < kNumCfg) << "tile config out of range";
      b->cfg[i] = cfg[i];
    }
    b->dirty = true;
    return 0;
  });
}

int lth_batch_num_items(lth_batch *b) {
  return guarded([&] {
    b->upload();
    return b->n_items;
  });
}

int lth_batch_run_plain(lth_batch *b, int cfg, void *stream) {
  return guarded([&] {
    ASSERT(cfg >= -1 && cfg < kNumCfg) << "tile config out of range";
    b->plain(cfg, stream);
    return 0;
  });
}

int lth_batch_run_grouped(lth_batch *b, int grid, void *stream) {
  return guarded([&] {
    b->grouped(grid, stream);
    return 0;
  });
}

int lth_batch_prepare(lth_batch *b, int grid) {
  return guarded([&] { return b->launch_grid(grid); });
}

int lth_batch_capture(lth_batch *b, int grid, void *stream) {
  return guarded([&] {
    ASSERT(stream) << "graph capture needs a created stream (not the null stream)";
    // uploads, module load, the occupancy query and the per-workgroup lists
    // happen outside capture
    grid = b->launch_grid(grid);
    if (b->graph) {
      lt_hip_graph_destroy(b->graph);
      b->graph = nullptr;
    }
    chk(lt_hip_graph_begin_capture(stream), "begin capture");
    b->grouped(grid, stream);
    chk(lt_hip_graph_end_capture(stream, &b->graph), "end capture");
    return 0;
  });
}

int lth_batch_replay(lth_batch *b, void *stream) {
  return guarded([&] {
    ASSERT(b->graph && !b->dirty) << "capture the batch first (lth_batch_capture)";
    chk(lt_hip_graph_launch(b->graph, stream), "graph launch");
    return 0;
  });
}

void lth_batch_free(lth_batch *b) { delete b; }

int lth_convert(const float *src, void *dst, int rows, int cols, int transpose, int dtype,
                void *stream) {
  return guarded([&] {
    ASSERT(dtype == 0 || dtype == 1) << "dtype must be 0 (bf16) or 1 (f16)";
    void *fn = gemm_module().fn(dtype == 0 ? "lt_cvt_bf16" : "lt_cvt_f16");
    const float *s = src;
    void *d = dst;
    int r = rows, c = cols, t = transpose;
    void *args[] = {&s, &d, &r, &c, &t};
    long n = (long)rows * cols;
    chk(lt_hip_launch(fn, (unsigned)((n + 255) / 256), 1, 1, 256, 1, 1, 0, stream, args),
        "convert launch");
    return 0;
  });
}

int lth_tile_config(int cfg, int *bm, int *bn) {
  if (cfg < 0 || cfg >= kNumCfg) return -1;
  *bm = kTileM[cfg];
  *bn = kTileN[cfg];
  return 0;
}

}  // extern "C"
