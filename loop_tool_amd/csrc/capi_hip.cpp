// hip backend extensions of the C ABI (see include/loop_tool_hip.h).
#include <algorithm>
#include <cstring>
#include <exception>
#include <string>

#include "hip/hip_backend.h"
#include "loop_tool/error.h"
#include "hip/tuned_store.h"
#include "loop_tool_hip.h"
#include "lt_hip.h"

struct ltc_compiled {
  std::shared_ptr<loop_tool::Compiled> cc;
};
struct ltc_looptree {
  loop_tool::LoopTree lt;
};

namespace {

thread_local std::string g_err;

const loop_tool::hip::HipCompiled *as_hip(const ltc_compiled *cc) {
  auto *h = dynamic_cast<const loop_tool::hip::HipCompiled *>(cc->cc.get());
  if (!h) throw std::runtime_error("not a hip Compiled");
  return h;
}

template <typename F>
int guarded(F f) {
  try {
    return f();
  } catch (const std::exception &e) {
    g_err = e.what();
  }
  return -1;
}

int copy_str(const std::string &s, char *buf, int cap) {
  if (buf && cap > 0) {
    size_t n = std::min(s.size(), (size_t)cap - 1);
    std::memcpy(buf, s.data(), n);
    buf[n] = '\0';
  }
  return (int)s.size();
}

}  // namespace

namespace loop_tool {
namespace hip {
// the message lth_last_error() reports, for entry points defined elsewhere
void set_lth_error(const std::string &msg) { g_err = msg; }
}  // namespace hip
}  // namespace loop_tool

extern "C" {

const char *lth_last_error(void) { return g_err.c_str(); }

int lth_is_hip(const ltc_compiled *cc) {
  return dynamic_cast<const loop_tool::hip::HipCompiled *>(cc->cc.get()) ? 1 : 0;
}

int lth_compiled_stream(const ltc_compiled *cc, void **stream) {
  return guarded([&] {
    *stream = as_hip(cc)->stream();
    return 0;
  });
}

int lth_num_stages(const ltc_compiled *cc) {
  return guarded([&] { return (int)as_hip(cc)->plan().stages.size(); });
}

int lth_stage_info(const ltc_compiled *cc, int i, char *name, int name_cap, char *kind,
                   int kind_cap, int64_t *bytes, int64_t *flops) {
  return guarded([&] {
    const auto &s = as_hip(cc)->plan().stages.at(i);
    copy_str(s.kernel, name, name_cap);
    copy_str(s.kind, kind, kind_cap);
    *bytes = s.bytes;
    *flops = s.flops;
    return 0;
  });
}

int lth_scratch_sizes(const ltc_compiled *cc, int64_t *out, int cap) {
  return guarded([&] {
    auto s = as_hip(cc)->scratch_sizes();
    for (int i = 0; i < (int)s.size() && i < cap; ++i) out[i] = s[i];
    return (int)s.size();
  });
}

int lth_tuned_set_dir(const char *dir) {
  return guarded([&] {
    loop_tool::hip::set_tuned_dir(dir ? dir : "");
    return 0;
  });
}

int lth_schedule_key(const ltc_looptree *program, char *buf, int cap) {
  return guarded([&] { return copy_str(loop_tool::hip::schedule_key(program->lt), buf, cap); });
}

int lth_tuned_put(const ltc_looptree *program, const ltc_looptree *schedule) {
  return guarded([&] {
    loop_tool::hip::tuned_put(program->lt, schedule->lt);
    return 0;
  });
}

ltc_looptree *lth_tuned_get(const ltc_looptree *program) {
  try {
    g_err.clear();
    auto t = loop_tool::hip::tuned_get(program->lt, nullptr);
    return t ? new ltc_looptree{*t} : nullptr;
  } catch (const std::exception &e) {
    g_err = e.what();
  }
  return nullptr;
}

int lth_run_on(const ltc_compiled *cc, void *const *io, int n_io, void *const *scratch,
               int n_scratch, void *stream) {
  return guarded([&] {
    as_hip(cc)->run_on(std::vector<void *>(io, io + n_io),
                       std::vector<void *>(scratch, scratch + n_scratch), stream);
    return 0;
  });
}

int lth_time(const ltc_compiled *cc, void *const *io, int n_io, int warmup, int iters,
             float *ms) {
  return guarded([&] {
    auto *h = as_hip(cc);
    std::vector<void *> mem(io, io + n_io);
    for (int i = 0; i < warmup; ++i) h->run(mem, false);
    void *s = h->stream();
    void *e0 = nullptr, *e1 = nullptr;
    auto chk = [](int rc) {
      if (rc) throw std::runtime_error(lt_hip_last_error());
    };
    chk(lt_hip_event_create(&e0));
    chk(lt_hip_event_create(&e1));
    chk(lt_hip_stream_synchronize(s));
    chk(lt_hip_event_record(e0, s));
    for (int i = 0; i < iters; ++i) h->run(mem, false);
    chk(lt_hip_event_record(e1, s));
    chk(lt_hip_event_synchronize(e1));
    float total = 0;
    chk(lt_hip_event_elapsed_ms(e0, e1, &total));
    lt_hip_event_destroy(e0);
    lt_hip_event_destroy(e1);
    *ms = total / (iters > 0 ? iters : 1);
    return 0;
  });
}

int lth_time_stage(const ltc_compiled *cc, void *const *io, int n_io, int stage, int warmup,
                   int iters, float *ms) {
  return guarded([&] {
    auto *h = as_hip(cc);
    std::vector<void *> mem(io, io + n_io);
    auto chk = [](int rc) {
      if (rc) throw std::runtime_error(lt_hip_last_error());
    };
    void *s = h->stream();
    h->run(mem, false);  // every stage once: scratch holds its real contents
    for (int i = 0; i < warmup; ++i) h->run_stage(mem, stage, s);
    void *e0 = nullptr, *e1 = nullptr;
    chk(lt_hip_event_create(&e0));
    chk(lt_hip_event_create(&e1));
    chk(lt_hip_stream_synchronize(s));
    chk(lt_hip_event_record(e0, s));
    for (int i = 0; i < iters; ++i) h->run_stage(mem, stage, s);
    chk(lt_hip_event_record(e1, s));
    chk(lt_hip_event_synchronize(e1));
    float total = 0;
    chk(lt_hip_event_elapsed_ms(e0, e1, &total));
    lt_hip_event_destroy(e0);
    lt_hip_event_destroy(e1);
    *ms = total / (iters > 0 ? iters : 1);
    return 0;
  });
}

}  // extern "C"

// ---- packs of compiled nests in one graph -------------------------------
struct lth_pack {
  struct Nest {
    std::shared_ptr<loop_tool::Compiled> keep;  // the Compiled outlives the graph
    const loop_tool::hip::HipCompiled *h = nullptr;
    std::vector<void *> io, scratch;
    double est = 0;  // estimated seconds: the plan's roofline time
  };
  // LTH_PACK_GROUP_MATMULS: the plain 16-bit matmul nests of one operand
  // type, run as their operands' conversions and one grouped launch
  struct Cvt {
    const float *src;
    void *dst;
    int rows, cols, transpose;
  };
  struct Group {
    lth_batch *b = nullptr;
    int dtype = 0, grid = 0;
    std::vector<Cvt> cvts;
    double est = 0;
  };
  std::vector<Nest> nests;
  std::vector<Group> groups;
  std::vector<void *> staged;           // the groups' 16-bit operands
  std::vector<void *> streams, events;  // extra lanes; fork and join events
  void *graph = nullptr;
  ~lth_pack() {
    if (loop_tool::hip::runtime_exiting()) return;
    if (graph) lt_hip_graph_destroy(graph);
    for (auto &g : groups)
      if (g.b) lth_batch_free(g.b);
    for (auto q : staged) lt_hip_free(q);
    for (auto &n : nests)
      for (auto q : n.scratch) lt_hip_free(q);
    for (auto e : events) lt_hip_event_destroy(e);
    for (auto st : streams) lt_hip_stream_destroy(st);
  }
};

extern "C" {

lth_pack *lth_pack_create(const ltc_compiled *const *cc, void *const *io, const int *n_io,
                          int n, int lanes, int flags, void *stream) {
  auto *p = new lth_pack();
  try {
    g_err.clear();
    ASSERT(n > 0 && cc && io && n_io) << "empty pack";
    ASSERT(lanes >= 1 && lanes <= 16) << "lanes must be in [1, 16]";
    ASSERT(stream) << "graph capture needs a created stream (not the null stream)";
    ASSERT((flags & ~LTH_PACK_GROUP_MATMULS) == 0) << "unknown pack flags " << flags;
    auto chk = [](int rc, const char *what) {
      if (rc) throw std::runtime_error(std::string(what) + ": " + lt_hip_last_error());
    };
    size_t off = 0;
    for (int i = 0; i < n; ++i) {
      lth_pack::Nest nest;
      nest.keep = cc[i]->cc;
      nest.h = as_hip(cc[i]);
      ASSERT(n_io[i] >= (int)nest.h->num_io())
          << "nest " << i << ": expected " << nest.h->num_io() << " device buffers";
      nest.io.assign(io + off, io + off + n_io[i]);
      off += n_io[i];
      nest.h->prepare();  // code objects load outside the capture
      const bool f32 = nest.h->plan().stages.back().kind == "contract_f32";
      nest.est = std::max(nest.h->algorithmic_bytes() / 8e12,
                          nest.h->algorithmic_flops() / (f32 ? 157.3e12 : 2.5e15));
      p->nests.push_back(std::move(nest));
    }
    // grouped matmul nests: 16-bit copies of their operands (A[M][K] and the
    // transposed Bt[N][K]) and one batch per operand type, made ready for a
    // launch that allocates and copies nothing
    std::vector<char> grouped(n, 0);
    if (flags & LTH_PACK_GROUP_MATMULS) {
      for (int dtype = 0; dtype < 2; ++dtype) {
        const char *kind = dtype ? "contract_f16" : "contract_bf16";
        lth_pack::Group g;
        g.dtype = dtype;
        std::vector<int> M, N, K;
        std::vector<void *> A16, Bt16, Cs;
        for (int i = 0; i < n; ++i) {
          const auto &pl = p->nests[i].h->plan();
          const auto &mm = pl.matmul;
          const int64_t ld = LTH_BATCH_PITCH(mm.K);
          if (!mm.valid || pl.stages.empty() || pl.stages.back().kind != kind ||
              mm.K % 8 || mm.M * ld * 2 >= (1LL << 31) || mm.N * ld * 2 >= (1LL << 31))
            continue;
          void *a = nullptr, *bt = nullptr;
          chk(lt_hip_malloc(&a, mm.M * ld * 2), "pack operand");
          p->staged.push_back(a);
          chk(lt_hip_malloc(&bt, mm.N * ld * 2), "pack operand");
          p->staged.push_back(bt);
          const auto &nio = p->nests[i].io;
          g.cvts.push_back({(const float *)nio[mm.a_input], a, (int)mm.M, (int)mm.K, 0});
          g.cvts.push_back({(const float *)nio[mm.b_input], bt, (int)mm.K, (int)mm.N, 1});
          M.push_back((int)mm.M);
          N.push_back((int)mm.N);
          K.push_back((int)mm.K);
          A16.push_back(a);
          Bt16.push_back(bt);
          Cs.push_back(nio[p->nests[i].h->num_inputs()]);
          g.est += p->nests[i].est;
          grouped[i] = 1;
        }
        if (M.empty()) continue;
        g.b = lth_batch_create((int)M.size(), M.data(), N.data(), K.data(), dtype);
        if (!g.b) throw std::runtime_error(std::string("pack batch: ") + lth_batch_last_error());
        p->groups.push_back(g);
        if (lth_batch_set_operands(g.b, A16.data(), Bt16.data(), Cs.data()) != 0 ||
            (p->groups.back().grid = lth_batch_prepare(g.b, 0)) < 0)
          throw std::runtime_error(std::string("pack batch: ") + lth_batch_last_error());
      }
    }
    // scratch of the nests that run their own plans
    for (int i = 0; i < n; ++i) {
      if (grouped[i]) continue;
      for (auto words : p->nests[i].h->scratch_sizes()) {
        void *b = nullptr;
        chk(lt_hip_malloc(&b, std::max<int64_t>(words, 1) * sizeof(float)), "pack scratch");
        p->nests[i].scratch.push_back(b);
      }
    }
    // lanes: longest first, each nest (or group, numbered n + g) to the lane
    // with the least estimated time
    std::vector<int> order;
    for (int i = 0; i < n; ++i)
      if (!grouped[i]) order.push_back(i);
    for (int g = 0; g < (int)p->groups.size(); ++g) order.push_back(n + g);
    auto est = [&](int i) { return i < n ? p->nests[i].est : p->groups[i - n].est; };
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return est(a) > est(b); });
    std::vector<double> load(lanes, 0.0);
    std::vector<std::vector<int>> lane(lanes);
    for (int i : order) {
      const int l = (int)(std::min_element(load.begin(), load.end()) - load.begin());
      lane[l].push_back(i);
      load[l] += est(i);
    }
    for (int l = 1; l < lanes; ++l) {
      void *st = nullptr;
      chk(lt_hip_stream_create(&st), "pack stream");
      p->streams.push_back(st);
    }
    for (int e = 0; e < lanes; ++e) {  // fork, then one join per extra lane
      void *ev = nullptr;
      chk(lt_hip_event_create(&ev), "pack event");
      p->events.push_back(ev);
    }
    chk(lt_hip_graph_begin_capture(stream), "begin capture");
    try {
      if (lanes > 1) {
        chk(lt_hip_event_record(p->events[0], stream), "fork");
        for (auto st : p->streams) chk(lt_hip_stream_wait_event(st, p->events[0]), "fork wait");
      }
      for (int l = 0; l < lanes; ++l) {
        void *st = l == 0 ? stream : p->streams[l - 1];
        for (int i : lane[l]) {
          if (i < n) {
            p->nests[i].h->run_on(p->nests[i].io, p->nests[i].scratch, st);
            continue;
          }
          const auto &g = p->groups[i - n];
          for (const auto &c : g.cvts)
            if (lth_convert(c.src, c.dst, c.rows, c.cols, c.transpose, g.dtype, st) != 0)
              throw std::runtime_error(std::string("pack convert: ") + lth_batch_last_error());
          if (lth_batch_run_grouped(g.b, g.grid, st) != 0)
            throw std::runtime_error(std::string("pack grouped launch: ") +
                                     lth_batch_last_error());
        }
      }
      for (int l = 1; l < lanes; ++l) {
        chk(lt_hip_event_record(p->events[l], p->streams[l - 1]), "join");
        chk(lt_hip_stream_wait_event(stream, p->events[l]), "join wait");
      }
    } catch (...) {
      void *g = nullptr;
      lt_hip_graph_end_capture(stream, &g);  // leave capture mode before reporting
      if (g) lt_hip_graph_destroy(g);
      throw;
    }
    chk(lt_hip_graph_end_capture(stream, &p->graph), "end capture");
    return p;
  } catch (const std::exception &e) {
    g_err = e.what();
    delete p;
    return nullptr;
  }
}

int lth_pack_replay(lth_pack *p, void *stream) {
  return guarded([&] {
    ASSERT(p && p->graph) << "no pack";
    if (lt_hip_graph_launch(p->graph, stream))
      throw std::runtime_error(std::string("graph launch: ") + lt_hip_last_error());
    return 0;
  });
}

void lth_pack_free(lth_pack *p) { delete p; }

}  // extern "C"
