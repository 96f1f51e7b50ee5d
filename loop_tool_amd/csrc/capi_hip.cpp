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
  std::vector<Nest> nests;
  std::vector<void *> streams, events;  // extra lanes; fork and join events
  void *graph = nullptr;
  ~lth_pack() {
    if (loop_tool::hip::runtime_exiting()) return;
    if (graph) lt_hip_graph_destroy(graph);
    for (auto &n : nests)
      for (auto p : n.scratch) lt_hip_free(p);
    for (auto e : events) lt_hip_event_destroy(e);
    for (auto st : streams) lt_hip_stream_destroy(st);
  }
};

extern "C" {

lth_pack *lth_pack_create(const ltc_compiled *const *cc, void *const *io, const int *n_io,
                          int n, int lanes, void *stream) {
  auto *p = new lth_pack();
  try {
    g_err.clear();
    ASSERT(n > 0 && cc && io && n_io) << "empty pack";
    ASSERT(lanes >= 1 && lanes <= 16) << "lanes must be in [1, 16]";
    ASSERT(stream) << "graph capture needs a created stream (not the null stream)";
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
      for (auto words : nest.h->scratch_sizes()) {
        void *b = nullptr;
        chk(lt_hip_malloc(&b, std::max<int64_t>(words, 1) * sizeof(float)), "pack scratch");
        nest.scratch.push_back(b);
      }
      const bool f32 = nest.h->plan().stages.back().kind == "contract_f32";
      nest.est = std::max(nest.h->algorithmic_bytes() / 8e12,
                          nest.h->algorithmic_flops() / (f32 ? 157.3e12 : 2.5e15));
      p->nests.push_back(std::move(nest));
    }
    // lanes: longest first, each nest to the lane with the least estimated time
    std::vector<int> order(n);
    for (int i = 0; i < n; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int b) { return p->nests[a].est > p->nests[b].est; });
    std::vector<double> load(lanes, 0.0);
    std::vector<std::vector<int>> lane(lanes);
    for (int i : order) {
      const int l = (int)(std::min_element(load.begin(), load.end()) - load.begin());
      lane[l].push_back(i);
      load[l] += p->nests[i].est;
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
        for (int i : lane[l]) p->nests[i].h->run_on(p->nests[i].io, p->nests[i].scratch, st);
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
