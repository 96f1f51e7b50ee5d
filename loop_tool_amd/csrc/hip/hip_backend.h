// HipCompiled: the Compiled object of the "hip" backend (see hip_backend.cpp).
#pragma once

#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "emit.h"
#include "loop_tool/backend.h"
#include "lt_hip.h"

namespace loop_tool {
namespace hip {

void ensure_runtime();
// true once the process has started exiting (device frees are skipped)
bool runtime_exiting();
std::shared_ptr<std::vector<char>> compile_source(const std::string &source,
                                                  const std::string &name,
                                                  bool *from_cache);
// drop a module source's code object from the in-process and disk caches
void evict_code(const std::string &source);
Knobs knobs_from(const LoopTree &lt, bool bf16_variant);

class HipCompiled : public Compiled {
 public:
  HipCompiled(const LoopTree &lt, bool bf16_variant);
  ~HipCompiled() override;

  void run(const std::vector<void *> &memory, bool sync = true) const override;
  std::string dump() const override;

  // Launch on a caller stream with caller-owned device buffers (inputs then
  // outputs) and scratch (scratch_sizes() floats each); nothing is
  // allocated, copied or synchronised.  Safe to call concurrently when every
  // caller brings its own scratch.
  void run_on(const std::vector<void *> &device_io, const std::vector<void *> &scratch,
              void *stream) const;
  std::vector<int64_t> scratch_sizes() const;
  // Launch only stage k of the plan on `stream` with device buffers (inputs
  // then outputs) and the calling thread's scratch of the Compiled (timing of
  // one kernel).
  void run_stage(const std::vector<void *> &device_io, int k, void *stream) const;

  // load the plan's code object now (before a graph capture, say)
  void prepare() const {
    std::lock_guard<std::mutex> g(mu_);
    load();
  }
  // true once the plan's code object is loaded on the device (by any run,
  // run_on, run_stage, stream() or prepare())
  bool loaded() const {
    std::lock_guard<std::mutex> g(mu_);
    return loaded_;
  }

  const Plan &plan() const { return plan_; }
  // the stored schedule this Compiled was lowered from (dump() shows it)
  std::string schedule_note;
  int64_t algorithmic_bytes() const { return algorithmic_bytes_; }
  int64_t algorithmic_flops() const { return algorithmic_flops_; }
  // the calling thread's stream of this Compiled (run() launches on it)
  void *stream() const { return context().stream; }
  size_t num_inputs() const;
  size_t num_io() const;

 private:
  // What run() uses on one calling thread: a stream, the plan's scratch
  // (two sets with overlap), device copies of host buffers and the overlap
  // stream and events.  Created on the thread's first call and used by that
  // thread only, so runs from different threads proceed concurrently, each
  // in order on its own stream.  A context lives as long as the Compiled
  // (threads that end leave theirs): one stream and one scratch set per
  // thread that ever ran it.
  struct RunContext {
    void *stream = nullptr;
    std::vector<void *> scratch, scratch2;
    std::unordered_map<size_t, std::pair<void *, size_t>> staging;
    void *side = nullptr;
    void *ev_staged[2] = {nullptr, nullptr};
    void *ev_done[2] = {nullptr, nullptr};
    uint64_t runs = 0;
  };
  RunContext &context() const;  // the calling thread's (loads the code object first)
  void load() const;            // with mu_ held
  void ensure_scratch(RunContext &c, int set = 0) const;
  void launch(const std::vector<void *> &device_buffers, void *stream, int first = 0,
              int last = -1) const;
  std::vector<void *> device_buffers(RunContext &c, const std::vector<void *> &io,
                                     int set = 0) const;

  Plan plan_;
  Knobs knobs_;
  mutable std::shared_ptr<std::vector<char>> code_;
  int64_t algorithmic_bytes_ = 0;
  int64_t algorithmic_flops_ = 0;

  mutable std::mutex mu_;  // guards loading and contexts_
  mutable bool loaded_ = false;
  mutable lt_hip_module module_ = nullptr;
  mutable std::vector<void *> functions_;
  // overlap of the staging stage with the previous run (see run())
  bool overlap_ = false;
  mutable std::unordered_map<std::thread::id, std::unique_ptr<RunContext>> contexts_;
};

}  // namespace hip
}  // namespace loop_tool
