// Launch-plan construction for the hip backend.
//
// Order of preference for each LoopTree:
//   1. contraction on MFMA (emit_contract.cpp)  - dense matmul / conv nests
//   2. row-cooperative reduction (emit_rowreduce.cpp) - softmax-like rows
//   3. pointwise lowering (emit_pointwise.cpp)  - everything else, exact
// The "hip:plain" annotation (or knobs.force_plain) selects 3 directly.
#include <sstream>

#include "emit.h"
#include "emit_internal.h"

namespace loop_tool {
namespace hip {

const char *kernel_prelude() {
  return R"(// generated by the loop_tool hip backend for gfx950
// (hipRTC provides the HIP device runtime implicitly)
typedef float lt_f4 __attribute__((ext_vector_type(4)));
typedef float lt_f2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ float lt_selmax(float a, float b) { return a > b ? a : b; }
__device__ __forceinline__ float lt_selmin(float a, float b) { return a < b ? a : b; }
)";
}

namespace {

// A reduction whose value is reused across extra loop vars (or by several
// consumers) is computed once into scratch by its own stage instead of being
// recomputed inside every consumer element.
std::map<IR::NodeRef, int> choose_materialized(const Dataflow &df) {
  std::map<IR::NodeRef, int> out;
  int next = 0;
  for (auto nr : df.topo) {
    const auto &n = df.nodes.at(nr);
    if (n.op == Operation::write || n.red_vars.empty()) continue;
    bool reuse = n.consumers > 1;
    for (const auto &p : df.nodes) {
      const auto &c = p.second;
      for (const auto &in : c.inputs) {
        if (in.src != nr) continue;
        if (c.vars.size() + c.red_vars.size() > n.vars.size()) reuse = true;
      }
    }
    if (reuse) out[nr] = next++;
  }
  return out;
}

}  // namespace

Plan build_plan(const Dataflow &df, const Knobs &knobs) {
  Plan plan;
  int n_in = df.num_inputs;
  for (int i = 0; i < n_in; ++i) {
    auto nr = df.ir().inputs()[i];
    plan.buffers.push_back({BufferSpec::INPUT, i, df.numel(df.nodes.at(nr).vars), nr});
  }
  for (size_t j = 0; j < df.writes.size(); ++j) {
    auto nr = df.writes[j];
    plan.buffers.push_back(
        {BufferSpec::OUTPUT, (int)j, df.numel(df.nodes.at(nr).vars), nr});
  }

  if (!knobs.force_plain) {
    if (try_contract(df, knobs, plan)) return plan;
    if (try_rowreduce(df, knobs, plan)) return plan;
  }

  auto mat = choose_materialized(df);
  std::map<int, std::string> input_args;
  for (int i = 0; i < n_in; ++i) input_args[i] = "in" + std::to_string(i);
  std::map<IR::NodeRef, int> scratch_buf;
  for (const auto &p : mat) {
    scratch_buf[p.first] = (int)plan.buffers.size();
    plan.buffers.push_back({BufferSpec::SCRATCH, p.second,
                            df.numel(df.nodes.at(p.first).vars), p.first});
  }

  std::stringstream src;
  src << kernel_prelude();
  std::stringstream desc;
  desc << "[hip] pointwise lowering, " << mat.size() << " materialised reductions\n";
  int block = knobs.block > 0 ? knobs.block : 256;
  // 32-bit index math when every buffer (and so every in-range index) fits
  bool int32 = true;
  for (const auto &b : plan.buffers) int32 &= b.numel < (1LL << 30);

  std::vector<IR::NodeRef> targets;
  for (auto nr : df.topo)
    if (mat.count(nr) || df.nodes.at(nr).op == Operation::write) targets.push_back(nr);

  int k = 0;
  const int block0 = block;
  for (auto target : targets) {
    block = block0;
    PointwiseEmitter em(df, mat, input_args);
    Stage st;
    st.kernel = "lt_pw" + std::to_string(k++);
    st.kind = "pointwise";
    // arguments: every input, every scratch buffer, then the target buffer
    std::stringstream sig;
    for (int i = 0; i < n_in; ++i) {
      sig << "const float* __restrict__ in" << i << ", ";
      st.buffers.push_back(i);
    }
    for (const auto &p : mat) {
      if (p.first == target) continue;
      sig << "const float* __restrict__ s" << p.second << ", ";
      st.buffers.push_back(scratch_buf.at(p.first));
    }
    int out_buf;
    const auto &tn = df.nodes.at(target);
    if (tn.op == Operation::write)
      out_buf = n_in + tn.output_index;
    else
      out_buf = scratch_buf.at(target);
    sig << "float* __restrict__ out";
    st.buffers.push_back(out_buf);
    st.writes.push_back(out_buf);
    // the schedule's split of the target's contiguous var: its innermost
    // extent (1, 2 or 4) sets the elements per lane, the loop above it
    // (64..1024 lanes, a multiple of 64) the workgroup size
    int vec = knobs.vec, tblock = block;
    const auto &tvars = df.nodes.at(target).vars;
    if (!tvars.empty()) {
      const int csym = df.sym(tvars.back()).id();
      const int64_t e = df.split_extent(target, csym), t = df.split_extent(target, csym, 1);
      if (vec == 0 && (e == 1 || e == 2 || e == 4)) vec = (int)e;
      if (!knobs.block_set && t >= 64 && t <= 1024 && t % 64 == 0) tblock = (int)t;
    }
    int64_t threads = 0;
    src << em.kernel(st.kernel, target, sig.str(), "out", tblock, &threads, vec, int32)
        << "\n";
    block = tblock;
    int64_t blocks = (threads + block - 1) / block;
    ASSERT(blocks < (1LL << 31)) << "grid too large";
    st.grid[0] = (unsigned)blocks;
    st.block[0] = block;
    plan.config["block"] = block;
    if (vec) plan.config["vec"] = vec;
    // algorithmic bytes: every input element read once, the target written once
    int64_t bytes = df.numel(df.nodes.at(target).vars) * 4;
    for (int i = 0; i < n_in; ++i) bytes += df.numel(df.nodes.at(df.ir().inputs()[i]).vars) * 4;
    st.bytes = bytes;
    plan.stages.push_back(st);
    desc << "  stage " << st.kernel << ": " << df.ir().dump(target) << " over "
         << threads << " lanes of " << block << " per workgroup" << (vec ? ", " + std::to_string(vec) + " elements per lane at most" : std::string()) << "\n";
  }
  plan.source = src.str();
  plan.description = desc.str();
  return plan;
}

}  // namespace hip
}  // namespace loop_tool
