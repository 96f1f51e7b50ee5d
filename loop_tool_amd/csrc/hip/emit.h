// HIP source emitters for the nest classes of the hip backend.
#pragma once

#include <map>
#include <string>
#include <vector>

#include "nest.h"

namespace loop_tool {
namespace hip {

// One device buffer of a plan: an IR input, an IR output, or backend scratch.
struct BufferSpec {
  enum Kind { INPUT, OUTPUT, SCRATCH } kind;
  int index;       // position among inputs/outputs, or scratch number
  int64_t numel;   // 4-byte words (float elements; 16-bit scratch holds two each)
  IR::NodeRef node;
};

// One kernel launch of a plan.
struct Stage {
  std::string kernel;          // extern "C" function name in the module
  std::string kind;            // "pointwise", "rowreduce", "contract", ...
  std::vector<int> buffers;    // plan buffer ids, in kernel argument order
  std::vector<int> writes;     // plan buffer ids the stage writes
  unsigned grid[3] = {1, 1, 1};
  unsigned block[3] = {256, 1, 1};
  unsigned shared_bytes = 0;
  int64_t bytes = 0;           // algorithmic HBM bytes (read + write)
  int64_t flops = 0;           // algorithmic flops
};

struct Plan {
  std::vector<BufferSpec> buffers;
  std::vector<Stage> stages;
  std::string source;          // all kernels of the plan, one module
  std::string description;     // human-readable summary (Compiled::dump)
  // the configuration the lowering chose (tile_m, tile_n, bk, vec, block,
  // ...), exported as the Compiled's int_properties
  std::map<std::string, int64_t> config;
  // set when the nest is a plain matmul: C[M][N] = sum_k A[M][K] * B[K][N]
  // over inputs a_input (A) and b_input (B), all three row-major with no
  // views, C the nest's only output (a pack can run such nests as one
  // grouped GEMM launch)
  struct Matmul {
    bool valid = false;
    int a_input = -1, b_input = -1;
    int64_t M = 0, N = 0, K = 0;
  } matmul;
};

// Knobs read from LoopTree annotations (see hip_backend.cpp).
struct Knobs {
  bool force_plain = false;    // "hip:plain" - only the pointwise lowering
  bool bf16 = false;           // "mfma_bf16" operand precision for contractions
  bool f16 = false;            // "mfma_f16"
  int block = 256;             // "block=N"
  bool block_set = false;      // block came from an annotation
  int vec = 0;                 // "vec=N" elements per thread (0 = auto)
  int tile_m = 0, tile_n = 0;  // "tile=MxN" for contractions (0 = auto)
  int bk = 0;                  // "bk=N": K tile of 16-bit contractions (32, 64)
  bool gather = false;         // "contract=gather": contractions without
                               // operand staging
  bool staged = false;         // "contract=staged": float contractions staged
                               // whatever their size
  bool igemm_plain = false;    // "igemm=plain": the single-stage tile instead of
                               // the software-pipelined one
  int waves = 0;               // "waves=2|4|8": waves per workgroup of the
                               // pipelined 16-bit tile (0 = auto)
  int lds_swz = -1;            // "lds=pad|swz|dl" (0, 1, 2): padded or XOR-permuted
                               // LDS rows of the register-staged pipelined tile,
                               // or the direct-to-LDS tile (-1 = auto: dl)
  int stage_blocked = -1;      // "stage=blocked|flat": k-blocked operand staging
                               // (-1 = auto: blocked when k_in spans several K tiles)
  int kb_outer = -1;           // "kb=outer|inner": order of the K blocks
  int streamk = -1;            // "streamk=0|1": float contractions share K tiles
                               // evenly over the workgroups (-1 = auto: when
                               // the tiles fill the workgroup slots unevenly)
  int overlap = -1;            // "overlap=0|1": run-ahead operand staging (off
                               // unless 1)
  int halo_direct = -1;        // "halo=direct|staged": the halo tile reads the
                               // activations' float source instead of a staged
                               // 16-bit copy (-1 = staged)
  int ksplit = -1;             // "ksplit=0|1": gathered contractions on 32x32
                               // tiles with K tiles split over the waves
                               // (-1 = auto: nests under 64 64x64 tiles)
  int stage_rows = 0;          // "srows=N": rows of a source plane per workgroup
                               // of the plane-staging pass (0 = auto: 256
                               // positions' worth)
  int igemm_halo = -1;         // "igemm=halo" (1) / "igemm=halo1" (2, one halo
                               // buffer) / "igemm=pipe" (0): staged 16-bit
                               // contractions load the Q rows of all taps of a
                               // K block once (-1 = auto: one buffer where it
                               // applies)
};

// Builds the whole launch plan for a LoopTree.
Plan build_plan(const Dataflow &df, const Knobs &knobs);

// Shared helpers for emitters.
std::string float_literal(double v);
std::string unary_expr(Operation op, const std::string &x);

}  // namespace hip
}  // namespace loop_tool
