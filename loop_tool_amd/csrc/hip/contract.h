// Contraction analysis shared by the MFMA lowerings of the hip backend
// (emit_contract.cpp: gathered operands; emit_igemm.cpp: staged operands).
#pragma once

#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "emit_internal.h"

namespace loop_tool {
namespace hip {

// One multiply operand: an IR input read through affine view indices.
struct Operand {
  int input_index = -1;
  std::vector<int64_t> sizes;                 // per dim
  std::vector<int64_t> strides;               // per dim (row-major)
  std::vector<std::map<int, int64_t>> coef;   // per dim: var sym id -> coefficient
  std::vector<int64_t> cst;                   // per dim constant
  std::vector<int64_t> lo, hi;                // per dim range of the index
  std::vector<int> pred_dims;                 // dims that may leave [0, size)
};

struct VarInfo {
  int sym;
  int64_t size;
  std::string name;
};

// D[b][i][j] = sum_k P[b][i][k] * Q[b][k][j] over the nest's vars:
// ops[po] = P (rows i), ops[qo] = Q (columns j, owns the output's contiguous var)
struct Contraction {
  Operand ops[2];
  int po = 0, qo = 1;
  std::vector<VarInfo> I, J, B, K;
  int64_t MD = 1, ND = 1, KD = 1, BD = 1;
  std::map<int, int64_t> ostride;  // output element stride per var sym id
  IR::NodeRef red_node = -1;       // the add reduction (its loops: the schedule)
};

// `lin` (a C expression) decomposed row-major over vars (the last fastest)
inline std::string decompose(const std::string &lin, const std::vector<VarInfo> &vars,
                             const std::string &indent) {
  std::stringstream s;
  if (vars.empty()) return "";
  s << indent << "{ long t_ = " << lin << ";\n";
  for (int d = (int)vars.size() - 1; d >= 0; --d) {
    if (d > 0)
      s << indent << "  " << vars[d].name << " = t_ % " << vars[d].size << "L; t_ /= "
        << vars[d].size << "L;\n";
    else
      s << indent << "  " << vars[d].name << " = t_;\n";
  }
  s << indent << "}\n";
  return s.str();
}

// sum of coef * var over `vars` as a C expression
inline std::string lin_expr(const std::map<int, int64_t> &coef,
                            const std::vector<VarInfo> &vars) {
  std::string out;
  for (const auto &v : vars) {
    auto it = coef.find(v.sym);
    if (it == coef.end() || it->second == 0) continue;
    out += (out.empty() ? "" : " + ") + std::string("(") + std::to_string(it->second) +
           "L * " + v.name + ")";
  }
  return out.empty() ? "0L" : out;
}

// Staged-operand lowering for 16-bit MFMA operands; false when the nest's
// operands cannot be staged (then the gathered lowering is used).
bool try_contract_staged(const Dataflow &df, const Contraction &c, const Knobs &knobs,
                         Plan &plan);

// Source of csrc/kernels/igemm_tile.h (embedded at build time)
const char *igemm_tile_source();

}  // namespace hip
}  // namespace loop_tool
