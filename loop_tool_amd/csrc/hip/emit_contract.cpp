// Contraction lowering (placeholder until implemented).
#include "emit_internal.h"

namespace loop_tool {
namespace hip {

bool try_contract(const Dataflow &, const Knobs &, Plan &) { return false; }

}  // namespace hip
}  // namespace loop_tool
