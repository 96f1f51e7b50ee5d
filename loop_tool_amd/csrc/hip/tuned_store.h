// Tuned-schedule store of the hip backend: the schedule the autotuner found
// for a program, kept on disk as IR text and looked up by Backend::compile
// (HipBackend::compile_impl) for every program it is handed.
//
// One file per program and target: <dir>/gfx950/<key>.ir, where <dir> is
// set_tuned_dir(), else $LT_HIP_TUNED_DIR, else
// $HOME/.cache/loop_tool_amd/tuned.  LT_HIP_TUNED=0 turns lookups off.
#pragma once

#include <atomic>
#include <memory>
#include <mutex>
#include <string>

#include "loop_tool/ir.h"

namespace loop_tool {
namespace hip {

void set_tuned_dir(const std::string &dir);
std::string tuned_dir();
// key of a program: hash of its serialized IR with var names replaced by
// their positions (hex)
std::string schedule_key(const LoopTree &lt);
// the stored schedule of `program`, or null when none is stored (or the
// stored one is not a schedule of this program); *key_out gets the key
std::unique_ptr<LoopTree> tuned_get(const LoopTree &program, std::string *key_out);
// store `schedule` as the tuned schedule of `program`; returns the key
std::string tuned_put(const LoopTree &program, const LoopTree &schedule);

}  // namespace hip
}  // namespace loop_tool
