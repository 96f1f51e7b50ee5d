// Tuned-schedule store of the hip backend (see tuned_store.h).
//
// A program is identified by its IR as the frontend hands it to the
// backend: the serialized IR text (loop_tool::serialize) with every var
// name replaced by its position, so that names numbered by the process that
// built the graph do not enter the key.  The key is the 64-bit FNV-1a hash of
// that text, in hex.  A stored schedule is the serialized IR of a LoopTree
// of the same program with other loop orders, splits, reuse and annotations;
// before it is used, its nodes, vars, var sizes, inputs and outputs are
// checked against the program's.
#include "tuned_store.h"

#include <cstdio>
#include <cstdlib>
#include <filesystem>
#include <fstream>
#include <sstream>

#include <unistd.h>

#include "loop_tool/compile.h"
#include "loop_tool/serialization.h"

namespace loop_tool {
namespace hip {

namespace {

std::mutex &dir_mutex() {
  static std::mutex m;
  return m;
}
std::string &dir_override() {
  static std::string d;
  return d;
}

uint64_t fnv1a(const std::string &s) {
  uint64_t h = 1469598103934665603ULL;
  for (unsigned char c : s) {
    h ^= c;
    h *= 1099511628211ULL;
  }
  return h;
}

bool is_name_char(char c) { return std::isalnum((unsigned char)c) || c == '_' || c == '\''; }

// `text` with every whole-word occurrence of the IR's var names replaced by
// "v<position>" (longest names first, so no name is a prefix of a later one)
std::string canonical_text(const IR &ir) {
  std::string text = serialize(ir);
  std::vector<std::pair<std::string, std::string>> names;
  auto vars = ir.vars();
  for (size_t i = 0; i < vars.size(); ++i)
    names.emplace_back(ir.var(vars[i]).name(), "v" + std::to_string(i));
  std::stable_sort(names.begin(), names.end(), [](const auto &a, const auto &b) {
    return a.first.size() > b.first.size();
  });
  for (const auto &n : names) {
    if (n.first.empty()) continue;
    std::string out;
    size_t pos = 0;
    while (true) {
      size_t hit = text.find(n.first, pos);
      if (hit == std::string::npos) break;
      const bool left = hit == 0 || !is_name_char(text[hit - 1]);
      const size_t end = hit + n.first.size();
      const bool right = end >= text.size() || !is_name_char(text[end]);
      out += text.substr(pos, hit - pos);
      out += left && right ? n.second : n.first;
      pos = end;
    }
    out += text.substr(pos);
    text = out;
  }
  return text;
}

// nodes, their ops, inputs and vars, the IR inputs and outputs, and every
// var's extent agree: `b` is a schedule of the program `a`
bool same_program(const LoopTree &a, const LoopTree &b) {
  const auto &x = a.ir, &y = b.ir;
  if (x.nodes() != y.nodes() || x.inputs() != y.inputs() || x.outputs() != y.outputs() ||
      x.vars() != y.vars())
    return false;
  for (auto nr : x.nodes()) {
    const auto &p = x.node(nr), &q = y.node(nr);
    if (p.op() != q.op() || p.inputs() != q.inputs() || p.vars() != q.vars()) return false;
  }
  Compiler ca(a), cb(b);
  for (auto v : x.vars()) {
    auto ia = ca.var_sizes.find(v), ib = cb.var_sizes.find(v);
    if ((ia == ca.var_sizes.end()) != (ib == cb.var_sizes.end())) return false;
    if (ia != ca.var_sizes.end() && ia->second != ib->second) return false;
  }
  return true;
}

std::string entry_path(const std::string &key) {
  auto d = tuned_dir();
  if (d.empty()) return "";
  return d + "/gfx950/" + key + ".ir";
}

}  // namespace

void set_tuned_dir(const std::string &dir) {
  std::lock_guard<std::mutex> g(dir_mutex());
  dir_override() = dir;
}

std::string tuned_dir() {
  {
    std::lock_guard<std::mutex> g(dir_mutex());
    if (!dir_override().empty()) return dir_override();
  }
  if (const char *d = std::getenv("LT_HIP_TUNED_DIR")) return d;
  if (const char *h = std::getenv("HOME")) return std::string(h) + "/.cache/loop_tool_amd/tuned";
  return "";
}

std::string schedule_key(const LoopTree &lt) {
  std::stringstream s;
  s << std::hex << fnv1a(canonical_text(lt.ir));
  return s.str();
}

std::unique_ptr<LoopTree> tuned_get(const LoopTree &program, std::string *key_out) {
  if (const char *off = std::getenv("LT_HIP_TUNED"))
    if (std::string(off) == "0") return nullptr;
  std::string key;
  try {
    key = schedule_key(program);
  } catch (const std::exception &) {
    return nullptr;  // an IR this library cannot serialize: compile as given
  }
  if (key_out) *key_out = key;
  const std::string path = entry_path(key);
  if (path.empty()) return nullptr;
  std::ifstream f(path);
  if (!f) return nullptr;
  std::stringstream text;
  text << f.rdbuf();
  std::unique_ptr<LoopTree> t;
  try {
    t = std::make_unique<LoopTree>(deserialize(text.str()));
  } catch (const std::exception &) {
    return nullptr;  // not a schedule this build reads: compile as given
  }
  if (!same_program(program, *t)) return nullptr;
  return t;
}

std::string tuned_put(const LoopTree &program, const LoopTree &schedule) {
  ASSERT(same_program(program, schedule))
      << "tuned schedule is not a schedule of the program (nodes, vars or sizes differ)";
  const std::string key = schedule_key(program);
  const std::string path = entry_path(key);
  ASSERT(!path.empty()) << "no tuned-schedule directory (set LT_HIP_TUNED_DIR or HOME)";
  std::error_code ec;
  std::filesystem::create_directories(std::filesystem::path(path).parent_path(), ec);
  ASSERT(!ec) << "cannot create " << std::filesystem::path(path).parent_path() << ": "
              << ec.message();
  static std::atomic<uint64_t> serial{0};
  const std::string tmp =
      path + ".tmp" + std::to_string((long)getpid()) + "_" + std::to_string(serial++);
  bool ok;
  {
    std::ofstream f(tmp);
    f << serialize(schedule.ir);
    f.close();
    ok = f.good();
  }
  if (!ok || std::rename(tmp.c_str(), path.c_str()) != 0) {
    std::remove(tmp.c_str());
    ASSERT(0) << "cannot write the tuned schedule " << path;
  }
  return key;
}

}  // namespace hip
}  // namespace loop_tool
