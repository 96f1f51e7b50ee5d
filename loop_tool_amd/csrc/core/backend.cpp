// Backend registry and Compiled helpers (see include/loop_tool/backend.h).
// Reference behaviour: src/core/backend.cpp:19-65.  The default backend is
// "hip" when it is registered (this library always registers it) and the
// first registered backend otherwise.
#include "loop_tool/backend.h"

#include <list>
#include <mutex>

#include "loop_tool/dynlib.h"

namespace loop_tool {

namespace {
std::mutex &registry_mutex() {
  static std::mutex m;
  return m;
}
std::unordered_map<std::string, std::shared_ptr<Backend>> &mutable_backends() {
  static std::unordered_map<std::string, std::shared_ptr<Backend>> backends;
  return backends;
}
// Plugin libraries stay mapped until the process ends: the registry holds
// backends whose code lives in them, and static destruction order between
// this list and the registry is not ours to choose.
std::list<DynamicLibrary> &loaded_libs() {
  static auto *libs = new std::list<DynamicLibrary>();
  return *libs;
}
}  // namespace

void Compiled::operator()(const std::vector<Tensor *> &tensors, bool sync) const {
  std::vector<void *> memory;
  for (auto *t : tensors) memory.emplace_back(t->data.address);
  run(memory, sync);
}

std::vector<void *> Compiled::allocate(std::vector<int64_t> &sizes) const {
  std::vector<void *> memory(sizes.size(), nullptr);
  for (size_t i = 0; i < sizes.size(); ++i)
    if (sizes[i] > 0) memory[i] = calloc(sizes[i], sizeof(float));
  return memory;
}

const std::unordered_map<std::string, std::shared_ptr<Backend>> &getBackends() {
  return mutable_backends();
}

void registerBackend(std::shared_ptr<Backend> backend) {
  std::lock_guard<std::mutex> guard(registry_mutex());
  mutable_backends()[backend->name()] = backend;
}

std::shared_ptr<Backend> &getDefaultBackend() {
  static std::shared_ptr<Backend> default_backend = [] {
    const auto &b = getBackends();
    ASSERT(!b.empty()) << "no backend registered";
    if (b.count("hip")) return b.at("hip");
    return b.begin()->second;
  }();
  return default_backend;
}

void setDefaultBackend(std::string backend) {
  ASSERT(getBackends().count(backend)) << "couldn't find backend " << backend;
  getDefaultBackend() = getBackends().at(backend);
}

void loadLibrary(std::string lib_name) {
  // dlopen runs the library's static RegisterBackend objects, which take the
  // registry lock themselves: hold only the list's own lock here
  static std::mutex libs_mutex;
  std::lock_guard<std::mutex> guard(libs_mutex);
  loaded_libs().emplace_back(lib_name.c_str(), true);
}

}  // namespace loop_tool
