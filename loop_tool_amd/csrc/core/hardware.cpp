// Hardware registry (see include/loop_tool/hardware.h).
// Reference behaviour: src/core/hardware.cpp:16-61 (CPU first, ids in
// registration order, default hardware id 0).
#include "loop_tool/hardware.h"

#include <mutex>

#include "loop_tool/tensor.h"

namespace loop_tool {

namespace {
std::mutex &hw_mutex() {
  static std::mutex m;
  return m;
}
std::vector<std::shared_ptr<Hardware>> &mutable_hardware() {
  static std::vector<std::shared_ptr<Hardware>> hw = {std::make_shared<CPUHardware>()};
  return hw;
}
}  // namespace

const std::vector<std::shared_ptr<Hardware>> &getHardware() {
  return mutable_hardware();
}

int getAvailableHardware() {
  int avail = 0;
  for (const auto &hw : getHardware())
    if (hw->count()) avail |= 1 << hw->id();
  return avail;
}

void registerHardware(std::shared_ptr<Hardware> hw) {
  std::lock_guard<std::mutex> guard(hw_mutex());
  hw->setId((int)getHardware().size());
  mutable_hardware().emplace_back(hw);
}

// The CPU backends of loop_tool are single threaded.
int availableCPUs() { return 1; }

int &getDefaultHardwareId() {
  static int id = 0;
  return id;
}

const std::shared_ptr<Hardware> &getDefaultHardware() {
  for (const auto &hw : getHardware())
    if (hw->id() == getDefaultHardwareId()) return hw;
  return getHardware().at(0);
}

void setDefaultHardwareId(int id) { getDefaultHardwareId() = id; }

Tensor::Tensor(size_t N, int hardware) : hardware_id(hardware) {
  data = getHardware().at(hardware_id)->alloc(N * sizeof(float));
  numel = N;
}

Tensor::~Tensor() { getHardware().at(hardware_id)->free(data); }

}  // namespace loop_tool
