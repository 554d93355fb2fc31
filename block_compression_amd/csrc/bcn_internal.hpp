// Host-side task description shared by the C ABI and the kernel launchers.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/bcn_mi355x.h"

namespace bcn {

struct Task {
  bcn_variant variant;
  uint32_t settings[19];          // bcn_bc7_settings or bcn_bc6h_settings words
  const uint8_t* texture;         // texel (0, texture_y_offset)
  bcn_source_format format;
  uint32_t pitch;                 // bytes per texture row
  uint32_t width, height;         // region size (multiples of 4)
  uint8_t* blocks;                // first block (buffer + blocks_offset rounded to words)
};

// Blocks per row / column of a task.
inline uint32_t blocks_x(const Task& t) { return (t.width + 3) / 4; }
inline uint32_t blocks_y(const Task& t) { return (t.height + 3) / 4; }

hipError_t launch_bc7(const Task& t, hipStream_t stream);
hipError_t launch_bc15(const Task& t, hipStream_t stream);
hipError_t launch_bc6h(const Task& t, hipStream_t stream);

// Launches one task with the per-format kernel.
hipError_t launch_task(const Task& t, hipStream_t stream);

// Restores the caller's current device on scope exit.
struct DeviceGuard {
  int prev = -1;
  bool ok = false;
  explicit DeviceGuard(int dev) {
    if (hipGetDevice(&prev) != hipSuccess) return;
    ok = prev == dev || hipSetDevice(dev) == hipSuccess;
  }
  ~DeviceGuard() {
    int cur = -1;
    if (prev >= 0 && hipGetDevice(&cur) == hipSuccess && cur != prev) (void)hipSetDevice(prev);
  }
};

}  // namespace bcn

// The task queue behind the C ABI's bcn_compressor handle.
struct bcn_compressor {
  int device;
  std::vector<bcn::Task> tasks;
  std::string last_error;
};
