// Occupancy caps through LDS: a launch that requests more dynamic LDS than its kernel uses leaves room for fewer
// workgroups per CU.  Host code shared by the sweep launches (sweep.hip) and the copy calibration
// (macroscopic.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstddef>

namespace lbm355 {

// Dynamic LDS (bytes) for a launch of `block_threads`-thread workgroups that leaves at most `waves` waves on each
// of a CU's 4 SIMDs (rounded up to whole workgroups: never fewer waves than that), given that the kernel itself
// uses kernel_lds bytes.  Returns kernel_lds when waves <= 0, when the kernel's own LDS already allows no more
// workgroups, or when the device limits cannot be read.
inline size_t lds_for_wave_cap(int waves, unsigned block_threads, size_t kernel_lds) {
  if (waves <= 0 || block_threads == 0) return kernel_lds;
  constexpr int MAX_DEV = 64;
  static std::atomic<int> lds_cu[MAX_DEV], lds_block[MAX_DEV];  // per device, 0 until queried
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= MAX_DEV) return kernel_lds;
  int cu = lds_cu[dev].load(std::memory_order_acquire);
  int blk = lds_block[dev].load(std::memory_order_relaxed);
  if (!cu) {
    if (hipDeviceGetAttribute(&cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev) != hipSuccess ||
        hipDeviceGetAttribute(&blk, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess || cu <= 0)
      return kernel_lds;
    lds_block[dev].store(blk, std::memory_order_relaxed);
    lds_cu[dev].store(cu, std::memory_order_release);
  }
  const int waves_per_block = (int)((block_threads + 63) / 64);
  const int blocks = std::max(1, (4 * waves + waves_per_block - 1) / waves_per_block);
  // midway between the sizes at which `blocks` and `blocks + 1` workgroups fit
  const size_t bytes = (((size_t)cu / blocks + (size_t)cu / (blocks + 1)) / 2) & ~(size_t)255;
  if (bytes <= kernel_lds || bytes > (size_t)blk) return kernel_lds;
  return bytes;
}

}  // namespace lbm355
