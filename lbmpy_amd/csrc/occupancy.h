// Occupancy caps through LDS: a launch that requests more dynamic LDS than its kernel uses leaves room for fewer
// workgroups per CU.  Host code shared by the sweep launches (sweep.hip, sweep_eso.hip) and the copy calibration
// (macroscopic.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstddef>
#include <cstdint>

#include "../../include/lbm355.h"

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

// Occupancy caps for the sweeps.  Most of them run faster with fewer waves resident per SIMD than their register
// use allows: each wave streams 2Q population planes, and past a few waves per SIMD more of them add no bandwidth.
// Waves per SIMD (0: no cap), found by sweeping the LDS request on MI355X at 256^3 (periodic, HIP events per
// launch), by pattern: pull, AA even, AA odd, EsoTwist, EsoPull, EsoPush (both parities), push.  Timings (ms,
// uncapped -> capped) of the D3Q27 cumulant FP32 kernels: pull 0.629 -> 0.603, AA even 0.619 -> 0.589, AA odd
// 0.635 -> 0.607, EsoTwist 0.634 -> 0.604, EsoPull 0.622 -> 0.599, EsoPush 0.617 -> 0.594, push 0.644 -> 0.605;
// of D3Q19 SRT / TRT / MRT: pull 0.425 -> 0.415, AA even 0.431 -> 0.415, AA odd 0.433 -> 0.420, the Eso patterns
// 0.423-0.434 -> 0.409-0.419, push 0.437 -> 0.421.  Caps that gained less than about 1.5% are left out.
// One wave per SIMD is slower everywhere.
enum CapKind { CAP_PULL, CAP_AA_EVEN, CAP_AA_ODD, CAP_ESOTWIST, CAP_ESOPULL, CAP_ESOPUSH, CAP_PUSH, CAP_KINDS };

struct WaveCaps {
  int family, stencil, store;
  int8_t waves[CAP_KINDS];
};

constexpr WaveCaps kWaveCaps[] = {
    {LBM355_FAMILY_CUMULANT, LBM355_D3Q27, LBM355_F32, {3, 3, 2, 2, 2, 2, 2}},
    {LBM355_FAMILY_CUMULANT, LBM355_D3Q27, LBM355_F64, {0, 2, 2, 0, 0, 0, 0}},  // AA: 1.337 / 1.344 -> 1.272 / 1.240
    {LBM355_FAMILY_CUMULANT, LBM355_D3Q27, LBM355_F16, {0, 4, 4, 0, 0, 0, 0}},  // AA: 0.334 / 0.329 -> 0.316 / 0.316
    {LBM355_FAMILY_MOMENT, LBM355_D3Q19, LBM355_F32, {4, 3, 3, 4, 4, 4, 3}},
    {LBM355_FAMILY_MOMENT, LBM355_D3Q27, LBM355_F32, {4, 3, 0, 0, 0, 0, 0}},  // pull 0.622->0.605, even 0.618->0.596
    {LBM355_FAMILY_CENTRAL, LBM355_D3Q19, LBM355_F32, {0, 0, 3, 3, 3, 0, 0}},  // AA odd 0.435 -> 0.415, Eso 1.7%
    {LBM355_FAMILY_CENTRAL, LBM355_D3Q27, LBM355_F32, {0, 0, 0, 2, 2, 0, 0}},  // EsoTwist 0.628 -> 0.603, EsoPull 2.6%
};

// The cap for one sweep launch.  `staged`: the kernel shifts x through LDS rows (the pull / AA-odd row kernel, the
// Eso / push row kernel).  The AA-even caps also pay with NoSlip faces, at 128^3 and 384^3, and at 512x256x256.
// The others were measured on periodic rows of at most 256 cells with the row kernels (the FP16 AA-odd kernel is
// never staged) and do not carry over: with fused NoSlip faces (kernels of other register use) the AA-odd caps
// cost 2-10%, and with 512-cell rows, where a 512-thread row block is alone on its CU at the cap, 15%.  So they
// apply only there.  The learned closure gains nothing from a cap.
// A launch's own waves_per_simd (1..8, or -1 for none) overrides the table: the best cap of the AA-odd kernels
// differs from one MI355X to another (D3Q27 cumulant FP32 at 256^3: 0.607 ms at 2 waves and 0.664 at 3 on one,
// 0.658 and 0.609 on another), and lbmpy_amd.sweep.tune_occupancy picks it by timing.
inline int sweep_wave_cap(const lbm355_sweep_desc* d, bool staged, bool walls, unsigned block_threads) {
  if (d->waves_per_simd != 0) return d->waves_per_simd > 0 ? d->waves_per_simd : 0;
  if (d->flags & LBM355_FLAG_LEARNED_CLOSURE) return 0;
  int kind;
  switch (d->pattern) {
    case LBM355_PULL: kind = CAP_PULL; break;
    case LBM355_AA_EVEN: kind = CAP_AA_EVEN; break;
    case LBM355_AA_ODD: kind = CAP_AA_ODD; break;
    case LBM355_ESOTWIST_EVEN: case LBM355_ESOTWIST_ODD: kind = CAP_ESOTWIST; break;
    case LBM355_ESOPULL_EVEN: case LBM355_ESOPULL_ODD: kind = CAP_ESOPULL; break;
    case LBM355_ESOPUSH_EVEN: case LBM355_ESOPUSH_ODD: kind = CAP_ESOPUSH; break;
    case LBM355_PUSH: kind = CAP_PUSH; break;
    default: return 0;
  }
  const bool f16 = d->store_dtype == LBM355_F16;
  if (kind != CAP_AA_EVEN) {
    if (walls || block_threads > 256) return 0;
    if (!staged && !(f16 && kind == CAP_AA_ODD)) return 0;
  }
  // the compute type each table row was measured with
  if (d->compute_dtype != (d->store_dtype == LBM355_F64 ? LBM355_F64 : LBM355_F32)) return 0;
  for (const WaveCaps& c : kWaveCaps)
    if (c.family == d->family && c.stencil == d->stencil && c.store == d->store_dtype) return c.waves[kind];
  return 0;
}

}  // namespace lbm355
