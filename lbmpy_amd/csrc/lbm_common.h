// Common device-side definitions of the MI355X (gfx950) stream-collide sweeps:
// stencil traits, storage conversions and the streaming/boundary addressing.
//
// Memory layout ("fzyx", as lbmpy's default field layout,
// lbmpy creationfunctions.py:522): PDF q of cell (x, y, z) lives at
//     base + q*stride_q + (z+g)*stride_z + (y+g)*stride_y + (x+g)
// with g ghost layers; x is contiguous so a wavefront of 64 lanes along x
// reads/writes 64 consecutive elements per population (coalesced).
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <stdint.h>

#include "../../include/lbm355.h"

// records the calling thread's last error (returned by lbm355_last_error) and returns `code`
int lbm355_set_error(int code, const char* msg);

namespace lbm355 {

// ---------------------------------------------------------------- stencils
// Direction order = lbmpy's 'walberla' ordering (stencils.py:133-135,194-198,221-227).
struct D2Q9 {
  static constexpr int D = 2, Q = 9;
  static constexpr int c[9][3] = {{0, 0, 0}, {0, 1, 0}, {0, -1, 0}, {-1, 0, 0}, {1, 0, 0},
                                  {-1, 1, 0}, {1, 1, 0}, {-1, -1, 0}, {1, -1, 0}};
  static constexpr int inv[9] = {0, 2, 1, 4, 3, 8, 7, 6, 5};
};
struct D3Q19 {
  static constexpr int D = 3, Q = 19;
  static constexpr int c[19][3] = {{0, 0, 0}, {0, 1, 0},  {0, -1, 0},  {-1, 0, 0}, {1, 0, 0},
                                   {0, 0, 1}, {0, 0, -1}, {-1, 1, 0},  {1, 1, 0},  {-1, -1, 0},
                                   {1, -1, 0}, {0, 1, 1}, {0, -1, 1},  {-1, 0, 1}, {1, 0, 1},
                                   {0, 1, -1}, {0, -1, -1}, {-1, 0, -1}, {1, 0, -1}};
  static constexpr int inv[19] = {0, 2, 1, 4, 3, 6, 5, 10, 9, 8, 7, 16, 15, 18, 17, 12, 11, 14, 13};
};
struct D3Q27 {
  static constexpr int D = 3, Q = 27;
  static constexpr int c[27][3] = {
      {0, 0, 0},   {0, 1, 0},  {0, -1, 0},  {-1, 0, 0},  {1, 0, 0},   {0, 0, 1},  {0, 0, -1},
      {-1, 1, 0},  {1, 1, 0},  {-1, -1, 0}, {1, -1, 0},  {0, 1, 1},   {0, -1, 1}, {-1, 0, 1},
      {1, 0, 1},   {0, 1, -1}, {0, -1, -1}, {-1, 0, -1}, {1, 0, -1},  {1, 1, 1},  {-1, 1, 1},
      {1, -1, 1},  {-1, -1, 1}, {1, 1, -1}, {-1, 1, -1}, {1, -1, -1}, {-1, -1, -1}};
  static constexpr int inv[27] = {0,  2,  1,  4,  3,  6,  5,  10, 9,  8,  7,  16, 15, 18,
                                  17, 12, 11, 14, 13, 26, 25, 24, 23, 22, 21, 20, 19};
};

// index of direction (cx, cy, cz) in stencil S, -1 if absent (compile-time)
template <class S>
__host__ __device__ constexpr int dir_index(int cx, int cy, int cz) {
  for (int i = 0; i < S::Q; ++i)
    if (S::c[i][0] == cx && S::c[i][1] == cy && S::c[i][2] == cz) return i;
  return -1;
}

// lattice weights of the rest state (maxwellian_equilibrium.py:24-61)
template <class S>
__host__ __device__ constexpr double weight(int i) {
  int n = (S::c[i][0] != 0) + (S::c[i][1] != 0) + (S::c[i][2] != 0);
  if (S::Q == 9) return n == 0 ? 4.0 / 9.0 : (n == 1 ? 1.0 / 9.0 : 1.0 / 36.0);
  if (S::Q == 19) return n == 0 ? 1.0 / 3.0 : (n == 1 ? 1.0 / 18.0 : 1.0 / 36.0);
  return n == 0 ? 8.0 / 27.0 : (n == 1 ? 2.0 / 27.0 : (n == 2 ? 1.0 / 54.0 : 1.0 / 216.0));
}

// ---------------------------------------------------------------- storage types
template <class TS, class TC> __device__ __forceinline__ TC load_as(const TS* p) { return (TC)(*p); }
template <> __device__ __forceinline__ float load_as<_Float16, float>(const _Float16* p) { return (float)(*p); }
template <class TS, class TC> __device__ __forceinline__ void store_as(TS* p, TC v) { *p = (TS)v; }

// ---------------------------------------------------------------- launch arguments
template <class TC>
struct SweepArgs {
  const void* src;            // PDF source (pull) / the single PDF array (AA)
  void* dst;                  // PDF destination (pull), == src for AA
  TC* rho_out;                // optional density output (interior-indexed, may be null)
  TC* u_out;                  // optional velocity output
  const TC* tab;              // method tables (collision-family specific)
  int nx, ny, nz, g;
  int64_t sy, sz, sq;         // element strides of the PDF arrays
  int64_t osy, osz, osc;      // element strides of the outputs (x stride 1)
  int face[3];                // LBM355_FACE_* per axis
  int flags;                  // LBM355_FLAG_*
  int nactive;                // moment family: active-moment count
  TC force[3];
  TC smag_cs2x18;             // 18 * Cs^2 (Smagorinsky)
  TC omega_shear;             // shear rate replaced by the subgrid model
  const void* closure;        // learned-closure parameter block (LBM355_FLAG_LEARNED_CLOSURE)
};

// ---------------------------------------------------------------- addressing
// Per axis: the coordinate of the neighbour at offset -1/0/+1 (periodic wrap folded in)
// and whether that neighbour is a NoSlip wall cell.
struct AxisNb {
  int lo, hi;        // coordinates (ghost-layer index space: interior 0..n-1 maps to g..n-1+g)
  bool wall_lo, wall_hi;
};

__device__ __forceinline__ AxisNb axis_neighbours(int x, int n, int g, int face) {
  AxisNb a;
  a.lo = x - 1 + g;
  a.hi = x + 1 + g;
  a.wall_lo = a.wall_hi = false;
  if (face == LBM355_FACE_PERIODIC) {
    if (x == 0) a.lo = n - 1 + g;
    if (x == n - 1) a.hi = g;
  } else if (face == LBM355_FACE_NOSLIP) {
    a.wall_lo = (x == 0);
    a.wall_hi = (x == n - 1);
  }
  return a;
}

}  // namespace lbm355
