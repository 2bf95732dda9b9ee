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

// ---------------------------------------------------------------- streaming patterns
// lbmpy's 'lehmann' direction orderings (stencils.py:146-149, 214-218, 242-248): +-pairs at positions
// (1, 2), (3, 4), ...; EsoPull / EsoPush treat the second direction of each pair ("even" position) and
// the first differently.
struct Lehmann2D {
  static constexpr int n = 9;
  static constexpr int c[9][3] = {{0, 0, 0}, {1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0},
                                  {1, 1, 0}, {-1, -1, 0}, {1, -1, 0}, {-1, 1, 0}};
};
struct Lehmann3D {
  static constexpr int n = 27;   // D3Q19 uses the first 19 entries
  static constexpr int c[27][3] = {
      {0, 0, 0},  {1, 0, 0},   {-1, 0, 0},  {0, 1, 0},  {0, -1, 0},  {0, 0, 1},   {0, 0, -1},
      {1, 1, 0},  {-1, -1, 0}, {1, 0, 1},   {-1, 0, -1}, {0, 1, 1},  {0, -1, -1}, {1, -1, 0},
      {-1, 1, 0}, {1, 0, -1},  {-1, 0, 1},  {0, 1, -1}, {0, -1, 1},  {1, 1, 1},   {-1, -1, -1},
      {1, 1, -1}, {-1, -1, 1}, {1, -1, 1},  {-1, 1, -1}, {-1, 1, 1}, {1, -1, -1}};
};

// true when direction i of S (walberla order) sits at an even position of the lehmann ordering
template <class S>
__host__ __device__ constexpr bool lehmann_even(int i) {
  for (int k = 0; k < (S::D == 2 ? Lehmann2D::n : S::Q); ++k) {
    const int* l = S::D == 2 ? Lehmann2D::c[k] : Lehmann3D::c[k];
    if (l[0] == S::c[i][0] && l[1] == S::c[i][1] && l[2] == S::c[i][2]) return k % 2 == 0;
  }
  return false;
}

// Where a kernel of pattern PAT reads (r, rslot) and writes (w, wslot) population i of the cell at x:
// cell offsets r, w in {-1, 0, 1}^3 (the per-pattern table in include/lbm355.h).
struct Access {
  int r[3], rslot, w[3], wslot;
};

__host__ __device__ constexpr int pos_part(int v) { return v > 0 ? v : 0; }

template <class S>
__host__ __device__ constexpr Access pattern_access(int pat, int i) {
  const int c0 = S::c[i][0], c1 = S::c[i][1], c2 = S::c[i][2];
  const int ii = S::inv[i];
  const bool e = lehmann_even<S>(i);
  Access a = {{0, 0, 0}, i, {0, 0, 0}, i};
  auto minus = [&](int(&o)[3]) { o[0] = -c0; o[1] = -c1; o[2] = -c2; };
  auto plus = [&](int(&o)[3]) { o[0] = c0; o[1] = c1; o[2] = c2; };
  switch (pat) {
    case LBM355_PULL:                  // read (-c, i), write (0, i)
      minus(a.r);
      break;
    case LBM355_PUSH:                  // read (0, i), write (+c, i)
      plus(a.w);
      break;
    case LBM355_AA_EVEN:               // read (0, i), write (0, inv)
      a.wslot = ii;
      break;
    case LBM355_AA_ODD:                // read (-c, inv), write (+c, i)
      minus(a.r);
      a.rslot = ii;
      plus(a.w);
      break;
    case LBM355_ESOTWIST_EVEN:         // read (max(-c,0), i), write (max(c,0), inv)
    case LBM355_ESOTWIST_ODD:          // read (max(-c,0), inv), write (max(c,0), i)
      a.r[0] = pos_part(-c0); a.r[1] = pos_part(-c1); a.r[2] = pos_part(-c2);
      a.w[0] = pos_part(c0); a.w[1] = pos_part(c1); a.w[2] = pos_part(c2);
      if (pat == LBM355_ESOTWIST_EVEN) a.wslot = ii; else a.rslot = ii;
      break;
    case LBM355_ESOPULL_EVEN:          // E: read (-c, i), write (0, inv); O: read (0, i), write (+c, inv)
    case LBM355_ESOPULL_ODD:           // E: read (-c, inv), write (0, i); O: read (0, inv), write (+c, i)
      if (e) minus(a.r); else plus(a.w);
      if (pat == LBM355_ESOPULL_EVEN) a.wslot = ii; else a.rslot = ii;
      break;
    case LBM355_ESOPUSH_EVEN:          // E: read (0, inv), write (+c, i); O: read (-c, inv), write (0, i)
    case LBM355_ESOPUSH_ODD:           // E: read (0, i), write (+c, inv); O: read (-c, i), write (0, inv)
      if (e) plus(a.w); else minus(a.r);
      if (pat == LBM355_ESOPUSH_EVEN) a.rslot = ii; else a.wslot = ii;
      break;
  }
  return a;
}

// the whole table of one pattern, for compile-time lookups in unrolled loops
template <class S, int PAT>
struct AccessTable {
  Access a[S::Q];
  __host__ __device__ constexpr AccessTable() : a() {
    for (int i = 0; i < S::Q; ++i) a[i] = pattern_access<S>(PAT, i);
  }
};

// the kernel that follows PAT (in-place patterns alternate EVEN and ODD)
__host__ __device__ constexpr int next_pattern(int pat) {
  return (pat == LBM355_PULL || pat == LBM355_PUSH) ? pat : (pat % 2 == 1 ? pat + 1 : pat - 1);
}
__host__ __device__ constexpr bool is_eso(int pat) { return pat >= LBM355_ESOTWIST_EVEN && pat <= LBM355_ESOPUSH_ODD; }
// patterns swept through the access table (sweep_eso.hip): the Eso patterns and push
__host__ __device__ constexpr bool table_pattern(int pat) { return is_eso(pat) || pat == LBM355_PUSH; }

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

// ---------------------------------------------------------------- arithmetic helpers
// 1 / x.  FP32: the hardware reciprocal (about 1 ulp) refined by one Newton step, 3 instructions instead of the
// 10 of an IEEE division; FP64: division.
template <class T>
__device__ __forceinline__ T recip(T x) {
  return T(1) / x;
}
template <>
__device__ __forceinline__ float recip<float>(float x) {
  const float r = __builtin_amdgcn_rcpf(x);
  return __builtin_fmaf(r, __builtin_fmaf(-x, r, 1.0f), r);
}

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
  // optional per-cell fields with the output strides (null when absent)
  TC* omega_out;              // Smagorinsky shear rate
  TC* nut_out;                // Smagorinsky eddy viscosity, nut_scale * omega_sgs * Pi
  const TC* rho_in;           // moment family: density input
  const TC* u_in;             // moment family: equilibrium-velocity input
  TC nut_scale;               // 3/2 Cs^2
  uint32_t seed;              // stochastic FP16 rounding: this launch's seed
};

// What a collision hands back to the cell beyond the populations: the Smagorinsky shear rate and Pi
// (sqrt(2 Pi_neq:Pi_neq) / rho), written to the optional per-cell outputs.
template <class TC>
struct CellExtras {
  TC omega_sgs, pi;
};

// ---------------------------------------------------------------- method tables
// The method table is read-only for a whole launch.  Read through the constant address space, its
// wave-uniform entries become scalar loads wherever they are read, also after stores of the same
// launch (a persistent loop's later tiles); through a generic pointer they would be per-lane loads.
template <class T>
using TablePtr = const __attribute__((address_space(4))) T*;
template <class T>
__device__ __forceinline__ TablePtr<T> table_ptr(const T* p) {
  return (TablePtr<T>)p;
}

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
