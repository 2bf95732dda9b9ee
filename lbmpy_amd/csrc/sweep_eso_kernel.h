// The sweep kernels of lbmpy's in-place Eso patterns and of push (per-element and row-staged) and their
// pickers, shared by the translation units that instantiate them (sweep_<family>_<compute type>.hip); the
// launch code is in sweep_eso.hip.
#pragma once
#include "sweep_cell.h"

namespace lbm355 {

template <class S, int FAM, int PAT, class TS, class TC>
__global__ void __launch_bounds__(256) sweep_eso_kernel(SweepArgs<TC> a) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= a.nx) return;
  sweep_cell<S, FAM, PAT, false, TS, TC>(a, x, blockIdx.y, blockIdx.z);
}

// x-shifted accesses of the table (compile-time counts, for the LDS rows of the row-staged kernel)
template <class S, int PAT>
__host__ __device__ constexpr int eso_shifted_reads() {
  int n = 0;
  for (int i = 0; i < S::Q; ++i) n += pattern_access<S>(PAT, i).r[0] != 0;
  return n;
}
template <class S, int PAT>
__host__ __device__ constexpr int eso_shifted_writes() {
  int n = 0;
  for (int i = 0; i < S::Q; ++i) n += pattern_access<S>(PAT, i).w[0] != 0;
  return n;
}

// Row-staged variant for a periodic x axis: the block spans one whole row (blockDim.x == nx), so the +-1 x
// shifts of the table go through LDS rows and every global access is an aligned, non-temporal row access.
// Every population a cell reads and writes lies in rows that only this block's cells touch (the in-place
// tables are per-cell race-free), so the two block barriers are all the synchronisation needed.
template <class S, int FAM, int PAT, class TS, class TC>
__global__ void __launch_bounds__(512) sweep_eso_rows_kernel(SweepArgs<TC> a) {
  constexpr int Q = S::Q, D = S::D;
  constexpr AccessTable<S, PAT> acc{};
  const int x = threadIdx.x, y = blockIdx.y, z = blockIdx.z;
  const int nx = a.nx;
  const AxisNb ay = axis_neighbours(y, a.ny, a.g, a.face[1]);
  AxisNb az = {0, 0, false, false};
  if (D == 3) az = axis_neighbours(z, a.nz, a.g, a.face[2]);
  const int sy = (int)a.sy, sz = (int)a.sz;
  const int X1 = x + a.g;
  const int Y[3] = {ay.lo * sy, (y + a.g) * sy, ay.hi * sy};
  const int Z[3] = {D == 3 ? az.lo * sz : 0, D == 3 ? (z + a.g) * sz : 0, D == 3 ? az.hi * sz : 0};
  const int XL[3] = {x == 0 ? nx - 1 : x - 1, x, x == nx - 1 ? 0 : x + 1};   // LDS row positions, periodic
  const TS* __restrict__ src = reinterpret_cast<const TS*>(a.src);
  TS* __restrict__ dst = reinterpret_cast<TS*>(a.dst);
  const int64_t sq = a.sq;
  extern __shared__ unsigned char lds_raw[];
  TS* tile = reinterpret_cast<TS*>(lds_raw);

  TC f[Q];
  int j = 0;
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    const Access& t = acc.a[i];
    const TS* p = src + t.rslot * sq + X1 + Y[1 + t.r[1]] + Z[1 + t.r[2]];
    if (t.r[0] != 0) {
      tile[j * nx + x] = __builtin_nontemporal_load(p);
      ++j;
    } else {
      f[i] = load_nt<TS, TC>(p);
    }
  }
  __syncthreads();
  j = 0;
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    const Access& t = acc.a[i];
    if (t.r[0] != 0) {
      f[i] = (TC)tile[j * nx + XL[1 + t.r[0]]];
      ++j;
    }
  }
  __syncthreads();

  collide_cell<S, FAM, TC, TS>(f, a, x, y, z);

  j = 0;
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    const Access& t = acc.a[i];
    if (t.w[0] != 0) {
      tile[j * nx + XL[1 + t.w[0]]] = (TS)f[i];
      ++j;
    } else {
      store_nt<TS, TC>(dst + t.wslot * sq + X1 + Y[1 + t.w[1]] + Z[1 + t.w[2]], f[i]);
    }
  }
  __syncthreads();
  j = 0;
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    const Access& t = acc.a[i];
    if (t.w[0] != 0) {
      __builtin_nontemporal_store(tile[j * nx + x], dst + t.wslot * sq + X1 + Y[1 + t.w[1]] + Z[1 + t.w[2]]);
      ++j;
    }
  }
}

template <class S, int FAM, class TS, class TC>
void* eso_pick_rows(int pat) {
  switch (pat) {
    case LBM355_ESOTWIST_EVEN: return (void*)sweep_eso_rows_kernel<S, FAM, LBM355_ESOTWIST_EVEN, TS, TC>;
    case LBM355_ESOTWIST_ODD: return (void*)sweep_eso_rows_kernel<S, FAM, LBM355_ESOTWIST_ODD, TS, TC>;
    case LBM355_ESOPULL_EVEN: return (void*)sweep_eso_rows_kernel<S, FAM, LBM355_ESOPULL_EVEN, TS, TC>;
    case LBM355_ESOPULL_ODD: return (void*)sweep_eso_rows_kernel<S, FAM, LBM355_ESOPULL_ODD, TS, TC>;
    case LBM355_ESOPUSH_EVEN: return (void*)sweep_eso_rows_kernel<S, FAM, LBM355_ESOPUSH_EVEN, TS, TC>;
    case LBM355_ESOPUSH_ODD: return (void*)sweep_eso_rows_kernel<S, FAM, LBM355_ESOPUSH_ODD, TS, TC>;
    case LBM355_PUSH: return (void*)sweep_eso_rows_kernel<S, FAM, LBM355_PUSH, TS, TC>;
  }
  return nullptr;
}

// the largest number of LDS rows any pattern of stencil S needs
template <class S>
constexpr int eso_max_rows() {
  int m = 0;
  for (int p = LBM355_ESOTWIST_EVEN; p <= LBM355_PUSH; ++p) {
    int r = 0, w = 0;
    for (int i = 0; i < S::Q; ++i) {
      r += pattern_access<S>(p, i).r[0] != 0;
      w += pattern_access<S>(p, i).w[0] != 0;
    }
    m = r > m ? r : m;
    m = w > m ? w : m;
  }
  return m;
}

template <class S, int FAM, class TS, class TC>
void* eso_pick_pattern(int pat) {
  switch (pat) {
    case LBM355_ESOTWIST_EVEN: return (void*)sweep_eso_kernel<S, FAM, LBM355_ESOTWIST_EVEN, TS, TC>;
    case LBM355_ESOTWIST_ODD: return (void*)sweep_eso_kernel<S, FAM, LBM355_ESOTWIST_ODD, TS, TC>;
    case LBM355_ESOPULL_EVEN: return (void*)sweep_eso_kernel<S, FAM, LBM355_ESOPULL_EVEN, TS, TC>;
    case LBM355_ESOPULL_ODD: return (void*)sweep_eso_kernel<S, FAM, LBM355_ESOPULL_ODD, TS, TC>;
    case LBM355_ESOPUSH_EVEN: return (void*)sweep_eso_kernel<S, FAM, LBM355_ESOPUSH_EVEN, TS, TC>;
    case LBM355_ESOPUSH_ODD: return (void*)sweep_eso_kernel<S, FAM, LBM355_ESOPUSH_ODD, TS, TC>;
    case LBM355_PUSH: return (void*)sweep_eso_kernel<S, FAM, LBM355_PUSH, TS, TC>;
  }
  return nullptr;
}

template <class S, int FAM, class TC>
void* eso_pick_store(int store, int pat, bool rows) {
  if (store == LBM355_F64)
    return rows ? eso_pick_rows<S, FAM, double, TC>(pat) : eso_pick_pattern<S, FAM, double, TC>(pat);
  if (store == LBM355_F32)
    return rows ? eso_pick_rows<S, FAM, float, TC>(pat) : eso_pick_pattern<S, FAM, float, TC>(pat);
  if constexpr (sizeof(TC) == 4) {
    if (store == LBM355_F16)
      return rows ? eso_pick_rows<S, FAM, _Float16, TC>(pat) : eso_pick_pattern<S, FAM, _Float16, TC>(pat);
  }
  return nullptr;
}

// Eso / push kernel pickers of each (family, compute type), one per translation unit; NULL when there is none.
#define LBM355_ESO_PICKER(NAME) void* NAME(int stencil, int store, int pat, bool rows)
LBM355_ESO_PICKER(eso_pick_moment_f64);
LBM355_ESO_PICKER(eso_pick_moment_f32);
LBM355_ESO_PICKER(eso_pick_central_f64);
LBM355_ESO_PICKER(eso_pick_central_f32);
LBM355_ESO_PICKER(eso_pick_cumulant_f64);
LBM355_ESO_PICKER(eso_pick_cumulant_f32);
LBM355_ESO_PICKER(eso_pick_cumulant_corrected_f64);
LBM355_ESO_PICKER(eso_pick_cumulant_corrected_f32);
LBM355_ESO_PICKER(eso_pick_cumulant_general_f64);
LBM355_ESO_PICKER(eso_pick_cumulant_general_f32);
LBM355_ESO_PICKER(eso_pick_cumulant_eq3_f64);
LBM355_ESO_PICKER(eso_pick_cumulant_eq3_f32);

}  // namespace lbm355
