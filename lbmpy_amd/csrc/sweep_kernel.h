// The pull / AA sweep kernel template and its per-(pattern, storage) pickers, shared by the translation
// units that instantiate it (sweep_<family>_<compute type>.hip, one per collision family and compute type, so
// that the library builds in parallel).  The C ABI and the launch code are in sweep.hip.
#pragma once
#include <type_traits>

#include "sweep_cell.h"

namespace lbm355 {

template <class S, int FAM, int PAT, bool WALLS, bool ROWLDS, class TS, class TC>
__global__ void __launch_bounds__(ROWLDS ? 512 : 256) sweep_kernel(SweepArgs<TC> a) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  if constexpr (!ROWLDS) {
    if (x >= a.nx) return;
    sweep_cell<S, FAM, PAT, WALLS, TS, TC>(a, x, blockIdx.y, blockIdx.z);
  } else {
    // The block spans one whole periodic row (blockDim.x == nx).  The +-1 x-shifts of pull / AA-odd
    // streaming go through LDS, so every global access is an aligned, contiguous, non-temporal one.
    constexpr int Q = S::Q, D = S::D;
    const int y = blockIdx.y;
    const int z = blockIdx.z;
    const AxisNb ax = axis_neighbours(x, a.nx, a.g, a.face[0]);
    const AxisNb ay = axis_neighbours(y, a.ny, a.g, a.face[1]);
    AxisNb az = {0, 0, false, false};
    if (D == 3) az = axis_neighbours(z, a.nz, a.g, a.face[2]);
    const int sy = (int)a.sy, sz = (int)a.sz;
    const int X1 = x + a.g;
    const int Y[3] = {ay.lo * sy, (y + a.g) * sy, ay.hi * sy};
    const int Z[3] = {D == 3 ? az.lo * sz : 0, D == 3 ? (z + a.g) * sz : 0, D == 3 ? az.hi * sz : 0};
    const int center = X1 + Y[1] + Z[1];
    // LDS row positions of x - 1 and x + 1 (periodic)
    const int XL[3] = {ax.lo - a.g, x, ax.hi - a.g};
    const TS* __restrict__ src = reinterpret_cast<const TS*>(a.src);
    TS* __restrict__ dst = reinterpret_cast<TS*>(a.dst);
    const int64_t sq = a.sq;
    extern __shared__ unsigned char lds_raw[];
    TS* tile = reinterpret_cast<TS*>(lds_raw);
    const int nx = a.nx;

    // walls are only in y / z here (x is periodic), so whether a block's row touches one is uniform
    const bool row_walls = WALLS && (ay.wall_lo || ay.wall_hi || az.wall_lo || az.wall_hi);

    TC f[Q];
    auto load = [&](auto walls) {
      constexpr bool W = decltype(walls)::value;
      int j = 0;
#pragma unroll
      for (int i = 0; i < Q; ++i) {
        const int cx = S::c[i][0], cy = S::c[i][1], cz = S::c[i][2];
        const int ii = S::inv[i];
        const int plane = (PAT == LBM355_PULL) ? i : ii;
        const int wplane = (PAT == LBM355_PULL) ? ii : i;
        bool wall = false;
        if (W)
          wall = (cy == 1 && ay.wall_lo) || (cy == -1 && ay.wall_hi) || (cz == 1 && az.wall_lo) ||
                 (cz == -1 && az.wall_hi);
        if (cx != 0) {
          if (wall)
            f[i] = load_nt<TS, TC>(src + wplane * sq + center);
          else
            tile[j * nx + x] = __builtin_nontemporal_load(src + plane * sq + X1 + Y[1 - cy] + Z[1 - cz]);
          ++j;
        } else {
          const TS* p = wall ? (src + wplane * sq + center) : (src + plane * sq + X1 + Y[1 - cy] + Z[1 - cz]);
          f[i] = load_nt<TS, TC>(p);
        }
      }
      __syncthreads();
      j = 0;
#pragma unroll
      for (int i = 0; i < Q; ++i) {
        const int cx = S::c[i][0], cy = S::c[i][1], cz = S::c[i][2];
        if (cx == 0) continue;
        bool wall = false;
        if (W)
          wall = (cy == 1 && ay.wall_lo) || (cy == -1 && ay.wall_hi) || (cz == 1 && az.wall_lo) ||
                 (cz == -1 && az.wall_hi);
        if (!wall) f[i] = (TC)tile[j * nx + XL[1 - cx]];
        ++j;
      }
      __syncthreads();
    };
    // AA odd writes f_i to x + c_i slot i; at a wall it is bounced into the cell's slot inv(i)
    auto store = [&](auto walls) {
      constexpr bool W = decltype(walls)::value;
      int j = 0;
#pragma unroll
      for (int i = 0; i < Q; ++i) {
        const int cx = S::c[i][0], cy = S::c[i][1], cz = S::c[i][2];
        const int ii = S::inv[i];
        bool wall = false;
        if (W)
          wall = (cy == -1 && ay.wall_lo) || (cy == 1 && ay.wall_hi) || (cz == -1 && az.wall_lo) ||
                 (cz == 1 && az.wall_hi);
        if (cx != 0) {
          if (wall)
            store_as<TS, TC>(dst + ii * sq + center, f[i]);
          else
            tile[j * nx + x] = (TS)f[i];
          ++j;
        } else {
          TS* p = wall ? (dst + ii * sq + center) : (dst + i * sq + X1 + Y[1 + cy] + Z[1 + cz]);
          store_nt<TS, TC>(p, f[i]);
        }
      }
      __syncthreads();
      j = 0;
#pragma unroll
      for (int i = 0; i < Q; ++i) {
        const int cx = S::c[i][0], cy = S::c[i][1], cz = S::c[i][2];
        if (cx == 0) continue;
        bool wall = false;
        if (W)
          wall = (cy == -1 && ay.wall_lo) || (cy == 1 && ay.wall_hi) || (cz == -1 && az.wall_lo) ||
                 (cz == 1 && az.wall_hi);
        if (!wall) __builtin_nontemporal_store(tile[j * nx + XL[1 - cx]], dst + i * sq + X1 + Y[1 + cy] + Z[1 + cz]);
        ++j;
      }
    };

    if (PAT == LBM355_AA_EVEN) {
#pragma unroll
      for (int i = 0; i < Q; ++i) f[i] = load_nt<TS, TC>(src + i * sq + center);
    } else if (row_walls) {
      load(std::true_type{});
    } else {
      load(std::false_type{});
    }

    collide_cell<S, FAM, TC, TS>(f, a, x, y, z);

    if (PAT == LBM355_PULL) {
#pragma unroll
      for (int i = 0; i < Q; ++i) store_nt<TS, TC>(dst + i * sq + center, f[i]);
    } else if (PAT == LBM355_AA_EVEN) {
#pragma unroll
      for (int i = 0; i < Q; ++i) store_nt<TS, TC>(dst + S::inv[i] * sq + center, f[i]);
    } else if (row_walls) {
      store(std::true_type{});
    } else {
      store(std::false_type{});
    }
  }
}

template <class S, int FAM, class TS, class TC>
void* pick_pattern(int pat, bool walls, bool rowlds) {
#define LBM_K(P, W, R) (void*)sweep_kernel<S, FAM, P, W, R, TS, TC>
  switch (pat) {
    case LBM355_PULL:
      if (rowlds) return walls ? LBM_K(LBM355_PULL, true, true) : LBM_K(LBM355_PULL, false, true);
      return walls ? LBM_K(LBM355_PULL, true, false) : LBM_K(LBM355_PULL, false, false);
    case LBM355_AA_EVEN:
      return LBM_K(LBM355_AA_EVEN, false, false);
    case LBM355_AA_ODD:
      if (rowlds) return walls ? LBM_K(LBM355_AA_ODD, true, true) : LBM_K(LBM355_AA_ODD, false, true);
      return walls ? LBM_K(LBM355_AA_ODD, true, false) : LBM_K(LBM355_AA_ODD, false, false);
  }
#undef LBM_K
  return nullptr;
}

template <class S, int FAM, class TC>
void* pick_store(int store, int pat, bool walls, bool rowlds) {
  if (store == LBM355_F64) return pick_pattern<S, FAM, double, TC>(pat, walls, rowlds);
  if (store == LBM355_F32) return pick_pattern<S, FAM, float, TC>(pat, walls, rowlds);
  if constexpr (sizeof(TC) == 4) {
    if (store == LBM355_F16) return pick_pattern<S, FAM, _Float16, TC>(pat, walls, rowlds);
  }
  return nullptr;
}

// Kernel pickers of each (family, compute type), one per translation unit; NULL when there is no kernel.
#define LBM355_PICKER(NAME) void* NAME(int stencil, int store, int pat, bool walls, bool rowlds)
LBM355_PICKER(pick_sweep_moment_f64);
LBM355_PICKER(pick_sweep_moment_f32);
LBM355_PICKER(pick_sweep_central_f64);
LBM355_PICKER(pick_sweep_central_f32);
LBM355_PICKER(pick_sweep_cumulant_f64);
LBM355_PICKER(pick_sweep_cumulant_f32);
LBM355_PICKER(pick_sweep_cumulant_corrected_f64);
LBM355_PICKER(pick_sweep_cumulant_corrected_f32);
LBM355_PICKER(pick_sweep_cumulant_general_f64);
LBM355_PICKER(pick_sweep_cumulant_general_f32);
LBM355_PICKER(pick_sweep_cumulant_eq3_f64);
LBM355_PICKER(pick_sweep_cumulant_eq3_f32);
LBM355_PICKER(pick_sweep_cumulant_learned_f32);
LBM355_PICKER(pick_sweep_cumulant_learned_eq3_f32);

}  // namespace lbm355
