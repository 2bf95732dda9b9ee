// Stream-collide sweep for FP16 storage with whole-row staging through LDS.
//
// With 2-byte populations, one element per lane gives 128-byte wavefront accesses, half of what the
// memory system moves efficiently.  Here a block spans one periodic row (blockDim.x == nx), and every
// population row the cell update reads or writes moves between HBM and LDS as 4-byte (half2)
// non-temporal accesses by the first nx/2 threads.  Each thread then takes its element from LDS, with
// the x shift of its direction applied there, collides through the shared per-cell code
// (collide_cell), and puts the results back into LDS rows that are stored the same way.  Measured on
// MI355X this is 2-5% faster than per-element FP16 access for the moment family and 9% slower for the
// D3Q27 cumulant (whose heavier collision hides the narrow accesses better at the higher occupancy of the
// per-element kernel), so only the moment family uses it.
//
// Rows per direction i (y / z walls are uniform over a row, x is periodic):
//   pull     read  src[i]      row (y - c_y, z - c_z), element x - c_x   (wall: src[inv i], own row, x)
//            write dst[i]      own row, element x
//   AA even  read  f[i]        own row, x;   write f[inv i] own row, x
//   AA odd   read  f[inv i]    row (y - c_y, z - c_z), element x - c_x   (wall: f[i], own row, x)
//            write f[i]        row (y + c_y, z + c_z), element x + c_x   (wall: f[inv i], own row, x)
// which is the addressing of sweep_cell (fieldaccess.py:66-75, 136-157; boundaryconditions.py:104-139).
#include <string>

#include "sweep_cell.h"

namespace lbm355 {

template <class S, int FAM, int PAT, bool WALLS, class TC>
__global__ void __launch_bounds__(512) sweep_h2_kernel(SweepArgs<TC> a) {
  constexpr int Q = S::Q, D = S::D;
  const int x = threadIdx.x;
  const int y = blockIdx.y, z = blockIdx.z;
  const int nx = a.nx, half_n = nx >> 1;
  const AxisNb ay = axis_neighbours(y, a.ny, 0, a.face[1]);
  AxisNb az = {0, 0, false, false};
  if (D == 3) az = axis_neighbours(z, a.nz, 0, a.face[2]);
  const int sy = (int)a.sy, sz = (int)a.sz;
  const int Y[3] = {ay.lo * sy, y * sy, ay.hi * sy};
  const int Z[3] = {D == 3 ? az.lo * sz : 0, D == 3 ? z * sz : 0, D == 3 ? az.hi * sz : 0};
  const int64_t sq = a.sq;
  const bool row_walls = WALLS && (ay.wall_lo || ay.wall_hi || az.wall_lo || az.wall_hi);
  const _Float16* __restrict__ src = reinterpret_cast<const _Float16*>(a.src);
  _Float16* __restrict__ dst = reinterpret_cast<_Float16*>(a.dst);
  extern __shared__ unsigned char lds_raw[];
  _Float16* tile = reinterpret_cast<_Float16*>(lds_raw);        // [Q][nx]
  uint32_t* tile2 = reinterpret_cast<uint32_t*>(lds_raw);       // [Q][nx / 2]

  auto rd_wall = [&](int i) {
    if (!WALLS || !row_walls) return false;
    const int cy = S::c[i][1], cz = S::c[i][2];
    return (cy == 1 && ay.wall_lo) || (cy == -1 && ay.wall_hi) || (cz == 1 && az.wall_lo) || (cz == -1 && az.wall_hi);
  };
  auto wr_wall = [&](int i) {
    if (!WALLS || !row_walls) return false;
    const int cy = S::c[i][1], cz = S::c[i][2];
    return (cy == -1 && ay.wall_lo) || (cy == 1 && ay.wall_hi) || (cz == -1 && az.wall_lo) || (cz == 1 && az.wall_hi);
  };

  // ---- stage the read rows (the first nx/2 threads, whole wavefronts since nx % 128 == 0)
  if (x < half_n) {
#pragma unroll
    for (int i = 0; i < Q; ++i) {
      const int cy = S::c[i][1], cz = S::c[i][2];
      int64_t row;
      if (PAT == LBM355_AA_EVEN) {
        row = i * sq + Y[1] + Z[1];
      } else {
        const bool w = rd_wall(i);
        const int plane = (PAT == LBM355_PULL) ? (w ? S::inv[i] : i) : (w ? i : S::inv[i]);
        row = plane * sq + (w ? Y[1] + Z[1] : Y[1 - cy] + Z[1 - cz]);
      }
      tile2[i * half_n + x] = __builtin_nontemporal_load(reinterpret_cast<const uint32_t*>(src + row) + x);
    }
  }
  __syncthreads();
  TC f[Q];
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    const int cx = S::c[i][0];
    int sx = 0;
    if (PAT != LBM355_AA_EVEN && !rd_wall(i)) sx = cx;
    int xs = x - sx;
    xs = xs < 0 ? xs + nx : (xs >= nx ? xs - nx : xs);
    f[i] = (TC)tile[i * nx + xs];
  }
  __syncthreads();

  collide_cell<S, FAM, TC>(f, a, x, y, z);

  // ---- write rows: LDS first (with the x shift), then whole rows
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    const int cx = S::c[i][0];
    int sx = 0;
    if (PAT == LBM355_AA_ODD && !wr_wall(i)) sx = cx;
    int xd = x + sx;
    xd = xd < 0 ? xd + nx : (xd >= nx ? xd - nx : xd);
    tile[i * nx + xd] = (_Float16)f[i];
  }
  __syncthreads();
  if (x < half_n) {
#pragma unroll
    for (int i = 0; i < Q; ++i) {
      const int cy = S::c[i][1], cz = S::c[i][2];
      int64_t row;
      if (PAT == LBM355_PULL) {
        row = i * sq + Y[1] + Z[1];
      } else if (PAT == LBM355_AA_EVEN) {
        row = S::inv[i] * sq + Y[1] + Z[1];
      } else {
        const bool w = wr_wall(i);
        row = w ? (S::inv[i] * sq + Y[1] + Z[1]) : (i * sq + Y[1 + cy] + Z[1 + cz]);
      }
      __builtin_nontemporal_store(tile2[i * half_n + x], reinterpret_cast<uint32_t*>(dst + row) + x);
    }
  }
}

template <class S, int FAM>
void* h2_pick_pattern(int pat, bool walls) {
  switch (pat) {
    case LBM355_PULL:
      return walls ? (void*)sweep_h2_kernel<S, FAM, LBM355_PULL, true, float>
                   : (void*)sweep_h2_kernel<S, FAM, LBM355_PULL, false, float>;
    case LBM355_AA_EVEN:
      return (void*)sweep_h2_kernel<S, FAM, LBM355_AA_EVEN, false, float>;
    case LBM355_AA_ODD:
      return walls ? (void*)sweep_h2_kernel<S, FAM, LBM355_AA_ODD, true, float>
                   : (void*)sweep_h2_kernel<S, FAM, LBM355_AA_ODD, false, float>;
  }
  return nullptr;
}

}  // namespace lbm355

using namespace lbm355;

// Launches the row-staged FP16 sweep when it applies; returns -1 (and launches nothing) when it does not:
// FP16 storage with FP32 arithmetic, moment family, no ghost layers, periodic x with
// nx % 128 == 0 and nx <= 512, even strides and a 4-byte aligned base, no density/velocity outputs and no
// subgrid model.  The descriptor has been validated by the caller.
int lbm355_launch_sweep_h2(const lbm355_sweep_desc* d, const void* src, void* dst, void* rho, void* u,
                           hipStream_t s) {
  if (d->store_dtype != LBM355_F16 || d->compute_dtype != LBM355_F32) return -1;
  if (d->ghost_layers != 0 || d->face[0] != LBM355_FACE_PERIODIC) return -1;
  if (d->nx % 128 != 0 || d->nx > 512) return -1;
  if ((d->stride_y | d->stride_z | d->stride_q) & 1) return -1;
  if ((((uintptr_t)src) | ((uintptr_t)dst)) & 3) return -1;
  if (rho || u || (d->flags & (LBM355_FLAG_SMAGORINSKY | LBM355_FLAG_LEARNED_CLOSURE))) return -1;
  const int D = d->stencil == LBM355_D2Q9 ? 2 : 3;
  bool walls = false;
  for (int k = 1; k < D; ++k) walls |= d->face[k] == LBM355_FACE_NOSLIP;
  void* k = nullptr;
  if (d->family == LBM355_FAMILY_MOMENT) {
    if (d->stencil == LBM355_D2Q9) k = h2_pick_pattern<D2Q9, LBM355_FAMILY_MOMENT>(d->pattern, walls);
    else if (d->stencil == LBM355_D3Q19) k = h2_pick_pattern<D3Q19, LBM355_FAMILY_MOMENT>(d->pattern, walls);
    else if (d->stencil == LBM355_D3Q27) k = h2_pick_pattern<D3Q27, LBM355_FAMILY_MOMENT>(d->pattern, walls);
  }
  if (!k) return -1;
  const int Q = d->stencil == LBM355_D2Q9 ? 9 : (d->stencil == LBM355_D3Q19 ? 19 : 27);
  SweepArgs<float> a = make_args<float>(d, src, dst, rho, u);
  const size_t lds = (size_t)Q * d->nx * 2;
  hipLaunchKernelGGL((void (*)(SweepArgs<float>))k, dim3(1, d->ny, d->nz), dim3(d->nx), lds, s, a);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return lbm355_set_error(LBM355_ERR_LAUNCH, (std::string("launch failed: ") + hipGetErrorString(e)).c_str());
  return LBM355_OK;
}
