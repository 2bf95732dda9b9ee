// Fused stream-collide sweeps of lbmpy's in-place Eso patterns (fieldaccess.py:160-316): EsoTwist,
// EsoPull and EsoPush, each an EVEN and an ODD kernel on one PDF array.  One thread per interior cell
// (the same grid as the pull / AA sweeps); each population is read from and written to the (cell
// offset, slot) of the pattern's access table (pattern_access in lbm_common.h), with periodic axes
// wrapped in the index math and ghost faces reaching into the ghost layer, where boundary links
// (lbm355_boundary_links) pick the populations up.  The collision is the shared per-cell code, so an
// Eso run gives the same numbers as a pull or AA run of the same method.  The two-array push pattern
// (fieldaccess.py:78-87: read slot i at x, write slot i at x + c_i) goes through the same table.
#include <string>

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

  collide_cell<S, FAM, TC>(f, a, x, y, z);

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

template <class TC>
void* eso_pick(const lbm355_sweep_desc* d, bool rows) {
  const int st = d->store_dtype, pat = d->pattern;
  if (d->family == LBM355_FAMILY_MOMENT || d->family == LBM355_FAMILY_CENTRAL) {
    const bool m = d->family == LBM355_FAMILY_MOMENT;
    if (d->stencil == LBM355_D2Q9)
      return m ? eso_pick_store<D2Q9, LBM355_FAMILY_MOMENT, TC>(st, pat, rows)
               : eso_pick_store<D2Q9, LBM355_FAMILY_CENTRAL, TC>(st, pat, rows);
    if (d->stencil == LBM355_D3Q19)
      return m ? eso_pick_store<D3Q19, LBM355_FAMILY_MOMENT, TC>(st, pat, rows)
               : eso_pick_store<D3Q19, LBM355_FAMILY_CENTRAL, TC>(st, pat, rows);
    if (d->stencil == LBM355_D3Q27)
      return m ? eso_pick_store<D3Q27, LBM355_FAMILY_MOMENT, TC>(st, pat, rows)
               : eso_pick_store<D3Q27, LBM355_FAMILY_CENTRAL, TC>(st, pat, rows);
  } else if (d->family == LBM355_FAMILY_CUMULANT && d->stencil == LBM355_D3Q27) {
    if (d->flags & (LBM355_FLAG_GALILEAN | LBM355_FLAG_FOURTH_ORDER))
      return eso_pick_store<D3Q27, FAMILY_CUMULANT_CORRECTED, TC>(st, pat, rows);
    return eso_pick_store<D3Q27, LBM355_FAMILY_CUMULANT, TC>(st, pat, rows);
  }
  return nullptr;
}

}  // namespace lbm355

using namespace lbm355;

int lbm355_launch_sweep_eso(const lbm355_sweep_desc* d, const void* src, void* dst, void* rho, void* u,
                            hipStream_t s) {
  // whole periodic rows of at most 512 cells (whole wavefronts) go through the row-staged kernel
  const size_t es = d->store_dtype == LBM355_F64 ? 8 : (d->store_dtype == LBM355_F32 ? 4 : 2);
  const int rows_needed = d->stencil == LBM355_D2Q9 ? eso_max_rows<D2Q9>()
                          : d->stencil == LBM355_D3Q19 ? eso_max_rows<D3Q19>() : eso_max_rows<D3Q27>();
  const size_t lds = (size_t)rows_needed * d->nx * es;
  const bool rows = d->face[0] == LBM355_FACE_PERIODIC && d->nx % 64 == 0 && d->nx <= 512 && lds <= 65536;
  void* k = d->compute_dtype == LBM355_F64 ? eso_pick<double>(d, rows) : eso_pick<float>(d, rows);
  if (!k) return lbm355_set_error(LBM355_ERR_UNSUPPORTED, "no Eso kernel for this stencil/family/dtype combination");
  const int bx = rows ? d->nx : (d->nx >= 256 ? 256 : ((d->nx + 63) / 64) * 64);
  const dim3 grid((d->nx + bx - 1) / bx, d->ny, d->nz), block(bx);
  const size_t shm = rows ? lds : 0;
  if (d->compute_dtype == LBM355_F64) {
    SweepArgs<double> a = make_args<double>(d, src, dst, rho, u);
    hipLaunchKernelGGL((void (*)(SweepArgs<double>))k, grid, block, shm, s, a);
  } else {
    SweepArgs<float> a = make_args<float>(d, src, dst, rho, u);
    hipLaunchKernelGGL((void (*)(SweepArgs<float>))k, grid, block, shm, s, a);
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    return lbm355_set_error(LBM355_ERR_LAUNCH, (std::string("launch failed: ") + hipGetErrorString(e)).c_str());
  return LBM355_OK;
}
