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
void* eso_pick_store(int store, int pat) {
  if (store == LBM355_F64) return eso_pick_pattern<S, FAM, double, TC>(pat);
  if (store == LBM355_F32) return eso_pick_pattern<S, FAM, float, TC>(pat);
  if constexpr (sizeof(TC) == 4) {
    if (store == LBM355_F16) return eso_pick_pattern<S, FAM, _Float16, TC>(pat);
  }
  return nullptr;
}

template <class TC>
void* eso_pick(const lbm355_sweep_desc* d) {
  const int st = d->store_dtype, pat = d->pattern;
  if (d->family == LBM355_FAMILY_MOMENT || d->family == LBM355_FAMILY_CENTRAL) {
    const bool m = d->family == LBM355_FAMILY_MOMENT;
    if (d->stencil == LBM355_D2Q9)
      return m ? eso_pick_store<D2Q9, LBM355_FAMILY_MOMENT, TC>(st, pat)
               : eso_pick_store<D2Q9, LBM355_FAMILY_CENTRAL, TC>(st, pat);
    if (d->stencil == LBM355_D3Q19)
      return m ? eso_pick_store<D3Q19, LBM355_FAMILY_MOMENT, TC>(st, pat)
               : eso_pick_store<D3Q19, LBM355_FAMILY_CENTRAL, TC>(st, pat);
    if (d->stencil == LBM355_D3Q27)
      return m ? eso_pick_store<D3Q27, LBM355_FAMILY_MOMENT, TC>(st, pat)
               : eso_pick_store<D3Q27, LBM355_FAMILY_CENTRAL, TC>(st, pat);
  } else if (d->family == LBM355_FAMILY_CUMULANT && d->stencil == LBM355_D3Q27) {
    if (d->flags & (LBM355_FLAG_GALILEAN | LBM355_FLAG_FOURTH_ORDER))
      return eso_pick_store<D3Q27, FAMILY_CUMULANT_CORRECTED, TC>(st, pat);
    return eso_pick_store<D3Q27, LBM355_FAMILY_CUMULANT, TC>(st, pat);
  }
  return nullptr;
}

}  // namespace lbm355

using namespace lbm355;

int lbm355_launch_sweep_eso(const lbm355_sweep_desc* d, const void* src, void* dst, void* rho, void* u,
                            hipStream_t s) {
  void* k = d->compute_dtype == LBM355_F64 ? eso_pick<double>(d) : eso_pick<float>(d);
  if (!k) return lbm355_set_error(LBM355_ERR_UNSUPPORTED, "no Eso kernel for this stencil/family/dtype combination");
  const int bx = d->nx >= 256 ? 256 : ((d->nx + 63) / 64) * 64;
  const dim3 grid((d->nx + bx - 1) / bx, d->ny, d->nz), block(bx);
  if (d->compute_dtype == LBM355_F64) {
    SweepArgs<double> a = make_args<double>(d, src, dst, rho, u);
    hipLaunchKernelGGL((void (*)(SweepArgs<double>))k, grid, block, 0, s, a);
  } else {
    SweepArgs<float> a = make_args<float>(d, src, dst, rho, u);
    hipLaunchKernelGGL((void (*)(SweepArgs<float>))k, grid, block, 0, s, a);
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    return lbm355_set_error(LBM355_ERR_LAUNCH, (std::string("launch failed: ") + hipGetErrorString(e)).c_str());
  return LBM355_OK;
}
