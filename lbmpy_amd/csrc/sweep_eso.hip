// Fused stream-collide sweeps of lbmpy's in-place Eso patterns (fieldaccess.py:160-316): EsoTwist,
// EsoPull and EsoPush, each an EVEN and an ODD kernel on one PDF array.  One thread per interior cell
// (the same grid as the pull / AA sweeps); each population is read from and written to the (cell
// offset, slot) of the pattern's access table (pattern_access in lbm_common.h), with periodic axes
// wrapped in the index math and ghost faces reaching into the ghost layer, where boundary links
// (lbm355_boundary_links) pick the populations up.  The collision is the shared per-cell code, so an
// Eso run gives the same numbers as a pull or AA run of the same method.  The two-array push pattern
// (fieldaccess.py:78-87: read slot i at x, write slot i at x + c_i) goes through the same table.
#include <string>

#include "occupancy.h"
#include "sweep_eso_kernel.h"

namespace lbm355 {

template <class TC>
void* eso_pick(const lbm355_sweep_desc* d, bool rows) {
  constexpr bool f64 = sizeof(TC) == 8;
  const int sten = d->stencil, st = d->store_dtype, pat = d->pattern;
  if (d->family == LBM355_FAMILY_MOMENT)
    return f64 ? eso_pick_moment_f64(sten, st, pat, rows) : eso_pick_moment_f32(sten, st, pat, rows);
  if (d->family == LBM355_FAMILY_CENTRAL)
    return f64 ? eso_pick_central_f64(sten, st, pat, rows) : eso_pick_central_f32(sten, st, pat, rows);
  if (d->family == LBM355_FAMILY_CUMULANT) {
    if (d->flags & (LBM355_FLAG_GALILEAN | LBM355_FLAG_FOURTH_ORDER))
      return f64 ? eso_pick_cumulant_corrected_f64(sten, st, pat, rows)
                 : eso_pick_cumulant_corrected_f32(sten, st, pat, rows);
    if (d->flags & LBM355_FLAG_GENERAL_CUMULANT)
      return f64 ? eso_pick_cumulant_general_f64(sten, st, pat, rows)
                 : eso_pick_cumulant_general_f32(sten, st, pat, rows);
    if (d->flags & LBM355_FLAG_HIGHER_ORDERS_RATE1)
      return f64 ? eso_pick_cumulant_eq3_f64(sten, st, pat, rows) : eso_pick_cumulant_eq3_f32(sten, st, pat, rows);
    return f64 ? eso_pick_cumulant_f64(sten, st, pat, rows) : eso_pick_cumulant_f32(sten, st, pat, rows);
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
  const size_t shm = lds_for_wave_cap(sweep_wave_cap(d, rows, false, bx), bx, rows ? lds : 0);
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
