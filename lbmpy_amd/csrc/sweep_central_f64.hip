// central-moment collisions, FP64 arithmetic: the sweep kernels of this family and compute type for every streaming pattern
// (templates in sweep_kernel.h and sweep_eso_kernel.h).
#include "sweep_eso_kernel.h"
#include "sweep_kernel.h"

namespace lbm355 {

void* pick_sweep_central_f64(int stencil, int store, int pat, bool walls, bool rowlds) {
  if (stencil == LBM355_D2Q9) return pick_store<D2Q9, LBM355_FAMILY_CENTRAL, double>(store, pat, walls, rowlds);
  if (stencil == LBM355_D3Q19) return pick_store<D3Q19, LBM355_FAMILY_CENTRAL, double>(store, pat, walls, rowlds);
  if (stencil == LBM355_D3Q27) return pick_store<D3Q27, LBM355_FAMILY_CENTRAL, double>(store, pat, walls, rowlds);
  return nullptr;
}

void* eso_pick_central_f64(int stencil, int store, int pat, bool rows) {
  if (stencil == LBM355_D2Q9) return eso_pick_store<D2Q9, LBM355_FAMILY_CENTRAL, double>(store, pat, rows);
  if (stencil == LBM355_D3Q19) return eso_pick_store<D3Q19, LBM355_FAMILY_CENTRAL, double>(store, pat, rows);
  if (stencil == LBM355_D3Q27) return eso_pick_store<D3Q27, LBM355_FAMILY_CENTRAL, double>(store, pat, rows);
  return nullptr;
}

}  // namespace lbm355
