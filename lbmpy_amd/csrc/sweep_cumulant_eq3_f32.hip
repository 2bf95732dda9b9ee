// D3Q27 cumulant collision with every cumulant of order >= 3 relaxed with rate 1 (set to equilibrium without
// being computed), FP32 arithmetic: the sweep kernels of this family and compute type for every streaming
// pattern (templates in sweep_kernel.h and sweep_eso_kernel.h).
#include "sweep_eso_kernel.h"
#include "sweep_kernel.h"

namespace lbm355 {

void* pick_sweep_cumulant_eq3_f32(int stencil, int store, int pat, bool walls, bool rowlds) {
  if (stencil != LBM355_D3Q27) return nullptr;
  return pick_store<D3Q27, FAMILY_CUMULANT_EQ3, float>(store, pat, walls, rowlds);
}

void* eso_pick_cumulant_eq3_f32(int stencil, int store, int pat, bool rows) {
  if (stencil != LBM355_D3Q27) return nullptr;
  return eso_pick_store<D3Q27, FAMILY_CUMULANT_EQ3, float>(store, pat, rows);
}

}  // namespace lbm355
