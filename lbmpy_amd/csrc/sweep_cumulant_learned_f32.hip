// D3Q27 cumulant collision with the learned closure (FP32 storage and arithmetic): the pull / AA sweep kernels of
// this family, with per-rate higher orders and with every order >= 3 at rate 1 (template in sweep_kernel.h).
#include "sweep_kernel.h"

namespace lbm355 {

void* pick_sweep_cumulant_learned_f32(int stencil, int store, int pat, bool walls, bool rowlds) {
  if (stencil != LBM355_D3Q27 || store != LBM355_F32) return nullptr;
  return pick_pattern<D3Q27, FAMILY_CUMULANT_LEARNED, float, float>(pat, walls, rowlds);
}

void* pick_sweep_cumulant_learned_eq3_f32(int stencil, int store, int pat, bool walls, bool rowlds) {
  if (stencil != LBM355_D3Q27 || store != LBM355_F32) return nullptr;
  return pick_pattern<D3Q27, FAMILY_CUMULANT_LEARNED_EQ3, float, float>(pat, walls, rowlds);
}

}  // namespace lbm355
