// Learned subgrid closure: a per-cell MLP 10 -> 32 -> 32 -> 1 (tanh) that maps the local
// non-equilibrium second moments of a cell to an eddy viscosity, replacing the Smagorinsky rate
// (turbulence_models.py:25-84) in the cumulant collision.  Weights are BF16, products run on the
// MFMA units (v_mfma_f32_32x32x8bf16_1k) with FP32 accumulation, one wavefront per 64 cells.
//
// Operand layout (cells on the N dimension).  For the 32x32x8 MFMA, lane l holds
//   A (M x K):  row l%32,    k = 4*(l/32) + t   (t = 0..3)
//   B (K x N):  column l%32, k = 4*(l/32) + t
//   C (M x N):  column l%32, rows 8*(v/4) + 4*(l/32) + v%4   (v = 0..15)
// A layer computed as  H^T [units x cells] = W^T [units x inputs] . X^T [inputs x cells]  leaves lane l
// holding, for cell column l%32, exactly the rows that the B operand of the next layer needs for its
// K blocks (K block kb = registers 4kb..4kb+3), so the hidden layers chain in registers.  Only the
// input layer needs one exchange between lanes l and l^32 (cell 32c + l%32 lives in lane 32c + l%32).
//
// Features of a cell (raw, before the per-feature normalisation x' = (x - shift) * scale):
//   q_ab = Pi_neq_ab / rho (xx, yy, zz, xy, xz, yz), Pi = sqrt(2 q:q), tr q, q:q, tr(q^3)
// Output: nu_t = nu_scale * softplus(o) - nu_shift;  omega = 1 / (1/omega_0 + 3 nu_t), with the relaxation
// time kept >= 0.5 + 3e-6 (omega < 2).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lbm355 {

constexpr int CLOSURE_IN = 10, CLOSURE_IN_PAD = 16, CLOSURE_H = 32;

// Device parameter block (the layout lbmpy_amd/closure.py packs; 16-byte aligned size).
struct ClosureParams {
  uint16_t w1t[CLOSURE_H][CLOSURE_IN_PAD];  // layer 1, [unit][input], BF16 (inputs 10..15 zero)
  uint16_t w2t[CLOSURE_H][CLOSURE_H];       // layer 2, [unit][input], BF16
  uint16_t w3[CLOSURE_H];                   // output layer, BF16
  float b1[CLOSURE_H], b2[CLOSURE_H];
  float b3;
  float nu_scale;
  float in_shift[CLOSURE_IN], in_scale[CLOSURE_IN];
  float nu_shift;                           // constant viscosity offset (calibrated), see closure_omega
  float pad;
};

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v16f __attribute__((ext_vector_type(16)));

// FP32 -> BF16 bits, round to nearest even (finite inputs)
// FP32 -> BF16 with round-to-nearest-even (one v_cvt_pk_bf16_f32 on gfx950; the same result as adding
// 0x7FFF plus the lowest kept bit and truncating, for every finite value)
__device__ __forceinline__ uint16_t f2bf(float f) { return __builtin_bit_cast(uint16_t, (__bf16)f); }
__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float(((uint32_t)b) << 16); }

__device__ __forceinline__ float softplus(float o) { return o > 20.f ? o : log1pf(expf(o)); }
__device__ __forceinline__ float sigmoid(float o) { return 1.f / (1.f + expf(-o)); }

// raw features from the non-equilibrium second moments Pi_neq (xx, yy, zz, xy, xz, yz) and rho
__device__ __forceinline__ void closure_features(const float (&p)[6], float rho, float (&x)[CLOSURE_IN]) {
  const float ir = 1.f / rho;
  const float q[6] = {p[0] * ir, p[1] * ir, p[2] * ir, p[3] * ir, p[4] * ir, p[5] * ir};
  const float qq = q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + 2.f * (q[3] * q[3] + q[4] * q[4] + q[5] * q[5]);
  // tr(q^3) of the symmetric matrix [[xx, xy, xz], [xy, yy, yz], [xz, yz, zz]]
  const float t3 = q[0] * q[0] * q[0] + q[1] * q[1] * q[1] + q[2] * q[2] * q[2] +
                   3.f * (q[3] * q[3] * (q[0] + q[1]) + q[4] * q[4] * (q[0] + q[2]) + q[5] * q[5] * (q[1] + q[2])) +
                   6.f * q[3] * q[4] * q[5];
#pragma unroll
  for (int i = 0; i < 6; ++i) x[i] = q[i];
  x[6] = sqrtf(2.f * qq);
  x[7] = q[0] + q[1] + q[2];
  x[8] = qq;
  x[9] = t3;
}

// tanh in FP32 for the hidden layers: sign(x) (1 - e) / (1 + e) with e = exp(-2|x|) from the hardware base-2
// exponential and reciprocal; absolute error about 2e-7, far below the BF16 step of the activations that the
// next layer reads.  Five plain instructions and two transcendental ones, no select.
__device__ __forceinline__ float tanh_fast(float x) {
  const float e = __builtin_amdgcn_exp2f(-2.885390081777927f * fabsf(x));   // 2 log2(e) |x|
  return copysignf((1.f - e) * __builtin_amdgcn_rcpf(1.f + e), x);
}

// A fragment (4 BF16) of a [32 x K] row-major BF16 matrix for K block kb
__device__ __forceinline__ v4s a_frag(const uint16_t* m, int ld, int kb, int lane) {
  const int row = lane & 31, k = 8 * kb + 4 * (lane >> 5);
  const uint16_t* p = m + row * ld + k;
  v4s r;
  r[0] = (short)p[0];
  r[1] = (short)p[1];
  r[2] = (short)p[2];
  r[3] = (short)p[3];
  return r;
}

// unit (row) held by C register v of `lane`
__device__ __forceinline__ int c_row(int v, int lane) { return 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3); }

// Wave-collective forward pass for the 64 cells of a wavefront (every lane must take part).
// x: this lane's raw features.  Returns this lane's output o.  If h1/h2 are given, the tanh
// activations of both cell blocks c = 0, 1 are kept in C layout (h[c][v]) for the backward pass.
__device__ __forceinline__ float mlp_forward_wave(const ClosureParams* __restrict__ P, const float (&x)[CLOSURE_IN],
                                                  float (*h1)[16], float (*h2)[16]) {
  const int lane = threadIdx.x & 63;
  const int half = lane >> 5;
  // own normalised BF16 features, packed two per dword (inputs 10..15 are zero)
  uint32_t own[8];
#pragma unroll
  for (int d = 0; d < 8; ++d) {
    const int f0 = 2 * d, f1 = 2 * d + 1;
    const uint32_t lo = f0 < CLOSURE_IN ? f2bf((x[f0] - P->in_shift[f0]) * P->in_scale[f0]) : 0u;
    const uint32_t hi = f1 < CLOSURE_IN ? f2bf((x[f1] - P->in_shift[f1]) * P->in_scale[f1]) : 0u;
    own[d] = lo | (hi << 16);
  }
  uint32_t other[8];
#pragma unroll
  for (int d = 0; d < 8; ++d) other[d] = (uint32_t)__shfl_xor((int)own[d], 32, 64);
  const v4s w1a[2] = {a_frag(&P->w1t[0][0], CLOSURE_IN_PAD, 0, lane), a_frag(&P->w1t[0][0], CLOSURE_IN_PAD, 1, lane)};
  v4s w2a[4];
#pragma unroll
  for (int kb = 0; kb < 4; ++kb) w2a[kb] = a_frag(&P->w2t[0][0], CLOSURE_H, kb, lane);
  float out = 0.f;
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    // features of cell 32c + lane%32: this lane's own (c == half) or the partner lane's; selected element by
    // element with compile-time indices so that both arrays stay in registers
    const bool mine = (c == half);
    v16f acc;
#pragma unroll
    for (int v = 0; v < 16; ++v) acc[v] = P->b1[c_row(v, lane)];
#pragma unroll
    for (int kb = 0; kb < 2; ++kb) {
      // k = 8kb + 4*half + t  ->  dwords 4kb + 2*half + {0, 1}
      const uint32_t o0 = half ? own[4 * kb + 2] : own[4 * kb], o1 = half ? own[4 * kb + 3] : own[4 * kb + 1];
      const uint32_t p0 = half ? other[4 * kb + 2] : other[4 * kb], p1 = half ? other[4 * kb + 3] : other[4 * kb + 1];
      const uint32_t d0 = mine ? o0 : p0, d1 = mine ? o1 : p1;
      v4s b;
      b[0] = (short)(d0 & 0xFFFF);
      b[1] = (short)(d0 >> 16);
      b[2] = (short)(d1 & 0xFFFF);
      b[3] = (short)(d1 >> 16);
      acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(w1a[kb], b, acc, 0, 0, 0);
    }
    v4s hb[4];
#pragma unroll
    for (int v = 0; v < 16; ++v) {
      const float h = tanh_fast(acc[v]);
      if (h1) h1[c][v] = h;
      hb[v >> 2][v & 3] = (short)f2bf(h);
    }
    v16f acc2;
#pragma unroll
    for (int v = 0; v < 16; ++v) acc2[v] = P->b2[c_row(v, lane)];
#pragma unroll
    for (int kb = 0; kb < 4; ++kb) acc2 = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(w2a[kb], hb[kb], acc2, 0, 0, 0);
    float part = 0.f;
#pragma unroll
    for (int v = 0; v < 16; ++v) {
      const float h = tanh_fast(acc2[v]);
      if (h2) h2[c][v] = h;
      part += bf2f(P->w3[c_row(v, lane)]) * h;
    }
    part += __shfl_xor(part, 32, 64);
    if (c == half) out = part + P->b3;
  }
  return out;
}

// shear relaxation rate from the network output: nu_t = nu_scale softplus(o) - nu_shift added to the
// molecular viscosity, relaxation time at least 0.5 + 3e-6
__device__ __forceinline__ float closure_omega(const ClosureParams* __restrict__ P, float o, float omega0) {
  const float tau = 1.f / omega0 + 3.f * (P->nu_scale * softplus(o) - P->nu_shift);
  return 1.f / fmaxf(tau, 0.500003f);
}

}  // namespace lbm355
