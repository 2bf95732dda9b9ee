// Central-moment collision (lbmpy methods/momentbased/centralmomentbasedmethod.py:305-394 with the
// binomial chimera transform of moment_transforms/centralmomenttransforms.py:168-483).
//
// lbmpy relaxes polynomial central moments K = P kappa:  K* = K + D (K_eq - K) + F_K.  The map
// f -> kappa is linear for a given u (kappa = T(u) f, T = binomial shift o raw moments), so
//     f* = f + T(u)^-1 [ P^-1 D P (kappa_eq - kappa) + P^-1 F_K ],
// which is how this kernel evaluates it: the change of the populations is computed directly,
// so no O(1) background value is ever subtracted (zero-centred storage keeps its precision).
//   dk     = kappa_eq - kappa  (27 monomial central moments, canonical index 9 ex + 3 ey + ez);
//            kappa = kt + kappa_w(u) with kt the central moments of the stored deviations and
//            kappa_w those of the lattice weights; kappa_eq - gamma comes from a polynomial table
//            in (drho, u), kappa_w - gamma = beta(u) in closed form (collide_cumulant.h).
//   relax  = r0 dk + sum_k c_k Pinv[:,k] (P[k,:] . dk)         (rank-one corrections, as for MRT)
//   force  Simple/Luo: T^-1 P^-1 F_K = L (the population force vector itself);
//          Guo:        (I - D/2) in polynomial space -> (1 - r0/2) L - 1/2 sum_k c_k T^-1 Pinv[:,k] (P[k,:] . T L)
//   df     = T(u)^-1 (relax + force part);  f* = f + df.
// Every stencil uses the chimera in both directions on the 3 x 3 x 3 grid of velocities, absent directions
// holding zeros.  For D3Q19 the raw moments of the eight monomials with all three exponents non-zero vanish on its
// directions, so a D3Q19 method's moments are the other 19; setting those eight raw moments to zero before the
// inverse chimera makes the grid's corners zero and leaves exactly the populations the 19 x 19 inverse gives.
#pragma once
#include "lbm_common.h"
#include "collide_cumulant.h"

namespace lbm355 {

struct CentralTab {
  static constexpr int EQ = 0,                 // [27][20]: (1, drho) x (1, ux, uy, uz, ux2, uy2, uz2, uxuy, uxuz, uyuz)
      R0 = 540, COEF = 541,                    // base rate, [27] active coefficients r_k - r0
      PROW = 568, PINV = 568 + 729,            // [27 active][27 monomials]
      L0 = 568 + 1458, LU = L0 + 27,           // force: L_i = L0_i + LU_i . u  (population space)
      FMODE = LU + 81,                         // 0 none, 1 Simple/Luo, 2 Guo
      GAUSS = FMODE + 1,                       // 1: kappa_eq = rho gamma, 2: the incompressible form
                                               // (LBM355_FLAG_GAUSSIAN_EQUILIBRIUM)
      SHEAR_ON = GAUSS + 1, SHEAR = GAUSS + 2, // 1, s: the closed-form rank-one sum (LBM355_FLAG_SHEAR_PROJECTOR)
      SIZE = GAUSS + 3;
};

template <class S, class T>
__device__ __forceinline__ void load_grid(const T (&f)[S::Q], T (&m)[3][3][3]) {
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b)
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const int i = dir_index<S>(a - 1, b - 1, c - 1);
        m[a][b][c] = i >= 0 ? f[i] : T(0);
      }
}

// populations on the 3x3x3 grid -> raw moments -> central moments about u (in place)
// (2-D stencils sit in the cz = 0 plane: the z passes then act as the identity on it)
template <class S, class T>
__device__ __forceinline__ void forward_central(T (&m)[3][3][3], const T (&u)[3]) {
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) chim_fwd(m[i][j][0], m[i][j][1], m[i][j][2]);
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int k = 0; k < 3; ++k) chim_fwd(m[i][0][k], m[i][1][k], m[i][2][k]);
#pragma unroll
  for (int j = 0; j < 3; ++j)
#pragma unroll
    for (int k = 0; k < 3; ++k) chim_fwd(m[0][j][k], m[1][j][k], m[2][j][k]);
#pragma unroll
  for (int j = 0; j < 3; ++j)
#pragma unroll
    for (int k = 0; k < 3; ++k) shift_fwd(m[0][j][k], m[1][j][k], m[2][j][k], u[0]);
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int k = 0; k < 3; ++k) shift_fwd(m[i][0][k], m[i][1][k], m[i][2][k], u[1]);
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) shift_fwd(m[i][j][0], m[i][j][1], m[i][j][2], u[2]);
}

// central moments about u -> populations (the inverse of forward_central), result on the grid
template <class S, class T>
__device__ __forceinline__ void backward_central(T (&m)[3][3][3], const T (&u)[3], T (&out)[S::Q]) {
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) shift_bwd(m[i][j][0], m[i][j][1], m[i][j][2], u[2]);
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int k = 0; k < 3; ++k) shift_bwd(m[i][0][k], m[i][1][k], m[i][2][k], u[1]);
#pragma unroll
  for (int j = 0; j < 3; ++j)
#pragma unroll
    for (int k = 0; k < 3; ++k) shift_bwd(m[0][j][k], m[1][j][k], m[2][j][k], u[0]);
  if constexpr (S::Q == 19) {
    // the raw moments of x y z, x^2 y z, ..., x^2 y^2 z^2 are zero for D3Q19 populations
#pragma unroll
    for (int a = 1; a < 3; ++a)
#pragma unroll
      for (int b = 1; b < 3; ++b)
#pragma unroll
        for (int c = 1; c < 3; ++c) m[a][b][c] = T(0);
  }
#pragma unroll
  for (int j = 0; j < 3; ++j)
#pragma unroll
    for (int k = 0; k < 3; ++k) chim_bwd(m[0][j][k], m[1][j][k], m[2][j][k]);
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int k = 0; k < 3; ++k) chim_bwd(m[i][0][k], m[i][1][k], m[i][2][k]);
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) chim_bwd(m[i][j][0], m[i][j][1], m[i][j][2]);
#pragma unroll
  for (int i = 0; i < S::Q; ++i) out[i] = m[S::c[i][0] + 1][S::c[i][1] + 1][S::c[i][2] + 1];
}

template <class T, int EX, int EY, int EZ>
__device__ __forceinline__ T beta_e(const T (&u)[3], const T (&u2)[3]) {
  return beta<EX, EY, EZ, T>(u, u2);
}

// Central moment about u of a rest Gaussian (c_s^2 = 1/3, centred at 0) of unit weight, truncated at second order
// in u: prod_a g(e_a, u_a) with g(0) = 1, g(1) = -u, g(2) = 1/3 + u^2 -- the drho part of the incompressible
// continuous equilibrium's central moments.
template <class T, int EX, int EY, int EZ>
__device__ __forceinline__ T rest_gauss(const T (&u)[3], const T (&u2)[3]) {
  constexpr int e[3] = {EX, EY, EZ};
  constexpr int n1 = (EX == 1) + (EY == 1) + (EZ == 1), n2 = (EX == 2) + (EY == 2) + (EZ == 2);
  if constexpr (n1 >= 3) {
    return T(0);
  } else {
    constexpr double third_n2 = n2 == 0 ? 1.0 : (n2 == 1 ? 1.0 / 3.0 : (n2 == 2 ? 1.0 / 9.0 : 1.0 / 27.0));
    T lin = T(1);
#pragma unroll
    for (int a = 0; a < 3; ++a)
      if (e[a] == 1) lin *= -u[a];
    T quad = T(third_n2);
    if constexpr (n1 == 0 && n2 > 0) {
      constexpr double c = n2 == 1 ? 1.0 : (n2 == 2 ? 1.0 / 3.0 : 1.0 / 9.0);
      T s = T(0);
#pragma unroll
      for (int a = 0; a < 3; ++a)
        if (e[a] == 2) s += u2[a];
      quad += T(c) * s;
    }
    return lin * quad;
  }
}

// EQUILIBRIUM: relax every central moment with rate 1 and no force terms, which maps a state to the
// method's equilibrium at its (rho, j) -- the macroscopic setter's use (macroscopic_value_kernels.py:29-45)
template <class S, class T, bool EQUILIBRIUM = false>
__device__ __forceinline__ void collide_central(T (&f)[S::Q], const SweepArgs<T>& a, T& rho_out_v, T (&u_out_v)[3]) {
  constexpr int Q = S::Q, D = S::D;
  using L = CentralTab;
  const TablePtr<T> __restrict__ tab = table_ptr(a.tab);

  T m[3][3][3];
  load_grid<S, T>(f, m);
  // density / velocity from the raw sums (computed again by the chimera below)
  T s = 0, j[3] = {0, 0, 0};
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    s += f[i];
#pragma unroll
    for (int d = 0; d < D; ++d)
      if (S::c[i][d] != 0) j[d] += (S::c[i][d] > 0) ? f[i] : -f[i];
  }
  const bool comp = a.flags & LBM355_FLAG_COMPRESSIBLE;
  const T drho = s;  // deviation storage (absolute input is shifted by the caller)
  const T rho = T(1) + drho;
  const T inv_ref = comp ? recip(rho) : T(1);
  T u[3] = {0, 0, 0};
#pragma unroll
  for (int d = 0; d < D; ++d) {
    const T jr = j[d] * inv_ref, sh = T(0.5) * a.force[d] * inv_ref;
    u[d] = (a.flags & LBM355_FLAG_EQ_SHIFT) ? jr + sh : jr;
    u_out_v[d] = (a.flags & LBM355_FLAG_FORCE) ? jr + sh : jr;
  }
  rho_out_v = rho;
  const T u2[3] = {u[0] * u[0], u[1] * u[1], u[2] * u[2]};
  const T um[10] = {T(1), u[0], u[1], u[2], u2[0], u2[1], u2[2], u[0] * u[1], u[0] * u[2], u[1] * u[2]};

  forward_central<S, T>(m, u);

  // dk = (kappa_eq - gamma) - beta - kt: from the table's polynomial in (drho, u), or in closed form
  // (host-checked): for the full continuous Maxwellian kappa_eq = rho gamma, i.e. kappa_eq - gamma = drho gamma;
  // for the incompressible one kappa_eq - gamma = drho rest_gauss(u)
  T dk[27];
#define LBM_DK(EX, EY, EZ)                                                                        \
  {                                                                                               \
    constexpr int e = 9 * EX + 3 * EY + EZ;                                                       \
    T eq = 0;                                                                                     \
    _Pragma("unroll") for (int t = 0; t < 10; ++t) eq += (tab[L::EQ + 20 * e + t] + tab[L::EQ + 20 * e + 10 + t] * drho) * um[t]; \
    dk[e] = eq - beta_e<T, EX, EY, EZ>(u, u2) - m[EX][EY][EZ];                                    \
  }
#define LBM_DKG(EX, EY, EZ)                                                                       \
  {                                                                                               \
    constexpr int e = 9 * EX + 3 * EY + EZ;                                                       \
    constexpr double g = gam_v<EX> * gam_v<EY> * gam_v<EZ>;                                       \
    if constexpr (g != 0.0)                                                                       \
      dk[e] = drho * T(g) - beta_e<T, EX, EY, EZ>(u, u2) - m[EX][EY][EZ];                         \
    else                                                                                          \
      dk[e] = -beta_e<T, EX, EY, EZ>(u, u2) - m[EX][EY][EZ];                                      \
  }
#define LBM_DKR(EX, EY, EZ)                                                                       \
  {                                                                                               \
    constexpr int e = 9 * EX + 3 * EY + EZ;                                                       \
    dk[e] = drho * rest_gauss<T, EX, EY, EZ>(u, u2) - beta_e<T, EX, EY, EZ>(u, u2) - m[EX][EY][EZ]; \
  }
#define LBM_DK_ALL(M)                                                                                  \
  M(0, 0, 0) M(0, 0, 1) M(0, 0, 2) M(0, 1, 0) M(0, 1, 1) M(0, 1, 2) M(0, 2, 0) M(0, 2, 1) M(0, 2, 2)   \
  M(1, 0, 0) M(1, 0, 1) M(1, 0, 2) M(1, 1, 0) M(1, 1, 1) M(1, 1, 2) M(1, 2, 0) M(1, 2, 1) M(1, 2, 2)   \
  M(2, 0, 0) M(2, 0, 1) M(2, 0, 2) M(2, 1, 0) M(2, 1, 1) M(2, 1, 2) M(2, 2, 0) M(2, 2, 1) M(2, 2, 2)
  if (D == 3 && (a.flags & LBM355_FLAG_GAUSSIAN_EQUILIBRIUM)) {
    if (tab[L::GAUSS] == T(2)) {
      LBM_DK_ALL(LBM_DKR)
    } else {
      LBM_DK_ALL(LBM_DKG)
    }
  } else {
    LBM_DK_ALL(LBM_DK)
  }
#undef LBM_DK_ALL
#undef LBM_DKR
#undef LBM_DKG
#undef LBM_DK

  const T r0 = EQUILIBRIUM ? T(1) : tab[L::R0];
  // the host-checked closed form of the rank-one sum: the conserved monomials (000, 100, 010, 001) relax with
  // rate 0 and the deviatoric second-order ones (the shear moments) with r0 + s
  const bool closed = !EQUILIBRIUM && D == 3 && (a.flags & LBM355_FLAG_SHEAR_PROJECTOR);
  const T s_sh = closed ? tab[L::SHEAR] : T(0);
  auto closed_sum = [&](const T (&x)[27], T scale, T (&acc)[27]) {
    // acc += scale * (-r0 x_conserved + s dev2(x))
    acc[0] -= scale * r0 * x[0];
    acc[9] -= scale * r0 * x[9];
    acc[3] -= scale * r0 * x[3];
    acc[1] -= scale * r0 * x[1];
    const T tr = (x[18] + x[6] + x[2]) * T(1.0 / 3.0);
    const T ss = scale * s_sh;
    acc[18] += ss * (x[18] - tr);
    acc[6] += ss * (x[6] - tr);
    acc[2] += ss * (x[2] - tr);
    acc[12] += ss * x[12];
    acc[10] += ss * x[10];
    acc[4] += ss * x[4];
  };
  T dkp[27];
#pragma unroll
  for (int e = 0; e < 27; ++e) dkp[e] = r0 * dk[e];
  if (closed) closed_sum(dk, T(1), dkp);
  for (int k = 0; k < ((EQUILIBRIUM || closed) ? 0 : a.nactive); ++k) {
    const TablePtr<T> prow = tab + L::PROW + 27 * k;
    const TablePtr<T> pinv = tab + L::PINV + 27 * k;
    T sk = 0;
#pragma unroll
    for (int e = 0; e < 27; ++e) sk += prow[e] * dk[e];
    sk *= tab[L::COEF + k];
#pragma unroll
    for (int e = 0; e < 27; ++e) dkp[e] += sk * pinv[e];
  }

  // force vector in population space (affine in u) and its Guo correction
  const int fmode = EQUILIBRIUM ? 0 : (int)tab[L::FMODE];
  T fl[Q];
  if (fmode != 0) {
#pragma unroll
    for (int i = 0; i < Q; ++i) {
      T v = tab[L::L0 + i];
#pragma unroll
      for (int d = 0; d < D; ++d) v += tab[L::LU + 3 * i + d] * u[d];
      fl[i] = v;
    }
    if (fmode == 2 && a.nactive > 0) {
      T kl[3][3][3];
      load_grid<S, T>(fl, kl);
      forward_central<S, T>(kl, u);
      if (closed) {
        T kf[27];
#pragma unroll
        for (int e = 0; e < 27; ++e) kf[e] = kl[e / 9][(e / 3) % 3][e % 3];
        closed_sum(kf, T(-0.5), dkp);
      }
      for (int k = 0; k < (closed ? 0 : a.nactive); ++k) {
        const TablePtr<T> prow = tab + L::PROW + 27 * k;
        const TablePtr<T> pinv = tab + L::PINV + 27 * k;
        T sk = 0;
#pragma unroll
        for (int e = 0; e < 27; ++e) sk += prow[e] * kl[e / 9][(e / 3) % 3][e % 3];
        sk *= T(-0.5) * tab[L::COEF + k];
#pragma unroll
        for (int e = 0; e < 27; ++e) dkp[e] += sk * pinv[e];
      }
    }
  }

#pragma unroll
  for (int e = 0; e < 27; ++e) m[e / 9][(e / 3) % 3][e % 3] = (D == 2 && e % 3 != 0) ? T(0) : dkp[e];
  T df[Q];
  backward_central<S, T>(m, u, df);
  const T fscale = fmode == 2 ? T(1) - T(0.5) * r0 : T(1);
#pragma unroll
  for (int i = 0; i < Q; ++i) f[i] += df[i] + (fmode != 0 ? fscale * fl[i] : T(0));
}

}  // namespace lbm355
