// D3Q27 cumulant collision (lbmpy methods/cumulantbased/cumulantbasedmethod.py:296-402 with the
// default polynomial cumulants of methods/default_moment_sets.py:106-145, CentralMoment forcing
// forcemodels.py:213-232, optional Smagorinsky turbulence_models.py:25-84).
//
// The algorithm is lbmpy's (raw moments -> central moments -> cumulants -> relax -> back), but
// every quantity is carried as its deviation from the rest-state Gaussian so that FP32 never
// holds an O(1) constant next to an O(u^2) signal:
//   f~        stored deviation PDFs (f - w)
//   kt_e      central moments of f~ about u            (binomial shift of raw moments of f~)
//   beta_e    central moments of w about u minus gamma  (closed form in u, no cancellation)
//   mu'_e     = (kt_e + beta_e - drho*gamma_e) / rho    normalised central-moment deviation
//   k'_e      cumulant deviations: k = gamma_k + k', gamma_k = 1/3 on the diagonal
// where gamma_e = prod_a {1, 0, 1/3}[e_a] are the Gaussian (c_s^2 = 1/3) moments.  The
// Gaussian has zero cumulants above second order, so the k'(mu') and mu'(k') relations carry
// no constant terms.  Cumulant equilibria (rho * Gaussian cumulants) are exactly the
// reference, so relaxation is C'* = (1 - r) C' for every non-conserved polynomial C.
// First-order central moments vanish after the "force before" step (kappa_100 = j - rho u + F/2
// = 0), which is used to drop them from the cumulant relations.
#pragma once
#include "lbm_common.h"
#include "closure.h"

namespace lbm355 {

// beta_e = prod_a (gamma_a + delta_a) - prod_a gamma_a, gamma = {1, 0, 1/3}, delta = {0, -u, u^2}
template <int E, class T>
__device__ __forceinline__ T gam() {
  return E == 0 ? T(1) : (E == 1 ? T(0) : T(1.0 / 3.0));
}
template <int E, class T>
__device__ __forceinline__ T del(T u, T u2) {
  return E == 0 ? T(0) : (E == 1 ? -u : u2);
}
template <int EX, int EY, int EZ, class T>
__device__ __forceinline__ T beta(const T (&u)[3], const T (&u2)[3]) {
  // (gx+dx)(gy+dy)(gz+dz) - gx gy gz = dx (gy+dy)(gz+dz) + gx [dy (gz+dz) + gy dz]
  const T dx = del<EX, T>(u[0], u2[0]), dy = del<EY, T>(u[1], u2[1]), dz = del<EZ, T>(u[2], u2[2]);
  const T gx = gam<EX, T>(), gy = gam<EY, T>(), gz = gam<EZ, T>();
  const T pz = EZ == 0 ? T(1) : gz + dz;
  const T py = EY == 0 ? T(1) : gy + dy;
  T yz = T(0);
  if (EY != 0) yz += dy * pz;
  if (EZ != 0) yz += gy * dz;
  T r = T(0);
  if (EX != 0) r += dx * (py * pz);
  if (EX != 1) r += gx * yz;
  return r;
}

// gamma_e of a monomial as a compile-time number
template <int E>
constexpr double gam_v = E == 0 ? 1.0 : (E == 1 ? 0.0 : 1.0 / 3.0);

// Central moment about u of the post-collision deviation from its normalised deviation v: rho v for f* - f_eq
// (NEQ), rho v + drho gamma_e - beta_e for f* - w.  The drho term is left out where gamma_e = 0 (an odd exponent).
template <int EX, int EY, int EZ, bool NEQ, class T>
__device__ __forceinline__ T post_moment(T v, T rho, T drho, const T (&u)[3], const T (&u2)[3]) {
  constexpr double g = gam_v<EX> * gam_v<EY> * gam_v<EZ>;
  if constexpr (NEQ)
    return rho * v;
  else if constexpr (g == 0.0)
    return rho * v - beta<EX, EY, EZ>(u, u2);
  else
    return rho * v + drho * T(g) - beta<EX, EY, EZ>(u, u2);
}
// the same for a deviation that vanishes (v = 0)
template <int EX, int EY, int EZ, bool NEQ, class T>
__device__ __forceinline__ T post_moment_zero(T drho, const T (&u)[3], const T (&u2)[3]) {
  constexpr double g = gam_v<EX> * gam_v<EY> * gam_v<EZ>;
  if constexpr (NEQ)
    return T(0);
  else if constexpr (g == 0.0)
    return -beta<EX, EY, EZ>(u, u2);
  else
    return drho * T(g) - beta<EX, EY, EZ>(u, u2);
}

// raw chimera along one axis: values at c = -1, 0, +1  ->  sums of c^0, c^1, c^2
template <class T>
__device__ __forceinline__ void chim_fwd(T& a, T& b, T& c) {
  const T t = c + a, d = c - a;
  a = t + b;
  b = d;
  c = t;
}
// inverse: moments (m0, m1, m2) -> values at c = 0 (in b), +1 (in c), -1 (in a)
template <class T>
__device__ __forceinline__ void chim_bwd(T& m0, T& m1, T& m2) {
  const T f0 = m0 - m2, fp = T(0.5) * (m2 + m1), fm = T(0.5) * (m2 - m1);
  m0 = fm;
  m1 = f0;
  m2 = fp;
}
// central shift by -u: raw (m0,m1,m2) -> central (k0,k1,k2)
template <class T>
__device__ __forceinline__ void shift_fwd(T& m0, T& m1, T& m2, T u) {
  const T k1 = m1 - u * m0;
  const T k2 = m2 - u * (m1 + k1);
  m1 = k1;
  m2 = k2;
}
// inverse shift by +u: central (k0,k1,k2) -> raw (m0,m1,m2)
template <class T>
__device__ __forceinline__ void shift_bwd(T& k0, T& k1, T& k2, T u) {
  const T m1 = k1 + u * k0;
  const T m2 = k2 + u * (k1 + m1);
  k1 = m1;
  k2 = m2;
}

// cumulant-table layout: 27 per-polynomial relaxation rates in the order of the default
// polynomial cumulant set (density, x, y, z, xy, xz, yz, x2-y2, x2-z2, x2+y2+z2, ...), then the
// fourth-order correction's parameters a, b (fourth_order_correction.py:70-94), its unlimited rates
// omega_3, omega_4, omega_5 (:127-139) and the limiter (zeros without the correction).
struct CumulantTab {
  static constexpr int RATES = 0, EXTRA = 27, BLOCKS = 33, SIZE = 116;
  static constexpr int A = 0, B = 1, W3 = 2, W4 = 3, W5 = 4, LIMITER = 5;
  // the relaxation as linear maps on the monomial cumulant deviations (post = block * pre), row-major, for
  // cumulant sets other than the default polynomial one: second order diagonal (200, 020, 002) and mixed
  // (110, 101, 011), third order (120, 102, 210, 012, 201, 021) and 111, fourth order (220, 202, 022) and
  // (211, 121, 112), fifth order (221, 212, 122), sixth 222
  static constexpr int B2D = BLOCKS, B2O = B2D + 9, B3 = B2O + 9, B111 = B3 + 36, B4D = B111 + 1, B4O = B4D + 9,
                       B5 = B4O + 9, B6 = B5 + 9;
};

// LEARNED: the shear rate comes from the learned closure (a separate kernel variant, so that the plain
// sweeps carry none of its MFMA register state).  CORRECTED: the variant that adds the Galilean
// correction (galilean_correction.py:56-91) and / or the fourth-order correction with its limited
// third-order rates (fourth_order_correction.py:97-237), per LBM355_FLAG_GALILEAN /
// LBM355_FLAG_FOURTH_ORDER.  Both corrections are written for normalised deviations: lbmpy's
// cumulants are rho-scaled (C = rho k, cumulanttransforms.py:181), so a correction term cor added
// to a post-collision polynomial cumulant enters here as cor / rho, and its derivative estimates,
// built from c_ab / rho = gamma_ab + mu'_ab, take the deviations mu'_ab directly.
// NEQ: the transforms act on f - f_eq(rho, u) instead of f~ = f - w.  In FP32 the populations of a flow at
// |u| ~ 0.05 carry O(1e-2) equilibrium parts: transforming them directly leaves rounding errors of ~1e-9 in the
// central moments, which in the deviation mu' = kappa / rho - gamma (an O(1e-6) non-equilibrium signal) become
// relative errors of ~1e-3 every step.  Subtracting the equilibrium first (its central moments are exactly
// rho * gamma, so mu' = kappa(f - f_eq) / rho) keeps every transformed quantity at the non-equilibrium scale;
// f_eq itself is evaluated in deviation form (no O(1) constant) and added back unchanged after the collision.
// Costs about 10% more instructions; the FP32-storage sweeps use it, FP16 storage (whose accuracy its rounding
// bounds) and FP64 arithmetic do not.
// GENERAL: a cumulant set other than lbmpy's default polynomial one; its relaxation comes as the linear maps
// of the table's BLOCKS (P^-1 (I - D) P restricted to each group of monomials), not as per-polynomial rates.
// EQ3: every cumulant of order >= 3 relaxes with rate 1 (lbmpy's default for a method given by its shear rate, and
// LBM355_FLAG_HIGHER_ORDERS_RATE1), so it is set to its equilibrium, zero, without being computed: the transforms
// and cumulant relations of orders >= 3 drop out before the collision, and after it the central moments of orders
// 3 and 5 are those of the equilibrium and those of orders 4 and 6 are products of the second-order cumulants.
template <class T, bool LEARNED = false, bool CORRECTED = false, bool NEQ = false, bool GENERAL = false,
          bool EQ3 = false>
__device__ __forceinline__ void collide_cumulant_d3q27(T (&f)[27], const SweepArgs<T>& a, T& rho_out_v,
                                                       T (&u_out_v)[3], CellExtras<T>& ex) {
  using S = D3Q27;
  const TablePtr<T> __restrict__ r = table_ptr(a.tab) + CumulantTab::RATES;
  const T hF[3] = {T(0.5) * a.force[0], T(0.5) * a.force[1], T(0.5) * a.force[2]};
  T m[3][3][3];  // [x][y][z]: first values at c=-1,0,+1, then exponents 0,1,2
#pragma unroll
  for (int cx = -1; cx <= 1; ++cx)
#pragma unroll
    for (int cy = -1; cy <= 1; ++cy)
#pragma unroll
      for (int cz = -1; cz <= 1; ++cz) m[cx + 1][cy + 1][cz + 1] = f[dir_index<S>(cx, cy, cz)];

  T drho, rho, irho, u[3];
  T feq[3][3][3];   // NEQ: equilibrium deviations f_eq - w, [x][y][z] at c = -1, 0, +1
  if constexpr (NEQ) {
    // density and momentum by partial sums (z, then y, then x)
    T s0[3][3], s1[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        const T t = m[i][j][2] + m[i][j][0];
        s0[i][j] = t + m[i][j][1];
        s1[i][j] = m[i][j][2] - m[i][j][0];
      }
    T q0[3], qy[3], qz[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      const T t = s0[i][2] + s0[i][0];
      q0[i] = t + s0[i][1];
      qy[i] = s0[i][2] - s0[i][0];
      qz[i] = (s1[i][2] + s1[i][0]) + s1[i][1];
    }
    const T t = q0[2] + q0[0];
    drho = t + q0[1];
    const T jx = q0[2] - q0[0], jy = (qy[2] + qy[0]) + qy[1], jz = (qz[2] + qz[0]) + qz[1];
    rho = T(1) + drho;
    irho = recip(rho);
    u[0] = (jx + hF[0]) * irho;
    u[1] = (jy + hF[1]) * irho;
    u[2] = (jz + hF[2]) * irho;
    // f_eq - w = rho prod_a (w1_a + d_a) - prod_a w1_a, with the 1-D weights w1 = (1/6, 2/3, 1/6) and
    // d(0) = -u^2, d(+-1) = (u^2 +- u) / 2; expanded as d_x P_yz + w1_x D_yz + drho P_x P_yz with
    // P = w1 + d, P_yz = P_y P_z and D_yz = P_y P_z - w1_y w1_z = d_y P_z + w1_y d_z (no cancellation)
    T d[3][3], P[3][3];
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
      const T h2 = T(0.5) * (u[ax] * u[ax]);
      d[ax][0] = h2 - T(0.5) * u[ax];
      d[ax][1] = -(h2 + h2);
      d[ax][2] = h2 + T(0.5) * u[ax];
      P[ax][0] = T(1.0 / 6.0) + d[ax][0];
      P[ax][1] = T(2.0 / 3.0) + d[ax][1];
      P[ax][2] = T(1.0 / 6.0) + d[ax][2];
    }
    constexpr T w1[3] = {T(1.0 / 6.0), T(2.0 / 3.0), T(1.0 / 6.0)};
#pragma unroll
    for (int j = 0; j < 3; ++j)
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const T pyz = P[1][j] * P[2][k];
        const T dyz = d[1][j] * P[2][k] + w1[j] * d[2][k];
        const T rpyz = drho * pyz;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
          feq[i][j][k] = d[0][i] * pyz + w1[i] * dyz + P[0][i] * rpyz;
          m[i][j][k] -= feq[i][j][k];
        }
      }
  }

  // raw moments (of f~, or of f - f_eq with NEQ) by the chimera transform (z, y, x passes)
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

  // density and velocity (conservedquantitycomputation.py:218-249, CentralMoment shift)
  if constexpr (!NEQ) {
    drho = m[0][0][0];
    rho = T(1) + drho;
    irho = recip(rho);
    u[0] = (m[1][0][0] + hF[0]) * irho;
    u[1] = (m[0][1][0] + hF[1]) * irho;
    u[2] = (m[0][0][1] + hF[2]) * irho;
  }
  rho_out_v = rho;
  u_out_v[0] = u[0];
  u_out_v[1] = u[1];
  u_out_v[2] = u[2];
  const T u2[3] = {u[0] * u[0], u[1] * u[1], u[2] * u[2]};

  // central moments of f~ about u
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

  // normalised central-moment deviations mu'
#define LBM_MU(EX, EY, EZ)                                                                                   \
  const T mu##EX##EY##EZ =                                                                                   \
      NEQ ? m[EX][EY][EZ] * irho                                                                             \
          : (m[EX][EY][EZ] + beta<EX, EY, EZ>(u, u2) - drho * (gam<EX, T>() * gam<EY, T>() * gam<EZ, T>())) * irho
  LBM_MU(2, 0, 0); LBM_MU(0, 2, 0); LBM_MU(0, 0, 2);
  LBM_MU(1, 1, 0); LBM_MU(1, 0, 1); LBM_MU(0, 1, 1);
  LBM_MU(1, 2, 0); LBM_MU(1, 0, 2); LBM_MU(2, 1, 0); LBM_MU(0, 1, 2); LBM_MU(2, 0, 1); LBM_MU(0, 2, 1);
  LBM_MU(1, 1, 1);
  LBM_MU(2, 2, 0); LBM_MU(2, 0, 2); LBM_MU(0, 2, 2);
  LBM_MU(2, 1, 1); LBM_MU(1, 2, 1); LBM_MU(1, 1, 2);
  LBM_MU(2, 2, 1); LBM_MU(2, 1, 2); LBM_MU(1, 2, 2);
  LBM_MU(2, 2, 2);
#undef LBM_MU
  constexpr T third = T(1.0 / 3.0);
  constexpr T ninth = T(1.0 / 9.0);

  // shear relaxation rate, possibly replaced by the Smagorinsky rate
  T rs[5] = {r[4], r[5], r[6], r[7], r[8]};
  if (LEARNED || (a.flags & LBM355_FLAG_SMAGORINSKY)) {
    // Pi_neq_ab = rho mu'_ab - (u_a F_b + u_b F_a) / 2  (second moment of f - f_eq)
    const T pxx = rho * mu200 - u[0] * a.force[0];
    const T pyy = rho * mu020 - u[1] * a.force[1];
    const T pzz = rho * mu002 - u[2] * a.force[2];
    const T pxy = rho * mu110 - T(0.5) * (u[0] * a.force[1] + u[1] * a.force[0]);
    const T pxz = rho * mu101 - T(0.5) * (u[0] * a.force[2] + u[2] * a.force[0]);
    const T pyz = rho * mu011 - T(0.5) * (u[1] * a.force[2] + u[2] * a.force[1]);
    T w = a.omega_shear;
    if (a.flags & LBM355_FLAG_SMAGORINSKY) {
      const T pi = sqrt(T(2) * (pxx * pxx + pyy * pyy + pzz * pzz + T(2) * (pxy * pxy + pxz * pxz + pyz * pyz))) * irho;
      const T tau0 = T(1) / a.omega_shear;
      w = T(2) / (tau0 + sqrt(a.smag_cs2x18 * pi + tau0 * tau0));
      ex.omega_sgs = w;
      ex.pi = pi;
    }
    if constexpr (LEARNED && sizeof(T) == 4) {
      // wave-collective: every lane of the wavefront reaches this point (the host requires nx % 64 == 0)
      const ClosureParams* P = reinterpret_cast<const ClosureParams*>(a.closure);
      const float p6[6] = {pxx, pyy, pzz, pxy, pxz, pyz};
      float feat[CLOSURE_IN];
      closure_features(p6, rho, feat);
      w = closure_omega(P, mlp_forward_wave(P, feat, nullptr, nullptr), a.omega_shear);
    }
    rs[0] = rs[1] = rs[2] = rs[3] = rs[4] = w;
  }

  // ---- cumulant deviations (second/third order equal the central moments)
  const T k220 = mu220 - (mu020 + mu200) * third - mu020 * mu200 - T(2) * mu110 * mu110;
  const T k202 = mu202 - (mu002 + mu200) * third - mu002 * mu200 - T(2) * mu101 * mu101;
  const T k022 = mu022 - (mu002 + mu020) * third - mu002 * mu020 - T(2) * mu011 * mu011;
  const T k211 = mu211 - mu011 * third - mu200 * mu011 - T(2) * mu101 * mu110;
  const T k121 = mu121 - mu101 * third - mu020 * mu101 - T(2) * mu011 * mu110;
  const T k112 = mu112 - mu110 * third - mu002 * mu110 - T(2) * mu011 * mu101;
  const T k221 = mu221 - (mu021 + mu201) * third - mu020 * mu201 - mu200 * mu021 - T(4) * mu110 * mu111 -
                 T(2) * (mu011 * mu210 + mu101 * mu120);
  const T k212 = mu212 - (mu012 + mu210) * third - mu002 * mu210 - mu200 * mu012 - T(4) * mu101 * mu111 -
                 T(2) * (mu011 * mu201 + mu102 * mu110);
  const T k122 = mu122 - (mu102 + mu120) * third - mu002 * mu120 - mu020 * mu102 - T(4) * mu011 * mu111 -
                 T(2) * (mu012 * mu110 + mu021 * mu101);
  const T k222 = mu222 - (mu202 + mu220 + mu022) * third + (mu200 + mu020 + mu002) * ninth -
                 (mu002 * mu220 + mu020 * mu202 + mu200 * mu022) +
                 T(2.0 / 3.0) * (mu002 * mu020 + mu002 * mu200 + mu020 * mu200) + T(2) * mu002 * mu020 * mu200 +
                 T(4) * (mu002 * mu110 * mu110 + mu200 * mu011 * mu011 + mu020 * mu101 * mu101) +
                 T(4.0 / 3.0) * (mu011 * mu011 + mu101 * mu101 + mu110 * mu110) + T(16) * mu011 * mu101 * mu110 -
                 T(4) * (mu011 * mu211 + mu101 * mu121 + mu110 * mu112) -
                 T(2) * (mu012 * mu210 + mu021 * mu201 + mu102 * mu120) - T(4) * mu111 * mu111;

  // third-order rates; the fourth-order correction limits them per cell with the x components of the
  // pre-collision third-order cumulants (fourth_order_correction.py:191-198, one value per group):
  // w + (1 - w) |C| / (rho lim + |C|), and with C = rho mu' the density cancels
  T w3[3] = {r[10], r[11], r[12]}, w4[3] = {r[13], r[14], r[15]};
  T w5 = r[16];
  if constexpr (CORRECTED) {
    if (a.flags & LBM355_FLAG_FOURTH_ORDER) {
      const TablePtr<T> __restrict__ x = table_ptr(a.tab) + CumulantTab::EXTRA;
      const T lim = x[CumulantTab::LIMITER];
      const T sp = fabs(mu120 + mu102), sm = fabs(mu120 - mu102), s1 = fabs(mu111);
      const T l3 = x[CumulantTab::W3] + (T(1) - x[CumulantTab::W3]) * sp / (lim + sp);
      const T l4 = x[CumulantTab::W4] + (T(1) - x[CumulantTab::W4]) * sm / (lim + sm);
      w3[0] = w3[1] = w3[2] = l3;
      w4[0] = w4[1] = w4[2] = l4;
      w5 = x[CumulantTab::W5] + (T(1) - x[CumulantTab::W5]) * s1 / (lim + s1);
    }
  }

  // ---- relaxation (equilibrium deviation is zero)
  T c200, c020, c002, c110, c101, c011, c120, c102, c210, c012, c201, c021, c111, c220, c202, c022, c211, c121,
      c112, c221, c212, c122, c222;
  if constexpr (GENERAL) {
    const TablePtr<T> __restrict__ b = table_ptr(a.tab);
    auto mv3 = [&](int o, T x0, T x1, T x2, T& y0, T& y1, T& y2) {
      y0 = b[o + 0] * x0 + b[o + 1] * x1 + b[o + 2] * x2;
      y1 = b[o + 3] * x0 + b[o + 4] * x1 + b[o + 5] * x2;
      y2 = b[o + 6] * x0 + b[o + 7] * x1 + b[o + 8] * x2;
    };
    mv3(CumulantTab::B2D, mu200, mu020, mu002, c200, c020, c002);
    mv3(CumulantTab::B2O, mu110, mu101, mu011, c110, c101, c011);
    const T x3[6] = {mu120, mu102, mu210, mu012, mu201, mu021};
    T y3[6];
#pragma unroll
    for (int i = 0; i < 6; ++i) {
      T acc = T(0);
#pragma unroll
      for (int j = 0; j < 6; ++j) acc += b[CumulantTab::B3 + 6 * i + j] * x3[j];
      y3[i] = acc;
    }
    c120 = y3[0];
    c102 = y3[1];
    c210 = y3[2];
    c012 = y3[3];
    c201 = y3[4];
    c021 = y3[5];
    c111 = b[CumulantTab::B111] * mu111;
    mv3(CumulantTab::B4D, k220, k202, k022, c220, c202, c022);
    mv3(CumulantTab::B4O, k211, k121, k112, c211, c121, c112);
    mv3(CumulantTab::B5, k221, k212, k122, c221, c212, c122);
    c222 = b[CumulantTab::B6] * k222;
  } else if constexpr (EQ3) {
    c110 = (T(1) - rs[0]) * mu110;
    c101 = (T(1) - rs[1]) * mu101;
    c011 = (T(1) - rs[2]) * mu011;
    const T d1 = (T(1) - rs[3]) * (mu200 - mu020);
    const T d2 = (T(1) - rs[4]) * (mu200 - mu002);
    const T tr = (T(1) - r[9]) * (mu200 + mu020 + mu002);
    c200 = (d1 + d2 + tr) * third;
    c020 = (tr - T(2) * d1 + d2) * third;
    c002 = (tr + d1 - T(2) * d2) * third;
    c120 = c102 = c210 = c012 = c201 = c021 = c111 = T(0);
    c220 = c202 = c022 = c211 = c121 = c112 = c221 = c212 = c122 = c222 = T(0);
  } else {
    c110 = (T(1) - rs[0]) * mu110;
    c101 = (T(1) - rs[1]) * mu101;
    c011 = (T(1) - rs[2]) * mu011;
    T d1 = (T(1) - rs[3]) * (mu200 - mu020);
    T d2 = (T(1) - rs[4]) * (mu200 - mu002);
    T tr = (T(1) - r[9]) * (mu200 + mu020 + mu002);
    T p1 = (T(1) - r[17]) * (k220 - T(2) * k202 + k022);
    T p2 = (T(1) - r[18]) * (k220 + k202 - T(2) * k022);
    T s4 = (T(1) - r[19]) * (k220 + k202 + k022);
    c211 = (T(1) - r[20]) * k211;
    c121 = (T(1) - r[21]) * k121;
    c112 = (T(1) - r[22]) * k112;
    if constexpr (CORRECTED) {
      // velocity-gradient estimates from the pre-collision second-order cumulants (shear rate w1 of
      // x2-y2, the Smagorinsky rate when that model is on; bulk rate w2)
      const T w1 = rs[3], w2 = r[9];
      const T dxu = -T(0.5) * w1 * (T(2) * mu200 - mu020 - mu002) - T(0.5) * w2 * (mu200 + mu020 + mu002);
      const T dyv = dxu + T(1.5) * w1 * (mu200 - mu020);
      const T dzw = dxu + T(1.5) * w1 * (mu200 - mu002);
      if (a.flags & LBM355_FLAG_GALILEAN) {
        // cor = -3 rho (1 - w/2) (u_a^2 D_a - u_b^2 D_b) on x2-y2, x2-z2; (1 - w2/2) and the sum on x2+y2+z2
        const T g1 = T(-3) * (T(1) - T(0.5) * w1), g2 = T(-3) * (T(1) - T(0.5) * w2);
        const T ex = u2[0] * dxu, ey = u2[1] * dyv, ez = u2[2] * dzw;
        d1 += g1 * (ex - ey);
        d2 += g1 * (ex - ez);
        tr += g2 * (ex + ey + ez);
      }
      if (a.flags & LBM355_FLAG_FOURTH_ORDER) {
        const TablePtr<T> __restrict__ x = table_ptr(a.tab) + CumulantTab::EXTRA;
        const T pre = T(1) / w1 - T(0.5);
        const T pa = pre * x[CumulantTab::A], pb = pre * x[CumulantTab::B];
        p1 += T(2.0 / 3.0) * pa * r[17] * (dxu - T(2) * dyv + dzw);
        p2 += T(2.0 / 3.0) * pa * r[17] * (dxu + dyv - T(2) * dzw);
        s4 -= T(4.0 / 3.0) * pa * r[19] * (dxu + dyv + dzw);
        // -1/3 pre w8 b (D_bc = -3 w1 mu'_bc)
        c211 += pb * r[20] * w1 * mu011;
        c121 += pb * r[20] * w1 * mu101;
        c112 += pb * r[20] * w1 * mu110;
      }
    }
    c200 = (d1 + d2 + tr) * third;
    c020 = (tr - T(2) * d1 + d2) * third;
    c002 = (tr + d1 - T(2) * d2) * third;
    const T ax = (T(1) - w3[0]) * (mu120 + mu102), bx = (T(1) - w4[0]) * (mu120 - mu102);
    const T ay = (T(1) - w3[1]) * (mu210 + mu012), by = (T(1) - w4[1]) * (mu210 - mu012);
    const T az = (T(1) - w3[2]) * (mu201 + mu021), bz = (T(1) - w4[2]) * (mu201 - mu021);
    c120 = T(0.5) * (ax + bx);
    c102 = T(0.5) * (ax - bx);
    c210 = T(0.5) * (ay + by);
    c012 = T(0.5) * (ay - by);
    c201 = T(0.5) * (az + bz);
    c021 = T(0.5) * (az - bz);
    c111 = (T(1) - w5) * mu111;
    c220 = (s4 + p1 + p2) * third;
    c202 = (s4 - p1) * third;
    c022 = (s4 - p2) * third;
    c221 = (T(1) - r[23]) * k221;
    c212 = (T(1) - r[24]) * k212;
    c122 = (T(1) - r[25]) * k122;
    c222 = (T(1) - r[26]) * k222;
  }

  // ---- central moments of the post-collision deviation f~* about u
  // NEQ: of the post-collision f* - f_eq.  Its zeroth and first moments are kept from before the collision, the
  // first ones plus the force F: exactly they are 0 and -F/2 -> +F/2, but in FP32 they also hold what the
  // rounded f_eq(rho, u) misses of the cell's mass and momentum (u = (j + F/2) / rho is rounded), so keeping them
  // conserves both to the rounding of the stored populations.  Rebuilding them from (rho, u) would instead add
  // that rounding error, which repeats from step to step in a slowly varying flow, to the momentum every step.
  if constexpr (NEQ) {
    m[1][0][0] += a.force[0];
    m[0][1][0] += a.force[1];
    m[0][0][1] += a.force[2];
  } else {
    m[0][0][0] = drho;
    m[1][0][0] = hF[0] + u[0];
    m[0][1][0] = hF[1] + u[1];
    m[0][0][1] = hF[2] + u[2];
  }
#define LBM_KT(EX, EY, EZ, V) m[EX][EY][EZ] = post_moment<EX, EY, EZ, NEQ>(V, rho, drho, u, u2)
#define LBM_KT0(EX, EY, EZ) m[EX][EY][EZ] = post_moment_zero<EX, EY, EZ, NEQ>(drho, u, u2)
  LBM_KT(2, 0, 0, c200); LBM_KT(0, 2, 0, c020); LBM_KT(0, 0, 2, c002);
  LBM_KT(1, 1, 0, c110); LBM_KT(1, 0, 1, c101); LBM_KT(0, 1, 1, c011);
  // normalised central-moment deviations of orders 4 and 6 from the cumulants (with EQ3 the cumulants of
  // orders >= 3 are zero and only the products of the second-order ones remain)
  if constexpr (EQ3) {
    const T n220 = (c020 + c200) * third + c020 * c200 + T(2) * c110 * c110;
    const T n202 = (c002 + c200) * third + c002 * c200 + T(2) * c101 * c101;
    const T n022 = (c002 + c020) * third + c002 * c020 + T(2) * c011 * c011;
    const T n211 = c011 * third + c200 * c011 + T(2) * c101 * c110;
    const T n121 = c101 * third + c020 * c101 + T(2) * c011 * c110;
    const T n112 = c110 * third + c002 * c110 + T(2) * c011 * c101;
    const T n222 = (c200 + c020 + c002) * ninth + (c002 * c020 + c002 * c200 + c020 * c200) * third +
                   c002 * c020 * c200 + T(2) * (c002 * c110 * c110 + c200 * c011 * c011 + c020 * c101 * c101) +
                   T(2.0 / 3.0) * (c011 * c011 + c101 * c101 + c110 * c110) + T(8) * c011 * c101 * c110;
    LBM_KT0(1, 2, 0); LBM_KT0(1, 0, 2); LBM_KT0(2, 1, 0); LBM_KT0(0, 1, 2); LBM_KT0(2, 0, 1); LBM_KT0(0, 2, 1);
    LBM_KT0(1, 1, 1);
    LBM_KT(2, 2, 0, n220); LBM_KT(2, 0, 2, n202); LBM_KT(0, 2, 2, n022);
    LBM_KT(2, 1, 1, n211); LBM_KT(1, 2, 1, n121); LBM_KT(1, 1, 2, n112);
    LBM_KT0(2, 2, 1); LBM_KT0(2, 1, 2); LBM_KT0(1, 2, 2);
    LBM_KT(2, 2, 2, n222);
  } else {
    const T n220 = c220 + (c020 + c200) * third + c020 * c200 + T(2) * c110 * c110;
    const T n202 = c202 + (c002 + c200) * third + c002 * c200 + T(2) * c101 * c101;
    const T n022 = c022 + (c002 + c020) * third + c002 * c020 + T(2) * c011 * c011;
    const T n211 = c211 + c011 * third + c200 * c011 + T(2) * c101 * c110;
    const T n121 = c121 + c101 * third + c020 * c101 + T(2) * c011 * c110;
    const T n112 = c112 + c110 * third + c002 * c110 + T(2) * c011 * c101;
    const T n221 = c221 + (c021 + c201) * third + c020 * c201 + c200 * c021 + T(4) * c110 * c111 +
                   T(2) * (c011 * c210 + c101 * c120);
    const T n212 = c212 + (c012 + c210) * third + c002 * c210 + c200 * c012 + T(4) * c101 * c111 +
                   T(2) * (c011 * c201 + c102 * c110);
    const T n122 = c122 + (c102 + c120) * third + c002 * c120 + c020 * c102 + T(4) * c011 * c111 +
                   T(2) * (c012 * c110 + c021 * c101);
    const T n222 = c222 + (c202 + c220 + c022) * third + (c200 + c020 + c002) * ninth +
                   (c002 * c220 + c020 * c202 + c200 * c022) +
                   (c002 * c020 + c002 * c200 + c020 * c200) * third + c002 * c020 * c200 +
                   T(2) * (c002 * c110 * c110 + c200 * c011 * c011 + c020 * c101 * c101) +
                   T(2.0 / 3.0) * (c011 * c011 + c101 * c101 + c110 * c110) + T(8) * c011 * c101 * c110 +
                   T(4) * (c011 * c211 + c101 * c121 + c110 * c112) +
                   T(2) * (c012 * c210 + c021 * c201 + c102 * c120) + T(4) * c111 * c111;
    LBM_KT(1, 2, 0, c120); LBM_KT(1, 0, 2, c102); LBM_KT(2, 1, 0, c210);
    LBM_KT(0, 1, 2, c012); LBM_KT(2, 0, 1, c201); LBM_KT(0, 2, 1, c021);
    LBM_KT(1, 1, 1, c111);
    LBM_KT(2, 2, 0, n220); LBM_KT(2, 0, 2, n202); LBM_KT(0, 2, 2, n022);
    LBM_KT(2, 1, 1, n211); LBM_KT(1, 2, 1, n121); LBM_KT(1, 1, 2, n112);
    LBM_KT(2, 2, 1, n221); LBM_KT(2, 1, 2, n212); LBM_KT(1, 2, 2, n122);
    LBM_KT(2, 2, 2, n222);
  }
#undef LBM_KT0
#undef LBM_KT

  // raw moments (shift by +u), then populations (inverse chimera)
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
  for (int cx = -1; cx <= 1; ++cx)
#pragma unroll
    for (int cy = -1; cy <= 1; ++cy)
#pragma unroll
      for (int cz = -1; cz <= 1; ++cz) {
        if constexpr (NEQ)
          f[dir_index<S>(cx, cy, cz)] = feq[cx + 1][cy + 1][cz + 1] + m[cx + 1][cy + 1][cz + 1];
        else
          f[dir_index<S>(cx, cy, cz)] = m[cx + 1][cy + 1][cz + 1];
      }
}

}  // namespace lbm355
