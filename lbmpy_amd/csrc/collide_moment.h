// Moment-family collision (SRT, TRT, MRT in population or raw-moment space).
//
// lbmpy's moment-based collision (methods/momentbased/momentbasedmethod.py:320-416) is
//     f* = f + M^-1 D (m_eq - M f) + force,
// with M the polynomial moment matrix and D = diag(relaxation rates).  When every
// moment polynomial has a definite parity (true for all default moment sets), the
// even and odd moments span the symmetric and antisymmetric population subspaces, so
//     M^-1 D M = r_e P_sym + r_o P_anti + sum_k (r_k - r_par(k)) M^-1 e_k e_k^T M,
// where r_e / r_o are "base" rates and the sum runs over the few moments whose rate
// differs from their parity's base rate ("active" moments, e.g. the 5 shear moments
// of an MRT).  Conserved moments never need the sum: their m_eq - m is a known
// constant (0 or F/2), which the host folds into the constant force column G0.
// So the kernel needs only
//     f*_i = f_i + r_e D+_i + r_o D-_i + sum_k s_k Minv[k][i] + G0_i + G_i . u,
//     D = f_eq - f,  s_k = coef_k * sum_i Mrow[k][i] D_i,
// SRT (r_e = r_o, no active moments) and TRT (r_e != r_o) are the same code.
//
// f_eq is the method's equilibrium in storage form (zero-centred or absolute):
//     f_eq_i = A_i + B_i drho + lam * (sum_a L_ia u_a + sum_a Qd_ia u_a^2 + sum_{a<b} Qo_iab u_a u_b)
// with lam = rho (compressible) or 1; the host checks that only the terms this
// code evaluates (L_ia with c_ia != 0, Qo_iab with c_ia c_ib != 0) are non-zero.
#pragma once
#include "lbm_common.h"

namespace lbm355 {

// table layout (compute-type elements); Q = stencil size
template <int Q>
struct MomentTab {
  static constexpr int A = 0, B = Q, L = 2 * Q, QD = 5 * Q, QO = 8 * Q, G0 = 11 * Q, G = 12 * Q, RATES = 15 * Q,
                       COEF = 15 * Q + 2, MROW = 16 * Q + 2, MINV = 16 * Q + 2 + Q * Q,
                       SHEAR_ON = 16 * Q + 2 + 2 * Q * Q, SHEAR = SHEAR_ON + 1, SIZE = SHEAR + Q;
};

// o: the cell's offset in the optional per-cell fields (density / velocity inputs).
template <class S, class T>
__device__ __forceinline__ void collide_moment(T (&f)[S::Q], const SweepArgs<T>& a, int64_t o, T& rho_out_v,
                                               T (&u_out_v)[3]) {
  constexpr int Q = S::Q;
  constexpr int D = S::D;
  using L = MomentTab<Q>;
  const TablePtr<T> __restrict__ tab = table_ptr(a.tab);

  T s = 0, j[3] = {0, 0, 0};
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    s += f[i];
#pragma unroll
    for (int d = 0; d < D; ++d)
      if (S::c[i][d] != 0) j[d] += (S::c[i][d] > 0) ? f[i] : -f[i];
  }
  const bool zc = a.flags & LBM355_FLAG_ZERO_CENTERED;
  const bool comp = a.flags & LBM355_FLAG_COMPRESSIBLE;
  T rho = zc ? T(1) + s : s;
  T drho = zc ? s : s - T(1);
  const T inv_ref = comp ? recip(rho) : T(1);
  T u[3] = {0, 0, 0};
#pragma unroll
  for (int d = 0; d < D; ++d) {
    const T jr = j[d] * inv_ref;
    const T shift = T(0.5) * a.force[d] * inv_ref;
    u[d] = (a.flags & LBM355_FLAG_EQ_SHIFT) ? jr + shift : jr;
    u_out_v[d] = (a.flags & LBM355_FLAG_FORCE) ? jr + shift : jr;
  }
  rho_out_v = rho;
  // density / velocity inputs (creationfunctions.py:662-677): the equilibrium and the force term take the given
  // values in place of the computed ones (the outputs above stay the PDFs' own, as lbmpy's output equations are)
  if (a.u_in) {
    if (a.rho_in) {
      rho = a.rho_in[o];
      drho = rho - T(1);
    }
#pragma unroll
    for (int d = 0; d < D; ++d) u[d] = a.u_in[o + d * a.osc];
  }
  const T lam = comp ? rho : T(1);

  // f_eq - f
  T dl[Q];
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    T p = 0;
#pragma unroll
    for (int d = 0; d < D; ++d) {
      if (S::c[i][d] != 0) p += tab[L::L + 3 * i + d] * u[d];
      p += tab[L::QD + 3 * i + d] * (u[d] * u[d]);
    }
    if (D == 3) {
      if (S::c[i][0] != 0 && S::c[i][1] != 0) p += tab[L::QO + 3 * i + 0] * (u[0] * u[1]);
      if (S::c[i][0] != 0 && S::c[i][2] != 0) p += tab[L::QO + 3 * i + 1] * (u[0] * u[2]);
      if (S::c[i][1] != 0 && S::c[i][2] != 0) p += tab[L::QO + 3 * i + 2] * (u[1] * u[2]);
    } else {
      if (S::c[i][0] != 0 && S::c[i][1] != 0) p += tab[L::QO + 3 * i + 0] * (u[0] * u[1]);
    }
    const T feq = tab[L::A + i] + tab[L::B + i] * drho + lam * p;
    dl[i] = feq - f[i];
  }

  // base rates on the symmetric / antisymmetric parts, plus affine force term:
  //   r_e (dl_i + dl_k) / 2 + r_o (dl_i - dl_k) / 2 = s dl_i + t dl_k,  s = (r_e + r_o) / 2, t = (r_e - r_o) / 2
  const T re = tab[L::RATES + 0], ro = tab[L::RATES + 1];
  const T sr = T(0.5) * (re + ro), tr = T(0.5) * (re - ro);
  T post[Q];
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    const int k = S::inv[i];
    T g = tab[L::G0 + i];
#pragma unroll
    for (int d = 0; d < D; ++d) g += tab[L::G + 3 * i + d] * u[d];
    if (k == i) {
      post[i] = f[i] + re * dl[i] + g;
    } else {
      post[i] = f[i] + sr * dl[i] + tr * dl[k] + g;
    }
  }
  // individually relaxed moments
  if constexpr (D == 3) {
    if (a.flags & LBM355_FLAG_SHEAR_PROJECTOR) {
      // they are exactly the shear moments, with one rate: their projector in closed form,
      //   post_i += s (9/2) w_i (c_i c_i) : Pi_dev,  Pi = sum_j c_j c_j (f_eq - f)_j,
      // with s (9/2) w_i from the table (SHEAR); the host checked the identity against the rank-one sum
      T pxx = 0, pyy = 0, pzz = 0, pxy = 0, pxz = 0, pyz = 0;
#pragma unroll
      for (int i = 0; i < Q; ++i) {
        const int cx = S::c[i][0], cy = S::c[i][1], cz = S::c[i][2];
        if (cx != 0) pxx += dl[i];
        if (cy != 0) pyy += dl[i];
        if (cz != 0) pzz += dl[i];
        if (cx * cy > 0) pxy += dl[i];
        if (cx * cy < 0) pxy -= dl[i];
        if (cx * cz > 0) pxz += dl[i];
        if (cx * cz < 0) pxz -= dl[i];
        if (cy * cz > 0) pyz += dl[i];
        if (cy * cz < 0) pyz -= dl[i];
      }
      const T t3 = (pxx + pyy + pzz) * T(1.0 / 3.0);
      const T dxx = pxx - t3, dyy = pyy - t3, dzz = pzz - t3;
      const T xy2 = pxy + pxy, xz2 = pxz + pxz, yz2 = pyz + pyz;
      const TablePtr<T> __restrict__ sh = tab + L::SHEAR;
#pragma unroll
      for (int i = 0; i < Q; ++i) {
        const int cx = S::c[i][0], cy = S::c[i][1], cz = S::c[i][2];
        if (cx == 0 && cy == 0 && cz == 0) continue;
        T q = 0;
        if (cx != 0) q += dxx;
        if (cy != 0) q += dyy;
        if (cz != 0) q += dzz;
        if (cx * cy > 0) q += xy2;
        if (cx * cy < 0) q -= xy2;
        if (cx * cz > 0) q += xz2;
        if (cx * cz < 0) q -= xz2;
        if (cy * cz > 0) q += yz2;
        if (cy * cz < 0) q -= yz2;
        post[i] += sh[i] * q;
      }
#pragma unroll
      for (int i = 0; i < Q; ++i) f[i] = post[i];
      return;
    }
  }
  // rank-one projector corrections
  for (int k = 0; k < a.nactive; ++k) {
    const TablePtr<T> mrow = tab + L::MROW + k * Q;
    const TablePtr<T> minv = tab + L::MINV + k * Q;
    T m = 0;
#pragma unroll
    for (int i = 0; i < Q; ++i) m += mrow[i] * dl[i];
    m *= tab[L::COEF + k];
#pragma unroll
    for (int i = 0; i < Q; ++i) post[i] += m * minv[i];
  }
#pragma unroll
  for (int i = 0; i < Q; ++i) f[i] = post[i];
}

}  // namespace lbm355
