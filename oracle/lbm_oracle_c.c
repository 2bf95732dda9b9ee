/* CPU FP64 oracle for the fused stream-collide sweep, plain C + OpenMP -- TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg may load this library; the
 * product path never does.  It is the C restatement of oracle/lbm_oracle.py for fully periodic
 * domains without ghost layers, used where the numpy oracle is too slow (64^3 runs of thousands of
 * steps, the timed CPU baseline).  The collision is data-driven: the host packs the method's
 * matrices and term lists (oracle/c_oracle.py), this file runs them per cell, in the order lbmpy
 * derives the sweep:
 *
 *  - density / velocity: methods/conservedquantitycomputation.py:218-249;
 *  - moment methods, population space:  f* = f + M^-1 D (m_eq - M f) + force
 *    (methods/momentbased/momentbasedmethod.py:389-412); raw-moment space:  m* = m + D (m_eq - m) + force_m,
 *    f* = M^-1 m*  (momentbasedmethod.py:359-387);
 *  - cumulant methods: monomial moments -> binomial shift by -u -> force half step -> cumulants
 *    by the Leibniz recursion of K' = C' K -> polynomial cumulants C = P c -> relax -> c* = P^-1 C* ->
 *    central moments -> force half step -> shift by +u -> populations
 *    (methods/cumulantbased/cumulantbasedmethod.py:296-402, moment_transforms/cumulanttransforms.py:46-220,
 *    moment_transforms/centralmomenttransforms.py:168-483, forcemodels.py:213-232);
 *  - force models Simple / Luo / Guo (forcemodels.py:179-316);
 *  - streaming: pull (fieldaccess.py:66-75) and AA even/odd (fieldaccess.py:136-157) with periodic
 *    wrap in the index arithmetic (the PeriodicTwoFieldsAccessor semantics, fieldaccess.py:90-133).
 *
 * Layout: pdf[q][z][y][x] (lbmpy's fzyx), no ghost layers; 2-D domains use nz = 1.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define QMAX 27
#define TMAX 64 /* equilibrium polynomial terms per moment */

enum { KIND_MOMENT_POP = 0, KIND_MOMENT_RAW = 1, KIND_CUMULANT = 2 };
enum { FM_NONE = 0, FM_SIMPLE = 1, FM_LUO = 2, FM_GUO = 3, FM_CENTRAL = 4 };

typedef struct oracle_method {
  int32_t kind, Q, D, force_model;
  int32_t zero_centered, compressible, eq_shift, guo_scalar; /* guo_scalar: SRT Guo = (1 - w/2) * Luo */
  int32_t c[QMAX][3];
  double w[QMAX];
  double F[3];
  double rates[QMAX];
  double M[QMAX][QMAX], Minv[QMAX][QMAX]; /* moment (poly, moment family) or monomial (cumulant) matrix */
  double bg[QMAX];                          /* background moments added for zero-centred storage */
  int32_t has_bg;
  /* equilibrium polynomials per moment (moment family: m_eq; cumulant: C_eq), terms over
     (rho, drho, u0, u1, u2); n_terms[i] < 0 marks "no equilibrium" (zeroth cumulant) */
  int32_t n_terms[QMAX];
  double coef[QMAX][TMAX];
  int8_t pow_[QMAX][TMAX][5];
  /* cumulant family */
  double P[QMAX][QMAX], Pinv[QMAX][QMAX];
  int32_t mono[QMAX][3];
  int32_t idx000, first_order[3];
  int32_t n_shift;                 /* kappa_e = sum over (i, r, coef, pw) of coef m_r s^pw */
  int32_t shift_i[256], shift_r[256], shift_pw[256][3];
  double shift_c[256];
  int32_t order[QMAX];             /* monomials sorted by total degree (0 first) */
  int32_t leib_start[QMAX + 1];    /* Leibniz terms of monomial k: [leib_start[k], leib_start[k+1]) */
  int32_t leib_rj[512], leib_rest[512];
  double leib_b[512];
} oracle_method;

static double eval_poly(const oracle_method* m, int i, double rho, double drho, const double* u) {
  const double v[5] = {rho, drho, u[0], u[1], u[2]};
  double out = 0.0;
  for (int t = 0; t < m->n_terms[i]; ++t) {
    double term = m->coef[i][t];
    for (int k = 0; k < 5; ++k)
      for (int p = 0; p < m->pow_[i][t][k]; ++p) term *= v[k];
    out += term;
  }
  return out;
}

/* Row-compressed copy of a QxQ matrix without its exact zeros.  Dropping a term 0 * x from a sum
   leaves the sum unchanged (finite x), so products equal the dense ones bit for bit. */
typedef struct csr {
  int start[QMAX + 1];
  int col[QMAX * QMAX];
  double val[QMAX * QMAX];
} csr;

typedef struct prepared {
  const struct oracle_method* m;
  csr M, Minv, P, Pinv;
} prepared;

static void compress(int Q, const double A[QMAX][QMAX], csr* out) {
  int k = 0;
  for (int i = 0; i < Q; ++i) {
    out->start[i] = k;
    for (int j = 0; j < Q; ++j)
      if (A[i][j] != 0.0) {
        out->col[k] = j;
        out->val[k] = A[i][j];
        ++k;
      }
  }
  out->start[Q] = k;
}

static void matvec(int Q, const csr* A, const double* x, double* y) {
  for (int i = 0; i < Q; ++i) {
    double s = 0.0;
    for (int k = A->start[i]; k < A->start[i + 1]; ++k) s += A->val[k] * x[A->col[k]];
    y[i] = s;
  }
}

/* Luo population force term w_i [F.(c_i - u)/cs2 + (c_i.F)(c_i.u)/cs2^2] (forcemodels.py:235-256) */
static void luo(const oracle_method* m, const double* u, double* fv) {
  const double cs2 = 1.0 / 3.0;
  const double Fu = m->F[0] * u[0] + m->F[1] * u[1] + m->F[2] * u[2];
  for (int i = 0; i < m->Q; ++i) {
    const double cF = m->c[i][0] * m->F[0] + m->c[i][1] * m->F[1] + m->c[i][2] * m->F[2];
    const double cu = m->c[i][0] * u[0] + m->c[i][1] * u[1] + m->c[i][2] * u[2];
    fv[i] = m->w[i] * ((cF - Fu) / cs2 + cF * cu / (cs2 * cs2));
  }
}

static void force_vector(const oracle_method* m, const double* u, double* fv) {
  if (m->force_model == FM_SIMPLE) {
    for (int i = 0; i < m->Q; ++i)
      fv[i] = m->w[i] * 3.0 * (m->c[i][0] * m->F[0] + m->c[i][1] * m->F[1] + m->c[i][2] * m->F[2]);
  } else {
    luo(m, u, fv);
  }
}

/* density and (equilibrium) velocity of one cell (conservedquantitycomputation.py:218-249) */
static void macroscopic(const oracle_method* m, const double* f, double* rho, double* drho, double* u) {
  double s = 0.0, j[3] = {0, 0, 0};
  for (int i = 0; i < m->Q; ++i) {
    s += f[i];
    for (int a = 0; a < 3; ++a) j[a] += m->c[i][a] * f[i];
  }
  *rho = m->zero_centered ? 1.0 + s : s;
  *drho = m->zero_centered ? s : s - 1.0;
  const double ref = m->compressible ? *rho : 1.0;
  for (int a = 0; a < 3; ++a) {
    u[a] = j[a] / ref;
    if (m->force_model != FM_NONE && m->eq_shift) u[a] += m->F[a] / (2.0 * ref);
  }
}

static void collide_moment(const prepared* pr, double* f) {
  const oracle_method* m = pr->m;
  const int Q = m->Q;
  double rho, drho, u[3], mom[QMAX], meq[QMAX], fv[QMAX], t[QMAX];
  macroscopic(m, f, &rho, &drho, u);
  for (int i = 0; i < Q; ++i) meq[i] = eval_poly(m, i, rho, drho, u);
  matvec(Q, &pr->M, f, mom);
  if (m->kind == KIND_MOMENT_RAW) {
    if (m->has_bg)
      for (int i = 0; i < Q; ++i) mom[i] += m->bg[i];
    for (int i = 0; i < Q; ++i) mom[i] += m->rates[i] * (meq[i] - mom[i]);
    if (m->force_model != FM_NONE) {
      force_vector(m, u, fv);
      matvec(Q, &pr->M, fv, t);
      for (int i = 0; i < Q; ++i) mom[i] += (m->force_model == FM_GUO ? 1.0 - 0.5 * m->rates[i] : 1.0) * t[i];
    }
    if (m->has_bg)
      for (int i = 0; i < Q; ++i) mom[i] -= m->bg[i];
    matvec(Q, &pr->Minv, mom, f);
    return;
  }
  for (int i = 0; i < Q; ++i) t[i] = m->rates[i] * (meq[i] - mom[i]);
  double df[QMAX];
  matvec(Q, &pr->Minv, t, df);
  for (int i = 0; i < Q; ++i) f[i] += df[i];
  if (m->force_model != FM_NONE) {
    force_vector(m, u, fv);
    if (m->force_model == FM_GUO) {
      if (m->guo_scalar) {
        for (int i = 0; i < Q; ++i) fv[i] *= 1.0 - m->rates[0] / 2.0;
      } else {
        double mf[QMAX];
        matvec(Q, &pr->M, fv, mf);
        for (int i = 0; i < Q; ++i) mf[i] *= 1.0 - 0.5 * m->rates[i];
        matvec(Q, &pr->Minv, mf, fv);
      }
    }
    for (int i = 0; i < Q; ++i) f[i] += fv[i];
  }
}

static void shift(const oracle_method* m, const double* in, const double* s, double* out) {
  for (int i = 0; i < m->Q; ++i) out[i] = 0.0;
  for (int k = 0; k < m->n_shift; ++k) {
    double t = m->shift_c[k] * in[m->shift_r[k]];
    for (int a = 0; a < 3; ++a)
      for (int p = 0; p < m->shift_pw[k][a]; ++p) t *= s[a];
    out[m->shift_i[k]] += t;
  }
}

static void collide_cumulant(const prepared* pr, double* f) {
  const oracle_method* m = pr->m;
  const int Q = m->Q;
  double rho, drho, u[3], mom[QMAX], kap[QMAX], L[QMAX], c[QMAX], C[QMAX], cp[QMAX], K[QMAX], t[QMAX];
  macroscopic(m, f, &rho, &drho, u);
  matvec(Q, &pr->M, f, mom);
  if (m->has_bg)
    for (int i = 0; i < Q; ++i) mom[i] += m->bg[i];
  const double mu[3] = {-u[0], -u[1], -u[2]};
  shift(m, mom, mu, kap);
  if (m->force_model != FM_NONE)
    for (int a = 0; a < m->D; ++a) kap[m->first_order[a]] += 0.5 * m->F[a];
  /* cumulants c_e = rho d^e [Xi.u + log K(Xi)] at 0 */
  const double k0 = kap[m->idx000];
  for (int o = 0; o < Q; ++o) {
    const int e = m->order[o];
    if (e == m->idx000) continue;
    double acc = kap[e];
    for (int k = m->leib_start[e]; k < m->leib_start[e + 1]; ++k)
      if (m->leib_rj[k] != e) acc -= m->leib_b[k] * L[m->leib_rj[k]] * kap[m->leib_rest[k]];
    L[e] = acc / k0;
  }
  for (int e = 0; e < Q; ++e) {
    const int deg = m->mono[e][0] + m->mono[e][1] + m->mono[e][2];
    if (deg == 0) c[e] = rho * log(k0);
    else if (deg == 1) {
      const int a = m->mono[e][0] ? 0 : (m->mono[e][1] ? 1 : 2);
      c[e] = rho * (u[a] + L[e]);
    } else c[e] = rho * L[e];
  }
  matvec(Q, &pr->P, c, C);
  for (int i = 0; i < Q; ++i)
    if (m->n_terms[i] >= 0) C[i] += m->rates[i] * (eval_poly(m, i, rho, drho, u) - C[i]);
  matvec(Q, &pr->Pinv, C, cp);
  /* central moments kappa_e = d^e exp(C(Xi) - Xi.u) at 0 with c -> c / rho */
  for (int e = 0; e < Q; ++e) {
    const int deg = m->mono[e][0] + m->mono[e][1] + m->mono[e][2];
    if (deg == 1) {
      const int a = m->mono[e][0] ? 0 : (m->mono[e][1] ? 1 : 2);
      L[e] = cp[e] / rho - u[a];
    } else if (deg > 1) L[e] = cp[e] / rho;
  }
  K[m->idx000] = k0;
  for (int o = 0; o < Q; ++o) {
    const int e = m->order[o];
    if (e == m->idx000) continue;
    double acc = 0.0;
    for (int k = m->leib_start[e]; k < m->leib_start[e + 1]; ++k)
      acc += m->leib_b[k] * L[m->leib_rj[k]] * K[m->leib_rest[k]];
    K[e] = acc;
  }
  if (m->force_model != FM_NONE)
    for (int a = 0; a < m->D; ++a) K[m->first_order[a]] += 0.5 * m->F[a];
  shift(m, K, u, t);
  if (m->has_bg)
    for (int i = 0; i < Q; ++i) t[i] -= m->bg[i];
  matvec(Q, &pr->Minv, t, f);
}

static void collide(const prepared* pr, double* f) {
  if (pr->m->kind == KIND_CUMULANT) collide_cumulant(pr, f);
  else collide_moment(pr, f);
}

static prepared* prepare(const oracle_method* m) {
  prepared* pr = (prepared*)malloc(sizeof(prepared));
  if (!pr) return NULL;
  pr->m = m;
  compress(m->Q, m->M, &pr->M);
  compress(m->Q, m->Minv, &pr->Minv);
  compress(m->Q, m->P, &pr->P);
  compress(m->Q, m->Pinv, &pr->Pinv);
  return pr;
}

static inline int64_t wrap(int64_t i, int64_t n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

/* One sweep over a fully periodic nx*ny*nz domain.
 * pattern 0: pull from src into dst;  1: AA EVEN (in place, src == dst);  2: AA ODD (in place). */
int oracle_sweep(const oracle_method* m, const double* src, double* dst, int32_t nx, int32_t ny, int32_t nz,
                 int32_t pattern, int32_t nthreads) {
  if (!m || !src || !dst || nx <= 0 || ny <= 0 || nz <= 0 || m->Q > QMAX) return -1;
  if (pattern != 0 && src != dst) return -1;
  const int64_t sy = nx, sz = (int64_t)nx * ny, sq = sz * nz;
  const int Q = m->Q;
  prepared* pr = prepare(m);
  if (!pr) return -2;
  int inv[QMAX];
  for (int i = 0; i < Q; ++i)
    for (int j = 0; j < Q; ++j)
      if (m->c[j][0] == -m->c[i][0] && m->c[j][1] == -m->c[i][1] && m->c[j][2] == -m->c[i][2]) inv[i] = j;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(static) collapse(2)
#endif
  for (int64_t z = 0; z < nz; ++z)
    for (int64_t y = 0; y < ny; ++y)
      for (int64_t x = 0; x < nx; ++x) {
        double f[QMAX];
        const int64_t here = z * sz + y * sy + x;
        for (int i = 0; i < Q; ++i) {
          const int64_t xm = wrap(x - m->c[i][0], nx), ym = wrap(y - m->c[i][1], ny), zm = wrap(z - m->c[i][2], nz);
          const int64_t back = zm * sz + ym * sy + xm;
          if (pattern == 0) f[i] = src[i * sq + back];
          else if (pattern == 1) f[i] = src[i * sq + here];
          else f[i] = src[inv[i] * sq + back];
        }
        collide(pr, f);
        for (int i = 0; i < Q; ++i) {
          const int64_t xp = wrap(x + m->c[i][0], nx), yp = wrap(y + m->c[i][1], ny), zp = wrap(z + m->c[i][2], nz);
          const int64_t fwd = zp * sz + yp * sy + xp;
          if (pattern == 0) dst[i * sq + here] = f[i];
          else if (pattern == 1) dst[inv[i] * sq + here] = f[i];
          else dst[i * sq + fwd] = f[i];
        }
      }
  free(pr);
  return 0;
}

/* collision only, on n cells stored as f[q * n + cell] (for the per-cell comparisons) */
int oracle_collide(const oracle_method* m, double* f, int64_t n) {
  if (!m || !f || n < 0) return -1;
  prepared* pr = prepare(m);
  if (!pr) return -2;
  for (int64_t k = 0; k < n; ++k) {
    double g[QMAX];
    for (int i = 0; i < m->Q; ++i) g[i] = f[i * n + k];
    collide(pr, g);
    for (int i = 0; i < m->Q; ++i) f[i * n + k] = g[i];
  }
  free(pr);
  return 0;
}

int64_t oracle_method_size(void) { return (int64_t)sizeof(oracle_method); }
