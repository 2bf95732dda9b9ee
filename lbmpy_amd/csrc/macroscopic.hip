// Macroscopic setter / getter on the GPU (lbmpy macroscopic_value_kernels.py:29-61), which the
// reference's LatticeBoltzmannStep runs on the CPU after copying the field back (lbstep.py:311-314,
// 444-449).
//
// setter: PDFs := equilibrium(rho, u - F/(2 rho_ref)) in storage form, written through the write
//         accessor of the *previous* time step (pull: slot i; an in-place pattern with previous EVEN:
//         that kernel's write accessor, e.g. AA slot inv(i) at x, EsoTwist slot inv(i) at x + max(c_i, 0)).
// getter: rho, u = j / rho_ref + F/(2 rho_ref) from the populations the last kernel wrote,
//         read through that kernel's write accessor (pull: slot i at x; AA EVEN: slot inv(i) at x;
//         AA ODD: slot i at x + c_i, bounced back into slot inv(i) at a wall; Eso: their tables).
// Neighbour accesses wrap on periodic axes and use the ghost layer on ghost faces.
// Both compute in FP64 regardless of the storage type.
#include "lbm_common.h"
#include "occupancy.h"
#include "collide_moment.h"
#include "collide_cumulant.h"
#include "collide_central.h"

namespace lbm355 {

// equilibrium of the moment family in storage form, from the FP64 method table
template <class S>
__device__ __forceinline__ void moment_equilibrium(const double* tab, double drho, double lam, const double (&u)[3],
                                                   double (&feq)[S::Q]) {
  using L = MomentTab<S::Q>;
#pragma unroll
  for (int i = 0; i < S::Q; ++i) {
    double p = 0;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      p += tab[L::L + 3 * i + d] * u[d];
      p += tab[L::QD + 3 * i + d] * (u[d] * u[d]);
    }
    p += tab[L::QO + 3 * i + 0] * (u[0] * u[1]) + tab[L::QO + 3 * i + 1] * (u[0] * u[2]) +
         tab[L::QO + 3 * i + 2] * (u[1] * u[2]);
    feq[i] = tab[L::A + i] + tab[L::B + i] * drho + lam * p;
  }
}

// D3Q27 cumulant equilibrium: backward chimera of rho * Gaussian central moments about u:
// f_i = rho * prod_a phi(c_ia, u_a), phi(0,u) = 2/3 - u^2, phi(+-1,u) = (1/3 + u^2 +- u) / 2
__device__ __forceinline__ double phi(int c, double u) {
  return c == 0 ? (2.0 / 3.0 - u * u) : 0.5 * (1.0 / 3.0 + u * u + c * u);
}

// Equilibrium populations in storage form (minus the weights when zero-centred) for density rho and equilibrium
// velocity u, no force shift applied, per collision family.
template <class S, int FAM>
__device__ __forceinline__ void equilibrium_storage(const double* tab, double rho, const double (&u)[3], int flags,
                                                    double (&feq)[S::Q]) {
  constexpr int Q = S::Q, D = S::D;
  const bool comp = flags & LBM355_FLAG_COMPRESSIBLE;
  if (FAM == LBM355_FAMILY_MOMENT) {
    moment_equilibrium<S>(tab, rho - 1.0, comp ? rho : 1.0, u, feq);
  } else if (FAM == LBM355_FAMILY_CENTRAL) {
    // a state with this density and momentum (j = rho_ref u) collided with every central moment relaxed
    // at rate 1 is the method's equilibrium; the collision works on deviations f - w
    SweepArgs<double> ca{};
    ca.tab = tab;
    ca.flags = comp ? LBM355_FLAG_COMPRESSIBLE : 0;
    const double ref_j = comp ? rho : 1.0;
#pragma unroll
    for (int i = 0; i < Q; ++i) {
      double cj = 0;
#pragma unroll
      for (int d = 0; d < D; ++d) cj += S::c[i][d] * ref_j * u[d];
      feq[i] = weight<S>(i) * ((rho - 1.0) + 3.0 * cj);
    }
    double r_out, u_out[3];
    collide_central<S, double, true>(feq, ca, r_out, u_out);
    if (!(flags & LBM355_FLAG_ZERO_CENTERED)) {
#pragma unroll
      for (int i = 0; i < Q; ++i) feq[i] += weight<S>(i);
    }
  } else {
    const bool zc = flags & LBM355_FLAG_ZERO_CENTERED;
#pragma unroll
    for (int i = 0; i < Q; ++i) {
      const double fi = rho * phi(S::c[i][0], u[0]) * phi(S::c[i][1], u[1]) * phi(S::c[i][2], u[2]);
      feq[i] = zc ? fi - weight<S>(i) : fi;
    }
  }
}

template <class S, int FAM, int PAT, class TS>
__global__ void set_equilibrium_kernel(TS* pdf, const double* rho_in, const double* u_in, const double* tab, int nx,
                                       int ny, int nz, int g, int64_t sy, int64_t sz, int64_t sq, int64_t osy,
                                       int64_t osz, int64_t osc, int flags, double fx, double fy, double fz,
                                       int face_x, int face_y, int face_z) {
  constexpr int Q = S::Q, D = S::D;
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= nx) return;
  const int y = blockIdx.y, z = blockIdx.z;
  const int64_t o = (int64_t)(x + g) + (int64_t)(y + g) * osy + (D == 3 ? (int64_t)(z + g) * osz : 0);
  const double rho = rho_in[o];
  const bool comp = flags & LBM355_FLAG_COMPRESSIBLE;
  const double ref = comp ? rho : 1.0;
  const double F[3] = {fx, fy, fz};
  double u[3] = {0, 0, 0};
#pragma unroll
  for (int d = 0; d < D; ++d) {
    u[d] = u_in[o + d * osc];
    if (flags & LBM355_FLAG_FORCE) u[d] -= 0.5 * F[d] / ref;
  }
  double feq[Q];
  equilibrium_storage<S, FAM>(tab, rho, u, flags, feq);
  const AxisNb ax = axis_neighbours(x, nx, g, face_x);
  const AxisNb ay = axis_neighbours(y, ny, g, face_y);
  AxisNb az = {0, 0, false, false};
  if (D == 3) az = axis_neighbours(z, nz, g, face_z);
  const int64_t X[3] = {ax.lo, x + g, ax.hi};
  const int64_t Y[3] = {ay.lo * sy, (y + g) * sy, ay.hi * sy};
  const int64_t Z[3] = {D == 3 ? az.lo * sz : 0, D == 3 ? (z + g) * sz : 0, D == 3 ? az.hi * sz : 0};
  constexpr AccessTable<S, PAT> acc{};
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    const Access& t = acc.a[i];
    pdf[t.wslot * sq + X[1 + t.w[0]] + Y[1 + t.w[1]] + Z[1 + t.w[2]]] = (TS)feq[i];
  }
}

template <class S, int PAT, class TS>
__global__ void get_macroscopic_kernel(const TS* pdf, double* rho_out, double* u_out, int nx, int ny, int nz, int g,
                                       int64_t sy, int64_t sz, int64_t sq, int64_t osy, int64_t osz, int64_t osc,
                                       int fx_mode, int fy_mode, int fz_mode, int flags, double fx, double fy,
                                       double fz) {
  constexpr int Q = S::Q, D = S::D;
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= nx) return;
  const int y = blockIdx.y, z = blockIdx.z;
  const AxisNb ax = axis_neighbours(x, nx, g, fx_mode);
  const AxisNb ay = axis_neighbours(y, ny, g, fy_mode);
  AxisNb az = {0, 0, false, false};
  if (D == 3) az = axis_neighbours(z, nz, g, fz_mode);
  const int64_t X[3] = {ax.lo, x + g, ax.hi};
  const int64_t Y[3] = {ay.lo * sy, (y + g) * sy, ay.hi * sy};
  const int64_t Z[3] = {D == 3 ? az.lo * sz : 0, D == 3 ? (z + g) * sz : 0, D == 3 ? az.hi * sz : 0};
  const int64_t center = X[1] + Y[1] + Z[1];
  double s = 0, j[3] = {0, 0, 0};
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    const int cx = S::c[i][0], cy = S::c[i][1], cz = S::c[i][2];
    double v;
    if (PAT != LBM355_AA_ODD) {
      const Access t = pattern_access<S>(PAT, i);
      v = (double)pdf[t.wslot * sq + X[1 + t.w[0]] + Y[1 + t.w[1]] + Z[1 + t.w[2]]];
    } else {
      const bool wall = (cx == -1 && ax.wall_lo) || (cx == 1 && ax.wall_hi) || (cy == -1 && ay.wall_lo) ||
                        (cy == 1 && ay.wall_hi) || (cz == -1 && az.wall_lo) || (cz == 1 && az.wall_hi);
      v = wall ? (double)pdf[S::inv[i] * sq + center] : (double)pdf[i * sq + X[1 + cx] + Y[1 + cy] + Z[1 + cz]];
    }
    s += v;
    j[0] += cx * v;
    j[1] += cy * v;
    j[2] += cz * v;
  }
  const double rho = (flags & LBM355_FLAG_ZERO_CENTERED) ? 1.0 + s : s;
  const double ref = (flags & LBM355_FLAG_COMPRESSIBLE) ? rho : 1.0;
  const double F[3] = {fx, fy, fz};
  const int64_t o = (int64_t)(x + g) + (int64_t)(y + g) * osy + (D == 3 ? (int64_t)(z + g) * osz : 0);
  if (rho_out) rho_out[o] = rho;
  if (u_out) {
#pragma unroll
    for (int d = 0; d < D; ++d) {
      double ud = j[d] / ref;
      if (flags & LBM355_FLAG_FORCE) ud += 0.5 * F[d] / ref;
      u_out[o + d * osc] = ud;
    }
  }
}

// Strain-rate tensor of one cell (macroscopic_value_kernels.py:67-85): the populations where the kernel `pat`
// wrote them (as get_macroscopic_kernel reads them, pattern chosen at run time: a diagnostic, not a hot loop),
// density and equilibrium velocity as the collision computes them, f_neq = f - f_eq(rho, u),
// Pi_ab = sum_i c_ia c_ib f_neq,i and S_ab = -1.5 (omega / rho_ref) Pi_ab, rho_ref = rho if compressible, else 1.
template <class S, int FAM, class TS>
__global__ void strain_rate_kernel(const TS* pdf, const double* tab, double* out, int pat, int nx, int ny, int nz,
                                   int g, int64_t sy, int64_t sz, int64_t sq, int64_t osy, int64_t osz, int64_t osc,
                                   int fx_mode, int fy_mode, int fz_mode, int flags, double fx, double fy, double fz,
                                   double omega) {
  constexpr int Q = S::Q, D = S::D;
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= nx) return;
  const int y = blockIdx.y, z = blockIdx.z;
  const AxisNb ax = axis_neighbours(x, nx, g, fx_mode);
  const AxisNb ay = axis_neighbours(y, ny, g, fy_mode);
  AxisNb az = {0, 0, false, false};
  if (D == 3) az = axis_neighbours(z, nz, g, fz_mode);
  const int64_t X[3] = {ax.lo, x + g, ax.hi};
  const int64_t Y[3] = {ay.lo * sy, (y + g) * sy, ay.hi * sy};
  const int64_t Z[3] = {D == 3 ? az.lo * sz : 0, D == 3 ? (z + g) * sz : 0, D == 3 ? az.hi * sz : 0};
  const int64_t center = X[1] + Y[1] + Z[1];
  double f[Q], s = 0, j[3] = {0, 0, 0};
  for (int i = 0; i < Q; ++i) {
    const int cx = S::c[i][0], cy = S::c[i][1], cz = S::c[i][2];
    double v;
    if (pat != LBM355_AA_ODD) {
      const Access t = pattern_access<S>(pat, i);
      v = (double)pdf[t.wslot * sq + X[1 + t.w[0]] + Y[1 + t.w[1]] + Z[1 + t.w[2]]];
    } else {
      const bool wall = (cx == -1 && ax.wall_lo) || (cx == 1 && ax.wall_hi) || (cy == -1 && ay.wall_lo) ||
                        (cy == 1 && ay.wall_hi) || (cz == -1 && az.wall_lo) || (cz == 1 && az.wall_hi);
      v = wall ? (double)pdf[S::inv[i] * sq + center] : (double)pdf[i * sq + X[1 + cx] + Y[1 + cy] + Z[1 + cz]];
    }
    f[i] = v;
    s += v;
    j[0] += cx * v;
    j[1] += cy * v;
    j[2] += cz * v;
  }
  const double rho = (flags & LBM355_FLAG_ZERO_CENTERED) ? 1.0 + s : s;
  const double ref = (flags & LBM355_FLAG_COMPRESSIBLE) ? rho : 1.0;
  const double F[3] = {fx, fy, fz};
  const bool shift = (flags & LBM355_FLAG_FORCE) && (flags & LBM355_FLAG_EQ_SHIFT);
  double u[3] = {0, 0, 0};
  for (int d = 0; d < D; ++d) u[d] = j[d] / ref + (shift ? 0.5 * F[d] / ref : 0.0);
  double feq[Q];
  equilibrium_storage<S, FAM>(tab, rho, u, flags, feq);
  double pi[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  for (int i = 0; i < Q; ++i) {
    const double n = f[i] - feq[i];
    for (int a = 0; a < D; ++a)
      for (int b = 0; b < D; ++b) pi[a][b] += S::c[i][a] * S::c[i][b] * n;
  }
  const double k = -1.5 * omega / ref;
  const int64_t o = (int64_t)(x + g) + (int64_t)(y + g) * osy + (D == 3 ? (int64_t)(z + g) * osz : 0);
  for (int a = 0; a < D; ++a)
    for (int b = 0; b < D; ++b) out[o + (a * D + b) * osc] = k * pi[a][b];
}

}  // namespace lbm355

using namespace lbm355;

namespace {
int fail2(int code, const char* m) { return lbm355_set_error(code, m); }

template <class S, int FAM, class TS>
int set_eq_launch(const lbm355_sweep_desc* d, void* pdf, const double* rho, const double* u, const double* tab,
                  int prev_pattern, hipStream_t s) {
  const int bx = d->nx >= 256 ? 256 : ((d->nx + 63) / 64) * 64;
  dim3 block(bx), grid((d->nx + bx - 1) / bx, d->ny, d->nz);
  decltype(&set_equilibrium_kernel<S, FAM, LBM355_PULL, TS>) k = nullptr;
  switch (prev_pattern) {
    case LBM355_PULL: k = set_equilibrium_kernel<S, FAM, LBM355_PULL, TS>; break;
    case LBM355_AA_EVEN: k = set_equilibrium_kernel<S, FAM, LBM355_AA_EVEN, TS>; break;
    case LBM355_AA_ODD: k = set_equilibrium_kernel<S, FAM, LBM355_AA_ODD, TS>; break;
    case LBM355_ESOTWIST_EVEN: k = set_equilibrium_kernel<S, FAM, LBM355_ESOTWIST_EVEN, TS>; break;
    case LBM355_ESOTWIST_ODD: k = set_equilibrium_kernel<S, FAM, LBM355_ESOTWIST_ODD, TS>; break;
    case LBM355_ESOPULL_EVEN: k = set_equilibrium_kernel<S, FAM, LBM355_ESOPULL_EVEN, TS>; break;
    case LBM355_ESOPULL_ODD: k = set_equilibrium_kernel<S, FAM, LBM355_ESOPULL_ODD, TS>; break;
    case LBM355_ESOPUSH_EVEN: k = set_equilibrium_kernel<S, FAM, LBM355_ESOPUSH_EVEN, TS>; break;
    case LBM355_ESOPUSH_ODD: k = set_equilibrium_kernel<S, FAM, LBM355_ESOPUSH_ODD, TS>; break;
    default: k = set_equilibrium_kernel<S, FAM, LBM355_PUSH, TS>; break;
  }
  hipLaunchKernelGGL(k, grid, block, 0, s, (TS*)pdf, rho, u, tab, d->nx, d->ny, d->nz, d->ghost_layers, d->stride_y,
                     d->stride_z, d->stride_q, d->out_stride_y, d->out_stride_z, d->out_stride_c, d->flags,
                     d->force[0], d->force[1], d->force[2], d->face[0], d->face[1], d->face[2]);
  return hipGetLastError() == hipSuccess ? 0 : fail2(LBM355_ERR_LAUNCH, "setter launch failed");
}

template <class S, class TS>
int get_launch(const lbm355_sweep_desc* d, const void* pdf, double* rho, double* u, int last_pattern,
               hipStream_t s) {
  const int bx = d->nx >= 256 ? 256 : ((d->nx + 63) / 64) * 64;
  dim3 block(bx), grid((d->nx + bx - 1) / bx, d->ny, d->nz);
  decltype(&get_macroscopic_kernel<S, LBM355_PULL, TS>) k = nullptr;
  switch (last_pattern) {
    case LBM355_PULL: k = get_macroscopic_kernel<S, LBM355_PULL, TS>; break;
    case LBM355_AA_EVEN: k = get_macroscopic_kernel<S, LBM355_AA_EVEN, TS>; break;
    case LBM355_AA_ODD: k = get_macroscopic_kernel<S, LBM355_AA_ODD, TS>; break;
    case LBM355_ESOTWIST_EVEN: k = get_macroscopic_kernel<S, LBM355_ESOTWIST_EVEN, TS>; break;
    case LBM355_ESOTWIST_ODD: k = get_macroscopic_kernel<S, LBM355_ESOTWIST_ODD, TS>; break;
    case LBM355_ESOPULL_EVEN: k = get_macroscopic_kernel<S, LBM355_ESOPULL_EVEN, TS>; break;
    case LBM355_ESOPULL_ODD: k = get_macroscopic_kernel<S, LBM355_ESOPULL_ODD, TS>; break;
    case LBM355_ESOPUSH_EVEN: k = get_macroscopic_kernel<S, LBM355_ESOPUSH_EVEN, TS>; break;
    case LBM355_ESOPUSH_ODD: k = get_macroscopic_kernel<S, LBM355_ESOPUSH_ODD, TS>; break;
    default: k = get_macroscopic_kernel<S, LBM355_PUSH, TS>; break;
  }
  hipLaunchKernelGGL(k, grid, block, 0, s, (const TS*)pdf, rho, u, d->nx, d->ny, d->nz, d->ghost_layers, d->stride_y,
                     d->stride_z, d->stride_q, d->out_stride_y, d->out_stride_z, d->out_stride_c, d->face[0],
                     d->face[1], d->face[2], d->flags, d->force[0], d->force[1], d->force[2]);
  return hipGetLastError() == hipSuccess ? 0 : fail2(LBM355_ERR_LAUNCH, "getter launch failed");
}

template <class S, int FAM>
int set_eq_dtype(const lbm355_sweep_desc* d, void* pdf, const double* rho, const double* u, const double* tab,
                 int prev, hipStream_t s) {
  if (d->store_dtype == LBM355_F64) return set_eq_launch<S, FAM, double>(d, pdf, rho, u, tab, prev, s);
  if (d->store_dtype == LBM355_F32) return set_eq_launch<S, FAM, float>(d, pdf, rho, u, tab, prev, s);
  return set_eq_launch<S, FAM, _Float16>(d, pdf, rho, u, tab, prev, s);
}

template <class S>
int get_dtype(const lbm355_sweep_desc* d, const void* pdf, double* rho, double* u, int last, hipStream_t s) {
  if (d->store_dtype == LBM355_F64) return get_launch<S, double>(d, pdf, rho, u, last, s);
  if (d->store_dtype == LBM355_F32) return get_launch<S, float>(d, pdf, rho, u, last, s);
  return get_launch<S, _Float16>(d, pdf, rho, u, last, s);
}

template <class S, int FAM>
int strain_launch(const lbm355_sweep_desc* d, const void* pdf, const double* tab, double* out, double omega,
                  int last, hipStream_t s) {
  const int bx = d->nx >= 256 ? 256 : ((d->nx + 63) / 64) * 64;
  dim3 block(bx), grid((d->nx + bx - 1) / bx, d->ny, d->nz);
  auto args = [&](auto kernel, auto* p) {
    hipLaunchKernelGGL(kernel, grid, block, 0, s, p, tab, out, last, d->nx, d->ny, d->nz, d->ghost_layers,
                       d->stride_y, d->stride_z, d->stride_q, d->out_stride_y, d->out_stride_z, d->out_stride_c,
                       d->face[0], d->face[1], d->face[2], d->flags, d->force[0], d->force[1], d->force[2], omega);
  };
  if (d->store_dtype == LBM355_F64) args(strain_rate_kernel<S, FAM, double>, (const double*)pdf);
  else if (d->store_dtype == LBM355_F32) args(strain_rate_kernel<S, FAM, float>, (const float*)pdf);
  else args(strain_rate_kernel<S, FAM, _Float16>, (const _Float16*)pdf);
  return hipGetLastError() == hipSuccess ? 0 : fail2(LBM355_ERR_LAUNCH, "strain-rate launch failed");
}
}  // namespace

extern "C" {

int lbm355_set_equilibrium(const lbm355_sweep_desc* d, void* pdf, const double* density, const double* velocity,
                           const double* table_f64, int32_t prev_pattern, void* stream) {
  if (!d || !pdf || !density || !velocity) return fail2(LBM355_ERR_ARG, "NULL argument");
  if (d->store_dtype < 0 || d->store_dtype > 2) return fail2(LBM355_ERR_ARG, "bad dtype");
  if (prev_pattern < LBM355_PULL || prev_pattern > LBM355_PUSH) return fail2(LBM355_ERR_ARG, "bad pattern");
  for (int a = 0; a < (d->stencil == LBM355_D2Q9 ? 2 : 3); ++a)
    if (d->face[a] < 0 || d->face[a] > 2 || (d->face[a] == LBM355_FACE_GHOST && d->ghost_layers < 1) ||
        ((table_pattern(prev_pattern) || prev_pattern == LBM355_AA_ODD) && d->face[a] == LBM355_FACE_NOSLIP))
      return fail2(LBM355_ERR_ARG, "bad face mode (fused NoSlip faces: the setter writes as after pull, push or an "
                                   "AA even kernel)");
  hipStream_t s = (hipStream_t)stream;
  if (d->family == LBM355_FAMILY_MOMENT) {
    if (!table_f64) return fail2(LBM355_ERR_ARG, "moment family needs the FP64 method table");
    if (d->stencil == LBM355_D2Q9)
      return set_eq_dtype<D2Q9, LBM355_FAMILY_MOMENT>(d, pdf, density, velocity, table_f64, prev_pattern, s);
    if (d->stencil == LBM355_D3Q19)
      return set_eq_dtype<D3Q19, LBM355_FAMILY_MOMENT>(d, pdf, density, velocity, table_f64, prev_pattern, s);
    if (d->stencil == LBM355_D3Q27)
      return set_eq_dtype<D3Q27, LBM355_FAMILY_MOMENT>(d, pdf, density, velocity, table_f64, prev_pattern, s);
  } else if (d->family == LBM355_FAMILY_CUMULANT && d->stencil == LBM355_D3Q27) {
    return set_eq_dtype<D3Q27, LBM355_FAMILY_CUMULANT>(d, pdf, density, velocity, table_f64, prev_pattern, s);
  } else if (d->family == LBM355_FAMILY_CENTRAL) {
    if (!table_f64) return fail2(LBM355_ERR_ARG, "central-moment family needs the FP64 method table");
    if (d->stencil == LBM355_D2Q9)
      return set_eq_dtype<D2Q9, LBM355_FAMILY_CENTRAL>(d, pdf, density, velocity, table_f64, prev_pattern, s);
    if (d->stencil == LBM355_D3Q19)
      return set_eq_dtype<D3Q19, LBM355_FAMILY_CENTRAL>(d, pdf, density, velocity, table_f64, prev_pattern, s);
    if (d->stencil == LBM355_D3Q27)
      return set_eq_dtype<D3Q27, LBM355_FAMILY_CENTRAL>(d, pdf, density, velocity, table_f64, prev_pattern, s);
  }
  return fail2(LBM355_ERR_UNSUPPORTED, "no setter for this stencil/family");
}

int lbm355_get_macroscopic(const lbm355_sweep_desc* d, const void* pdf, double* density, double* velocity,
                           int32_t last_pattern, void* stream) {
  if (!d || !pdf) return fail2(LBM355_ERR_ARG, "NULL argument");
  if (d->store_dtype < 0 || d->store_dtype > 2) return fail2(LBM355_ERR_ARG, "bad dtype");
  if (last_pattern < LBM355_PULL || last_pattern > LBM355_PUSH) return fail2(LBM355_ERR_ARG, "bad pattern");
  if (table_pattern(last_pattern))
    for (int a = 0; a < (d->stencil == LBM355_D2Q9 ? 2 : 3); ++a)
      if (d->face[a] == LBM355_FACE_NOSLIP) return fail2(LBM355_ERR_ARG, "Eso patterns have no fused NoSlip faces");
  hipStream_t s = (hipStream_t)stream;
  if (d->stencil == LBM355_D2Q9) return get_dtype<D2Q9>(d, pdf, density, velocity, last_pattern, s);
  if (d->stencil == LBM355_D3Q19) return get_dtype<D3Q19>(d, pdf, density, velocity, last_pattern, s);
  if (d->stencil == LBM355_D3Q27) return get_dtype<D3Q27>(d, pdf, density, velocity, last_pattern, s);
  return fail2(LBM355_ERR_ARG, "unknown stencil");
}

int lbm355_get_strain_rate(const lbm355_sweep_desc* d, const void* pdf, const double* table_f64, double* strain_rate,
                           double omega_shear, int32_t last_pattern, void* stream) {
  if (!d || !pdf || !strain_rate || !table_f64) return fail2(LBM355_ERR_ARG, "NULL argument");
  if (d->store_dtype < 0 || d->store_dtype > 2) return fail2(LBM355_ERR_ARG, "bad dtype");
  if (last_pattern < LBM355_PULL || last_pattern > LBM355_PUSH) return fail2(LBM355_ERR_ARG, "bad pattern");
  for (int a = 0; a < (d->stencil == LBM355_D2Q9 ? 2 : 3); ++a)
    if (d->face[a] < 0 || d->face[a] > 2 || (d->face[a] == LBM355_FACE_GHOST && d->ghost_layers < 1) ||
        (table_pattern(last_pattern) && d->face[a] == LBM355_FACE_NOSLIP))
      return fail2(LBM355_ERR_ARG, "bad face mode (Eso patterns: periodic or ghost faces)");
  hipStream_t s = (hipStream_t)stream;
  if (d->family == LBM355_FAMILY_MOMENT) {
    if (d->stencil == LBM355_D2Q9)
      return strain_launch<D2Q9, LBM355_FAMILY_MOMENT>(d, pdf, table_f64, strain_rate, omega_shear, last_pattern, s);
    if (d->stencil == LBM355_D3Q19)
      return strain_launch<D3Q19, LBM355_FAMILY_MOMENT>(d, pdf, table_f64, strain_rate, omega_shear, last_pattern, s);
    if (d->stencil == LBM355_D3Q27)
      return strain_launch<D3Q27, LBM355_FAMILY_MOMENT>(d, pdf, table_f64, strain_rate, omega_shear, last_pattern, s);
  } else if (d->family == LBM355_FAMILY_CUMULANT && d->stencil == LBM355_D3Q27) {
    return strain_launch<D3Q27, LBM355_FAMILY_CUMULANT>(d, pdf, table_f64, strain_rate, omega_shear, last_pattern, s);
  } else if (d->family == LBM355_FAMILY_CENTRAL) {
    if (d->stencil == LBM355_D2Q9)
      return strain_launch<D2Q9, LBM355_FAMILY_CENTRAL>(d, pdf, table_f64, strain_rate, omega_shear, last_pattern, s);
    if (d->stencil == LBM355_D3Q19)
      return strain_launch<D3Q19, LBM355_FAMILY_CENTRAL>(d, pdf, table_f64, strain_rate, omega_shear, last_pattern, s);
    if (d->stencil == LBM355_D3Q27)
      return strain_launch<D3Q27, LBM355_FAMILY_CENTRAL>(d, pdf, table_f64, strain_rate, omega_shear, last_pattern, s);
  }
  return fail2(LBM355_ERR_UNSUPPORTED, "no strain-rate getter for this stencil/family");
}

int64_t lbm355_desc_size(void) { return (int64_t)sizeof(lbm355_sweep_desc); }

}  // extern "C"

// ------------------------------------------------------------------ bandwidth / counter calibration
namespace {
template <int NT>
__global__ void __launch_bounds__(256) stream_copy_kernel(const float* __restrict__ src, float* __restrict__ dst,
                                                          int64_t n) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  if (NT)
    __builtin_nontemporal_store(__builtin_nontemporal_load(src + t), dst + t);
  else
    dst[t] = src[t];
}

// Q planes per thread, the sweep's access pattern without the collision: thread t loads element t of every
// plane, then stores them all (in place when src == dst, as an AA-even sweep does).
template <int P, typename T, int NT>
__global__ void __launch_bounds__(256) plane_copy_kernel(const T* src, T* dst, int64_t n,
                                                         int64_t stride) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  T v[P];
#pragma unroll
  for (int q = 0; q < P; ++q) v[q] = NT ? __builtin_nontemporal_load(src + q * stride + t) : src[q * stride + t];
#pragma unroll
  for (int q = 0; q < P; ++q) {
    if (NT)
      __builtin_nontemporal_store(v[q], dst + q * stride + t);
    else
      dst[q * stride + t] = v[q];
  }
}

template <int P, typename T>
int plane_copy_launch(const void* src, void* dst, int64_t n, int64_t stride, int32_t nt, int32_t waves,
                      hipStream_t s) {
  const int64_t blocks = (n + 255) / 256;
  const size_t lds = lbm355::lds_for_wave_cap(waves, 256, 0);
  if (nt)
    hipLaunchKernelGGL((plane_copy_kernel<P, T, 1>), dim3(blocks), dim3(256), lds, s, (const T*)src, (T*)dst, n,
                       stride);
  else
    hipLaunchKernelGGL((plane_copy_kernel<P, T, 0>), dim3(blocks), dim3(256), lds, s, (const T*)src, (T*)dst, n,
                       stride);
  return hipGetLastError() == hipSuccess ? LBM355_OK : lbm355_set_error(LBM355_ERR_LAUNCH, "plane copy launch failed");
}
}  // namespace

extern "C" int lbm355_plane_copy(int32_t dtype, int32_t n_planes, const void* src, void* dst, int64_t plane_elems,
                                 int64_t plane_stride, int32_t nontemporal, int32_t waves_per_simd, void* stream) {
  if (!src || !dst || plane_elems < 0 || plane_stride < plane_elems)
    return lbm355_set_error(LBM355_ERR_ARG, "bad plane-copy arguments (plane_stride must be >= plane_elems)");
  if (waves_per_simd < 0 || waves_per_simd > 8)
    return lbm355_set_error(LBM355_ERR_ARG, "plane copy: waves_per_simd must be 0 (no cap) to 8");
  if (n_planes != 19 && n_planes != 27)
    return lbm355_set_error(LBM355_ERR_UNSUPPORTED, "plane copy: n_planes must be 19 or 27");
  if (plane_elems == 0) return LBM355_OK;
  hipStream_t s = (hipStream_t)stream;
  const int nt = nontemporal ? 1 : 0;
  const int w = waves_per_simd;
  switch (dtype) {
    case LBM355_F64:
      return n_planes == 19 ? plane_copy_launch<19, double>(src, dst, plane_elems, plane_stride, nt, w, s)
                            : plane_copy_launch<27, double>(src, dst, plane_elems, plane_stride, nt, w, s);
    case LBM355_F32:
      return n_planes == 19 ? plane_copy_launch<19, float>(src, dst, plane_elems, plane_stride, nt, w, s)
                            : plane_copy_launch<27, float>(src, dst, plane_elems, plane_stride, nt, w, s);
    case LBM355_F16:
      return n_planes == 19 ? plane_copy_launch<19, _Float16>(src, dst, plane_elems, plane_stride, nt, w, s)
                            : plane_copy_launch<27, _Float16>(src, dst, plane_elems, plane_stride, nt, w, s);
    default:
      return lbm355_set_error(LBM355_ERR_ARG, "plane copy: unknown dtype");
  }
}

namespace {
// VALU issue ceiling: 8 independent FP32 multiply-add chains of `iters` steps per thread (8 * iters v_fma_f32
// per wave), their sum stored so that nothing is dropped.  x -> 0.999 x + 0.001 keeps the values bounded.
__global__ void __launch_bounds__(256) valu_probe_kernel(float* __restrict__ out, int64_t n, int32_t iters) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  float x[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) x[k] = (float)(t & 255) * 1e-3f + (float)k;
  for (int i = 0; i < iters; ++i) {
#pragma unroll
    for (int k = 0; k < 8; ++k) x[k] = fmaf(x[k], 0.999f, 1e-3f);
  }
  float sum = 0.f;
#pragma unroll
  for (int k = 0; k < 8; ++k) sum += x[k];
  out[t] = sum;
}
}  // namespace

extern "C" int lbm355_valu_probe(float* out, int64_t n_threads, int32_t iters, void* stream) {
  if (!out || n_threads < 0 || iters < 0) return lbm355_set_error(LBM355_ERR_ARG, "bad VALU-probe arguments");
  if (n_threads == 0) return LBM355_OK;
  const int64_t blocks = (n_threads + 255) / 256;
  hipLaunchKernelGGL(valu_probe_kernel, dim3(blocks), dim3(256), 0, (hipStream_t)stream, out, n_threads, iters);
  return hipGetLastError() == hipSuccess ? LBM355_OK : lbm355_set_error(LBM355_ERR_LAUNCH, "VALU probe launch failed");
}

extern "C" int lbm355_stream_copy(const void* src, void* dst, int64_t n_floats, int32_t nontemporal, void* stream) {
  if (!src || !dst || n_floats < 0) return lbm355_set_error(LBM355_ERR_ARG, "bad copy arguments");
  if (n_floats == 0) return LBM355_OK;
  const int64_t blocks = (n_floats + 255) / 256;
  if (nontemporal)
    hipLaunchKernelGGL(stream_copy_kernel<1>, dim3(blocks), dim3(256), 0, (hipStream_t)stream, (const float*)src,
                       (float*)dst, n_floats);
  else
    hipLaunchKernelGGL(stream_copy_kernel<0>, dim3(blocks), dim3(256), 0, (hipStream_t)stream, (const float*)src,
                       (float*)dst, n_floats);
  return hipGetLastError() == hipSuccess ? LBM355_OK : lbm355_set_error(LBM355_ERR_LAUNCH, "copy launch failed");
}
