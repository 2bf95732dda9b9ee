// Fused stream-collide sweep for MI355X (gfx950) and its C ABI (include/lbm355.h).
//
// One thread per interior cell, x fastest (a 64-lane wavefront covers 64 consecutive cells of
// one row, so each population is read and written as contiguous 256 B (FP32) segments).
// Blocks span up to 256 cells of a row; the grid is (row blocks, ny, nz).  Streaming:
//   pull    read  src[i][x - c_i],        write dst[i][x]           (fieldaccess.py:66-75)
//   AA even read  f[i][x],                write f[inv i][x]         (fieldaccess.py:136-145)
//   AA odd  read  f[inv i][x - c_i],      write f[i][x + c_i]       (fieldaccess.py:148-157)
// (the in-place Eso patterns are in sweep_eso.hip)
// Periodic axes wrap inside the index math (PeriodicTwoFieldsAccessor semantics,
// fieldaccess.py:90-133), NoSlip faces bounce back in place of the separate boundary kernel
// (boundaries/boundaryconditions.py:104-139): a pull read from a wall cell returns
// src[inv i][x]; AA odd reads src[i][x] and redirects the write to f[inv i][x].
#include <string>
#include <type_traits>
#include <cstring>
#include <cstdio>

#include "occupancy.h"
#include "sweep_kernel.h"

namespace lbm355 {

// ------------------------------------------------------------------ NoSlip index list
template <class S, int PAT, class T>
__global__ void noslip_links_kernel(T* pdf, int64_t sy, int64_t sz, int64_t sq, const int32_t* links, int64_t n) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int32_t* l = links + 4 * t;
  const int x = l[0], y = l[1], z = l[2], d = l[3];
  if (d < 0 || d >= S::Q) return;
  int cx = 0, cy = 0, cz = 0, id = 0;
#pragma unroll
  for (int i = 0; i < S::Q; ++i)
    if (i == d) {
      cx = S::c[i][0];
      cy = S::c[i][1];
      cz = S::c[i][2];
      id = S::inv[i];
    }
  const int64_t cell = x + y * sy + (S::D == 3 ? z * sz : 0);
  const int64_t nb = (x + cx) + (y + cy) * sy + (S::D == 3 ? (z + cz) * sz : 0);
  if (PAT == LBM355_PULL) {
    pdf[id * sq + nb] = pdf[d * sq + cell];
  } else if (PAT == LBM355_AA_EVEN) {
    pdf[d * sq + nb] = pdf[id * sq + cell];
  } else {
    pdf[id * sq + cell] = pdf[d * sq + nb];
  }
}

// ------------------------------------------------------------------ general boundary links
struct BoundaryArgs {
  int64_t sy, sz, sq;
  int n[3], g, periodic[3], fused[3];
  int zc, comp, eq_shift, adapt;
  int normal[3], mirror_axis;
  double velocity[3];
  double density;
  double force[3];
  const double* eq;
};

__device__ __forceinline__ int wrap_axis(int c, int n, int g, int periodic) {
  if (!periodic) return c;
  int k = c - g;
  k = k < 0 ? k + n : (k >= n ? k - n : k);
  return k + g;
}

// f_out(i) of the cell at (cx, cy, cz) after kernel PAT, and f_in(i) of the kernel that follows
template <class S, int PAT>
__device__ __forceinline__ int64_t bc_out(const BoundaryArgs& b, int i, int cx, int cy, int cz) {
  int slot = i;
  if (PAT == LBM355_AA_ODD) {
    const int tx = cx + S::c[i][0], ty = cy + S::c[i][1], tz = cz + S::c[i][2];
    // a population the odd sweep bounced at a fused NoSlip face already sits in the cell's slot inv(i)
    const bool bounced = (b.fused[0] && (tx < b.g || tx >= b.g + b.n[0])) ||
                         (b.fused[1] && (ty < b.g || ty >= b.g + b.n[1])) ||
                         (S::D == 3 && b.fused[2] && (tz < b.g || tz >= b.g + b.n[2]));
    if (bounced) {
      slot = S::inv[i];
    } else {
      cx = tx;
      cy = ty;
      cz = tz;
    }
  } else {
    // the write accessor of the kernel that ran last
    const Access t = pattern_access<S>(PAT, i);
    slot = t.wslot;
    cx += t.w[0];
    cy += t.w[1];
    cz += t.w[2];
  }
  return slot * b.sq + wrap_axis(cx, b.n[0], b.g, b.periodic[0]) + wrap_axis(cy, b.n[1], b.g, b.periodic[1]) * b.sy +
         (S::D == 3 ? wrap_axis(cz, b.n[2], b.g, b.periodic[2]) * b.sz : 0);
}
// whether bc_out(b, i, cx, cy, cz) addresses a cell of the array (ghost layers included): the outflow kinds read
// one cell back along a normal the caller supplies, so a normal pointing into the fluid must not read outside
template <class S, int PAT>
__device__ __forceinline__ bool bc_out_in_array(const BoundaryArgs& b, int i, int cx, int cy, int cz) {
  if (PAT == LBM355_AA_ODD) {
    cx += S::c[i][0];
    cy += S::c[i][1];
    cz += S::c[i][2];
  } else {
    const Access t = pattern_access<S>(PAT, i);
    cx += t.w[0];
    cy += t.w[1];
    cz += t.w[2];
  }
  const int c[3] = {cx, cy, cz};
  for (int a = 0; a < S::D; ++a)
    if (!b.periodic[a] && (c[a] < 0 || c[a] >= b.n[a] + 2 * b.g)) return false;
  return true;
}

// the read accessor of the kernel that runs next
template <class S, int PAT>
__device__ __forceinline__ int64_t bc_in(const BoundaryArgs& b, int i, int cx, int cy, int cz) {
  const Access t = pattern_access<S>(next_pattern(PAT), i);
  const int slot = t.rslot;
  cx += t.r[0];
  cy += t.r[1];
  cz += t.r[2];
  return slot * b.sq + wrap_axis(cx, b.n[0], b.g, b.periodic[0]) + wrap_axis(cy, b.n[1], b.g, b.periodic[1]) * b.sy +
         (S::D == 3 ? wrap_axis(cz, b.n[2], b.g, b.periodic[2]) * b.sz : 0);
}

template <class S, int PAT, class T, int KIND>
__global__ void boundary_links_kernel(T* pdf, const int32_t* links, int64_t n, double* link_vel, BoundaryArgs b) {
  using TC = typename std::conditional<sizeof(T) == 8, double, float>::type;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int32_t* l = links + 4 * t;
  const int x = l[0], y = l[1], z = l[2], d = l[3];
  if (d < 0 || d >= S::Q) return;
  int id = 0, cdx = 0, cdy = 0, cdz = 0;
#pragma unroll
  for (int i = 0; i < S::Q; ++i)
    if (i == d) {
      id = S::inv[i];
      cdx = S::c[i][0];
      cdy = S::c[i][1];
      cdz = S::c[i][2];
    }
  if (KIND == LBM355_BC_SIMPLE_OUTFLOW || KIND == LBM355_BC_OUTFLOW_INIT || KIND == LBM355_BC_OUTFLOW) {
    if (!bc_out_in_array<S, PAT>(b, id, x + cdx - b.normal[0], y + cdy - b.normal[1], z + cdz - b.normal[2])) return;
  }
  if (KIND == LBM355_BC_SIMPLE_OUTFLOW) {
    // copy from the fluid cell one step back along the normal from the boundary cell: x + c_d - n
    pdf[bc_in<S, PAT>(b, id, x, y, z)] =
        pdf[bc_out<S, PAT>(b, id, x + cdx - b.normal[0], y + cdy - b.normal[1], z + cdz - b.normal[2])];
    return;
  }
  if (KIND == LBM355_BC_OUTFLOW_INIT || KIND == LBM355_BC_OUTFLOW) {
    // the fluid cell one step back along the normal from the boundary cell: x + c_d - n
    const T v = pdf[bc_out<S, PAT>(b, id, x + cdx - b.normal[0], y + cdy - b.normal[1], z + cdz - b.normal[2])];
    double* st = link_vel + 2 * t;
    if (KIND == LBM355_BC_OUTFLOW_INIT) {
      st[0] = (double)v;
      st[1] = (double)v;
    } else {
      const TC c = (TC)0.57735026918962576451;   // sqrt(1/3) dt/dx
      const TC inter = (TC)st[0] * c + (TC(1) - c) * (TC)st[1];
      pdf[bc_in<S, PAT>(b, id, x, y, z)] = (T)inter;
      st[0] = (double)v;
      st[1] = (double)inter;
    }
    return;
  }
  const int64_t target = bc_in<S, PAT>(b, id, x, y, z);
  if (KIND == LBM355_BC_FREESLIP_LINKS) {
    // per-link normal n (wall to fluid, entries -1/0/1) and reference direction r; x + c_d + n and the accessor
    // offset of r stay within one cell of x (n cancels c_d on wall axes, r on the others)
    const double* q = link_vel + 4 * t;
    const int r = (int)q[3];
    if (r < 0 || r >= S::Q) return;
    pdf[target] = pdf[bc_out<S, PAT>(b, r, x + cdx + (int)q[0], y + cdy + (int)q[1], z + cdz + (int)q[2])];
    return;
  }
  if (KIND == LBM355_BC_NOSLIP) {
    pdf[target] = pdf[bc_out<S, PAT>(b, d, x, y, z)];
  } else if (KIND == LBM355_BC_UBB) {
    TC u[3] = {(TC)b.velocity[0], (TC)b.velocity[1], (TC)b.velocity[2]};
    if (link_vel) {
      u[0] = (TC)link_vel[3 * t];
      u[1] = (TC)link_vel[3 * t + 1];
      u[2] = (TC)link_vel[3 * t + 2];
    }
    TC rho = TC(1);
    if (b.density > 0) {
      rho = (TC)b.density;
    } else if (b.comp) {
      TC s = 0;
#pragma unroll
      for (int i = 0; i < S::Q; ++i) s += (TC)pdf[bc_out<S, PAT>(b, i, x, y, z)];
      rho = b.zc ? TC(1) + s : s;
    }
    if (b.adapt) {
      const TC ref = b.comp ? rho : TC(1);
#pragma unroll
      for (int a = 0; a < 3; ++a) u[a] -= TC(0.5) * (TC)b.force[a] / ref;
    }
    const TC cu = (TC)cdx * u[0] + (TC)cdy * u[1] + (TC)cdz * u[2];
    TC w = 0;
#pragma unroll
    for (int i = 0; i < S::Q; ++i)
      if (i == d) w = (TC)weight<S>(i);
    pdf[target] = (T)((TC)pdf[bc_out<S, PAT>(b, d, x, y, z)] - TC(6) * w * cu * rho);
  } else if (KIND == LBM355_BC_FIXED_DENSITY) {
    // 2 * (symmetric part of the equilibrium at rho_w and the fluid cell's velocity) - f_out(d); the
    // velocity comes from the fluid cell's populations with rho replaced by rho_w
    TC j[3] = {0, 0, 0};
#pragma unroll
    for (int i = 0; i < S::Q; ++i) {
      const TC v = (TC)pdf[bc_out<S, PAT>(b, i, x, y, z)];
#pragma unroll
      for (int a = 0; a < 3; ++a)
        if (S::c[i][a] != 0) j[a] += S::c[i][a] > 0 ? v : -v;
    }
    const TC rho_w = (TC)b.density;
    const TC ref = b.comp ? rho_w : TC(1);
    TC u[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) u[a] = j[a] / ref + (b.eq_shift ? (TC)b.force[a] / (TC(2) * ref) : TC(0));
    const double* e = b.eq + 8 * d;
    const TC quad = (TC)e[2] * u[0] * u[0] + (TC)e[3] * u[1] * u[1] + (TC)e[4] * u[2] * u[2] +
                    (TC)e[5] * u[0] * u[1] + (TC)e[6] * u[0] * u[2] + (TC)e[7] * u[1] * u[2];
    const TC eq_sym = (TC)e[0] + (TC)e[1] * (rho_w - TC(1)) + (b.comp ? rho_w : TC(1)) * quad;
    pdf[target] = (T)(TC(2) * eq_sym - (TC)pdf[bc_out<S, PAT>(b, d, x, y, z)]);
  } else {
    // specular reflection at the half-way plane: mirror d along the wall normal, take the inverse
    int mc[3] = {cdx, cdy, cdz};
    mc[b.mirror_axis] = -mc[b.mirror_axis];
    int m = 0;
#pragma unroll
    for (int i = 0; i < S::Q; ++i)
      if (S::c[i][0] == -mc[0] && S::c[i][1] == -mc[1] && S::c[i][2] == -mc[2]) m = i;
    pdf[target] = pdf[bc_out<S, PAT>(b, m, x + cdx + b.normal[0], y + cdy + b.normal[1], z + cdz + b.normal[2])];
  }
}

// ------------------------------------------------------------------ FP64 conserved sums
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__global__ void zero4_kernel(double* out) {
  if (threadIdx.x < 4) out[threadIdx.x] = 0.0;
}

template <class S, class TS>
__global__ void __launch_bounds__(256) reduce_kernel(const TS* pdf, int nx, int ny, int nz, int g, int64_t sy,
                                                     int64_t sz, int64_t sq, int zc, double* out) {
  constexpr int Q = S::Q;
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y, z = blockIdx.z;
  double r = 0, jx = 0, jy = 0, jz = 0;
  if (x < nx) {
    const int64_t c = (x + g) + (y + g) * sy + (S::D == 3 ? (z + g) * sz : 0);
#pragma unroll
    for (int i = 0; i < Q; ++i) {
      const double v = (double)pdf[i * sq + c];
      r += v;
      if (S::c[i][0]) jx += S::c[i][0] > 0 ? v : -v;
      if (S::c[i][1]) jy += S::c[i][1] > 0 ? v : -v;
      if (S::c[i][2]) jz += S::c[i][2] > 0 ? v : -v;
    }
    if (zc) r += 1.0;
  }
  r = wave_sum(r);
  jx = wave_sum(jx);
  jy = wave_sum(jy);
  jz = wave_sum(jz);
  __shared__ double part[4][4];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) {
    part[wid][0] = r;
    part[wid][1] = jx;
    part[wid][2] = jy;
    part[wid][3] = jz;
  }
  __syncthreads();
  if (threadIdx.x < 4) {
    const int nw = (blockDim.x + 63) >> 6;
    double s = 0;
    for (int w = 0; w < nw; ++w) s += part[w][threadIdx.x];
    atomicAdd(out + threadIdx.x, s);
  }
}

// Count of interior populations that are NaN or infinite (failure detection): per-wave ballot counts, one
// 64-bit atomic per wave.
__global__ void zero1_kernel(unsigned long long* out) {
  if (threadIdx.x == 0) *out = 0ull;
}

template <class S, class TS>
__global__ void __launch_bounds__(256) nonfinite_kernel(const TS* pdf, int nx, int g, int64_t sy, int64_t sz,
                                                        int64_t sq, unsigned long long* out) {
  constexpr int Q = S::Q;
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y, z = blockIdx.z;
  unsigned n = 0;
  if (x < nx) {
    const int64_t c = (x + g) + (y + g) * sy + (S::D == 3 ? (z + g) * sz : 0);
#pragma unroll
    for (int i = 0; i < Q; ++i) {
      const TS v = pdf[i * sq + c];
      bool fin;
      if constexpr (sizeof(TS) == 8)
        fin = __builtin_isfinite(v);
      else
        fin = __builtin_isfinite((float)v);   // exact for FP16 and FP32
      n += fin ? 0u : 1u;
    }
  }
  unsigned long long total = 0;
#pragma unroll
  for (int b = 0; b < 5; ++b) total += (unsigned long long)__popcll(__ballot(n & (1u << b))) << b;
  if ((threadIdx.x & 63) == 0 && total) atomicAdd(out, total);
}

}  // namespace lbm355

// ====================================================================== host side
using namespace lbm355;

namespace {
thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}
}  // namespace

int lbm355_set_error(int code, const char* msg) { return fail(code, msg); }

namespace {

template <class TC>
using KernelFn = void (*)(SweepArgs<TC>);

struct Launch {
  dim3 grid, block;
};

Launch geometry(int nx, int ny, int nz, bool rowlds = false) {
  int bx = rowlds ? nx : (nx >= 256 ? 256 : ((nx + 63) / 64) * 64);
  Launch l;
  l.block = dim3(bx, 1, 1);
  l.grid = dim3((nx + bx - 1) / bx, ny, nz);
  return l;
}

size_t elem_size(int dtype) { return dtype == LBM355_F64 ? 8 : (dtype == LBM355_F32 ? 4 : 2); }

int xshift_dirs(int stencil) { return stencil == LBM355_D2Q9 ? 6 : (stencil == LBM355_D3Q19 ? 10 : 18); }

// LDS row shifting applies when a block can span a whole periodic row
// Not for FP16 storage: with 2-byte populations a wavefront access moves 128 bytes, and the barriers around the LDS
// shift keep too few of them in flight.  The direct kernel, whose x-shifted loads are served partly from cache,
// ran the D3Q27 cumulant AA-odd sweep at 256^3 in 0.441 ms against 0.487 ms on MI355X.  With both at their best
// occupancy the LDS kernel's odd step is 3% faster (0.313 against 0.322 ms), but the even step after it 2% slower
// (0.322 against 0.316 ms), and D3Q19 FP16 loses 10% with it: the direct kernel stays.
bool use_rowlds(const lbm355_sweep_desc* d, size_t* lds_bytes) {
  if (d->store_dtype == LBM355_F16) return false;
  if (d->pattern == LBM355_AA_EVEN || table_pattern(d->pattern) || d->face[0] != LBM355_FACE_PERIODIC) return false;
  if (d->nx % 64 != 0 || d->nx > 512) return false;
  const size_t bytes = (size_t)xshift_dirs(d->stencil) * d->nx * elem_size(d->store_dtype);
  if (bytes > 65536) return false;
  *lds_bytes = bytes;
  return true;
}

// Dynamic LDS to request for a sweep launch: the kernel's own (kernel_lds), or more where a cap applies.
size_t launch_lds(const lbm355_sweep_desc* d, bool rowlds, bool walls, size_t kernel_lds, unsigned block_threads) {
  return lds_for_wave_cap(sweep_wave_cap(d, rowlds, walls, block_threads), block_threads, kernel_lds);
}

template <class TC>
void* pick_kernel(const lbm355_sweep_desc* d, bool walls, bool rowlds) {
  constexpr bool f64 = sizeof(TC) == 8;
  const int st = d->stencil, sd = d->store_dtype, p = d->pattern;
  if (d->family == LBM355_FAMILY_MOMENT)
    return f64 ? pick_sweep_moment_f64(st, sd, p, walls, rowlds) : pick_sweep_moment_f32(st, sd, p, walls, rowlds);
  if (d->family == LBM355_FAMILY_CENTRAL)
    return f64 ? pick_sweep_central_f64(st, sd, p, walls, rowlds) : pick_sweep_central_f32(st, sd, p, walls, rowlds);
  if (d->family == LBM355_FAMILY_CUMULANT) {
    if (d->flags & LBM355_FLAG_LEARNED_CLOSURE)
      return f64 ? nullptr
                 : (d->flags & LBM355_FLAG_HIGHER_ORDERS_RATE1)
                       ? pick_sweep_cumulant_learned_eq3_f32(st, sd, p, walls, rowlds)
                       : pick_sweep_cumulant_learned_f32(st, sd, p, walls, rowlds);
    if (d->flags & (LBM355_FLAG_GALILEAN | LBM355_FLAG_FOURTH_ORDER))
      return f64 ? pick_sweep_cumulant_corrected_f64(st, sd, p, walls, rowlds)
                 : pick_sweep_cumulant_corrected_f32(st, sd, p, walls, rowlds);
    if (d->flags & LBM355_FLAG_GENERAL_CUMULANT)
      return f64 ? pick_sweep_cumulant_general_f64(st, sd, p, walls, rowlds)
                 : pick_sweep_cumulant_general_f32(st, sd, p, walls, rowlds);
    if (d->flags & LBM355_FLAG_HIGHER_ORDERS_RATE1)
      return f64 ? pick_sweep_cumulant_eq3_f64(st, sd, p, walls, rowlds)
                 : pick_sweep_cumulant_eq3_f32(st, sd, p, walls, rowlds);
    return f64 ? pick_sweep_cumulant_f64(st, sd, p, walls, rowlds) : pick_sweep_cumulant_f32(st, sd, p, walls, rowlds);
  }
  return nullptr;
}

int stencil_q(int s) { return s == LBM355_D2Q9 ? 9 : (s == LBM355_D3Q19 ? 19 : (s == LBM355_D3Q27 ? 27 : 0)); }
int stencil_d(int s) { return s == LBM355_D2Q9 ? 2 : 3; }

int validate(const lbm355_sweep_desc* d) {
  if (!d) return fail(LBM355_ERR_ARG, "desc is NULL");
  const int Q = stencil_q(d->stencil);
  if (!Q) return fail(LBM355_ERR_ARG, "unknown stencil");
  const int D = stencil_d(d->stencil);
  if (d->nx <= 0 || d->ny <= 0 || d->nz <= 0) return fail(LBM355_ERR_ARG, "domain sizes must be positive");
  if (D == 2 && d->nz != 1) return fail(LBM355_ERR_ARG, "2-D stencil needs nz == 1");
  if (d->ghost_layers < 0 || d->ghost_layers > 1) return fail(LBM355_ERR_ARG, "ghost_layers must be 0 or 1");
  if (d->ny > 65535 || d->nz > 65535) return fail(LBM355_ERR_ARG, "ny and nz must be <= 65535");
  for (int a = 0; a < D; ++a) {
    if (d->face[a] < 0 || d->face[a] > 2) return fail(LBM355_ERR_ARG, "bad face mode");
    if (d->face[a] == LBM355_FACE_GHOST && d->ghost_layers < 1)
      return fail(LBM355_ERR_ARG, "FACE_GHOST needs a ghost layer");
  }
  const int64_t g2 = 2 * d->ghost_layers;
  if (d->stride_y < d->nx + g2) return fail(LBM355_ERR_ARG, "stride_y smaller than a row");
  if (D == 3 && d->stride_z < d->stride_y * (d->ny + g2)) return fail(LBM355_ERR_ARG, "stride_z smaller than a plane");
  const int64_t plane = D == 3 ? d->stride_z * (d->nz + g2) : d->stride_y * (d->ny + g2);
  if (d->stride_q < plane) return fail(LBM355_ERR_ARG, "stride_q smaller than one population");
  if (plane >= (int64_t(1) << 31)) return fail(LBM355_ERR_ARG, "one population must hold < 2^31 elements");
  if (d->compute_dtype != LBM355_F64 && d->compute_dtype != LBM355_F32)
    return fail(LBM355_ERR_ARG, "compute dtype must be F64 or F32");
  if (d->store_dtype < 0 || d->store_dtype > 2) return fail(LBM355_ERR_ARG, "bad store dtype");
  if (d->pattern < LBM355_PULL || d->pattern > LBM355_PUSH) return fail(LBM355_ERR_ARG, "bad pattern");
  if (table_pattern(d->pattern)) {
    for (int a = 0; a < D; ++a)
      if (d->face[a] == LBM355_FACE_NOSLIP)
        return fail(LBM355_ERR_UNSUPPORTED, "Eso and push patterns take walls as boundary links, not fused NoSlip faces");
    if (d->flags & LBM355_FLAG_LEARNED_CLOSURE)
      return fail(LBM355_ERR_UNSUPPORTED, "the learned closure runs with pull or AA streaming");
  }
  if (!d->tables) return fail(LBM355_ERR_ARG, "method table pointer is NULL");
  if (d->waves_per_simd < -1 || d->waves_per_simd > 8)
    return fail(LBM355_ERR_ARG, "waves_per_simd must be -1 (no cap), 0 (the library's table) or 1..8");
  if ((d->family == LBM355_FAMILY_MOMENT || d->family == LBM355_FAMILY_CENTRAL) &&
      (d->n_active < 0 || d->n_active > Q))
    return fail(LBM355_ERR_ARG, "n_active out of range");
  return LBM355_OK;
}

template <class TC>
int launch_sweep(const lbm355_sweep_desc* d, const void* src, void* dst, void* rho, void* u, hipStream_t s) {
  if (table_pattern(d->pattern)) return lbm355_launch_sweep_eso(d, src, dst, rho, u, s);
  bool walls = false;
  for (int a = 0; a < stencil_d(d->stencil); ++a) walls |= d->face[a] == LBM355_FACE_NOSLIP;
  size_t lds = 0;
  const bool rowlds = use_rowlds(d, &lds);
  void* k = pick_kernel<TC>(d, walls, rowlds);
  if (!k) return fail(LBM355_ERR_UNSUPPORTED, "no kernel for this stencil/family/dtype combination");
  SweepArgs<TC> a = make_args<TC>(d, src, dst, rho, u);
  Launch l = geometry(d->nx, d->ny, d->nz, rowlds);
  hipLaunchKernelGGL((KernelFn<TC>)k, l.grid, l.block, launch_lds(d, rowlds, walls, rowlds ? lds : 0, l.block.x), s, a);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(LBM355_ERR_LAUNCH, std::string("launch failed: ") + hipGetErrorString(e));
  return LBM355_OK;
}

template <class S, class T>
int launch_links(int pattern, void* pdf, int64_t sy, int64_t sz, int64_t sq, const int32_t* links, int64_t n,
                 hipStream_t s) {
  const int bs = 256;
  const int64_t nb = (n + bs - 1) / bs;
  if (pattern == LBM355_PULL)
    hipLaunchKernelGGL((noslip_links_kernel<S, LBM355_PULL, T>), dim3(nb), dim3(bs), 0, s, (T*)pdf, sy, sz, sq, links, n);
  else if (pattern == LBM355_AA_EVEN)
    hipLaunchKernelGGL((noslip_links_kernel<S, LBM355_AA_EVEN, T>), dim3(nb), dim3(bs), 0, s, (T*)pdf, sy, sz, sq, links,
                       n);
  else
    hipLaunchKernelGGL((noslip_links_kernel<S, LBM355_AA_ODD, T>), dim3(nb), dim3(bs), 0, s, (T*)pdf, sy, sz, sq, links,
                       n);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(LBM355_ERR_LAUNCH, std::string("launch failed: ") + hipGetErrorString(e));
  return LBM355_OK;
}

template <class S>
int links_dtype(int dtype, int pattern, void* pdf, int64_t sy, int64_t sz, int64_t sq, const int32_t* links,
                int64_t n, hipStream_t s) {
  if (dtype == LBM355_F64) return launch_links<S, double>(pattern, pdf, sy, sz, sq, links, n, s);
  if (dtype == LBM355_F32) return launch_links<S, float>(pattern, pdf, sy, sz, sq, links, n, s);
  if (dtype == LBM355_F16) return launch_links<S, _Float16>(pattern, pdf, sy, sz, sq, links, n, s);
  return fail(LBM355_ERR_ARG, "bad dtype");
}

template <class S, int PAT, class T>
int launch_bc_kind(const lbm355_boundary_desc* d, T* pdf, const int32_t* links, int64_t n, double* vel,
                   const BoundaryArgs& b, hipStream_t s) {
  const int grid = (int)((n + 255) / 256);
  if (d->kind == LBM355_BC_NOSLIP)
    hipLaunchKernelGGL((boundary_links_kernel<S, PAT, T, LBM355_BC_NOSLIP>), dim3(grid), dim3(256), 0, s, pdf, links, n,
                       vel, b);
  else if (d->kind == LBM355_BC_UBB)
    hipLaunchKernelGGL((boundary_links_kernel<S, PAT, T, LBM355_BC_UBB>), dim3(grid), dim3(256), 0, s, pdf, links, n, vel,
                       b);
  else if (d->kind == LBM355_BC_FIXED_DENSITY)
    hipLaunchKernelGGL((boundary_links_kernel<S, PAT, T, LBM355_BC_FIXED_DENSITY>), dim3(grid), dim3(256), 0, s, pdf,
                       links, n, vel, b);
  else if (d->kind == LBM355_BC_OUTFLOW)
    hipLaunchKernelGGL((boundary_links_kernel<S, PAT, T, LBM355_BC_OUTFLOW>), dim3(grid), dim3(256), 0, s, pdf, links,
                       n, vel, b);
  else if (d->kind == LBM355_BC_OUTFLOW_INIT)
    hipLaunchKernelGGL((boundary_links_kernel<S, PAT, T, LBM355_BC_OUTFLOW_INIT>), dim3(grid), dim3(256), 0, s, pdf,
                       links, n, vel, b);
  else if (d->kind == LBM355_BC_SIMPLE_OUTFLOW)
    hipLaunchKernelGGL((boundary_links_kernel<S, PAT, T, LBM355_BC_SIMPLE_OUTFLOW>), dim3(grid), dim3(256), 0, s, pdf,
                       links, n, vel, b);
  else if (d->kind == LBM355_BC_FREESLIP_LINKS)
    hipLaunchKernelGGL((boundary_links_kernel<S, PAT, T, LBM355_BC_FREESLIP_LINKS>), dim3(grid), dim3(256), 0, s, pdf,
                       links, n, vel, b);
  else
    hipLaunchKernelGGL((boundary_links_kernel<S, PAT, T, LBM355_BC_FREESLIP>), dim3(grid), dim3(256), 0, s, pdf, links,
                       n, vel, b);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(LBM355_ERR_LAUNCH, std::string("launch failed: ") + hipGetErrorString(e));
  return LBM355_OK;
}

template <class S, class T>
int launch_bc(const lbm355_boundary_desc* d, void* pdf, const int32_t* links, int64_t n, double* vel,
              const BoundaryArgs& b, hipStream_t s) {
#define LBM_BC(P) \
  case P:          \
    return launch_bc_kind<S, P, T>(d, (T*)pdf, links, n, vel, b, s);
  switch (d->pattern) {
    LBM_BC(LBM355_PULL)
    LBM_BC(LBM355_AA_EVEN)
    LBM_BC(LBM355_AA_ODD)
    LBM_BC(LBM355_ESOTWIST_EVEN)
    LBM_BC(LBM355_ESOTWIST_ODD)
    LBM_BC(LBM355_ESOPULL_EVEN)
    LBM_BC(LBM355_ESOPULL_ODD)
    LBM_BC(LBM355_ESOPUSH_EVEN)
    LBM_BC(LBM355_ESOPUSH_ODD)
    LBM_BC(LBM355_PUSH)
  }
#undef LBM_BC
  return fail(LBM355_ERR_ARG, "bad pattern");
}

template <class S>
int launch_bc_dtype(const lbm355_boundary_desc* d, void* pdf, const int32_t* links, int64_t n, double* vel,
                    const BoundaryArgs& b, hipStream_t s) {
  if (d->store_dtype == LBM355_F64) return launch_bc<S, double>(d, pdf, links, n, vel, b, s);
  if (d->store_dtype == LBM355_F32) return launch_bc<S, float>(d, pdf, links, n, vel, b, s);
  return launch_bc<S, _Float16>(d, pdf, links, n, vel, b, s);
}

template <class S, class TS>
int launch_reduce(const lbm355_sweep_desc* d, const void* pdf, double* out, hipStream_t s) {
  hipLaunchKernelGGL(zero4_kernel, dim3(1), dim3(64), 0, s, out);
  Launch l = geometry(d->nx, d->ny, d->nz);
  hipLaunchKernelGGL((reduce_kernel<S, TS>), l.grid, l.block, 0, s, (const TS*)pdf, d->nx, d->ny, d->nz,
                     d->ghost_layers, d->stride_y, d->stride_z, d->stride_q,
                     (d->flags & LBM355_FLAG_ZERO_CENTERED) ? 1 : 0, out);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(LBM355_ERR_LAUNCH, std::string("launch failed: ") + hipGetErrorString(e));
  return LBM355_OK;
}

template <class S, class TS>
int launch_nonfinite(const lbm355_sweep_desc* d, const void* pdf, unsigned long long* out, hipStream_t s) {
  hipLaunchKernelGGL(zero1_kernel, dim3(1), dim3(64), 0, s, out);
  Launch l = geometry(d->nx, d->ny, d->nz);
  hipLaunchKernelGGL((nonfinite_kernel<S, TS>), l.grid, l.block, 0, s, (const TS*)pdf, d->nx, d->ghost_layers,
                     d->stride_y, d->stride_z, d->stride_q, out);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(LBM355_ERR_LAUNCH, std::string("launch failed: ") + hipGetErrorString(e));
  return LBM355_OK;
}

template <class S>
int nonfinite_dtype(const lbm355_sweep_desc* d, const void* pdf, unsigned long long* out, hipStream_t s) {
  if (d->store_dtype == LBM355_F64) return launch_nonfinite<S, double>(d, pdf, out, s);
  if (d->store_dtype == LBM355_F32) return launch_nonfinite<S, float>(d, pdf, out, s);
  if (d->store_dtype == LBM355_F16) return launch_nonfinite<S, _Float16>(d, pdf, out, s);
  return fail(LBM355_ERR_ARG, "bad dtype");
}

template <class S>
int reduce_dtype(const lbm355_sweep_desc* d, const void* pdf, double* out, hipStream_t s) {
  if (d->store_dtype == LBM355_F64) return launch_reduce<S, double>(d, pdf, out, s);
  if (d->store_dtype == LBM355_F32) return launch_reduce<S, float>(d, pdf, out, s);
  if (d->store_dtype == LBM355_F16) return launch_reduce<S, _Float16>(d, pdf, out, s);
  return fail(LBM355_ERR_ARG, "bad dtype");
}
}  // namespace

int lbm355_validate_desc(const lbm355_sweep_desc* d) { return validate(d); }

extern "C" {

int lbm355_sweep(const lbm355_sweep_desc* d, const void* pdf_src, void* pdf_dst, void* density_out,
                 void* velocity_out, void* stream) {
  int rc = validate(d);
  if (rc) return rc;
  if (!pdf_src || !pdf_dst) return fail(LBM355_ERR_ARG, "PDF pointers must not be NULL");
  const bool two_arrays = d->pattern == LBM355_PULL || d->pattern == LBM355_PUSH;
  if (!two_arrays && pdf_src != pdf_dst)
    return fail(LBM355_ERR_ARG, "in-place patterns (AA, Eso) need pdf_src == pdf_dst");
  if (two_arrays && pdf_src == pdf_dst) return fail(LBM355_ERR_ARG, "pull and push need two arrays");
  if (d->family == LBM355_FAMILY_CUMULANT && d->stencil != LBM355_D3Q27)
    return fail(LBM355_ERR_UNSUPPORTED, "cumulant sweep is implemented for D3Q27");
  if ((d->flags & LBM355_FLAG_SMAGORINSKY) && d->family != LBM355_FAMILY_CUMULANT)
    return fail(LBM355_ERR_UNSUPPORTED, "Smagorinsky is implemented for the cumulant sweep");
  if (d->flags & (LBM355_FLAG_GALILEAN | LBM355_FLAG_FOURTH_ORDER)) {
    if (d->family != LBM355_FAMILY_CUMULANT)
      return fail(LBM355_ERR_UNSUPPORTED, "the Galilean / fourth-order corrections are D3Q27 cumulant options");
    if (d->flags & LBM355_FLAG_LEARNED_CLOSURE)
      return fail(LBM355_ERR_UNSUPPORTED, "the cumulant corrections are not combined with the learned closure");
    if ((d->flags & LBM355_FLAG_FOURTH_ORDER) && (d->flags & LBM355_FLAG_SMAGORINSKY))
      return fail(LBM355_ERR_UNSUPPORTED, "the fourth-order correction is not combined with the Smagorinsky model");
  }
  if (d->flags & LBM355_FLAG_LEARNED_CLOSURE) {
    if (d->family != LBM355_FAMILY_CUMULANT || d->compute_dtype != LBM355_F32 || d->store_dtype != LBM355_F32)
      return fail(LBM355_ERR_UNSUPPORTED, "the learned closure runs in the FP32 cumulant sweep (FP32 storage)");
    if (!d->closure) return fail(LBM355_ERR_ARG, "closure parameter pointer is NULL");
    if (d->nx % 64) return fail(LBM355_ERR_UNSUPPORTED, "the learned closure needs nx % 64 == 0 (whole wavefronts)");
  }
  if ((d->flags & LBM355_FLAG_GENERAL_CUMULANT) &&
      (d->family != LBM355_FAMILY_CUMULANT ||
       (d->flags & (LBM355_FLAG_SMAGORINSKY | LBM355_FLAG_LEARNED_CLOSURE | LBM355_FLAG_GALILEAN |
                    LBM355_FLAG_FOURTH_ORDER))))
    return fail(LBM355_ERR_UNSUPPORTED, "cumulant sets other than the default polynomial one run without subgrid "
                                        "models and corrections");
  if ((d->flags & LBM355_FLAG_HIGHER_ORDERS_RATE1) &&
      (d->family != LBM355_FAMILY_CUMULANT ||
       (d->flags & (LBM355_FLAG_GENERAL_CUMULANT | LBM355_FLAG_GALILEAN | LBM355_FLAG_FOURTH_ORDER))))
    return fail(LBM355_ERR_UNSUPPORTED, "LBM355_FLAG_HIGHER_ORDERS_RATE1 is for the default cumulant set without "
                                        "the corrections");
  if ((d->flags & LBM355_FLAG_SHEAR_PROJECTOR) &&
      !((d->family == LBM355_FAMILY_MOMENT || d->family == LBM355_FAMILY_CENTRAL) && d->stencil != LBM355_D2Q9))
    return fail(LBM355_ERR_UNSUPPORTED, "LBM355_FLAG_SHEAR_PROJECTOR is for the moment and central-moment families on "
                                        "3-D stencils");
  if ((d->flags & LBM355_FLAG_GAUSSIAN_EQUILIBRIUM) &&
      !(d->family == LBM355_FAMILY_CENTRAL && d->stencil != LBM355_D2Q9))
    return fail(LBM355_ERR_UNSUPPORTED, "LBM355_FLAG_GAUSSIAN_EQUILIBRIUM is for the central-moment family on 3-D "
                                        "stencils");
  if ((d->flags & LBM355_FLAG_STOCHASTIC_ROUNDING) && (d->store_dtype != LBM355_F16 || d->compute_dtype != LBM355_F32))
    return fail(LBM355_ERR_ARG, "stochastic rounding is for FP16 storage with FP32 arithmetic");
  if ((d->omega_out || d->eddy_viscosity_out) && !(d->flags & LBM355_FLAG_SMAGORINSKY))
    return fail(LBM355_ERR_ARG, "the shear-rate and eddy-viscosity outputs belong to the Smagorinsky model");
  if (d->density_in && !d->velocity_in)
    return fail(LBM355_ERR_ARG, "a density input needs a velocity input as well");
  if (d->velocity_in && d->family != LBM355_FAMILY_MOMENT)
    return fail(LBM355_ERR_UNSUPPORTED, "density / velocity inputs are implemented for the moment family");
  if (density_out || velocity_out || d->omega_out || d->eddy_viscosity_out || d->velocity_in) {
    // the per-cell fields span the PDF array's spatial extent (ghost layers included) with x stride 1
    const int D = stencil_d(d->stencil);
    const int64_t g2 = 2 * d->ghost_layers;
    if (d->out_stride_y < d->nx + g2) return fail(LBM355_ERR_ARG, "out_stride_y smaller than a row");
    if (D == 3 && d->out_stride_z < d->out_stride_y * (d->ny + g2))
      return fail(LBM355_ERR_ARG, "out_stride_z smaller than a plane");
    const int64_t volume = D == 3 ? d->out_stride_z * (d->nz + g2) : d->out_stride_y * (d->ny + g2);
    if ((velocity_out || d->velocity_in) && d->out_stride_c < volume)
      return fail(LBM355_ERR_ARG, "out_stride_c smaller than one component");
  }
  hipStream_t s = (hipStream_t)stream;
  if (d->compute_dtype == LBM355_F64) return launch_sweep<double>(d, pdf_src, pdf_dst, density_out, velocity_out, s);
  return launch_sweep<float>(d, pdf_src, pdf_dst, density_out, velocity_out, s);
}

int64_t lbm355_table_elements(int32_t stencil, int32_t family) {
  const int Q = stencil_q(stencil);
  if (!Q) return fail(LBM355_ERR_ARG, "unknown stencil"), -1;
  if (family == LBM355_FAMILY_MOMENT)
    return Q == 9 ? MomentTab<9>::SIZE : (Q == 19 ? MomentTab<19>::SIZE : MomentTab<27>::SIZE);
  if (family == LBM355_FAMILY_CUMULANT) {
    if (stencil != LBM355_D3Q27) return fail(LBM355_ERR_UNSUPPORTED, "the cumulant sweep is D3Q27 only"), -1;
    return CumulantTab::SIZE;
  }
  if (family == LBM355_FAMILY_CENTRAL) return CentralTab::SIZE;
  return fail(LBM355_ERR_ARG, "unknown collision family"), -1;
}

int lbm355_noslip_links(int32_t stencil, int32_t pattern, int32_t dtype, void* pdf, int64_t stride_y,
                        int64_t stride_z, int64_t stride_q, const int32_t* links, int64_t n_links, void* stream) {
  if (n_links == 0) return LBM355_OK;
  if (!pdf || !links || n_links < 0) return fail(LBM355_ERR_ARG, "bad pointer or link count");
  if (pattern < 0 || pattern > 2) return fail(LBM355_ERR_ARG, "bad pattern");
  hipStream_t s = (hipStream_t)stream;
  if (stencil == LBM355_D2Q9) return links_dtype<D2Q9>(dtype, pattern, pdf, stride_y, stride_z, stride_q, links, n_links, s);
  if (stencil == LBM355_D3Q19)
    return links_dtype<D3Q19>(dtype, pattern, pdf, stride_y, stride_z, stride_q, links, n_links, s);
  if (stencil == LBM355_D3Q27)
    return links_dtype<D3Q27>(dtype, pattern, pdf, stride_y, stride_z, stride_q, links, n_links, s);
  return fail(LBM355_ERR_ARG, "unknown stencil");
}

int lbm355_boundary_links(const lbm355_boundary_desc* d, void* pdf, const int32_t* links, int64_t n_links,
                          double* link_velocity, void* stream) {
  if (!d) return fail(LBM355_ERR_ARG, "desc is NULL");
  if (n_links == 0) return LBM355_OK;
  if (!pdf || !links || n_links < 0) return fail(LBM355_ERR_ARG, "bad pointer or link count");
  if (d->pattern < LBM355_PULL || d->pattern > LBM355_PUSH) return fail(LBM355_ERR_ARG, "bad pattern");
  if (table_pattern(d->pattern) && (d->fused_noslip[0] || d->fused_noslip[1] || d->fused_noslip[2]))
    return fail(LBM355_ERR_ARG, "fused NoSlip faces exist with pull and AA streaming only");
  if (d->store_dtype < 0 || d->store_dtype > 2) return fail(LBM355_ERR_ARG, "bad dtype");
  if (d->kind < LBM355_BC_NOSLIP || d->kind > LBM355_BC_SIMPLE_OUTFLOW) return fail(LBM355_ERR_ARG, "bad boundary kind");
  if (d->kind == LBM355_BC_FREESLIP_LINKS && !link_velocity)
    return fail(LBM355_ERR_ARG, "FreeSlip with per-link normals needs its link data (normal, reference direction)");
  if ((d->kind == LBM355_BC_OUTFLOW || d->kind == LBM355_BC_OUTFLOW_INIT) && !link_velocity)
    return fail(LBM355_ERR_ARG, "the outflow boundary needs its per-link state array");
  if (d->kind == LBM355_BC_FIXED_DENSITY && (!d->equilibrium || !(d->density > 0)))
    return fail(LBM355_ERR_ARG, "FixedDensity needs a positive density and the equilibrium table");
  const int D = stencil_d(d->stencil);
  if (!stencil_q(d->stencil)) return fail(LBM355_ERR_ARG, "unknown stencil");
  if (d->nx <= 0 || d->ny <= 0 || d->nz <= 0 || d->ghost_layers < 0) return fail(LBM355_ERR_ARG, "bad sizes");
  BoundaryArgs b;
  b.sy = d->stride_y;
  b.sz = d->stride_z;
  b.sq = d->stride_q;
  b.n[0] = d->nx;
  b.n[1] = d->ny;
  b.n[2] = d->nz;
  b.g = d->ghost_layers;
  for (int a = 0; a < 3; ++a) {
    b.periodic[a] = (a < D) ? (d->periodic[a] != 0) : 0;
    b.fused[a] = (a < D) ? (d->fused_noslip[a] != 0) : 0;
    b.normal[a] = (a < D) ? d->normal[a] : 0;
    b.velocity[a] = d->velocity[a];
  }
  b.zc = (d->flags & LBM355_FLAG_ZERO_CENTERED) ? 1 : 0;
  b.comp = (d->flags & LBM355_FLAG_COMPRESSIBLE) ? 1 : 0;
  b.density = d->density;
  b.eq_shift = (d->flags & LBM355_FLAG_EQ_SHIFT) ? 1 : 0;
  b.adapt = (d->flags & LBM355_FLAG_ADAPT_VELOCITY) ? 1 : 0;
  for (int a = 0; a < 3; ++a) b.force[a] = (a < D) ? d->force[a] : 0.0;
  b.eq = d->equilibrium;
  b.mirror_axis = 0;
  if (d->kind == LBM355_BC_FREESLIP) {
    int nz = 0;
    for (int a = 0; a < D; ++a)
      if (b.normal[a] != 0) {
        if (b.normal[a] != 1 && b.normal[a] != -1) return fail(LBM355_ERR_ARG, "FreeSlip normal entries must be -1, 0, 1");
        b.mirror_axis = a;
        ++nz;
      }
    if (nz != 1) return fail(LBM355_ERR_ARG, "FreeSlip needs a normal along one axis");
  }
  hipStream_t s = (hipStream_t)stream;
  if (d->stencil == LBM355_D2Q9) return launch_bc_dtype<D2Q9>(d, pdf, links, n_links, link_velocity, b, s);
  if (d->stencil == LBM355_D3Q19) return launch_bc_dtype<D3Q19>(d, pdf, links, n_links, link_velocity, b, s);
  return launch_bc_dtype<D3Q27>(d, pdf, links, n_links, link_velocity, b, s);
}

int lbm355_reduce_conserved(const lbm355_sweep_desc* d, const void* pdf, double* out, void* stream) {
  if (!d || !pdf || !out) return fail(LBM355_ERR_ARG, "NULL argument");
  hipStream_t s = (hipStream_t)stream;
  if (d->stencil == LBM355_D2Q9) return reduce_dtype<D2Q9>(d, pdf, out, s);
  if (d->stencil == LBM355_D3Q19) return reduce_dtype<D3Q19>(d, pdf, out, s);
  if (d->stencil == LBM355_D3Q27) return reduce_dtype<D3Q27>(d, pdf, out, s);
  return fail(LBM355_ERR_ARG, "unknown stencil");
}

int lbm355_count_nonfinite(const lbm355_sweep_desc* d, const void* pdf, uint64_t* out, void* stream) {
  if (!d || !pdf || !out) return fail(LBM355_ERR_ARG, "NULL argument");
  if (d->nx <= 0 || d->ny <= 0 || d->nz <= 0) return fail(LBM355_ERR_ARG, "bad interior size");
  hipStream_t s = (hipStream_t)stream;
  unsigned long long* o = reinterpret_cast<unsigned long long*>(out);
  if (d->stencil == LBM355_D2Q9) return nonfinite_dtype<D2Q9>(d, pdf, o, s);
  if (d->stencil == LBM355_D3Q19) return nonfinite_dtype<D3Q19>(d, pdf, o, s);
  if (d->stencil == LBM355_D3Q27) return nonfinite_dtype<D3Q27>(d, pdf, o, s);
  return fail(LBM355_ERR_ARG, "unknown stencil");
}

const char* lbm355_last_error(void) { return g_err.c_str(); }

int32_t lbm355_version(void) { return 9; }

#ifndef LBM355_SOURCE_HASH
#define LBM355_SOURCE_HASH "unknown"
#endif
const char* lbm355_source_hash(void) { return LBM355_SOURCE_HASH; }

}  // extern "C"
