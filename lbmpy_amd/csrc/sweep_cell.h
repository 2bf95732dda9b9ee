// Per-cell pieces of the stream-collide sweep shared by every kernel that updates cells: the
// single-domain sweeps (sweep.hip) and the batched ensemble sweep (batch.hip).  Both call
// collide_cell / sweep_cell, so a cell updated in a batch runs exactly the instructions of the
// same cell updated on its own.
#pragma once
#include <type_traits>

#include "lbm_common.h"
#include "collide_moment.h"
#include "collide_cumulant.h"
#include "collide_central.h"

namespace lbm355 {

template <class S, int FAM, class TC>
struct Collider;

template <class S, class TC>
struct Collider<S, LBM355_FAMILY_MOMENT, TC> {
  static constexpr bool deviation_form = false;
  template <bool NEQ>
  static __device__ __forceinline__ void run(TC (&f)[S::Q], const SweepArgs<TC>& a, int64_t o, TC& rho, TC (&u)[3],
                                             CellExtras<TC>&) {
    collide_moment<S, TC>(f, a, o, rho, u);
  }
};

// kernel-variant family code of the cumulant sweep with the learned closure (internal, not in the ABI)
constexpr int FAMILY_CUMULANT_LEARNED = 0x102;

template <class TC>
struct Collider<D3Q27, FAMILY_CUMULANT_LEARNED, TC> {
  static constexpr bool deviation_form = true;
  template <bool NEQ>
  static __device__ __forceinline__ void run(TC (&f)[27], const SweepArgs<TC>& a, int64_t, TC& rho, TC (&u)[3],
                                             CellExtras<TC>& ex) {
    collide_cumulant_d3q27<TC, true, false, NEQ>(f, a, rho, u, ex);
  }
};

// the learned-closure sweep when every cumulant of order >= 3 relaxes with rate 1
constexpr int FAMILY_CUMULANT_LEARNED_EQ3 = 0x106;

template <class TC>
struct Collider<D3Q27, FAMILY_CUMULANT_LEARNED_EQ3, TC> {
  static constexpr bool deviation_form = true;
  template <bool NEQ>
  static __device__ __forceinline__ void run(TC (&f)[27], const SweepArgs<TC>& a, int64_t, TC& rho, TC (&u)[3],
                                             CellExtras<TC>& ex) {
    collide_cumulant_d3q27<TC, true, false, NEQ, false, true>(f, a, rho, u, ex);
  }
};

// kernel-variant family code of the cumulant sweep with the Galilean / fourth-order corrections (internal)
constexpr int FAMILY_CUMULANT_CORRECTED = 0x103;

template <class TC>
struct Collider<D3Q27, FAMILY_CUMULANT_CORRECTED, TC> {
  static constexpr bool deviation_form = true;
  template <bool NEQ>
  static __device__ __forceinline__ void run(TC (&f)[27], const SweepArgs<TC>& a, int64_t, TC& rho, TC (&u)[3],
                                             CellExtras<TC>& ex) {
    collide_cumulant_d3q27<TC, false, true, NEQ>(f, a, rho, u, ex);
  }
};

// kernel-variant family code of the cumulant sweep for cumulant sets other than the default polynomial one
constexpr int FAMILY_CUMULANT_GENERAL = 0x104;

template <class TC>
struct Collider<D3Q27, FAMILY_CUMULANT_GENERAL, TC> {
  static constexpr bool deviation_form = true;
  template <bool NEQ>
  static __device__ __forceinline__ void run(TC (&f)[27], const SweepArgs<TC>& a, int64_t, TC& rho, TC (&u)[3],
                                             CellExtras<TC>& ex) {
    collide_cumulant_d3q27<TC, false, false, NEQ, true>(f, a, rho, u, ex);
  }
};

// kernel-variant family code of the cumulant sweep when every cumulant of order >= 3 relaxes with rate 1
constexpr int FAMILY_CUMULANT_EQ3 = 0x105;

template <class TC>
struct Collider<D3Q27, FAMILY_CUMULANT_EQ3, TC> {
  static constexpr bool deviation_form = true;
  template <bool NEQ>
  static __device__ __forceinline__ void run(TC (&f)[27], const SweepArgs<TC>& a, int64_t, TC& rho, TC (&u)[3],
                                             CellExtras<TC>& ex) {
    collide_cumulant_d3q27<TC, false, false, NEQ, false, true>(f, a, rho, u, ex);
  }
};

template <class TC>
struct Collider<D3Q27, LBM355_FAMILY_CUMULANT, TC> {
  static constexpr bool deviation_form = true;
  template <bool NEQ>
  static __device__ __forceinline__ void run(TC (&f)[27], const SweepArgs<TC>& a, int64_t, TC& rho, TC (&u)[3],
                                             CellExtras<TC>& ex) {
    collide_cumulant_d3q27<TC, false, false, NEQ>(f, a, rho, u, ex);
  }
};

// number of directions with a non-zero x component (those are shifted through LDS)
template <class S>
__host__ __device__ constexpr int n_xshift() {
  int n = 0;
  for (int i = 0; i < S::Q; ++i) n += S::c[i][0] != 0;
  return n;
}

// Non-temporal accesses for aligned, streamed-once data (each PDF is read and written exactly
// once per sweep); misaligned (x +- 1) accesses keep the default policy so that the partially
// used cache lines are served from cache to the neighbouring wavefront.
template <class TS, class TC>
__device__ __forceinline__ TC load_nt(const TS* p) {
  return (TC)__builtin_nontemporal_load(p);
}
template <class TS, class TC>
__device__ __forceinline__ void store_nt(TS* p, TC v) {
  __builtin_nontemporal_store((TS)v, p);
}

// ROWLDS: the block spans one whole row and x is periodic (blockDim.x == nx).  The +-1
// x-shifts of pull / AA-odd streaming are then done in LDS, so every global load and store
// is an aligned, contiguous, non-temporal 64-lane access:
//   read : tile[j][x] = row(y - c_y, z - c_z)[x];  f_i(x) = tile[j][(x - c_x) mod nx]
//   write: tile[j][x] = f_i(x);  row(y + c_y, z + c_z)[x] = tile[j][(x - c_x) mod nx]
template <class S, class TC>
struct Collider<S, LBM355_FAMILY_CENTRAL, TC> {
  static constexpr bool deviation_form = true;
  template <bool NEQ>
  static __device__ __forceinline__ void run(TC (&f)[S::Q], const SweepArgs<TC>& a, int64_t, TC& rho, TC (&u)[3],
                                             CellExtras<TC>&) {
    collide_central<S, TC>(f, a, rho, u);
  }
};

// 32-bit integer hash (xorshift-multiply): mixes a cell index with a launch seed
__device__ __forceinline__ uint32_t mix32(uint32_t v) {
  v ^= v >> 16;
  v *= 0x7FEB352Du;
  v ^= v >> 15;
  v *= 0x846CA68Bu;
  v ^= v >> 16;
  return v;
}

// Stochastic rounding of a cell's populations to FP16 values (kept in TC, so that the store's conversion is
// exact): each value goes to one of its two neighbouring FP16 numbers with probability proportional to its
// closeness, gfx950's v_cvt_sr_f16_f32, which compares the high bits of its random operand with the dropped
// mantissa bits (measured: a value 1/8, 2/8, ... 7/8 of an FP16 ulp above an FP16 number rounds up with those
// frequencies when bits 19..31 of the operand are uniform, and never when only bits 0..12 vary).  The operand of
// direction i is the low 24 bits of the cell's hash times an odd 24-bit constant of i (one full-rate v_mul_u32_u24),
// whose high bits are well mixed.
template <class S, class TC>
__device__ __forceinline__ void round_stochastic_f16(TC (&f)[S::Q], uint32_t seed, int x, int y, int z, int nx,
                                                     int ny) {
  typedef _Float16 half2_t __attribute__((ext_vector_type(2)));
  const uint32_t cell = (uint32_t)x + (uint32_t)nx * ((uint32_t)y + (uint32_t)ny * (uint32_t)z);
  const uint32_t h = mix32(cell * 0x9E3779B9u + seed);
#pragma unroll
  for (int i = 0; i < S::Q; ++i) {
    const uint32_t k = (0x5BD1E9u + 0x3C6EF2u * (uint32_t)i) & 0xFFFFFFu;   // odd (bit 0 of both terms: 1, 0)
    const uint32_t r = __umul24(h, k);
    const half2_t zero = {0, 0};
    const half2_t v = __builtin_amdgcn_cvt_sr_f16_f32(zero, (float)f[i], r, false);
    f[i] = (TC)(float)v[0];
  }
}

// Collision of one cell's populations plus the optional per-cell outputs (density, velocity, the Smagorinsky
// rate and eddy viscosity).  Shared by every sweep kernel (single domain, row-LDS and batched), so all of them
// run the same arithmetic.  TS is the storage type: FP16 storage may round stochastically.
template <class S, int FAM, class TC, class TS = TC>
__device__ __forceinline__ void collide_cell(TC (&f)[S::Q], const SweepArgs<TC>& a, int x, int y, int z) {
  constexpr int Q = S::Q, D = S::D;
  const bool absolute = !(a.flags & LBM355_FLAG_ZERO_CENTERED);
  if (Collider<S, FAM, TC>::deviation_form && absolute) {
#pragma unroll
    for (int i = 0; i < Q; ++i) f[i] -= TC(weight<S>(i));
  }
  // offset of the cell in the optional per-cell fields
  const int64_t o = (int64_t)(x + a.g) + (int64_t)(y + a.g) * a.osy + (D == 3 ? (int64_t)(z + a.g) * a.osz : 0);
  TC rho, u[3];
  CellExtras<TC> ex = {a.omega_shear, TC(0)};
  // the cumulant collision transforms f - f_eq for FP32 storage with FP32 arithmetic (collide_cumulant.h)
  constexpr bool NEQ = std::is_same<TC, float>::value && std::is_same<TS, float>::value;
  Collider<S, FAM, TC>::template run<NEQ>(f, a, o, rho, u, ex);
  if (Collider<S, FAM, TC>::deviation_form && absolute) {
#pragma unroll
    for (int i = 0; i < Q; ++i) f[i] += TC(weight<S>(i));
  }
  if (a.rho_out) a.rho_out[o] = rho;
  if (a.u_out) {
#pragma unroll
    for (int d = 0; d < D; ++d) a.u_out[o + d * a.osc] = u[d];
  }
  if (a.omega_out) a.omega_out[o] = ex.omega_sgs;
  if (a.nut_out) a.nut_out[o] = a.nut_scale * ex.omega_sgs * ex.pi;
  if constexpr (std::is_same<TS, _Float16>::value) {
    if (a.flags & LBM355_FLAG_STOCHASTIC_ROUNDING) round_stochastic_f16<S, TC>(f, a.seed, x, y, z, a.nx, a.ny);
  }
}

// Loads and stores of element `off` (0 <= off < 2^31) of a population plane.  With 2-byte storage the byte offset
// fits in 32 bits (a plane holds < 2^31 elements), so they are raw buffer accesses: the plane's base lives in a
// buffer descriptor that scalar instructions build, one 32-bit vector offset serves every plane, and no population
// needs 64-bit address arithmetic.  Wider storage uses ordinary pointers.  NT: non-temporal.
constexpr int BUFFER_DWORD3 = 0x00020000;   // raw (untyped) buffer access on gfx9 family targets

template <class TS, class TC, bool NT>
__device__ __forceinline__ TC plane_load(const TS* plane, int off) {
  if constexpr (sizeof(TS) == 2) {
    const __amdgpu_buffer_rsrc_t r =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<TS*>(plane), (short)0, -1, BUFFER_DWORD3);
    const unsigned short bits = __builtin_amdgcn_raw_buffer_load_b16(r, (int)((uint32_t)off * 2u), 0, NT ? 2 : 0);
    TS v;
    __builtin_memcpy(&v, &bits, 2);
    return (TC)v;
  } else {
    return NT ? load_nt<TS, TC>(plane + off) : load_as<TS, TC>(plane + off);
  }
}
template <class TS, class TC, bool NT>
__device__ __forceinline__ void plane_store(TS* plane, int off, TC value) {
  if constexpr (sizeof(TS) == 2) {
    const __amdgpu_buffer_rsrc_t r = __builtin_amdgcn_make_buffer_rsrc(plane, (short)0, -1, BUFFER_DWORD3);
    const TS v = (TS)value;
    unsigned short bits;
    __builtin_memcpy(&bits, &v, 2);
    __builtin_amdgcn_raw_buffer_store_b16(bits, r, (int)((uint32_t)off * 2u), 0, NT ? 2 : 0);
  } else {
    if (NT)
      store_nt<TS, TC>(plane + off, value);
    else
      store_as<TS, TC>(plane + off, value);
  }
}

// One cell of the sweep with direct (non-LDS) addressing: load with streaming, collide, store.
// With WALLS, only cells next to a NoSlip face take the bounce-back addressing; the test is per cell and
// the wall-free addressing is the same code as without WALLS, so interior cells pay nothing for walls.
template <class S, int FAM, int PAT, bool WALLS, class TS, class TC>
__device__ __forceinline__ void sweep_cell(const SweepArgs<TC>& a, int x, int y, int z) {
  constexpr int Q = S::Q, D = S::D;
  const AxisNb ax = axis_neighbours(x, a.nx, a.g, a.face[0]);
  const AxisNb ay = axis_neighbours(y, a.ny, a.g, a.face[1]);
  AxisNb az = {0, 0, false, false};
  if (D == 3) az = axis_neighbours(z, a.nz, a.g, a.face[2]);
  // in-plane element offsets of the 3 x positions, 3 rows and 3 planes around the cell
  const int sy = (int)a.sy, sz = (int)a.sz;
  const int X[3] = {ax.lo, x + a.g, ax.hi};
  const int Y[3] = {ay.lo * sy, (y + a.g) * sy, ay.hi * sy};
  const int Z[3] = {D == 3 ? az.lo * sz : 0, D == 3 ? (z + a.g) * sz : 0, D == 3 ? az.hi * sz : 0};
  const int center = X[1] + Y[1] + Z[1];
  const TS* __restrict__ src = reinterpret_cast<const TS*>(a.src);
  TS* __restrict__ dst = reinterpret_cast<TS*>(a.dst);
  const int64_t sq = a.sq;
  const bool near_wall =
      WALLS && (ax.wall_lo || ax.wall_hi || ay.wall_lo || ay.wall_hi || az.wall_lo || az.wall_hi);

  TC f[Q];
  if constexpr (table_pattern(PAT)) {
    // Eso patterns and push: every access from the pattern's table (periodic wrap or ghost layer in X/Y/Z).
    // In place, a cell reads exactly the locations it writes, so one thread per cell needs no synchronisation.
    // Accesses without an x offset are row-aligned and non-temporal.
    constexpr AccessTable<S, PAT> acc{};
#pragma unroll
    for (int i = 0; i < Q; ++i) {
      const Access& t = acc.a[i];
      const TS* p = src + t.rslot * sq + X[1 + t.r[0]] + Y[1 + t.r[1]] + Z[1 + t.r[2]];
      f[i] = t.r[0] == 0 ? load_nt<TS, TC>(p) : load_as<TS, TC>(p);
    }
    collide_cell<S, FAM, TC, TS>(f, a, x, y, z);
#pragma unroll
    for (int i = 0; i < Q; ++i) {
      const Access& t = acc.a[i];
      TS* p = dst + t.wslot * sq + X[1 + t.w[0]] + Y[1 + t.w[1]] + Z[1 + t.w[2]];
      if (t.w[0] == 0)
        store_nt<TS, TC>(p, f[i]);
      else
        store_as<TS, TC>(p, f[i]);
    }
    return;
  }
  // pull reads src[i] at x - c_i; AA odd reads slot inv(i) there.  At a wall the
  // bounce-back value sits in the cell itself: pull slot inv(i), AA odd slot i.
  auto load = [&](auto walls) {
    constexpr bool W = decltype(walls)::value;
#pragma unroll
    for (int i = 0; i < Q; ++i) {
      const int cx = S::c[i][0], cy = S::c[i][1], cz = S::c[i][2];
      const int ii = S::inv[i];
      const int plane = (PAT == LBM355_PULL) ? i : ii;
      const int wplane = (PAT == LBM355_PULL) ? ii : i;
      bool wall = false;
      if (W)
        wall = (cx == 1 && ax.wall_lo) || (cx == -1 && ax.wall_hi) || (cy == 1 && ay.wall_lo) ||
               (cy == -1 && ay.wall_hi) || (cz == 1 && az.wall_lo) || (cz == -1 && az.wall_hi);
      f[i] = wall ? plane_load<TS, TC, false>(src + wplane * sq, center)
                  : plane_load<TS, TC, false>(src + plane * sq, X[1 - cx] + Y[1 - cy] + Z[1 - cz]);
    }
  };
  // AA odd writes f_i to x + c_i slot i; at a wall it is bounced into the cell's slot inv(i)
  auto store = [&](auto walls) {
    constexpr bool W = decltype(walls)::value;
#pragma unroll
    for (int i = 0; i < Q; ++i) {
      const int cx = S::c[i][0], cy = S::c[i][1], cz = S::c[i][2];
      const int ii = S::inv[i];
      bool wall = false;
      if (W)
        wall = (cx == -1 && ax.wall_lo) || (cx == 1 && ax.wall_hi) || (cy == -1 && ay.wall_lo) ||
               (cy == 1 && ay.wall_hi) || (cz == -1 && az.wall_lo) || (cz == 1 && az.wall_hi);
      if (wall)
        plane_store<TS, TC, false>(dst + ii * sq, center, f[i]);
      else
        plane_store<TS, TC, false>(dst + i * sq, X[1 + cx] + Y[1 + cy] + Z[1 + cz], f[i]);
    }
  };

  if (PAT == LBM355_AA_EVEN) {
#pragma unroll
    for (int i = 0; i < Q; ++i) f[i] = plane_load<TS, TC, true>(src + i * sq, center);
  } else if (near_wall) {
    load(std::true_type{});
  } else {
    load(std::false_type{});
  }

  collide_cell<S, FAM, TC, TS>(f, a, x, y, z);

  if (PAT == LBM355_PULL) {
#pragma unroll
    for (int i = 0; i < Q; ++i) plane_store<TS, TC, true>(dst + i * sq, center, f[i]);
  } else if (PAT == LBM355_AA_EVEN) {
#pragma unroll
    for (int i = 0; i < Q; ++i) plane_store<TS, TC, true>(dst + S::inv[i] * sq, center, f[i]);
  } else if (near_wall) {
    store(std::true_type{});
  } else {
    store(std::false_type{});
  }
}

// Kernel arguments from a public descriptor (include/lbm355.h).
template <class TC>
inline SweepArgs<TC> make_args(const lbm355_sweep_desc* d, const void* src, void* dst, void* rho, void* u) {
  SweepArgs<TC> a;
  a.src = src;
  a.dst = dst;
  a.rho_out = (TC*)rho;
  a.u_out = (TC*)u;
  a.tab = (const TC*)d->tables;
  a.nx = d->nx;
  a.ny = d->ny;
  a.nz = d->nz;
  a.g = d->ghost_layers;
  a.sy = d->stride_y;
  a.sz = d->stride_z;
  a.sq = d->stride_q;
  a.osy = d->out_stride_y;
  a.osz = d->out_stride_z;
  a.osc = d->out_stride_c;
  for (int i = 0; i < 3; ++i) {
    a.face[i] = d->face[i];
    a.force[i] = (TC)d->force[i];
  }
  a.flags = d->flags;
  a.nactive = d->n_active;
  a.smag_cs2x18 = (TC)(18.0 * d->smagorinsky_constant * d->smagorinsky_constant);
  a.omega_shear = (TC)d->omega_shear;
  a.closure = d->closure;
  a.omega_out = (TC*)d->omega_out;
  a.nut_out = (TC*)d->eddy_viscosity_out;
  a.rho_in = (const TC*)d->density_in;
  a.u_in = (const TC*)d->velocity_in;
  a.nut_scale = (TC)(1.5 * d->smagorinsky_constant * d->smagorinsky_constant);
  a.seed = d->rounding_seed;
  return a;
}

}  // namespace lbm355

// checks a sweep descriptor (sizes, strides, dtypes, table pointer); defined in sweep.hip
int lbm355_validate_desc(const lbm355_sweep_desc* d);

// sweeps of the Eso patterns and push (sweep_eso.hip); the descriptor has been validated
int lbm355_launch_sweep_eso(const lbm355_sweep_desc* d, const void* src, void* dst, void* rho, void* u,
                            hipStream_t s);
