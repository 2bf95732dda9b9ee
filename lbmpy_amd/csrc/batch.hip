// Batched ensemble sweep: many independent domains (members) of one stencil / collision family /
// streaming pattern / storage type, each with its own sizes, faces, force and method table,
// advanced by ONE launch per time step.
//
// Work split: every member's interior cells, in memory order (x fastest), are cut into tiles of
// 64 cells, one wavefront each; tiles are numbered consecutively over the members.  The grid is
// persistent: wave w of G handles tiles w, w + G, w + 2G, ... until all are done, so any grid size
// covers any ensemble and every wave reaches the end of its loop.  A wave reads its tile's member from
// a tile -> member map (one wave-uniform load), decodes (x, y, z) with 32-bit arithmetic (a member
// holds < 2^31 cells) and runs sweep_cell, the same per-cell function as the single-domain sweep
// (sweep_cell.h): a member's cells get bit-identical results to a single-domain run of that member.
//
// Optional XCD-aware order: MI355X dispatches consecutive workgroups round-robin over its 8 XCDs
// (each with its own L2).  With `xcd_order`, the workgroups of one XCD take one contiguous
// stretch of the tile sequence, so the rows one tile's neighbour reads touch are fetched into the
// same L2 that the neighbouring tiles are processed on.
#include <mutex>
#include <string>
#include <vector>

#include "sweep_cell.h"

namespace lbm355 {

#define CONST_AS __attribute__((address_space(4)))
#define GLOBAL_AS __attribute__((address_space(1)))

// a generic PDF pointer re-tagged as global memory: accesses through it become global (not flat) ones
__device__ __forceinline__ void* global_space(void* p) { return (void*)(GLOBAL_AS char*)p; }

// field by field: a struct is not copied out of the constant address space as a whole
template <class TC>
__device__ __forceinline__ SweepArgs<TC> load_args(const CONST_AS SweepArgs<TC>* p) {
  SweepArgs<TC> a;
  a.src = p->src;
  a.dst = p->dst;
  a.rho_out = p->rho_out;
  a.u_out = p->u_out;
  a.tab = p->tab;
  a.nx = p->nx;
  a.ny = p->ny;
  a.nz = p->nz;
  a.g = p->g;
  a.sy = p->sy;
  a.sz = p->sz;
  a.sq = p->sq;
  a.osy = p->osy;
  a.osz = p->osz;
  a.osc = p->osc;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    a.face[k] = p->face[k];
    a.force[k] = p->force[k];
  }
  a.flags = p->flags;
  a.nactive = p->nactive;
  a.smag_cs2x18 = p->smag_cs2x18;
  a.omega_shear = p->omega_shear;
  a.closure = p->closure;
  // batch_create rejects members with per-cell auxiliary fields
  a.omega_out = nullptr;
  a.nut_out = nullptr;
  a.rho_in = nullptr;
  a.u_in = nullptr;
  a.nut_scale = TC(0);
  a.seed = 0;
  return a;
}

template <class TC>
struct BatchEntry {
  SweepArgs<TC> a;
  int64_t first_tile;   // index of this member's first tile
  int64_t ncells;       // interior cells
};

// FP32 occupancy target: 6 waves per SIMD for the moment family up to 19 directions (the compiler
// otherwise keeps more VGPRs than the argument loads and index decoding need; 80 hold everything without
// spills), 4 for the cumulant collision (128 VGPRs), 3 for the D3Q27 moment family and for the central-moment
// family beyond D2Q9 (its chimera transforms need up to 168 VGPRs); FP64 arithmetic takes what it needs.
// Every instance compiles without spills.
template <class S, int FAM, int PAT, class TS, class TC>
__host__ __device__ constexpr int batch_min_waves() {
  if (sizeof(TC) == 8) return 1;
  if (FAM == LBM355_FAMILY_MOMENT) return S::Q <= 19 ? 6 : 3;
  if (FAM == LBM355_FAMILY_CENTRAL) return S::Q <= 9 ? 4 : 3;
  return 4;
}

template <class S, int FAM, int PAT, class TS, class TC>
__global__ void __launch_bounds__(256)
__attribute__((amdgpu_waves_per_eu(batch_min_waves<S, FAM, PAT, TS, TC>(), 8)))
batch_kernel(const BatchEntry<TC>* __restrict__ members, const int32_t* __restrict__ tile_member, int64_t ntiles,
             int swap, int xcd_order) {
  const int waves_per_block = blockDim.x >> 6;
  int b = blockIdx.x;
  if (xcd_order && (gridDim.x & 7) == 0) b = (blockIdx.x & 7) * (gridDim.x >> 3) + (blockIdx.x >> 3);
  const int64_t wave = (int64_t)b * waves_per_block + (threadIdx.x >> 6);
  const int64_t nwaves = (int64_t)gridDim.x * waves_per_block;
  const int lane = threadIdx.x & 63;
  // The member table and the tile map stay unchanged during the launch: read through the constant
  // address space they become wave-uniform scalar loads, also after the stores of earlier tiles (the
  // collision reads the method tables the same way, table_ptr).  The PDF pointers are global.
  const CONST_AS BatchEntry<TC>* mem = (const CONST_AS BatchEntry<TC>*)members;
  const CONST_AS int32_t* map = (const CONST_AS int32_t*)tile_member;
  for (int64_t tile = wave; tile < ntiles; tile += nwaves) {
    const int m = __builtin_amdgcn_readfirstlane(map[tile]);     // tile is uniform over the wave
    const CONST_AS BatchEntry<TC>* e = mem + m;
    const uint32_t local = (uint32_t)(tile - e->first_tile) * 64u + (uint32_t)lane;
    if (local < (uint32_t)e->ncells) {
      SweepArgs<TC> a = load_args(&e->a);
      a.src = global_space(const_cast<void*>(a.src));
      a.dst = global_space(a.dst);
      if (swap) {
        void* t = const_cast<void*>(a.src);
        a.src = a.dst;
        a.dst = t;
      }
      const uint32_t nx = (uint32_t)a.nx, ny = (uint32_t)a.ny;
      const uint32_t row = local / nx;
      const int x = (int)(local - row * nx);
      const uint32_t zz = row / ny;
      const int y = (int)(row - zz * ny);
      sweep_cell<S, FAM, PAT, true, TS, TC>(a, x, y, (int)zz);
    }
  }
}

template <class TC>
using BatchFn = void (*)(const BatchEntry<TC>*, const int32_t*, int64_t, int, int);

template <class S, int FAM, class TS, class TC>
void* batch_pick_pattern(int pat) {
  switch (pat) {
    case LBM355_PULL: return (void*)batch_kernel<S, FAM, LBM355_PULL, TS, TC>;
    case LBM355_AA_EVEN: return (void*)batch_kernel<S, FAM, LBM355_AA_EVEN, TS, TC>;
    case LBM355_AA_ODD: return (void*)batch_kernel<S, FAM, LBM355_AA_ODD, TS, TC>;
  }
  return nullptr;
}

template <class S, int FAM, class TC>
void* batch_pick_store(int store, int pat) {
  if constexpr (sizeof(TC) == 8) {
    if (store == LBM355_F64) return batch_pick_pattern<S, FAM, double, TC>(pat);
  } else {
    if (store == LBM355_F32) return batch_pick_pattern<S, FAM, float, TC>(pat);
    if (store == LBM355_F16) return batch_pick_pattern<S, FAM, _Float16, TC>(pat);
  }
  return nullptr;
}

template <class TC>
void* batch_pick(const lbm355_sweep_desc* d, int pattern) {
  if (d->family == LBM355_FAMILY_MOMENT) {
    if (d->stencil == LBM355_D2Q9) return batch_pick_store<D2Q9, LBM355_FAMILY_MOMENT, TC>(d->store_dtype, pattern);
    if (d->stencil == LBM355_D3Q19) return batch_pick_store<D3Q19, LBM355_FAMILY_MOMENT, TC>(d->store_dtype, pattern);
    if (d->stencil == LBM355_D3Q27) return batch_pick_store<D3Q27, LBM355_FAMILY_MOMENT, TC>(d->store_dtype, pattern);
  } else if (d->family == LBM355_FAMILY_CUMULANT && d->stencil == LBM355_D3Q27) {
    if (d->flags & LBM355_FLAG_HIGHER_ORDERS_RATE1)
      return batch_pick_store<D3Q27, FAMILY_CUMULANT_EQ3, TC>(d->store_dtype, pattern);
    return batch_pick_store<D3Q27, LBM355_FAMILY_CUMULANT, TC>(d->store_dtype, pattern);
  } else if (d->family == LBM355_FAMILY_CENTRAL) {
    if (d->stencil == LBM355_D2Q9) return batch_pick_store<D2Q9, LBM355_FAMILY_CENTRAL, TC>(d->store_dtype, pattern);
    if (d->stencil == LBM355_D3Q19) return batch_pick_store<D3Q19, LBM355_FAMILY_CENTRAL, TC>(d->store_dtype, pattern);
    if (d->stencil == LBM355_D3Q27) return batch_pick_store<D3Q27, LBM355_FAMILY_CENTRAL, TC>(d->store_dtype, pattern);
  }
  return nullptr;
}

}  // namespace lbm355

using namespace lbm355;

struct lbm355_batch {
  int32_t stencil, family, store_dtype, compute_dtype, pattern_kind;  // pattern_kind: PULL or AA
  int32_t variant_flags;  // the kernel-selecting flags the members share (LBM355_FLAG_HIGHER_ORDERS_RATE1)
  int32_t nmembers;
  int64_t ntiles;
  int64_t ncells;
  void* entries;          // device array of BatchEntry<TC>
  int32_t* tile_member;   // device array: member of each 64-cell tile
  int device;
};

namespace {
int berr(int code, const std::string& msg) { return lbm355_set_error(code, msg.c_str()); }

// The member table: each member's kernel arguments, its first tile and its cell count.  Returns the tile and
// cell totals through `tiles` / `cells`.
template <class TC>
int fill_entries(const lbm355_sweep_desc* d, const void* const* src, void* const* dst, int32_t n, BatchEntry<TC>* out,
                 int64_t* tiles_out, int64_t* cells_out) {
  int64_t tiles = 0, cells = 0;
  for (int k = 0; k < n; ++k) {
    out[k].a = make_args<TC>(&d[k], src[k], dst[k], nullptr, nullptr);
    out[k].first_tile = tiles;
    out[k].ncells = (int64_t)d[k].nx * d[k].ny * d[k].nz;
    if (out[k].ncells >= (int64_t(1) << 31)) return berr(LBM355_ERR_ARG, "a batch member must hold < 2^31 cells");
    tiles += (out[k].ncells + 63) / 64;
    cells += out[k].ncells;
  }
  *tiles_out = tiles;
  *cells_out = cells;
  return LBM355_OK;
}

template <class TC>
int upload(const lbm355_sweep_desc* d, const void* const* src, void* const* dst, int32_t n, lbm355_batch* b) {
  std::vector<BatchEntry<TC>> host(n);
  int64_t tiles = 0, cells = 0;
  int rc = fill_entries<TC>(d, src, dst, n, host.data(), &tiles, &cells);
  if (rc) return rc;
  std::vector<int32_t> map;
  map.reserve((size_t)tiles);
  for (int k = 0; k < n; ++k) map.insert(map.end(), (size_t)((host[k].ncells + 63) / 64), k);
  void* dev = nullptr;
  void* dmap = nullptr;
  hipError_t e = hipMalloc(&dev, sizeof(BatchEntry<TC>) * (size_t)n);
  if (e == hipSuccess) e = hipMalloc(&dmap, sizeof(int32_t) * map.size());
  if (e == hipSuccess) e = hipMemcpy(dev, host.data(), sizeof(BatchEntry<TC>) * (size_t)n, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(dmap, map.data(), sizeof(int32_t) * map.size(), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    if (dev) (void)hipFree(dev);
    if (dmap) (void)hipFree(dmap);
    return berr(LBM355_ERR_LAUNCH, std::string("batch table upload failed: ") + hipGetErrorString(e));
  }
  b->entries = dev;
  b->tile_member = (int32_t*)dmap;
  b->ntiles = tiles;
  b->ncells = cells;
  return LBM355_OK;
}
}  // namespace

extern "C" {

}  // extern "C"

namespace {
// The members must share stencil, family, pattern kind and dtypes, and use only what the batched kernel runs.
// Sets *kind to LBM355_PULL or LBM355_AA_EVEN.
int check_members(const lbm355_sweep_desc* descs, const void* const* pdf_src, void* const* pdf_dst, int32_t n_members,
                  int* kind) {
  if (!descs || !pdf_src || !pdf_dst || n_members <= 0) return berr(LBM355_ERR_ARG, "bad batch arguments");
  const lbm355_sweep_desc& d0 = descs[0];
  const int kind0 = d0.pattern == LBM355_PULL ? LBM355_PULL : LBM355_AA_EVEN;
  *kind = kind0;
  for (int k = 0; k < n_members; ++k) {
    const lbm355_sweep_desc& d = descs[k];
    int rc = lbm355_validate_desc(&d);
    if (rc) return rc;
    if (d.stencil != d0.stencil || d.family != d0.family || d.store_dtype != d0.store_dtype ||
        d.compute_dtype != d0.compute_dtype || (d.pattern == LBM355_PULL ? LBM355_PULL : LBM355_AA_EVEN) != kind0)
      return berr(LBM355_ERR_ARG, "batch members must share stencil, family, streaming pattern and dtypes");
    if ((d.flags ^ d0.flags) & LBM355_FLAG_HIGHER_ORDERS_RATE1)
      return berr(LBM355_ERR_ARG, "batch members must agree on LBM355_FLAG_HIGHER_ORDERS_RATE1");
    if (d.flags & (LBM355_FLAG_SMAGORINSKY | LBM355_FLAG_LEARNED_CLOSURE))
      return berr(LBM355_ERR_UNSUPPORTED, "subgrid models are not batched");
    if (table_pattern(d.pattern)) return berr(LBM355_ERR_UNSUPPORTED, "batched ensembles stream with pull or AA");
    if (d.flags & (LBM355_FLAG_GALILEAN | LBM355_FLAG_FOURTH_ORDER))
      return berr(LBM355_ERR_UNSUPPORTED, "the cumulant corrections are not batched");
    if (d.flags & LBM355_FLAG_STOCHASTIC_ROUNDING)
      return berr(LBM355_ERR_UNSUPPORTED, "stochastic rounding is not batched");
    if (d.flags & LBM355_FLAG_GENERAL_CUMULANT)
      return berr(LBM355_ERR_UNSUPPORTED, "only the default polynomial cumulant set is batched");
    if (d.omega_out || d.eddy_viscosity_out || d.density_in || d.velocity_in)
      return berr(LBM355_ERR_UNSUPPORTED, "per-cell subgrid outputs and density / velocity inputs are not batched");
    if (!pdf_src[k] || !pdf_dst[k]) return berr(LBM355_ERR_ARG, "NULL member PDF pointer");
    if (kind0 == LBM355_PULL && pdf_src[k] == pdf_dst[k]) return berr(LBM355_ERR_ARG, "pull needs two arrays");
    if (kind0 != LBM355_PULL && pdf_src[k] != pdf_dst[k]) return berr(LBM355_ERR_ARG, "AA needs src == dst");
  }
  void* probe = d0.compute_dtype == LBM355_F64 ? batch_pick<double>(&d0, kind0) : batch_pick<float>(&d0, kind0);
  if (!probe) return berr(LBM355_ERR_UNSUPPORTED, "no batched kernel for this stencil/family/dtype combination");
  return LBM355_OK;
}

// Checks the launch shape and step kind of a batched step.
int check_launch(int32_t pattern_kind, int32_t step_kind, int32_t threads, int32_t blocks) {
  if (threads != 64 && threads != 128 && threads != 192 && threads != 256)
    return berr(LBM355_ERR_ARG, "threads per block must be 64, 128, 192 or 256");
  if (blocks <= 0) return berr(LBM355_ERR_ARG, "blocks must be positive");
  if (pattern_kind == LBM355_PULL && step_kind != 0 && step_kind != 1)
    return berr(LBM355_ERR_ARG, "pull step_kind is 0 (src->dst) or 1 (dst->src)");
  if (pattern_kind != LBM355_PULL && step_kind != LBM355_AA_EVEN && step_kind != LBM355_AA_ODD)
    return berr(LBM355_ERR_ARG, "AA step_kind is LBM355_AA_EVEN or LBM355_AA_ODD");
  return LBM355_OK;
}

// The launch of one batched step (shared by the allocating and the allocation-free entry points).
int launch_batch(int32_t stencil, int32_t family, int32_t store_dtype, int32_t compute_dtype, int32_t pattern_kind,
                 int32_t variant_flags, const void* entries, const int32_t* tile_member, int64_t ntiles,
                 int32_t step_kind, int32_t threads, int32_t blocks, int32_t xcd_order, hipStream_t s) {
  int rc = check_launch(pattern_kind, step_kind, threads, blocks);
  if (rc) return rc;
  const int pattern = pattern_kind == LBM355_PULL ? LBM355_PULL : step_kind;
  const int swap = pattern_kind == LBM355_PULL ? step_kind : 0;
  lbm355_sweep_desc d{};
  d.stencil = stencil;
  d.family = family;
  d.store_dtype = store_dtype;
  d.flags = variant_flags;
  void* k = compute_dtype == LBM355_F64 ? batch_pick<double>(&d, pattern) : batch_pick<float>(&d, pattern);
  if (!k) return berr(LBM355_ERR_UNSUPPORTED, "no batched kernel");
  const int64_t max_blocks = (ntiles * 64 + threads - 1) / threads;
  const int grid = (int)(blocks < max_blocks ? blocks : max_blocks);
  if (compute_dtype == LBM355_F64)
    hipLaunchKernelGGL((BatchFn<double>)k, dim3(grid), dim3(threads), 0, s, (const BatchEntry<double>*)entries,
                       tile_member, ntiles, swap, xcd_order);
  else
    hipLaunchKernelGGL((BatchFn<float>)k, dim3(grid), dim3(threads), 0, s, (const BatchEntry<float>*)entries,
                       tile_member, ntiles, swap, xcd_order);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return berr(LBM355_ERR_LAUNCH, std::string("launch failed: ") + hipGetErrorString(e));
  return LBM355_OK;
}
}  // namespace

extern "C" {

int lbm355_batch_create(const lbm355_sweep_desc* descs, const void* const* pdf_src, void* const* pdf_dst,
                        int32_t n_members, lbm355_batch** out) {
  if (!out) return berr(LBM355_ERR_ARG, "bad batch arguments");
  *out = nullptr;
  int kind0 = 0;
  int rc0 = check_members(descs, pdf_src, pdf_dst, n_members, &kind0);
  if (rc0) return rc0;
  const lbm355_sweep_desc& d0 = descs[0];
  lbm355_batch* b = new lbm355_batch();
  b->stencil = d0.stencil;
  b->family = d0.family;
  b->store_dtype = d0.store_dtype;
  b->compute_dtype = d0.compute_dtype;
  b->pattern_kind = kind0;
  b->variant_flags = d0.flags & LBM355_FLAG_HIGHER_ORDERS_RATE1;
  b->nmembers = n_members;
  (void)hipGetDevice(&b->device);
  int rc = d0.compute_dtype == LBM355_F64 ? upload<double>(descs, pdf_src, pdf_dst, n_members, b)
                                          : upload<float>(descs, pdf_src, pdf_dst, n_members, b);
  if (rc) {
    delete b;
    return rc;
  }
  *out = b;
  return LBM355_OK;
}

int lbm355_batch_sweep(const lbm355_batch* b, int32_t step_kind, int32_t threads, int32_t blocks,
                       int32_t xcd_order, void* stream) {
  if (!b) return berr(LBM355_ERR_ARG, "batch is NULL");
  return launch_batch(b->stencil, b->family, b->store_dtype, b->compute_dtype, b->pattern_kind, b->variant_flags,
                      b->entries, b->tile_member, b->ntiles, step_kind, threads, blocks, xcd_order,
                      (hipStream_t)stream);
}

int lbm355_batch_info(const lbm355_batch* b, int64_t* n_tiles, int64_t* n_cells) {
  if (!b) return berr(LBM355_ERR_ARG, "batch is NULL");
  if (n_tiles) *n_tiles = b->ntiles;
  if (n_cells) *n_cells = b->ncells;
  return LBM355_OK;
}

}  // extern "C"

// ---- allocation-free batched step ---------------------------------------------------------------------------
// lbm355_batch_sweep_static keeps the member table and the tile -> member map in `scratch`, a file-scope device
// array every GPU the library runs on holds until the process ends, instead of the device memory
// lbm355_batch_create allocates.  Layout: the member table (up to STATIC_MAX_MEMBERS entries of the larger entry
// type) at offset 0, the tile map (up to STATIC_MAX_TILES int32) after it.
namespace {
constexpr int32_t STATIC_MAX_MEMBERS = 1024;
constexpr int64_t STATIC_MAX_TILES = int64_t(1) << 21;            // 2^21 tiles of 64 cells: 134,217,728 cells
constexpr size_t STATIC_TABLE_BYTES = STATIC_MAX_MEMBERS * sizeof(BatchEntry<double>);
constexpr size_t STATIC_MAP_OFFSET = (STATIC_TABLE_BYTES + 255) / 256 * 256;
constexpr size_t STATIC_SCRATCH_BYTES = STATIC_MAP_OFFSET + STATIC_MAX_TILES * sizeof(int32_t);
static_assert(sizeof(BatchEntry<double>) >= sizeof(BatchEntry<float>), "the table is sized for the larger entry");
}  // namespace

static __device__ __align__(256) unsigned char scratch[STATIC_SCRATCH_BYTES];

namespace {
constexpr int STATIC_MAX_DEVICES = 64;
std::mutex g_static_mutex;                                  // guards the host staging table and the pointers below
void* g_scratch_ptr[STATIC_MAX_DEVICES] = {};               // device address of `scratch` on each GPU
alignas(256) unsigned char g_static_table[STATIC_TABLE_BYTES];   // host staging of the member table

// map[t] = the member whose tiles contain tile t (the last member with first_tile <= t)
template <class TC>
__global__ void __launch_bounds__(256) static_tile_map_kernel(const BatchEntry<TC>* __restrict__ entries,
                                                              int32_t n, int32_t* __restrict__ map, int64_t ntiles) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ntiles) return;
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (entries[mid].first_tile <= t)
      lo = mid;
    else
      hi = mid - 1;
  }
  map[t] = lo;
}

// The device address of `scratch` on the current device, looked up once per device (hipGetSymbolAddress: the
// array's name is a host-side stand-in, not a device address).  Called with g_static_mutex held.
int scratch_address(void** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return berr(LBM355_ERR_LAUNCH, std::string("hipGetDevice failed: ") + hipGetErrorString(e));
  if (dev < 0 || dev >= STATIC_MAX_DEVICES) return berr(LBM355_ERR_UNSUPPORTED, "device index beyond 63");
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return berr(LBM355_ERR_LAUNCH, std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return LBM355_OK;
}

template <class TC>
int static_step(const lbm355_sweep_desc* descs, const void* const* src, void* const* dst, int32_t n, int kind,
                int32_t step_kind, int32_t threads, int32_t blocks, int32_t xcd_order, hipStream_t s) {
  BatchEntry<TC>* table = reinterpret_cast<BatchEntry<TC>*>(g_static_table);
  int64_t tiles = 0, cells = 0;
  int rc = fill_entries<TC>(descs, src, dst, n, table, &tiles, &cells);
  if (rc) return rc;
  if (tiles > STATIC_MAX_TILES)
    return berr(LBM355_ERR_ARG, "lbm355_batch_sweep_static: the members hold more than 2^21 tiles of 64 cells "
                                "(134,217,728 cells); use lbm355_batch_create");
  void* base = nullptr;
  rc = scratch_address(&base);
  if (rc) return rc;
  BatchEntry<TC>* dev_table = reinterpret_cast<BatchEntry<TC>*>(base);
  int32_t* dev_map = reinterpret_cast<int32_t*>(static_cast<unsigned char*>(base) + STATIC_MAP_OFFSET);
  // stream-ordered after the previous step's reads of the scratch; the host staging is reused only after the
  // copy has finished
  hipError_t e = hipMemcpyAsync(dev_table, table, sizeof(BatchEntry<TC>) * (size_t)n, hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess)
    return berr(LBM355_ERR_LAUNCH, std::string("member-table copy failed: ") + hipGetErrorString(e));
  hipLaunchKernelGGL(static_tile_map_kernel<TC>, dim3((unsigned)((tiles + 255) / 256)), dim3(256), 0, s, dev_table,
                     n, dev_map, tiles);
  e = hipGetLastError();
  if (e != hipSuccess) return berr(LBM355_ERR_LAUNCH, std::string("tile-map launch failed: ") + hipGetErrorString(e));
  const lbm355_sweep_desc& d0 = descs[0];
  return launch_batch(d0.stencil, d0.family, d0.store_dtype, d0.compute_dtype, kind,
                      d0.flags & LBM355_FLAG_HIGHER_ORDERS_RATE1, dev_table, dev_map, tiles,
                      step_kind, threads, blocks, xcd_order, s);
}
}  // namespace

extern "C" {

int lbm355_batch_sweep_static(const lbm355_sweep_desc* descs, const void* const* pdf_src, void* const* pdf_dst,
                              int32_t n_members, int32_t step_kind, int32_t threads, int32_t blocks,
                              int32_t xcd_order, void* stream) {
  int kind = 0;
  int rc = check_members(descs, pdf_src, pdf_dst, n_members, &kind);
  if (rc) return rc;
  rc = check_launch(kind, step_kind, threads, blocks);
  if (rc) return rc;
  if (n_members > STATIC_MAX_MEMBERS)
    return berr(LBM355_ERR_ARG, "lbm355_batch_sweep_static takes at most 1024 members; use lbm355_batch_create");
  std::lock_guard<std::mutex> lock(g_static_mutex);
  hipStream_t s = (hipStream_t)stream;
  if (descs[0].compute_dtype == LBM355_F64)
    return static_step<double>(descs, pdf_src, pdf_dst, n_members, kind, step_kind, threads, blocks, xcd_order, s);
  return static_step<float>(descs, pdf_src, pdf_dst, n_members, kind, step_kind, threads, blocks, xcd_order, s);
}

int lbm355_batch_destroy(lbm355_batch* b) {
  if (!b) return LBM355_OK;
  int cur = 0;
  (void)hipGetDevice(&cur);
  (void)hipSetDevice(b->device);
  hipError_t e = hipFree(b->entries);
  hipError_t e2 = hipFree(b->tile_member);
  if (e == hipSuccess) e = e2;
  (void)hipSetDevice(cur);
  delete b;
  if (e != hipSuccess) return berr(LBM355_ERR_LAUNCH, std::string("hipFree failed: ") + hipGetErrorString(e));
  return LBM355_OK;
}

}  // extern "C"
