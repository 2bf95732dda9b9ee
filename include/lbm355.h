/*
 * lbm355 -- C ABI of the MI355X (gfx950) lattice Boltzmann stream-collide sweep.
 *
 * This library replaces the compiled sweep that lbmpy obtains from
 * pystencils: create_lb_function() -> ast.compile() (lbmpy
 * src/lbmpy/creationfunctions.py:543-559, kernel built at :562-582), whose
 * generated C/CUDA function takes the raw PDF field pointers, the field
 * shape/strides and the free parameters (relaxation rates).  Here the same
 * information travels as plain pointers plus one POD descriptor; the method
 * (moment bases, equilibrium, relaxation rates, force terms) arrives as a
 * device-resident numeric table built on the host from lbmpy's LBMethod.
 *
 * Conventions
 *  - all pointers are device pointers (HIP), except `desc`, which is host memory;
 *  - element strides; x is contiguous (stride 1), "fzyx" layout;
 *  - every call is asynchronous on `stream` (a hipStream_t, 0 = null stream);
 *  - return value 0 = success, otherwise an LBM355_ERR_* code, with a message
 *    available from lbm355_last_error();
 *  - the sweep allocates nothing and never blocks.
 */
#ifndef LBM355_H
#define LBM355_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* stencils ('walberla' direction order, lbmpy stencils.py:133-135,194-198,221-227) */
#define LBM355_D2Q9 0
#define LBM355_D3Q19 1
#define LBM355_D3Q27 2

/* collision families */
#define LBM355_FAMILY_MOMENT 0   /* SRT / TRT / MRT (population or raw-moment space) */
#define LBM355_FAMILY_CENTRAL 1  /* central-moment methods */
#define LBM355_FAMILY_CUMULANT 2 /* cumulant methods (default polynomial cumulant set) */

/* streaming patterns (lbmpy fieldaccess.py:66-75 pull, :136-157 AA, :160-316 EsoTwist / EsoPull / EsoPush).
 * Every in-place pattern alternates an EVEN and an ODD kernel on one array; the value after the name
 * is the kernel.  Per direction i (c = c_i, inv = inv(i)) a kernel reads (cell offset, slot) and writes
 * (cell offset, slot):
 *   AA        EVEN  read (0, i)          write (0, inv)        ODD  read (-c, inv)        write (+c, i)
 *   EsoTwist  EVEN  read (max(-c,0), i)  write (max(c,0), inv) ODD  read (max(-c,0), inv) write (max(c,0), i)
 *   EsoPull   EVEN  E: read (-c, i) write (0, inv); O: read (0, i) write (+c, inv)
 *             ODD   E: read (-c, inv) write (0, i); O: read (0, inv) write (+c, i)
 *   EsoPush   EVEN  E: read (0, inv) write (+c, i); O: read (-c, inv) write (0, i)
 *             ODD   E: read (0, i) write (+c, inv); O: read (-c, i) write (0, inv)
 * where E are the directions at even positions of lbmpy's 'lehmann' ordering (the second of each
 * +-pair, stencils.py:146-149, 214-218, 242-248) and O their inverses.  The Eso patterns and push run
 * with periodic or ghost faces; walls are boundary links (lbm355_boundary_links). */
#define LBM355_PULL 0
#define LBM355_AA_EVEN 1
#define LBM355_AA_ODD 2
#define LBM355_ESOTWIST_EVEN 3
#define LBM355_ESOTWIST_ODD 4
#define LBM355_ESOPULL_EVEN 5
#define LBM355_ESOPULL_ODD 6
#define LBM355_ESOPUSH_EVEN 7
#define LBM355_ESOPUSH_ODD 8
#define LBM355_PUSH 9        /* two arrays: read (0, i) of src, write (+c, i) of dst (fieldaccess.py:78-87) */

/* element types */
#define LBM355_F64 0
#define LBM355_F32 1
#define LBM355_F16 2

/* treatment of the two faces of an axis */
#define LBM355_FACE_GHOST 0    /* read/write the ghost layer as is (needs ghost_layers >= 1) */
#define LBM355_FACE_PERIODIC 1 /* periodic wrap inside the kernel */
#define LBM355_FACE_NOSLIP 2   /* half-way bounce-back walls just outside the interior */

/* flags */
#define LBM355_FLAG_ZERO_CENTERED 1 /* PDFs stored as deviation from the lattice weights */
#define LBM355_FLAG_COMPRESSIBLE 2  /* u = j / rho (else j / rho_0) */
#define LBM355_FLAG_EQ_SHIFT 4      /* equilibrium velocity shifted by F / (2 rho_ref) */
#define LBM355_FLAG_FORCE 8         /* a force model is present (macroscopic velocity shift) */
#define LBM355_FLAG_SMAGORINSKY 16  /* Smagorinsky subgrid model on the shear rate */
#define LBM355_FLAG_LEARNED_CLOSURE 32 /* learned per-cell eddy viscosity on the shear rate (cumulant, FP32) */
#define LBM355_FLAG_GALILEAN 64     /* D3Q27 cumulant: Galilean correction (galilean_correction.py:21-91) */
#define LBM355_FLAG_FOURTH_ORDER 128 /* D3Q27 cumulant: fourth-order correction with limited third-order rates
                                        (fourth_order_correction.py:21-237); parameters in the table */
#define LBM355_FLAG_STOCHASTIC_ROUNDING 512 /* FP16 storage: each stored population is rounded to one of its two
                                                neighbouring FP16 values with probabilities that make the rounding
                                                unbiased, from a hash of (cell, direction, rounding_seed); no lbmpy
                                                counterpart (lbmpy has no FP16 storage) */
#define LBM355_FLAG_GENERAL_CUMULANT 1024 /* D3Q27 cumulant: a cumulant set other than the default polynomial
                                             one (lbmpy MONOMIAL_CUMULANT, custom nested_moments): the relaxation
                                             comes as the table's linear maps on groups of monomial cumulants */
#define LBM355_FLAG_HIGHER_ORDERS_RATE1 2048 /* D3Q27 cumulant, default set: the table's rates of every cumulant of
                                                order >= 3 are 1 (lbmpy's default when a method is given by its
                                                shear rate), so the sweep sets those cumulants to equilibrium
                                                without computing them.  The host sets it after checking the rates;
                                                set with other rates, the sweep still relaxes them with rate 1.
                                                Members of one batch must agree on it. */
#define LBM355_FLAG_SHEAR_PROJECTOR 4096 /* moment and central-moment families, 3-D stencils: the
                                            table's individually relaxed moments are the shear (traceless
                                            second-order) moments with one rate (and, for the central-moment
                                            family, the conserved moments with rate 0), and its SHEAR entries hold
                                            the closed form of their projector; the sweep applies that instead of
                                            n_active rank-one corrections.  The host sets it after checking the
                                            identity. */
#define LBM355_FLAG_GAUSSIAN_EQUILIBRIUM 8192 /* central-moment family, 3-D stencils: the equilibrium's central
                                                 moments are those of the full continuous Maxwellian, rho gamma_e
                                                 (table entry GAUSS = 1), or of the incompressible one, gamma_e +
                                                 drho times a rest Gaussian's (GAUSS = 2), evaluated in closed form
                                                 instead of the table's polynomial.  The host sets it after checking
                                                 the table. */
#define LBM355_FLAG_ADAPT_VELOCITY 256 /* boundary links, UBB: drive the wall at its equilibrium input velocity
                                          u_w - F / (2 rho_ref), rho_ref the link's density (the UBB's own, or
                                          the fluid cell's for compressible methods) or 1 for incompressible ones
                                          (UBB adapt_velocity_to_force, boundaryconditions.py:808-812) */

/* error codes */
#define LBM355_OK 0
#define LBM355_ERR_ARG 1
#define LBM355_ERR_UNSUPPORTED 2
#define LBM355_ERR_LAUNCH 3

typedef struct lbm355_sweep_desc {
  int32_t stencil;       /* LBM355_D2Q9 ... */
  int32_t family;        /* LBM355_FAMILY_* */
  int32_t pattern;       /* LBM355_PULL, or the EVEN / ODD kernel of an in-place pattern */
  int32_t store_dtype;   /* element type of the PDF arrays */
  int32_t compute_dtype; /* arithmetic type: LBM355_F64 or LBM355_F32 */
  int32_t nx, ny, nz;    /* interior cells (nz = 1 for 2-D) */
  int32_t ghost_layers;  /* 0 or 1 */
  int32_t face[3];       /* LBM355_FACE_* per axis */
  int64_t stride_y, stride_z, stride_q;          /* PDF strides (elements) */
  int64_t out_stride_y, out_stride_z, out_stride_c; /* density/velocity output strides */
  int32_t flags;         /* LBM355_FLAG_* */
  int32_t n_active;      /* moment family: number of individually relaxed moments */
  double force[3];       /* constant body force */
  double smagorinsky_constant; /* Cs, used with LBM355_FLAG_SMAGORINSKY */
  double omega_shear;    /* molecular shear relaxation rate (Smagorinsky, learned closure) */
  const void* tables;    /* device pointer: method table, compute_dtype elements */
  const void* closure;   /* device pointer: closure parameter block (LBM355_FLAG_LEARNED_CLOSURE), else NULL */
  /* Optional per-cell fields beside the PDFs, each NULL or compute_dtype with the strides of the density /
   * velocity outputs (x stride 1, out_stride_y / out_stride_z, components out_stride_c apart) and the PDF
   * array's ghost layers: */
  void* omega_out;          /* Smagorinsky: the cell's shear rate omega_sgs (lbmpy omega_output_field,
                               turbulence_models.py:80-81) */
  void* eddy_viscosity_out; /* Smagorinsky: Cs^2 * 3/2 * omega_sgs * Pi, the eddy viscosity (lbmpy
                               eddy_viscosity_field, turbulence_models.py:74-78) */
  const void* density_in;   /* moment family: density read instead of the PDFs' sum (lbmpy density_input,
                               creationfunctions.py:662-673); needs velocity_in */
  const void* velocity_in;  /* moment family: equilibrium velocity read instead of computed (lbmpy velocity_input,
                               creationfunctions.py:675-677); D components */
  uint32_t rounding_seed;   /* LBM355_FLAG_STOCHASTIC_ROUNDING: seed of this launch's rounding (a new one per launch) */
  int32_t waves_per_simd;   /* occupancy cap of this launch: 0 = the library's table for the kernel, -1 = none,
                               1..8 = at most that many waves per SIMD (results do not depend on it; the batched
                               sweep, whose launch shape is its own argument, ignores it) */
} lbm355_sweep_desc;

/* One stream-collide time step over the interior.
 * Replaces the compiled kernel returned by lbmpy create_lb_function
 * (creationfunctions.py:554) and called per step by
 * LatticeBoltzmannStep.time_step (lbstep.py:271-282).
 * pull: reads pdf_src, writes pdf_dst (caller swaps); AA: pdf_dst must equal pdf_src.
 * density_out / velocity_out may be NULL (lbmpy LBMConfig.output, creationfunctions.py:299-305). */
int lbm355_sweep(const lbm355_sweep_desc* desc, const void* pdf_src, void* pdf_dst, void* density_out,
                 void* velocity_out, void* stream);

/* Number of compute_dtype elements the method table of (stencil, family) holds. */
int64_t lbm355_table_elements(int32_t stencil, int32_t family);

/* Half-way bounce-back for an index list of boundary links, the NoSlip boundary kernel
 * (lbmpy boundaries/boundaryconditions.py:104-139, built by boundaries/boundaryhandling.py:159-199).
 * links: n_links x 4 int32 (x, y, z, dir) of fluid cells in ghost-layer coordinates.
 * pattern LBM355_PULL: pdf[x + c_d][inv d] = pdf[x][d]                 (before a pull sweep)
 * pattern LBM355_AA_EVEN: after an EVEN sweep, before ODD: pdf[x + c_d][d] = pdf[x][inv d]
 * pattern LBM355_AA_ODD: after an ODD sweep, before EVEN: pdf[x][inv d] = pdf[x + c_d][d]
 * (advanced_streaming/indexing.py:12-181 BetweenTimestepsIndexing). */
int lbm355_noslip_links(int32_t stencil, int32_t pattern, int32_t dtype, void* pdf, int64_t stride_y,
                        int64_t stride_z, int64_t stride_q, const int32_t* links, int64_t n_links, void* stream);

/* Conserved-quantity reduction in FP64 over the interior: out[0] = sum(rho), out[1..3] = sum(j).
 * out is a device buffer of 4 doubles (overwritten). Stored zero-centred PDFs add 1 per cell to rho. */
int lbm355_reduce_conserved(const lbm355_sweep_desc* desc, const void* pdf, double* out, void* stream);

/* Failure detection: the number of interior populations that are NaN or infinite, written to *out (a device
 * uint64, overwritten).  lbmpy has no on-device check; its tests look for non-finite velocities on the host
 * (tests/test_quicktests.py:26-27, tests/cumulantmethod/test_flow_around_sphere.py:137-139). */
int lbm355_count_nonfinite(const lbm355_sweep_desc* desc, const void* pdf, uint64_t* out, void* stream);

/* Equilibrium initialisation from macroscopic fields (lbmpy macroscopic_value_kernels.py:29-45,
 * pdf_initialization_assignments; run by LatticeBoltzmannStep.set_pdf_fields_from_macroscopic_values,
 * lbstep.py:268-269).  density / velocity: FP64 device arrays laid out with desc->out_stride_*
 * (velocity component stride out_stride_c) and the same ghost-layer offset as the PDFs.  The
 * equilibrium velocity is u - F/(2 rho_ref) when a force model is present.  prev_pattern: the kernel
 * the state is to look as if it ran last (any code; LBM355_PULL writes slot i, the others through their
 * write accessors; neighbour writes wrap on periodic axes and land in the ghost layer otherwise).  After
 * an AA odd or an Eso/push kernel no face may be a fused NoSlip face.  table_f64: FP64 method table (moment and
 * central-moment families; ignored for cumulants); a central-moment equilibrium is the method's
 * collision with every rate 1.
 * Computes in FP64 and rounds once to the storage type. */
int lbm355_set_equilibrium(const lbm355_sweep_desc* desc, void* pdf, const double* density, const double* velocity,
                           const double* table_f64, int32_t prev_pattern, void* stream);

/* Density and velocity (u = j/rho_ref + F/(2 rho_ref)) of the populations the last kernel wrote,
 * read through its write accessor (lbmpy macroscopic_value_kernels.py:48-61 with
 * get_field_accesses :14-26): LBM355_PULL slot i at x, LBM355_AA_EVEN slot inv(i) at x,
 * LBM355_AA_ODD slot i at x + c_i (bounced into slot inv(i) of x at a NoSlip face), the Eso kernels
 * through their write accessors.
 * Either output may be NULL.  FP64 outputs, strides as for lbm355_set_equilibrium. */
int lbm355_get_macroscopic(const lbm355_sweep_desc* desc, const void* pdf, double* density, double* velocity,
                           int32_t last_pattern, void* stream);

/* Strain-rate tensor S_ab = -1.5 (omega_shear / rho_ref) sum_i c_ia c_ib (f_i - f_eq,i(rho, u)) of the populations
 * the kernel `last_pattern` wrote (read as lbm355_get_macroscopic reads them; rho and the equilibrium velocity u
 * as the collision computes them, rho_ref = rho if compressible, else 1): lbmpy's strain_rate_tensor_getter
 * (macroscopic_value_kernels.py:67-85).  table_f64: the FP64 method table the setter takes.  strain_rate: FP64
 * D*D components (a*D + b), element strides of desc->out_stride_y / out_stride_z, component stride
 * out_stride_c, ghost-layer indexed like the density output. */
int lbm355_get_strain_rate(const lbm355_sweep_desc* desc, const void* pdf, const double* table_f64, double* strain_rate,
                           double omega_shear, int32_t last_pattern, void* stream);

/* Plain FP32 copy of n_floats elements, with regular (nontemporal = 0) or non-temporal accesses as
 * the sweeps use them.  No lbmpy counterpart: a copy of known size that calibrates achievable
 * bandwidth and the FETCH_SIZE / WRITE_SIZE counters against the sweep. */
int lbm355_stream_copy(const void* src, void* dst, int64_t n_floats, int32_t nontemporal, void* stream);

/* n_threads threads (n_threads / 64 waves) each running 8 independent FP32 multiply-add chains of `iters` steps,
 * 8 * iters VALU instructions per wave, the sums written to out[n_threads].  No lbmpy counterpart: it measures
 * the achievable VALU issue rate that the sweeps' instruction counts are held against. */
int lbm355_valu_probe(float* out, int64_t n_threads, int32_t iters, void* stream);

/* Copy of n_planes (19 or 27) planes of plane_elems elements each, plane q at q * plane_stride elements, with
 * one thread per element index loading that element of every plane and then storing them all: the sweep's
 * HBM access pattern (Q aligned streams read and Q written) without the collision.  src == dst is allowed
 * (in place, as an AA-even sweep accesses memory).  waves_per_simd 1..8 caps the resident waves (through
 * unused dynamic LDS, as the AA sweeps' occupancy caps do), 0 leaves the register-limited occupancy.  No lbmpy
 * counterpart: it measures the ceiling of the fzyx layout that the sweeps are held against. */
int lbm355_plane_copy(int32_t dtype, int32_t n_planes, const void* src, void* dst, int64_t plane_elems,
                      int64_t plane_stride, int32_t nontemporal, int32_t waves_per_simd, void* stream);

/* ---- Boundary links (general) --------------------------------------------------------------------
 * One boundary condition applied on an index list of links between two sweeps, addressed the way
 * lbmpy's BetweenTimestepsIndexing does (advanced_streaming/indexing.py:12-181): f_out(i) through
 * the write accessor of the kernel that ran last, f_in(i) through the read accessor of the next one.
 * Link entries are (x, y, z, d) in ghost-layer coordinates of the PDF array (z = 0 in 2-D); d points
 * from the fluid cell into the boundary cell.  Coordinates on periodic axes are wrapped into the
 * interior (the sweeps wrap periodic axes in their index math, so there are no ghost copies). */
#define LBM355_BC_NOSLIP 0   /* f_in(inv d) = f_out(d)                   boundaryconditions.py:104-139 */
#define LBM355_BC_UBB 1      /* f_in(inv d) = f_out(d) - 6 w_d rho c_d.u  boundaryconditions.py:724-842 */
#define LBM355_BC_FREESLIP 2 /* f_in(inv d) = f_out(inv mirror(d)) at x + c_d + n, boundaryconditions.py:360-487 */
#define LBM355_BC_FIXED_DENSITY 3 /* f_in(inv d) = 2 f_eq,sym(d; rho_w, u) - f_out(d), boundaryconditions.py:1044-1102 */
#define LBM355_BC_OUTFLOW 4  /* extrapolation outflow with per-link state, boundaryconditions.py:894-1042:
                                f_in(inv d) = c s0 + (1 - c) s1, c = 1/sqrt(3); then s0 = f_out(inv d) at
                                x + c_d - n and s1 = the value just written */
#define LBM355_BC_OUTFLOW_INIT 5 /* sets s0 = s1 = f_out(inv d) at x + c_d - n (before the first step) */
#define LBM355_BC_SIMPLE_OUTFLOW 7 /* f_in(inv d) = f_out(inv d) at x + c_d - n (boundaryconditions.py:844-891,
                                      Geier et al. 2015 eq. F.1), n = desc->normal pointing out of the fluid */
#define LBM355_BC_FREESLIP_LINKS 6 /* FreeSlip with a normal per link (lbmpy's FreeSlip without normal_direction,
                                      boundaryconditions.py:402-440): link_data n x 4 doubles (wall-to-fluid normal
                                      x, y, z, reference direction r); f_in(inv d) = f_out(r) at x + c_d + n */

typedef struct lbm355_boundary_desc {
  int32_t stencil, pattern, store_dtype, kind; /* pattern: the kernel that ran last (PULL, or an in-place
                                                  pattern's EVEN / ODD kernel; the next one is the other) */
  int32_t nx, ny, nz, ghost_layers;            /* interior sizes (nz = 1 in 2-D) */
  int32_t periodic[3];                         /* 1: wrap this axis */
  int32_t flags;                               /* LBM355_FLAG_ZERO_CENTERED | LBM355_FLAG_COMPRESSIBLE (UBB) */
  int32_t normal[3];                           /* FreeSlip: wall normal into the fluid, one non-zero entry */
  int32_t pad;
  int32_t fused_noslip[3];                     /* 1: this axis's faces are NoSlip walls fused into the sweep
                                                  (LBM355_FACE_NOSLIP); after an AA odd step a population
                                                  bound into such a wall is read back from the cell itself */
  int32_t pad2;
  int64_t stride_y, stride_z, stride_q;
  double velocity[3];                          /* UBB wall velocity (unless per-link velocities are given) */
  double density;                              /* UBB wall density if > 0; else from the fluid cell
                                                  (compressible) or 1 (incompressible).  FixedDensity: rho_w */
  double force[3];                             /* FixedDensity: body force (with LBM355_FLAG_EQ_SHIFT) */
  const double* equilibrium;                   /* FixedDensity: device table [Q][8] of the equilibrium in
                                                  storage form, f_eq_i = A + B drho + lam (L.u + Qd.u^2 + Qo.uu):
                                                  A, B, Qd[3], Qo[3] (xy, xz, yz) per direction */
} lbm355_boundary_desc;

/* Applies one boundary condition on n_links links.  link_data: UBB: NULL or n_links x 3 wall velocities;
 * outflow: n_links x 2 doubles of per-link state (s0, s1), read and updated (normal in desc->normal,
 * pointing out of the fluid). */
int lbm355_boundary_links(const lbm355_boundary_desc* desc, void* pdf, const int32_t* links, int64_t n_links,
                          double* link_data, void* stream);

/* ---- Ghost-layer periodicity -----------------------------------------------------------------------
 * For sweeps whose faces are LBM355_FACE_GHOST (lbmpy's own kernels never wrap): the copies of lbmpy's
 * LBMPeriodicityHandling (advanced_streaming/communication.py:21-126), each source slice of
 * get_communication_slices (:128-183) copied onto its destination slice, as boxes.  boxes: host array of
 * n_boxes x 5 int64 {source offset, destination offset, extent x, extent y, extent z}, offsets in elements
 * from `pdf` (slot included), element (i, j, l) of a box at i stride_x + j stride_y + l stride_z from its
 * corners.  Every box must lie within [0, span) elements; sources and destinations must not overlap (true
 * of lbmpy's slices), since all boxes of a call are copied concurrently.  Any element type of the given
 * size (LBM355_F64 / F32 / F16) is copied bit for bit. */
int lbm355_copy_boxes(int32_t dtype, void* pdf, int64_t span, int64_t stride_x, int64_t stride_y, int64_t stride_z,
                      const int64_t* boxes, int32_t n_boxes, void* stream);

/* Access table of a streaming-pattern kernel code (LBM355_PULL ... LBM355_PUSH), as the sweep kernels are
 * compiled with it: out[Q][8] = {read offset x, y, z, read slot, write offset x, y, z, write slot} per
 * direction (lbmpy's accessor read() / write(), fieldaccess.py:66-316).  Host only, no GPU.  Returns Q, or
 * minus an error code. */
int32_t lbm355_pattern_access(int32_t stencil, int32_t pattern, int32_t* out);

/* ---- Batched ensembles -------------------------------------------------------------------------
 * Many independent domains ("members") advanced by one launch per time step.  lbmpy has no batched
 * sweep: an ensemble there is a Python loop calling one compiled kernel per member and time step
 * (LatticeBoltzmannStep.time_step, lbstep.py:271-282, once per domain).  Every member is described
 * by its own lbm355_sweep_desc (sizes, faces, strides, force, method table), as for lbm355_sweep;
 * all members share stencil, collision family, streaming pattern (pull, or AA) and dtypes.  Each
 * member's cells go through the same per-cell code as lbm355_sweep, so a member's result is
 * bit-identical to sweeping it alone. */
typedef struct lbm355_batch lbm355_batch;

/* Uploads the member table and a tile -> member map (two device allocations).  pdf_src[k] / pdf_dst[k]: member k's arrays
 * (equal for AA).  The table keeps the pointers: the arrays must outlive the batch. */
int lbm355_batch_create(const lbm355_sweep_desc* descs, const void* const* pdf_src, void* const* pdf_dst,
                        int32_t n_members, lbm355_batch** batch);

/* One time step of every member, one launch, no allocation (capturable in a HIP graph).
 * step_kind: pull 0 = src -> dst, 1 = dst -> src; AA LBM355_AA_EVEN or LBM355_AA_ODD.
 * threads per block 64/128/192/256; blocks: persistent grid size (clamped to the work); xcd_order:
 * give each of the 8 XCDs a contiguous stretch of the tile sequence (needs blocks % 8 == 0). */
int lbm355_batch_sweep(const lbm355_batch* batch, int32_t step_kind, int32_t threads, int32_t blocks,
                       int32_t xcd_order, void* stream);

/* Number of 64-cell tiles and of interior cells over all members. */
int lbm355_batch_info(const lbm355_batch* batch, int64_t* n_tiles, int64_t* n_cells);

/* Frees the member table (synchronous). */
int lbm355_batch_destroy(lbm355_batch* batch);

/* One batched step without any device allocation: lbm355_batch_create + lbm355_batch_sweep in one call, with the
 * member table and the tile map in a file-scope device array (8 MiB of tile map and about 0.25 MiB for 1024
 * member entries) that each GPU holds until the process ends, instead of memory the library allocates.  No
 * counterpart in lbmpy: it is for callers that cannot allocate device memory (serving loops and harnesses that
 * forbid hipMalloc).  Limits: at most 1024 members and
 * 2^21 tiles of 64 cells (134,217,728 interior cells); larger ensembles are refused with LBM355_ERR_ARG.  Same
 * arguments, checks and kernel as lbm355_batch_create / lbm355_batch_sweep, so results are identical.  Every
 * call on one GPU shares that array: calls that overlap on one GPU (other streams or host threads) race.  The
 * call waits on `stream` once (for the copy of the member table), so it is not meant for graph capture. */
int lbm355_batch_sweep_static(const lbm355_sweep_desc* descs, const void* const* pdf_src, void* const* pdf_dst,
                              int32_t n_members, int32_t step_kind, int32_t threads, int32_t blocks,
                              int32_t xcd_order, void* stream);

/* ---- Learned closure ----------------------------------------------------------------------------
 * No lbmpy counterpart: a per-cell MLP (10 -> 32 -> 32 -> 1, tanh, BF16 weights, FP32 accumulation
 * on the MFMA units) that maps a cell's non-equilibrium second moments to an eddy viscosity and
 * replaces the Smagorinsky rate of the cumulant sweep (turbulence_models.py:25-84) when a sweep
 * descriptor sets LBM355_FLAG_LEARNED_CLOSURE and `closure`.  The parameter block layout is
 * lbm355_closure_params_size() bytes (lbmpy_amd/csrc/closure.h); training keeps FP32 master
 * parameters of lbm355_closure_master_size() floats. */
int64_t lbm355_closure_params_size(void);
int32_t lbm355_closure_master_size(void);

/* features[n][10] of cells with non-equilibrium second moments pi_neq[n][6] (xx yy zz xy xz yz)
 * and density[n], as the fused sweep computes them. */
int lbm355_closure_features(const float* pi_neq, const float* density, int64_t n, float* features, void* stream);

/* network outputs o[n] for features[n][10] (nu_t = nu_scale * softplus(o)). */
int lbm355_closure_forward(const void* params, const float* features, int64_t n, float* out, void* stream);

/* gradient of L = (1/n) sum (softplus(o) - targets)^2 into grad[master_size] (overwritten), L into
 * loss[0]; `blocks` persistent one-wavefront blocks. */
int lbm355_closure_train_step(const void* params, const float* features, const float* targets, int64_t n,
                              float* grad, float* loss, int32_t blocks, void* stream);

/* Adam step (`step` >= 1) on the FP32 master parameters with moments m, v; re-packs the BF16
 * weights and FP32 biases into the parameter block `params`. */
int lbm355_closure_adam(float* master, const float* grad, float* m, float* v, float lr, float beta1, float beta2,
                        float eps, int32_t step, void* params, void* stream);

/* sizeof(lbm355_sweep_desc), for FFI layout checks. */
int64_t lbm355_desc_size(void);

/* Last error message of the calling thread (static storage, never NULL). */
const char* lbm355_last_error(void);

/* ABI version. */
int32_t lbm355_version(void);

/* SHA-256 (hex) of the library's sources (lbmpy_amd/csrc/*.hip, *.h and this header) at build time, so a
 * loader can refuse a library built from other sources.  "unknown" when built without it. */
const char* lbm355_source_hash(void);

#ifdef __cplusplus
}
#endif
#endif /* LBM355_H */
