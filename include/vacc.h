/*
 * vacc.h -- C ABI of libvacc, the MI355X (gfx950) velocity-space acceleration
 * library.  Drop-in for the per-cell acceleration path of fmihpc/vlasiator:
 * the Lorentz-force transform and SLICE-3D intersections, the three 1-D
 * semi-Lagrangian PLM/PPM/PQM column remaps over sparse 4x4x4 velocity blocks
 * with inline block add/remove, and the sparsity-driven block adjust.
 *
 * Plain C types only.  Every function returns 0 on success, a positive status
 * bitmask for recoverable per-cell conditions (VACC_STATUS_*), or a negative
 * error code (VACC_E_*).  The library never aborts.
 *
 * Reference interfaces replaced (file:line in fmihpc/vlasiator):
 *   vacc_compute_cell_intersections  <- compute_cell_intersections(SpatialCell*,popID,map_order,dt,timer)
 *                                       vlasovsolver/cpu_acc_intersections.cpp:44
 *                                       (with compute_acceleration_transformation,
 *                                       vlasovsolver/cpu_acc_transform.cpp:66)
 *   vacc_accelerate_cells            <- cpu_accelerate_cells(mpiGrid,cells,popID,map_order)
 *                                       vlasovsolver/cpu_acc_semilag.cpp:45 (and its GPU twin
 *                                       gpu_accelerate_cells, vlasovsolver/gpu_acc_semilag.cpp:52)
 *                                       and, on a one-cell context,
 *                                       cpu_accelerate_cell(SpatialCell*,popID,map_order)
 *                                       vlasovsolver/cpu_acc_semilag.cpp:92
 *   vacc_map_1d                      <- map_1d(SpatialCell*,popID,I,di,dj,dk,dim)
 *                                       vlasovsolver/cpu_acc_map.cpp:121
 *   vacc_adjust_blocks               <- SpatialCell::adjustSingleCellVelocityBlocks(popID,doDeleteEmpty)
 *                                       spatial_cells/spatial_cell_cpu.cpp:275 (adjust_velocity_blocks :179)
 *   vacc_moments                     <- calculateMoments_V / blockVelocityFirstMoments
 *                                       vlasovsolver/arch_moments.cpp:342, arch_moments.h:83
 *   vacc_max_v_dt                    <- updateAccelerationMaxdt, vlasovsolver/cpu_acc_transform.cpp:41
 *   vacc_upload_cells/vacc_download_cells
 *                                    <- the SpatialCell velocity mesh + block container state
 *                                       (spatial_cells/velocity_mesh_cpu.h:47, velocity_block_container.h:59)
 */
#ifndef VACC_H
#define VACC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VACC_WID 4
#define VACC_WID3 64
#define VACC_INVALID_ID 0xFFFFFFFFu

/* reconstruction (Makefile ACC_SEMILAG_PLM / _PPM / _PQM) */
enum { VACC_PLM = 0, VACC_PPM = 1, VACC_PQM = 2 };

/* Storage of the distribution function in HBM.  F32: 64 floats per block.  BF16 / F16: 64
 * 16-bit codes plus one float scale per block, f = scale * code; the remap reads and writes
 * this form and computes in FP32.  A block is encoded as: m = max|f| = mant * 2^p,
 * code = RNE(f * 2^-p), scale = S_f / S_code (the block's FP32 sums, so the block mass is
 * kept) when that ratio is within [2^(p-1), 2^(p+1)], else 2^p.  Upload and download convert
 * from and to FP32. */
enum { VACC_STORE_F32 = 0, VACC_STORE_BF16 = 1, VACC_STORE_F16 = 2 };

/* per-cell status bits (OR-ed over cells in return values) */
#define VACC_STATUS_WALL 1u        /* target blocks within wall margin (cpu_acc_map.cpp:325-338) */
#define VACC_STATUS_DK 2u          /* non-positive intersection_dk, or a deposit span (dv + 3|di| + 3|dj|)/dk
                                      above 118 cells: cell passed through, not remapped */
#define VACC_STATUS_CAPACITY 4u    /* block pool capacity exceeded: step not performed */

/* errors */
#define VACC_E_ARG (-1)
#define VACC_E_HIP (-2)
#define VACC_E_NOMEM (-3)
#define VACC_E_STATE (-4)

/* Velocity mesh of one population (velocity_mesh_parameters.h:57-65). */
typedef struct vacc_mesh {
   uint32_t grid_length[3];   /* blocks per axis, each <= 65535 */
   double min_limits[3];
   double max_limits[3];
} vacc_mesh;

/* Particle species and sparsity parameters (particle_species.h:45-57, object_wrapper.cpp:120-134). */
typedef struct vacc_species {
   double mass;
   double charge;
   double sparse_min_value;          /* <pop>_sparse.minValue */
   int32_t sparse_add_width_v;       /* <pop>_sparse.blockAddWidthV */
   /* <pop>_sparse.dynamicAlgorithm: 0 the static minValue; 1 or 2 a per-cell threshold set at
    * every adjust (SpatialCell::updateSparseMinValue, spatial_cell_cpu.cpp:699-727), linear
    * in the cell's density (1) or block count (2) through (dynamicBulkValue1,
    * dynamicMinValue1) and (dynamicBulkValue2, dynamicMinValue2), clamped to
    * [dynamicMinValue1, dynamicMinValue2].  The threshold is the content test of the adjust
    * and of the remap's written blocks, and the reconstruction threshold of the next remaps
    * (cpu_acc_map.cpp:539-547).  The density is that of the population's current f. */
   int32_t sparse_dynamic_algorithm;
   double sparse_dynamic_bulk1, sparse_dynamic_bulk2;   /* dynamicBulkValue1, dynamicBulkValue2 */
   double sparse_dynamic_min1, sparse_dynamic_min2;     /* dynamicMinValue1, dynamicMinValue2 */
   /* <pop>_sparse.conserve_mass: after every adjust the cell's f is scaled by (sum of f before)
    * / (sum after), block_adjust_cpu.cpp:96-113 (the 16-bit forms scale each block's scale) */
   int32_t sparse_conserve_mass;
   int32_t reserved;
} vacc_species;

typedef struct vacc_options {
   int32_t method;            /* VACC_PLM / VACC_PPM / VACC_PQM */
   int32_t storage;           /* VACC_STORE_* */
   int32_t wall_margin;       /* bailout.velocity_space_wall_block_margin */
   int32_t ohm_gradpe_term;   /* fieldsolver.ohmGradPeTerm */
   double hall_minimum_rhoq;  /* P::hallMinimumRhoq */
   double max_rotation_deg;   /* vlasovsolver.maxSlAccelerationRotation */
   int32_t fast_arith;        /* 0: remap arithmetic rounds exactly as the reference CPU path
                                 (bit-identical f); 1: the tolerance form: reconstructions in
                                 deviations from the cell average with fused multiply-adds and
                                 hardware reciprocal/sqrt, and a flux-form deposit from the FP64
                                 intersections (f within 5e-6 of an FP64 run per cell, L-inf
                                 relative to the cell's max and L1; block lists unchanged).
                                 Cells whose deposits outrun the tolerance kernel's target ring,
                                 and columns where the reference's 8-lane maximum rule binds,
                                 take the exact arithmetic. */
   int32_t use_graph;         /* 1: vacc_step captures its launches in a hipGraph per state
                                 variant (map order, list/pool parity, options) on first use
                                 and replays it afterwards; results are identical.  Not used
                                 while kernel profiling is on. */
   int32_t sums_from_f;       /* tolerance form only.  0: the last remap pass of an acceleration
                                 (vacc_step, vacc_accelerate_cells, vacc_accelerate_populations)
                                 takes FP32 sums of each block as it stores it (sum f and the
                                 in-block index sums), and the density/velocity moments and the
                                 block adjust's removed mass that follow are formed from them
                                 (relative error of order 1e-7) instead of re-reading f; 1: both
                                 are always summed from f.  The exact form always sums f. */
} vacc_options;

/* Per-cell field input of the transform (cpu_acc_transform.cpp:71-115). */
typedef struct vacc_cell_fields {
   double B[3];          /* BGB*VOL + PERB*VOL */
   double dB[6];         /* dPERBXVOLdy, dPERBXVOLdz, dPERBYVOLdx, dPERBYVOLdz, dPERBZVOLdx, dPERBZVOLdy */
   double V[3];          /* bulk velocity V_V (used when the moments are not taken from f) */
   double rhoq;          /* RHOQ_V */
   double EgradPe[3];    /* E*GRADPE */
   double dt;            /* subcycle dt of the cell (Population::subcycleDt) */
} vacc_cell_fields;

typedef struct vacc_ctx vacc_ctx;

/* Context: one population's mesh and species; owns device pools and a stream. */
int vacc_create(const vacc_mesh* mesh, const vacc_species* species, const vacc_options* opts, vacc_ctx** out);
int vacc_destroy(vacc_ctx* ctx);
const char* vacc_last_error(void);

/* Upload ncells cells, packed: cell c holds blocks [offsets[c], offsets[c+1]) of gids and
 * f (64 floats per block, in-block index i + 4*j + 16*k).  Replaces the context's cells.
 * Blocks of a cell must have distinct valid GIDs. */
int vacc_upload_cells(vacc_ctx* ctx, uint32_t ncells, const uint64_t* offsets,
                      const uint32_t* gids, const float* f);
/* Total blocks held as of the last call that changed the block lists (no device sync). */
int vacc_total_blocks(vacc_ctx* ctx, uint64_t* total);
/* Number of cells and blocks per cell (counts[ncells]) currently held. */
int vacc_get_counts(vacc_ctx* ctx, uint32_t* ncells, uint32_t* counts);
/* Download: offsets[ncells+1] written; gids/f sized by the caller from vacc_get_counts.
 * Blocks of each cell are returned in ascending GID order. */
int vacc_download_cells(vacc_ctx* ctx, uint64_t* offsets, uint32_t* gids, float* f);

/* Per-cell intersections (12 doubles per cell: x {I,di,dj,dk}, y {...}, z {...}),
 * as stored in Population::intersection_* (spatial_cell_cpu.hpp:89-91). */
int vacc_set_intersections(vacc_ctx* ctx, const double* isect);
int vacc_get_intersections(vacc_ctx* ctx, double* isect);
/* Transform + intersections for every cell on the device (FP64).  If use_moments != 0 the
 * pivot bulk velocity is the first moment of the current f (calculateMoments_V), otherwise
 * fields[c].V.  fields: ncells entries, kept resident for later calls; NULL reuses the
 * fields of the last vacc_set_fields / vacc_compute_cell_intersections call. */
int vacc_compute_cell_intersections(vacc_ctx* ctx, const vacc_cell_fields* fields, uint32_t map_order,
                                    int32_t use_moments);
/* Host FP64 transform + intersections of a single cell (no device work). */
int vacc_cell_intersections_host(const vacc_mesh* mesh, const vacc_species* species, const vacc_options* opts,
                                 const vacc_cell_fields* fields, uint32_t map_order, double out[12]);
/* max_v_dt of a cell (cpu_acc_transform.cpp:41-55). */
double vacc_max_v_dt(const vacc_species* species, const vacc_options* opts, const double B[3]);

/* One 1-D remap of every cell along dim with the stored intersections. */
int vacc_map_1d(vacc_ctx* ctx, uint32_t dim);
/* Three remaps in map_order (0 XYZ, 1 YZX, 2 ZXY) for every cell of the context.  A
 * context holding many cells is the batched form (every kernel covers all cells); a
 * context holding one cell is the per-SpatialCell form of cpu_accelerate_cell. */
int vacc_accelerate_cells(vacc_ctx* ctx, uint32_t map_order);
/* Single-cell sparsity adjust of every cell (doDeleteEmptyBlocks = true).  rho_loss may be
 * NULL; otherwise ncells FP64 values of the removed mass (RHOLOSSADJUST increments). */
int vacc_adjust_blocks(vacc_ctx* ctx, double* rho_loss);
/* Sparsity adjust with spatial neighbours: SpatialCell::adjust_velocity_blocks(spatial_neighbors,
 * popID, true), spatial_cells/spatial_cell_cpu.cpp:179-273, as driven by adjustVelocityBlocks
 * (grid.cpp:775-814).  Besides the cell's content blocks and their velocity neighbours, every
 * block with content in a spatial neighbour is kept (or added, zero) as it is, without velocity
 * dilation (:212-222).  Neighbours are given as cells of this context, nbr_off[ncells+1] into
 * nbr[] (their content as of the current f), and/or as caller-supplied content GIDs,
 * extra_off[ncells+1] into extra_gids[] (e.g. remote neighbours).  Either pair may be NULL.
 * rho_loss as in vacc_adjust_blocks.  The block storage grows when the neighbours add more
 * blocks than the reserve holds. */
int vacc_adjust_blocks_spatial(vacc_ctx* ctx, const uint32_t* nbr_off, const uint32_t* nbr, const uint64_t* extra_off,
                               const uint32_t* extra_gids, double* rho_loss);
/* Launch-shape autotuner (no counterpart in the reference): times the remap of each pass
 * dimension on the current state with 8, 32, 128, 512, 1024, 2048 and 4096 workgroups per compute
 * unit and, in the tolerance form (fast_arith = 1), workgroups of 64, 128 or 256 lanes (4, 8
 * or 16 work units of 16 lanes, i.e. columns or column sets, per workgroup), twice each, and
 * keeps the fastest shape per dimension for later steps.  The trial passes write scratch
 * buffers only, so the state is unchanged; results of later steps do not depend on the
 * choice.  mult_out (3 entries, may be NULL) receives the chosen workgroups per CU. */
int vacc_autotune(vacc_ctx* ctx, uint32_t* mult_out);
/* The remap launch shape per pass dimension (no counterpart in the reference): workgroups per
 * compute unit and lanes per workgroup (3 entries each, either may be NULL). */
int vacc_get_remap_shape(vacc_ctx* ctx, uint32_t* wg_per_cu, uint32_t* lanes_per_wg);
/* Block lookup: index[i] = position of block gids[i] within cell cells[i]'s block list, in
 * the ascending-GID order that vacc_download_cells returns (the local ID after a download),
 * or VACC_INVALID_ID when the cell holds no such block or cells[i] >= ncells
 * (VelocityMesh::getLocalID, spatial_cells/velocity_mesh_cpu.h:371-377, over the velocity
 * mesh's GID -> LID map, open_bucket_hashtable.h:286).  The device keeps every cell's list in
 * ascending GID order after an adjust (a remap leaves pass order; the call then reorders the
 * lists, blocks unchanged), so the map is a bisection of the resident list.  Host arrays of
 * nq entries; same state afterwards. */
int vacc_find_blocks(vacc_ctx* ctx, uint64_t nq, const uint32_t* cells, const uint32_t* gids, uint32_t* index);
/* Block lookup that allocates nothing.  It has no counterpart in the reference.  It is for
 * callers that cannot allocate device memory: tight serving loops, and harnesses that forbid
 * hipMalloc.  Same arguments, answers and state afterwards as vacc_find_blocks, for at most
 * VACC_FIND_NOALLOC_MAX queries; a larger nq returns VACC_E_ARG.  The queries and answers are
 * staged in one fixed device array per GPU (3 * VACC_FIND_NOALLOC_MAX words, 768 KiB, held
 * until the process ends) instead of a buffer allocated per call.  Every context on one GPU
 * shares that array, so calls that overlap in time on one GPU (from several threads or
 * contexts) race; calls made one after another do not. */
#define VACC_FIND_NOALLOC_MAX 65536u
int vacc_find_blocks_noalloc(vacc_ctx* ctx, uint64_t nq, const uint32_t* cells, const uint32_t* gids,
                             uint32_t* index);
/* FP64 moments per cell: out[4*c+0] = n, [1..3] = n*V (arch_moments.h:83-108).  They are
 * kept resident until f changes, so a following use_moments transform does not recompute
 * them. */
int vacc_moments(vacc_ctx* ctx, double* out);

/* Pressure tensor per cell (the second moments, arch_moments.h:123-150 with
 * arch_moments.cpp:474-489): out[6*c+0..5] = mass * sum f (v - V)_a (v - V)_b dV^3 for
 * (a,b) = xx, yy, zz, yz, xz, xy (P_11, P_22, P_33, P_23, P_13, P_12), FP64, v at cell centres.
 * pivot: 3 doubles per cell, the bulk velocity over all species (VX_V, VY_V, VZ_V), or NULL
 * for this population's own bulk velocity from its first moments. */
int vacc_pressure(vacc_ctx* ctx, const double* pivot, double* out);
/* Make the per-cell transform inputs resident (ncells entries; the dt of each entry is the
 * cell's subcycle dt).  Replaced by the next call; cleared by an upload. */
int vacc_set_fields(vacc_ctx* ctx, const vacc_cell_fields* fields);
/* One acceleration subcycle of every cell with a single host synchronisation:
 * calculateAcceleration(popID, ..., acceleratedCells, dt), vlasovsolver/vlasovmover.cpp:410 --
 * moments when use_moments != 0 (calculateMoments_V, :420), transform + intersections from
 * the resident fields, the three remaps in map_order (cpu_accelerate_cells, :462) and, when
 * adjust != 0, the sparsity adjust (adjustVelocityBlocks, :473).  rho_loss as in
 * vacc_adjust_blocks (may be NULL).  Returns the OR of the per-cell status bits. */
int vacc_step(vacc_ctx* ctx, uint32_t map_order, int32_t use_moments, int32_t adjust, double* rho_loss);
/* Acceleration subcycling of one population (calculateAcceleration(mpiGrid, dt),
 * vlasovsolver/vlasovmover.cpp:481-565, and its subcycle step :410-475).
 * vacc_subcycle_count: per cell max_v_dt = |T_gyro| * max_rotation_deg / 360 from the resident
 * fields' B (updateAccelerationMaxdt, cpu_acc_transform.cpp:41-55) and the subcycle count
 * max(ceil(|dt| / max_v_dt), 1) (getAccelerationSubcycles, vlasovmover.cpp:398-401), kept
 * resident; *nmax receives the maximum over the context's cells.  Across processes the caller
 * takes the maximum of nmax (the reference's MPI_Allreduce, vlasovmover.cpp:544) and runs that
 * many steps.
 * vacc_subcycle_step: subcycle `step` (0-based) of every cell with step < its count, the others
 * unchanged: moments pivot (use_moments != 0), transform and intersections with the cell's
 * subcycle dt -- max_v_dt, or what is left of |dt| on its last subcycle, with the sign of dt
 * (vlasovmover.cpp:429-456) -- the three remaps in map_order (the same order for every
 * subcycle of a step, :424-426), and, when adjust != 0, the single-cell sparsity adjust of
 * every cell (:472-474, and the final adjust :556; adjusting a cell that was not accelerated
 * changes nothing).  rho_loss as in vacc_adjust_blocks.  One host synchronisation.
 * vacc_accelerate_subcycled: count, then every step with adjust; *nsub (may be NULL) receives
 * the step count, rho_loss (may be NULL) the removed mass summed over the steps. */
int vacc_subcycle_count(vacc_ctx* ctx, double dt, uint32_t* nmax);
int vacc_subcycle_step(vacc_ctx* ctx, uint32_t step, uint32_t map_order, int32_t use_moments, int32_t adjust,
                       double* rho_loss);
int vacc_accelerate_subcycled(vacc_ctx* ctx, double dt, uint32_t map_order, int32_t use_moments, uint32_t* nsub,
                              double* rho_loss);
/* One acceleration step of several populations of the same cells (one context each, npop <=
 * 8, e.g. protons and He2+): calculateAcceleration(mpiGrid, dt), vlasovmover.cpp:481-565, with
 * one subcycle per population.  Populations are advanced in order; with use_moments != 0 each
 * one's transform pivots on the plasma moments of all populations as they stand
 * (calculateMoments_V, arch_moments.cpp:342-440: V = sum m nV / sum m n, RHOQ = sum q n).
 * Then its three remaps and, when adjust != 0, its adjust.  rho_loss may be NULL, or hold npop
 * pointers (each NULL or ncells doubles).  One host synchronisation; returns the OR of the
 * per-cell status bits of all populations. */
int vacc_accelerate_populations(vacc_ctx* const* ctxs, uint32_t npop, uint32_t map_order, int32_t use_moments,
                                int32_t adjust, double* const* rho_loss);
/* Synthetic initial state, the project helper used by tests and benchmarks: mirrors
 * Project::setVelocitySpace + MultiPeak::fillPhaseSpace (projects/project.cpp:239-263,
 * projects/MultiPeak/MultiPeak.cpp:127-222) followed by the initial single-cell adjust
 * (grid.cpp:289).  Every block of the candidate box bbox[6*c..] = {x0,y0,z0,x1,y1,z1}
 * (inclusive block indices) is sampled at cell centres as the float sum of the cell's
 * peaks; blocks with content and their sparse_add_width_v neighbours are kept with their
 * sampled values.  Replaces the context's cells; peaks holds max_peaks entries per cell. */
typedef struct vacc_peak {
   int32_t kind;       /* 0: Maxwellian (projects/project.h:66-72), 1: gyrotropic ring */
   int32_t reserved;
   double n;           /* number density */
   double V[3];        /* centre (Maxwellian) or ring centre velocity */
   double T;           /* temperature (Maxwellian) */
   double axis[3];     /* unit ring axis */
   double vperp;       /* ring speed perpendicular to axis */
   double vth;         /* ring thermal speed */
} vacc_peak;
int vacc_init_peaks(vacc_ctx* ctx, uint32_t ncells, uint32_t max_peaks, const vacc_peak* peaks,
                    const int32_t* npeaks, const int32_t* bbox);
/* Kernel timing with HIP events recorded on the context's stream (no counterpart in the
 * reference, whose phiprof timers measure host wall time).  Between vacc_profile_begin and
 * vacc_profile_read every remap launch (k_remap), column build (k_columns + scan) and block
 * adjust is bracketed by events; the block counts each remap read and wrote are logged. */
typedef struct vacc_profile {
   uint64_t remap_launches;
   double remap_ms;           /* summed k_remap durations */
   uint64_t remap_src_blocks; /* summed source blocks over the remap launches */
   uint64_t remap_tgt_blocks; /* summed target blocks */
   uint64_t build_launches;
   double build_ms;           /* k_columns + scan */
   uint64_t adjust_launches;
   double adjust_ms;          /* adjust count + scan + write */
   uint64_t transform_launches;
   double transform_ms;       /* moments + transform/intersections */
   uint64_t trans_launches;
   double trans_ms;           /* spatial translation (include/vacc_trans.h), all its kernels */
   uint64_t remap_sum_blocks; /* target blocks of the remaps that also stored block sums */
} vacc_profile;
int vacc_profile_begin(vacc_ctx* ctx);
int vacc_profile_read(vacc_ctx* ctx, vacc_profile* out);
/* Wait for the context's stream; returns the OR of per-cell status bits since the last call. */
int vacc_sync(vacc_ctx* ctx);
/* Per-cell status words since the last vacc_sync (ncells entries). */
int vacc_get_status(vacc_ctx* ctx, uint32_t* status);

#ifdef __cplusplus
}
#endif
#endif
