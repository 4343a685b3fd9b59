/*
 * vacc_oracle.h -- CPU oracle for the velocity-space acceleration (ACC) path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * CPU algorithm (fmihpc/vlasiator, VEC_FALLBACK_GENERIC, VECL=8, Realf=float,
 * Real=double, no FP contraction).  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it, and only as the checker or the
 * CPU baseline.  The product (vlasiator_amd/) never links or calls it.
 *
 * Each function cites the reference file:line it follows.  Pinning status of
 * each part (what independent check holds it) is listed in oracle/README.md.
 */
#ifndef VACC_ORACLE_H
#define VACC_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VO_WID 4
#define VO_WID2 16
#define VO_WID3 64
#define VO_MAX_BLOCKS_PER_DIM 256        /* common.h:72 */
#define VO_INVALID 0xFFFFFFFFu            /* definitions.h:63-66 */

enum { VO_PLM = 0, VO_PPM = 1, VO_PQM = 2 };

/* Velocity mesh geometry (velocity_mesh_parameters.cpp:137-163). */
typedef struct vo_mesh {
   uint32_t len[3];   /* blocks per dimension */
   double vmin[3];
   double vmax[3];
} vo_mesh;

/* cell size along dim: (vmax-vmin)/len/WID (velocity_mesh_parameters.cpp:153-163) */
double vo_cell_size(const vo_mesh* m, int dim);

/* ---------------- reconstruction coefficients (one lane) ---------------- */
/* v points at the padded column data; k is the index of the source cell. */
void vo_plm_coeff(const float* v, int k, float threshold, float a[2]);
void vo_ppm_coeff(const float* v, int k, float threshold, float a[3]);
void vo_pqm_coeff(const float* v, int k, float threshold, float a[5]);
/* Batched helpers for tests: stencils is n x 9 values centred at index 4. */
void vo_coeff_batch(int method, const float* stencils, int n, float threshold, float* out);

/* ---------------- column sort (cpu_acc_sort_blocks.cpp:46-155) ---------------- */
/* Outputs: sorted GIDs (n), column offsets/lengths (<= n), set column offsets/counts (<= n). */
int vo_sort_columns(const vo_mesh* m, const uint32_t* gids, uint32_t n, int dim,
                    uint32_t* sorted, uint32_t* col_off, uint32_t* col_len, uint32_t* n_cols,
                    uint32_t* set_col_off, uint32_t* set_ncols, uint32_t* n_sets);

/* ---------------- one 1-D remap (cpu_acc_map.cpp:121-636) ---------------- */
/* Returns 0 on success, -1 if cap_out is too small (n_out then holds the required size).
 * Output blocks are emitted per column set in increasing k.  wall_flag is set to 1 when
 * a target column reaches within wall_margin blocks of the mesh wall (cpu_acc_map.cpp:325-338). */
int vo_map_1d(const vo_mesh* m, uint32_t n_in, const uint32_t* gids_in, const float* f_in,
              double I, double di, double dj, double dk, int dim, int method,
              double min_value, int wall_margin,
              uint32_t cap_out, uint32_t* n_out, uint32_t* gids_out, float* f_out, int* wall_flag);

/* Full 3-D step of one cell: three map_1d calls in map_order (cpu_acc_semilag.cpp:92-130).
 * intersections[12] = {x: I,di,dj,dk, y: ..., z: ...}. In-place on a malloc'd cell. */
typedef struct vo_cell {
   uint32_t n;
   uint32_t cap;
   uint32_t* gids;
   float* f;
} vo_cell;
int  vo_cell_init(vo_cell* c, uint32_t n, const uint32_t* gids, const float* f);
void vo_cell_free(vo_cell* c);
int  vo_accelerate_cell(const vo_mesh* m, vo_cell* c, const double intersections[12], int map_order,
                        int method, double min_value, int wall_margin, int* wall_flag);

/* ---------------- sparsity / block adjust (spatial_cell_cpu.cpp:179-316) ---------------- */
/* Single-cell adjust with doDeleteEmptyBlocks=true.  Output sorted by GID.  The removed
 * mass is accumulated in the order of the no-content list reversed (:227-267). */
int vo_adjust(const vo_mesh* m, uint32_t n_in, const uint32_t* gids, const float* f,
              double min_value, int add_width,
              uint32_t cap, uint32_t* n_out, uint32_t* gids_out, float* f_out, double* rho_loss);
/* The same with the spatial neighbours' content blocks: extra_gids (n_extra, any order, may
 * repeat) are kept and added as they are, without velocity dilation (:212-222); this is
 * SpatialCell::adjust_velocity_blocks(spatial_neighbors, popID, true), :179-273. */
int vo_adjust_spatial(const vo_mesh* m, uint32_t n_in, const uint32_t* gids, const float* f,
                      double min_value, int add_width, uint32_t n_extra, const uint32_t* extra_gids,
                      uint32_t cap, uint32_t* n_out, uint32_t* gids_out, float* f_out, double* rho_loss);
/* Sparsity threshold of a cell, SpatialCell::updateSparseMinValue (spatial_cell_cpu.cpp:699-727):
 * algorithm 0 -> min_value; 1 (x = rho) or 2 (x = nblocks) -> k x + b on the line through
 * (bulk1, min1) and (bulk2, min2), below min1 -> min1, above min2 -> min2. */
double vo_sparse_min_value(int algorithm, double min_value, double bulk1, double bulk2, double min1, double min2,
                           double rho, uint32_t nblocks);
/* sparse_conserve_mass (block_adjust_cpu.cpp:96-113): pre = sum of f_pre, post = sum of
 * f_post (Real = double sums in block order), and, if post != 0, f_post *= pre / post with
 * each product rounded to Realf. */
void vo_conserve_mass(uint32_t n_pre, const float* f_pre, uint32_t n_post, float* f_post);
/* content flag per block: any f >= min_value (spatial_cell_cpu.cpp:290-316) */
void vo_content_flags(uint32_t n, const float* f, double min_value, uint8_t* flags);

/* ---------------- moments (arch_moments.h:83-108) ---------------- */
void vo_moments(const vo_mesh* m, uint32_t n, const uint32_t* gids, const float* f, double out[4]);
/* second moments (arch_moments.h:123-150): out = mass * sum f (v - V)_a (v - V)_b DV3 for
 * xx, yy, zz, yz, xz, xy, v at cell centres, summed per value in block order */
void vo_pressure(const vo_mesh* m, uint32_t n, const uint32_t* gids, const float* f, const double V[3],
                 double mass, double out[6]);

/* ---------------- spatial translation (TRANS_SEMILAG_PPM) ---------------- */
#define VO_TRANS_STENCIL 2   /* VLASOV_STENCIL_WIDTH for PPM, definitions.h:78-84 */
/* PPM coefficients on non-uniform widths h (cpu_1d_ppm_nonuniform.hpp:40-62); h and v are
 * indexed like the stencil, source cell k. */
void vo_ppm_nu_coeff(const float* h, const float* v, int k, float threshold, float a[3]);
/* n stencils of 5 values (source cell at 2), one set of 5 widths: out n x 3 */
void vo_ppm_nu_batch(const float* h5, const float* stencils, int n, float threshold, float* out);
/* trans_map_1d_amr over the given pencils (cpu_trans_map_amr.cpp:214-441); f in place */
int vo_trans_map_1d(const vo_mesh* m, uint32_t ncells, const uint64_t* offsets, const uint32_t* gids, float* f,
                    uint32_t npencils, const uint32_t* plen, const uint64_t* pstart, const uint32_t* pcell,
                    const float* pdz, const float* pratio, int dim, float dt, float threshold);

/* ---------------- transform and intersections (FP64) ---------------- */
/* cpu_acc_transform.cpp:66-153.  fields: B[3], dB[6] = {dBXdy,dBXdz,dBYdx,dBYdz,dBZdx,dBZdy},
 * V[3] bulk velocity, rhoq, EgradPe[3].  Output: 3x4 affine, row-major {L00 L01 L02 t0; ...}. */
typedef struct vo_fields {
   double B[3];
   double dB[6];
   double V[3];
   double rhoq;
   double EgradPe[3];
} vo_fields;
void vo_transform(const vo_fields* fl, double mass, double charge, double dt,
                  double hall_min_rhoq, int ohm_gradpe, double out[12]);
/* Affine inverse of a 3x4 transform (Transform::inverse for Affine mode). */
void vo_affine_inverse(const double in[12], double out[12]);
/* cpu_acc_intersections.cpp:44-450: out[12] indexed by dimension (x,y,z) x (I,di,dj,dk). */
void vo_intersections(const vo_mesh* m, const double fwd[12], int map_order, double out[12]);
/* cpu_acc_transform.cpp:41-55: max_v_dt */
double vo_max_v_dt(const double B[3], double mass, double charge, double max_rotation_deg);

/* ---------------- synthetic initial condition ---------------- */
/* Sum of tri-Maxwellians (projects/project.h:66-72, MultiPeak.cpp:127-222) plus an optional
 * gyrotropic ring, sampled at cell centres over a bounding box of candidate blocks, then the
 * single-cell adjust keeps content blocks and their neighbours. */
typedef struct vo_peak {
   int kind;          /* 0: Maxwellian, 1: ring about axis */
   double n;          /* density */
   double V[3];       /* centre velocity */
   double T;          /* temperature (Maxwellian) */
   double axis[3];    /* ring axis (unit) */
   double vperp;      /* ring perpendicular speed */
   double vth;        /* ring thermal speed */
} vo_peak;
double vo_peak_value(const vo_peak* p, double mass, double vx, double vy, double vz);
int vo_init_cell(const vo_mesh* m, const vo_peak* peaks, int npeaks, double mass, double min_value,
                 int add_width, uint32_t cap, uint32_t* n_out, uint32_t* gids_out, float* f_out);

/* ---------------- batched OpenMP driver (cpu baseline) ---------------- */
/* Accelerates ncells cells with schedule(dynamic,1) as cpu_acc_semilag.cpp:66.  Cells are
 * packed: offsets[ncells+1] into gids/f; outputs written to per-cell malloc'd storage and
 * copied back into out arrays with out_offsets (cap total).  Returns total output blocks or -1. */
int64_t vo_accelerate_cells(const vo_mesh* m, int ncells, const uint32_t* offsets,
                            const uint32_t* gids, const float* f, const double* intersections,
                            int map_order, int method, double min_value, int wall_margin,
                            int do_adjust, int add_width, int nthreads,
                            uint64_t cap, uint32_t* out_offsets, uint32_t* gids_out, float* f_out,
                            int* wall_flags);

#ifndef VO_REALF_DOUBLE
/* ---------------- 16-bit block storage (the library's BF16 / FP16 forms) ---------------- */
enum { VO_ST_F32 = 0, VO_ST_BF16 = 1, VO_ST_F16 = 2 };
uint16_t vo_f32_to_bf16(float x);
float vo_bf16_to_f32(uint16_t h);
uint16_t vo_f32_to_f16(float x);
float vo_f16_to_f32(uint16_t h);
float vo_block_csum(const float x[64]);
void vo_encode_block(const float x[64], int storage, uint16_t code[64], float* scale);
void vo_decode_block(const uint16_t code[64], float scale, int storage, float out[64]);
/* element-wise RNE conversion to the 16-bit code */
void vo_convert16(int storage, const float* x, uint32_t n, uint16_t* out);
/* encode then decode every block in place (the stored values of the 16-bit forms) */
void vo_requantize(uint32_t nblocks, float* f, int storage);
#endif

#ifdef __cplusplus
}
#endif
#endif
