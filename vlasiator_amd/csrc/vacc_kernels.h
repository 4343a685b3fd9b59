// vacc_kernels.h -- kernel argument blocks and launch wrappers of libvacc.
#pragma once

#include "vacc_common.h"

namespace vacc {

// A pool of block data: FP32 (64 floats per block) or 16-bit codes (64 per block) plus one
// FP32 scale per block (vacc_codec.h).  storage: 0 FP32, 1 BF16, 2 FP16.
struct Pool {
   void* f;
   float* scale;
   int storage;
};

// Work unit of the remap (one column, or a whole column set whose columns' target ranges
// share a block), 16 bytes.
//   x: b0 | b1 << 16   transverse block indices in the pass frame
//   y: record index of the first column (col_rec)
//   z: number of columns | COPY_FLAG | EXACT_FLAG | WIDE_FLAG
//   w: source blocks of the unit (its work)
// Column record, 16 bytes.
//   x: first block index along the pass axis | length << 16
//   y: position of the first block in the sorted source list (sorted_src)
//   z: first target block | last target block << 16
//   w: output slot of the first target block (cell-local)
constexpr uint32_t COPY_FLAG = 0x80000000u;
// The tolerance-form remap (k_remap_lean) leaves a unit to the exact kernel when one of its
// columns starts with every lane of an 8-lane group at or below the Lagrangian base (the
// reference's maximum-lane rule, cpu_acc_map.cpp:515-529, then bounds the deposits by the
// group's first lane), or when its cell's deposits over a source block could outrun the ring.
constexpr uint32_t EXACT_FLAG = 0x40000000u;
// Units of cells whose deposit span exceeds MAX_SPAN_CELLS (up to MAX_SPAN_WIDE) go to the
// wide instance of the exact kernel, in either arithmetic form.
constexpr uint32_t WIDE_FLAG = 0x20000000u;
constexpr uint32_t UNIT_NCOL_MASK = 0x1FFFFFFFu;
constexpr int LEN_BINS = 256;   // length bins of the work-unit ordering
// The last two bins hold the units marked EXACT_FLAG and then those marked WIDE_FLAG, so that
// the work list ends with them: the tolerance-form kernel takes the list up to EXACT_BIN's
// start, the exact kernel the EXACT_BIN units (or, in the exact form, everything before
// WIDE_BIN), and the wide instance the WIDE_BIN units.
constexpr int EXACT_BIN = LEN_BINS - 2;
constexpr int WIDE_BIN = LEN_BINS - 1;
// RemapArgs::range of an exact-kernel launch
enum { RANGE_ALL = 0, RANGE_EXACT = 1, RANGE_WIDE = 2 };

struct ColumnArgs {
   const uint32_t* cell_off;   // [ncells+1] input list offsets
   const uint32_t* gid;        // input list GIDs
   const uint32_t* src;        // input list -> f pool index (INVALID = zero block)
   const double* isect;        // [ncells*12]
   uint32_t* sorted_src;       // [N] at cell_off
   uint4* unit_rec;            // [N] at cell_off
   uint4* col_rec;             // [N] at cell_off
   uint32_t* n_units;          // [ncells]
   uint32_t* len_hist;         // [LEN_BINS] work units per length bin (zeroed before the launch)
   uint32_t* n_targets;        // [ncells]
   uint32_t* status;           // [ncells + 1] OR-ed status bits; entry ncells ORs all cells
   uint32_t ncells;
   uint32_t len[3];
   GidSplit gs;                // block coordinates of a GID (from len)
   float dv;                   // (float) cell size along the pass axis
   float vmin;                 // (float) mesh minimum along the pass axis
   int wall_margin;
   const uint8_t* active;      // [ncells] or null: cells with 0 pass through unchanged (subcycling)
   int fast;                   // tolerance form: mark the units the exact kernel must take
   uint32_t* cell_hist;        // [LEN_BINS * ncells] the cell's work units per length bin (bin-major)
};

struct ScanArgs {
   const uint32_t* a;      // [n] counts
   const uint32_t* b;      // [n] counts (may be null)
   uint32_t* a_off;        // [n+1] exclusive scan
   uint32_t* b_off;        // [n+1]
   uint32_t n;
   uint32_t cap_a;         // capacity check on the a total (0 = none)
   uint32_t* flags;        // device flags word: bit VACC_STATUS_CAPACITY set on overflow
   const uint32_t* hist;   // [LEN_BINS] optional histogram to scan into hist_cursor
   uint32_t* hist_cursor;
};

struct OrderArgs {
   const uint32_t* cell_off;
   const uint32_t* unit_base;   // [ncells+1]
   const uint4* unit_rec;
   const uint32_t* hist_cursor; // [LEN_BINS] first list position of each length bin
   const uint32_t* cell_hist;   // [LEN_BINS * ncells] exclusive prefix over cells of the bin's units
   uint32_t* perm;              // [total units] unit record indices (unit_rec), longest first
   uint32_t* perm_cell;         // [total units] the cell of each perm entry
   const uint32_t* flags;
   uint32_t ncells;
};

struct RemapArgs {
   const uint32_t* cell_off;    // input list offsets (record base of each cell)
   const uint32_t* unit_base;   // [ncells+1] global work-unit index base
   const uint32_t* out_off;     // [ncells+1] output block offsets
   const uint32_t* perm;        // [total units] unit record indices in processing order
   const uint32_t* perm_cell;   // [total units] cell of each entry of perm
   const uint4* unit_rec;
   const uint4* col_rec;
   const uint32_t* sorted_src;
   const double* isect;
   Pool in, out;                // source and target pools
   uint32_t* gid_out;
   uint32_t* src_out;
   uint8_t* content_out;
   uint32_t* status;            // [ncells + 1], entry ncells ORs all cells
   const uint32_t* flags;       // global flags (capacity overflow -> skip)
   uint32_t ncells;
   uint32_t len[3];
   float dv, vmin;
   float threshold;             // (float) sparse min value, reconstruction threshold
   double min_value;            // content test threshold
   const double* minval;        // [ncells] sparsity threshold per cell (velocityBlockMinValue)
   double dv_d, vmin_d;         // cell size and mesh minimum along the pass axis (FP64)
   int range;                   // exact kernel: RANGE_ALL (every unit before WIDE_BIN's),
                                // RANGE_EXACT (EXACT_BIN's units), RANGE_WIDE (WIDE_BIN's units)
   const uint32_t* len_hist;    // [LEN_BINS] units per length bin; [EXACT_BIN] = EXACT_FLAG units,
                                // [WIDE_BIN] = WIDE_FLAG units
   float4* blk_stats;           // null, or per output block (at its slot) the sums of its stored
                                // values: S = sum f, and sum i f, sum j f, sum k f over the
                                // in-block indices (i, j, k) = 0..3
};

// LDS bitmap words per chunk of rows (k_columns, k_adjust)
constexpr int ADJ_BM_WORDS = 5120;

struct AdjustArgs {
   const uint32_t* cell_off;   // input list
   const uint32_t* gid;
   const uint32_t* src;
   const uint8_t* content;     // per input block, or null -> computed from f
   Pool pool;                  // pool the src indices point into
   const uint32_t* out_off;    // [ncells+1] (write phase)
   uint32_t* n_keep;           // [ncells] (count phase)
   uint32_t* gid_out;          // (write phase)
   uint32_t* src_out;          // (write phase)
   double* rho_loss;           // [ncells] (count phase, may be null)
   const float4* blk_stats;    // null, or per pool block its sums (the remap's blk_stats): the
                               // count phase then takes a removed block's mass from them
   const uint32_t* flags;      // global flags (capacity overflow -> the write phase is skipped)
   uint32_t ncells;
   uint32_t len[3];
   GidSplit gs;                // block coordinates of a GID (from len)
   double min_value;
   double dv3;
   int add_width;              // < 0: keep exactly the existing blocks (re-sort only)
   // spatial neighbours (null: none): cells of this list whose content blocks are kept
   // (nbr_off [ncells+1] into nbr), and caller GIDs to keep (ext_off [ncells+1] into ext_gid)
   const uint32_t* nbr_off;
   const uint32_t* nbr;
   const uint64_t* ext_off;
   const uint32_t* ext_gid;
   // per cell (null: not kept): the count phase's keep bitmap (ADJ_BM_WORDS words) and box
   // (8 ints: the block box lo x, y, z, hi x, y, z, and 1 when the bitmap was stored) for a
   // cell whose box fits one LDS chunk; the write phase reads them instead of building them
   uint32_t* keep_bm;
   int* keep_box;
};

struct MomentArgs {
   const uint32_t* cell_off;
   const uint32_t* gid;
   const uint32_t* src;
   Pool pool;
   double* out;                // [ncells*4]
   uint32_t ncells;
   uint32_t len[3];
   GidSplit gs;                // block coordinates of a GID (from len)
   double vmin[3];
   double cell[3];
   const float4* blk_stats;    // null, or per pool block its sums (the remap's blk_stats): the
                               // moments are then formed from them instead of from f
};

// synthetic initial state (vacc_init.hip)
struct InitArgs {
   const vacc_peak* peaks;
   const int32_t* npeaks;
   const int32_t* bbox;
   const uint32_t* bm_off;   // [ncells+1] word offsets of the per-cell content bitmaps
   uint32_t* content_bm;
   uint32_t* n_keep;
   const uint32_t* out_off;
   uint32_t* gid_out;
   uint32_t* src_out;
   float* f_out;
   uint32_t ncells;
   uint32_t max_peaks;
   uint32_t len[3];
   double vmin[3];
   double cell[3];
   double mass;
   double min_value;
   int add_width;
};

void launch_init(const InitArgs& a, bool write_phase, hipStream_t s);
int init_box_words_max();
void launch_content(const uint32_t* src, const Pool& pool, uint8_t* content, uint32_t n, double mv, hipStream_t s);

void launch_columns(int dim, const ColumnArgs& a, hipStream_t s);
void launch_scan(const ScanArgs& a, hipStream_t s);
void launch_order(const OrderArgs& a, hipStream_t s);
void launch_bin_scan(uint32_t* cell_hist, uint32_t ncells, hipStream_t s);
// grid: grid_mult persistent workgroups per CU
// grid_mult workgroups per compute unit; wgs: lanes per workgroup of the tolerance-form kernel
// (64, 128 or 256: 4, 8 or 16 work units); the exact kernel always runs 256-lane workgroups
void launch_remap(int dim, int method, bool fast, RemapArgs a, int grid_mult, int wgs, hipStream_t s);
void launch_adjust(const AdjustArgs& a, bool write_phase, hipStream_t s);
void launch_moments(const MomentArgs& a, hipStream_t s);
// plasma pivot from the moments of up to MAX_POP populations holding the same cells
constexpr int MAX_POP = 8;
struct PlasmaArgs {
   const double* mom[MAX_POP];   // [4*ncells] per population: n, nVx, nVy, nVz
   double mass[MAX_POP];
   double charge[MAX_POP];
   int npop;
   uint32_t ncells;
   double* out;                  // [4*ncells]: VX_V, VY_V, VZ_V, RHOQ_V
};
void launch_plasma(const PlasmaArgs& a, hipStream_t s);
// use_moments: 0 fields V and rhoq; 1 this population's moments; 2 plasma pivot buffer
// Subcycles of one acceleration step (vlasovmover.cpp:398-456): per cell max_v_dt from the
// resident fields' B (cpu_acc_transform.cpp:41-55) and the subcycle count
// max(ceil(|dt| / max_v_dt), 1); the context's maximum goes to *nmax (atomicMax).
void launch_subcycle_count(const TransformIn* fields, uint32_t ncells, double mass, double charge,
                           double max_rotation_deg, double dt, double* maxdt, uint32_t* nsub, uint32_t* nmax,
                           hipStream_t s);
// Subcycle `step`: active[c] = step < nsub[c]; subdt[c] = max_v_dt, or what is left of |dt| on
// the cell's last subcycle, with the sign of dt (0 for inactive cells).
void launch_subcycle_step(const double* maxdt, const uint32_t* nsub, uint32_t ncells, uint32_t step, double dt,
                          double* subdt, uint8_t* active, hipStream_t s);
// Sparsity threshold of every cell, SpatialCell::updateSparseMinValue (spatial_cell_cpu.cpp:699-727):
// algorithm 1 takes x = the density moments[4c], 2 the cell's block count from off[].
struct SparseArgs {
   int algorithm;
   double min_value, bulk1, bulk2, min1, min2;
};
void launch_sparse_min(const SparseArgs& p, const double* moments, const uint32_t* off, double* minval,
                       uint32_t ncells, hipStream_t s);
// content flag of every block of a list with the cell's threshold (one workgroup per cell)
void launch_content_cells(const uint32_t* off, const uint32_t* src, const Pool& pool, const double* minval,
                          uint8_t* content, uint32_t ncells, hipStream_t s);
// FP64 sum of f over each cell's blocks
void launch_cell_sums(const uint32_t* off, const uint32_t* src, const Pool& pool, double* out, uint32_t ncells,
                      hipStream_t s);
// sparse_conserve_mass (block_adjust_cpu.cpp:96-113): f *= pre[c] / post[c] where post != 0,
// each product rounded to FP32 (16-bit forms: the block scale)
void launch_cell_scale(const uint32_t* off, const uint32_t* src, const Pool& pool, const double* pre,
                       const double* post, uint32_t ncells, hipStream_t s);
void launch_fill_f64(double* dst, double v, uint32_t n, hipStream_t s);
// Second moments (arch_moments.h:123-150): out[6c..] = mass * sum f (v - V)_a (v - V)_b DV3 for
// (a,b) = xx, yy, zz, yz, xz, xy about the pivot pivot[3c..] (FP64 sums)
void launch_pressure(const MomentArgs& a, const double* pivot, double mass, double* out, hipStream_t s);
// dt_override: [ncells] or null; replaces fields[c].dt.
void launch_transform(const MeshInfo& m, const TransformIn* fields, const double* dt_override, const double* moments, int use_moments,
                      double mass, double charge, double hall_min_rhoq, int ohm_gradpe, int map_order,
                      double* isect_out, uint32_t ncells, hipStream_t s);
void launch_gather(const uint32_t* src, const Pool& in, const Pool& out, uint32_t n, hipStream_t s);
// FP32 blocks <-> 16-bit pool (one wave per block)
void launch_encode(const float* f, const Pool& out, uint32_t n, hipStream_t s);
void launch_decode(const Pool& in, float* f, uint32_t n, hipStream_t s);
void launch_iota(uint32_t* dst, uint32_t n, hipStream_t s);
void launch_fill_u32(uint32_t* dst, uint32_t v, uint32_t n, hipStream_t s);
void launch_find(const uint32_t* off, const uint32_t* gid, uint32_t ncells, const uint32_t* qcell, const uint32_t* qgid,
                 uint32_t* out, uint32_t nq, hipStream_t s);

int device_cus();   // compute units of the current device

// ---------------------------------------------------------------------------
// Spatial translation (vacc_trans.hip)
// ---------------------------------------------------------------------------
// Width-only scalars of the h4 face estimate on non-uniform widths for one face, with
// h0..h3 = h[k-2..k+1] (cpu_face_estimates.hpp:179-191): the float and double
// sub-expressions of the reference, evaluated once per face on the host.
struct TransFace {
   double D1;          // 1 / (h0 + h1 + h2 + h3)
   double DA;          // 1 / (h0 + h1 + h2) + 1 / (h1 + h2 + h3)
   double DB;          // h0 + 2 h1
   double DC;          // 2 h2 + h3
   float SA;           // (h0 + h1)(h2 + h3) / (h1 + h2)
   float SB;           // h2 (h2 + h3) / ((h0 + h1 + h2)(h0 + h1))
   float SC;           // h1 (h0 + h1) / ((h1 + h2 + h3)(h2 + h3))
   float h1, h2;
   float pad[3];
};
// One contribution to a target cell: from source position i of a pencil (stencil i-2..i+2)
// into position i-1 (kind 2), i (kind 1) or i+1 (kind 0).
constexpr int TRANS_MAXN = 12;   // distinct stencil cells per target cell
struct TransEntry {
   TransFace fl, fr;   // faces k = i and k = i + 1
   float dz_i;         // source width
   float dz_t;         // target width (kinds 0 and 2)
   float ratio;        // target ratio
   uint8_t st[5];      // stencil cells as slots of the target cell's table (0xFF: none)
   uint8_t kind;
   uint8_t pad[2];
   uint32_t pad2[3];   // 160 bytes, a multiple of 16
};
struct TransCell {
   uint32_t e0;        // first entry
   uint32_t ne;        // entries (0: the cell is not written and keeps its values)
   uint32_t nn;        // stencil cells in nbr
   uint32_t nbr[TRANS_MAXN];
   uint32_t walked;    // 1: written by the pencil walk (k_trans_walk), skipped by k_trans
};
// Pencil walk: a segment of one pencil whose walked cells are all written by one wave per
// velocity block.  A walked cell receives every contribution of this pencil from this
// segment's sources, at one window position or at two that are at least three apart (a
// periodic pencil's wrap); a cell that several pencils write at one run position each, or
// whose two positions lie in different segments, is walked in launch levels, part after part.
// Window positions 0..n-1 are consecutive pencil positions; sources are window positions
// js..je, with js = 2 and je = n - 3; walked positions lie in js-1 .. je+1.
constexpr int WALK_MAXN = 64;    // window positions per segment (one lane each)
constexpr uint32_t WALK_W = 1u;           // flags: the position's cell is written by the segment
constexpr uint32_t WALK_SAVE_SHIFT = 1;   // flags bits 1-2: finished sum saved in carry 1 or 2
constexpr uint32_t WALK_INIT_SHIFT = 3;   // flags bits 3-4: sum starts from carry 1 or 2, or
constexpr uint32_t WALK_INIT_STORED = 3;  // from the block an earlier launch level stored (FP32)
struct WalkSeg {
   uint32_t pos0;      // first window position in the walk position and face arrays
   uint32_t n;         // window positions
   uint32_t js, je;    // first and last source
};
struct WalkPos {
   uint32_t cell;      // context cell (VACC_INVALID_ID: none)
   float dz;           // width along the dimension
   float ratio;        // target ratio
   uint32_t flags;     // WALK_W, save and init carry slots
};
struct TransArgs {
   const uint32_t* off;    // [ncells+1] list offsets (GIDs ascending within each cell)
   const uint32_t* gid;
   const uint32_t* src;    // list -> input pool index (INVALID = zero block)
   Pool in;
   Pool out;               // compact output pool (block b of the list at b), same storage form
   const TransCell* cells;
   const TransEntry* entries;
   const uint32_t* rest;   // [nrest] cells that k_trans writes or copies (walked == 0)
   uint32_t nrest;
   uint32_t ncells;
   uint32_t nblocks;
   int dim;
   uint32_t len[3];
   float dvz, vz_min;      // (float) cell size and mesh minimum along dim
   float dt;
   float thr;              // reconstruction threshold
   float scale;            // (float)(1 / thr)
};
void launch_trans(const TransArgs& a, hipStream_t s);
struct WalkArgs {
   TransArgs t;            // lists, pools and velocity constants (cells, entries, rest unused)
   const WalkSeg* seg;
   const WalkPos* pos;
   const TransFace* face;  // left face of each window position (indexed like pos)
   const uint32_t* bounds; // [nseg][wps + 1] GID bounds of each wave's share (k_walk_bounds)
   uint32_t nseg;
   uint32_t wps;           // waves per segment
   uint32_t nmesh;         // velocity mesh blocks
   uint32_t bin_shift;     // GID >> bin_shift < 4096: bins of the share histogram
};
// the share bounds are computed into `bounds` first ([nseg][wps + 1] words)
void launch_trans_walk(const WalkArgs& a, uint32_t* bounds, hipStream_t s);

}  // namespace vacc
