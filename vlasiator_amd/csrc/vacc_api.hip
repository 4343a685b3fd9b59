// vacc_api.hip -- host side of libvacc: device pools, pass orchestration and the C ABI
// declared in include/vacc.h.
//
// State of a context between calls (all on the device):
//   list L[lcur]: per-cell offsets, block GIDs and, per block, the index of its 64 values
//                 in pool F[fcur] (INVALID = a block created by adjust, all zero);
//   pool F[fcur]: block values.
// A remap pass reads L[lcur]/F[fcur] and writes a compact L[1-lcur]/F[1-fcur]; the block
// adjust writes a new list L[1-lcur] that still points into F[fcur].
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/vacc_trans.h"
#include "vacc_blocks_io.h"
#include "vacc_common.h"
#include "vacc_kernels.h"

using namespace vacc;

static thread_local std::string g_err;

static int set_err(int code, const std::string& msg) {
   g_err = msg;
   return code;
}

#define HIPCHK(expr)                                                                          \
   do {                                                                                       \
      hipError_t e_ = (expr);                                                                 \
      if (e_ != hipSuccess) return set_err(VACC_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
   } while (0)

struct vacc_ctx {
   int device = 0;             // the device current at vacc_create; every call runs on it
   vacc_mesh mesh;
   vacc_species sp;
   vacc_options opt;
   MeshInfo mi;
   hipStream_t stream = nullptr;
   uint32_t ncells = 0;
   uint32_t cell_cap = 0;      // capacity of per-cell arrays
   size_t cap = 0;             // block capacity of lists, pools and pass scratch
   int lcur = 0, fcur = 0;
   uint64_t nblocks = 0;       // total blocks as of the last synchronisation point
   bool content_valid = false;
   bool moments_valid = false; // `moments` holds the moments of the current list
   // blk_stats holds the block sums of pool F[fcur] (written by the last remap pass of a
   // tolerance-form step, vacc_options.sums_from_f = 0); the moments and the adjust's removed
   // mass are then formed from them
   bool stats_valid = false;
   bool fields_valid = false;  // `fields` holds the caller's transform inputs
   // lists and pools
   uint32_t* off[2] = {nullptr, nullptr};
   uint32_t* gid[2] = {nullptr, nullptr};
   uint32_t* src[2] = {nullptr, nullptr};
   void* f[2] = {nullptr, nullptr};      // block data: FP32, or 16-bit codes
   float* scl[2] = {nullptr, nullptr};   // per-block scales of the 16-bit forms
   uint8_t* content = nullptr;
   float4* blk_stats = nullptr;          // per pool block (S, sum i f, sum j f, sum k f), or null
   // pass scratch
   uint32_t* sorted_src = nullptr;
   uint4* unit_rec = nullptr;
   uint4* col_rec = nullptr;
   uint32_t* perm = nullptr;
   uint32_t* perm_cell = nullptr;
   uint32_t* n_units = nullptr;
   uint32_t* n_tgt = nullptr;
   uint32_t* unit_base = nullptr;
   uint32_t* len_hist = nullptr;     // [2*LEN_BINS]: histogram, cursors
   uint32_t* status = nullptr;
   uint32_t* flags = nullptr;
   double* isect = nullptr;
   TransformIn* fields = nullptr;
   double* moments = nullptr;
   double* pivot = nullptr;          // [4*ncells] plasma pivot of a multi-population step
   double* rho_loss = nullptr;
   // subcycling (vacc_subcycle_count / vacc_subcycle_step): per cell max_v_dt, subcycle count,
   // the current subcycle's dt and whether the cell takes part in it
   double* sub_maxdt = nullptr;
   double* sub_dt = nullptr;
   uint32_t* sub_n = nullptr;
   uint8_t* active = nullptr;
   uint32_t* sub_nmax = nullptr;     // device word: maximum subcycle count over the cells
   // per-cell sparsity threshold (velocityBlockMinValue) and the sums of sparse_conserve_mass
   double* minval = nullptr;
   double* csum_pre = nullptr;
   double* csum_post = nullptr;
   // the adjust count phase's keep bitmap and box per cell, read by its write phase
   uint32_t* keep_bm = nullptr;
   int* keep_box = nullptr;
   uint32_t* cell_hist = nullptr;   // [LEN_BINS * ncells] work units per (length bin, cell)
   bool sub_valid = false;           // sub_maxdt / sub_n hold the counts of the last vacc_subcycle_count
   bool subcycling = false;          // launches inside vacc_subcycle_step use sub_dt and active
   double sub_dt_total = 0.0;
   uint32_t sub_max = 0;
   hipEvent_t mom_ev = nullptr;      // moments of this population are ready (multi-population step)
   // spatial-neighbour adjust inputs (device copies, grown on demand)
   uint32_t* d_nbr_off = nullptr;
   uint32_t* d_nbr = nullptr;
   uint64_t* d_ext_off = nullptr;
   uint32_t* d_ext = nullptr;
   size_t nbr_cap = 0, ext_cap = 0, off_cap = 0, ext_off_cap = 0;
   // captured steps (vacc_options.use_graph): one executable graph per state variant; gen
   // changes whenever buffers move or the cell set changes, which retires every graph
   struct StepGraph {
      uint64_t gen;
      uint32_t map_order;
      int use_moments, adjust, lcur, fcur;
      bool moments_in, content_in, stats_in;
      hipGraphExec_t exec;
      int lcur_out, fcur_out;
      bool moments_out, content_out, sorted_out, stats_out;
   };
   std::vector<StepGraph> graphs;
   uint64_t gen = 0;
   int remap_mult[3] = {128, 128, 128};   // remap workgroups per CU, per pass dimension (vacc_autotune)
   int remap_wgs[3] = {256, 256, 256};    // lanes per workgroup of the tolerance-form remap (vacc_autotune)
   int remap_mult_s[3] = {128, 128, 128}; // the same for the passes that store block sums (their own
   int remap_wgs_s[3] = {256, 256, 256};  // kernel instance, with more registers in scratch)
   uint32_t* h_sum = nullptr;        // pinned host words: capacity flag, block total, status OR
   // streamed block transfer (vacc_blocks_io.h): two pinned chunks, their copy events and,
   // for the 16-bit forms, a device FP32 stage of two chunks; allocated on first use
   float* io_pin[2] = {nullptr, nullptr};
   float* io_dev = nullptr;
   hipEvent_t io_ev[2] = {nullptr, nullptr};
   // block lists in ascending GID order within each cell (after a block adjust or relist)
   bool sorted = false;
   uint64_t cell_epoch = 0;         // changes with every new cell set
   // spatial translation schedule per dimension (vacc_trans_set_pencils)
   struct TransDim {
      TransCell* cells = nullptr;
      TransEntry* entries = nullptr;
      uint32_t* rest = nullptr;    // cells that are not walked
      uint32_t* bounds = nullptr;  // walk share bounds, [nseg][wps + 1]
      size_t bounds_cap = 0;
      WalkSeg* seg = nullptr;      // pencil walk segments, their window positions and faces
      WalkPos* pos = nullptr;
      TransFace* face = nullptr;
      size_t cells_cap = 0, entries_cap = 0, rest_cap = 0, seg_cap = 0, pos_cap = 0, face_cap = 0;
      uint32_t nrest = 0;
      uint32_t nseg = 0;
      uint32_t wps = 1;            // walk waves per segment
      std::vector<uint32_t> level_first;   // walk launch levels: first segment of each, and the end
      uint64_t epoch = ~0ull;
   } trans[3];
   // event timing
   struct Rec {
      hipEvent_t a, b;
      int kind;   // 0 remap, 1 build, 2 adjust, 3 transform, 4 translation
   };
   bool prof = false;
   std::vector<Rec> recs;
   std::vector<hipEvent_t> ev_free;
   uint32_t* prof_counts = nullptr;   // [2*PROF_LOG] input/output totals of each remap
   uint32_t nlog = 0;
   std::vector<uint8_t> prof_sums;    // [PROF_LOG] the logged remap stored block sums (16 B per target block)
};

static constexpr uint32_t PROF_LOG = 16384;

static hipEvent_t take_event(vacc_ctx* c) {
   if (!c->ev_free.empty()) {
      hipEvent_t e = c->ev_free.back();
      c->ev_free.pop_back();
      return e;
   }
   hipEvent_t e = nullptr;
   (void)hipEventCreate(&e);
   return e;
}

// Bracket a region of launches on the context stream with events when profiling.
struct ProfScope {
   vacc_ctx* c;
   int kind;
   hipEvent_t a = nullptr;
   ProfScope(vacc_ctx* c_, int kind_) : c(c_), kind(kind_) {
      if (c->prof) {
         a = take_event(c);
         (void)hipEventRecord(a, c->stream);
      }
   }
   ~ProfScope() {
      if (c->prof && a) {
         hipEvent_t b = take_event(c);
         (void)hipEventRecord(b, c->stream);
         c->recs.push_back({a, b, kind});
      }
   }
};

static void drop_graphs(vacc_ctx* c) {
   for (auto& g : c->graphs) (void)hipGraphExecDestroy(g.exec);
   c->graphs.clear();
   c->gen++;
}

static void free_blocks(vacc_ctx* c) {
   drop_graphs(c);
   for (int i = 0; i < 2; ++i) {
      if (c->gid[i]) (void)hipFree(c->gid[i]);
      if (c->src[i]) (void)hipFree(c->src[i]);
      if (c->f[i]) (void)hipFree(c->f[i]);
      if (c->scl[i]) (void)hipFree(c->scl[i]);
      c->scl[i] = nullptr;
      c->gid[i] = c->src[i] = nullptr;
      c->f[i] = nullptr;
   }
   if (c->content) (void)hipFree(c->content);
   if (c->blk_stats) (void)hipFree(c->blk_stats);
   c->blk_stats = nullptr;
   c->stats_valid = false;
   if (c->sorted_src) (void)hipFree(c->sorted_src);
   if (c->unit_rec) (void)hipFree(c->unit_rec);
   if (c->col_rec) (void)hipFree(c->col_rec);
   if (c->perm) (void)hipFree(c->perm);
   if (c->perm_cell) (void)hipFree(c->perm_cell);
   c->perm_cell = nullptr;
   c->content = nullptr;
   c->sorted_src = nullptr;
   c->unit_rec = c->col_rec = nullptr;
   c->perm = nullptr;
   c->cap = 0;
}

static void free_cells(vacc_ctx* c) {
   for (int i = 0; i < 2; ++i) {
      if (c->off[i]) (void)hipFree(c->off[i]);
      c->off[i] = nullptr;
   }
   uint32_t** u32s[] = {&c->n_units, &c->n_tgt, &c->unit_base, &c->status, &c->len_hist};
   for (auto p : u32s) {
      if (*p) (void)hipFree(*p);
      *p = nullptr;
   }
   if (c->isect) (void)hipFree(c->isect);
   if (c->fields) (void)hipFree(c->fields);
   if (c->moments) (void)hipFree(c->moments);
   if (c->pivot) (void)hipFree(c->pivot);
   if (c->rho_loss) (void)hipFree(c->rho_loss);
   void* subs[] = {c->sub_maxdt, c->sub_dt, c->sub_n, c->active, c->sub_nmax, c->minval, c->csum_pre, c->csum_post,
                   c->keep_bm, c->keep_box, c->cell_hist};
   for (void* p : subs)
      if (p) (void)hipFree(p);
   c->keep_bm = nullptr;
   c->keep_box = nullptr;
   c->cell_hist = nullptr;
   c->sub_maxdt = c->sub_dt = nullptr;
   c->sub_n = c->sub_nmax = nullptr;
   c->minval = c->csum_pre = c->csum_post = nullptr;
   c->active = nullptr;
   c->sub_valid = false;
   c->isect = nullptr;
   c->fields = nullptr;
   c->moments = nullptr;
   c->pivot = nullptr;
   c->rho_loss = nullptr;
   c->cell_cap = 0;
}

static int alloc_cells(vacc_ctx* c, uint32_t n) {
   drop_graphs(c);   // the cell count or per-cell buffers change
   if (n <= c->cell_cap && c->off[0]) return 0;
   free_cells(c);
   const size_t m = (size_t)n + 1;
   for (int i = 0; i < 2; ++i) HIPCHK(hipMalloc(&c->off[i], m * sizeof(uint32_t)));
   HIPCHK(hipMalloc(&c->n_units, m * sizeof(uint32_t)));
   HIPCHK(hipMalloc(&c->n_tgt, m * sizeof(uint32_t)));
   HIPCHK(hipMalloc(&c->unit_base, m * sizeof(uint32_t)));
   HIPCHK(hipMalloc(&c->len_hist, 2 * LEN_BINS * sizeof(uint32_t)));
   HIPCHK(hipMalloc(&c->status, m * sizeof(uint32_t)));
   HIPCHK(hipMemset(c->status, 0, m * sizeof(uint32_t)));
   HIPCHK(hipMalloc(&c->isect, m * 12 * sizeof(double)));
   HIPCHK(hipMemset(c->isect, 0, m * 12 * sizeof(double)));
   HIPCHK(hipMalloc(&c->fields, m * sizeof(TransformIn)));
   HIPCHK(hipMalloc(&c->moments, m * 4 * sizeof(double)));
   HIPCHK(hipMalloc(&c->pivot, m * 4 * sizeof(double)));
   HIPCHK(hipMalloc(&c->rho_loss, m * sizeof(double)));
   HIPCHK(hipMalloc(&c->sub_maxdt, m * sizeof(double)));
   HIPCHK(hipMalloc(&c->sub_dt, m * sizeof(double)));
   HIPCHK(hipMalloc(&c->sub_n, m * sizeof(uint32_t)));
   HIPCHK(hipMalloc(&c->active, m));
   HIPCHK(hipMalloc(&c->sub_nmax, sizeof(uint32_t)));
   HIPCHK(hipMalloc(&c->minval, m * sizeof(double)));
   HIPCHK(hipMalloc(&c->csum_pre, m * sizeof(double)));
   HIPCHK(hipMalloc(&c->csum_post, m * sizeof(double)));
   HIPCHK(hipMalloc(&c->keep_bm, (size_t)n * ADJ_BM_WORDS * sizeof(uint32_t)));
   HIPCHK(hipMalloc(&c->keep_box, m * 8 * sizeof(int)));
   HIPCHK(hipMalloc(&c->cell_hist, (size_t)LEN_BINS * n * sizeof(uint32_t)));
   c->cell_cap = n;
   return 0;
}

static size_t elem_bytes(const vacc_ctx* c) { return c->opt.storage == VACC_STORE_F32 ? sizeof(float) : sizeof(uint16_t); }

static Pool pool(const vacc_ctx* c, int i) { return Pool{c->f[i], c->scl[i], c->opt.storage}; }

// the tolerance form's block sums (blk_stats) are allocated with the block storage
static bool use_block_sums(const vacc_ctx* c) { return c->opt.fast_arith && !c->opt.sums_from_f; }

// (Re)allocate block storage for `want` blocks, preserving the current list and pool.  On
// failure the new buffers are freed and the old ones are kept.
static int ensure_blocks(vacc_ctx* c, size_t want) {
   if (want <= c->cap && c->gid[0]) return 0;
   if (want > 0xF0000000ull) return set_err(VACC_E_NOMEM, "block count exceeds 32-bit list indices");
   const size_t ncap = std::max<size_t>(want, 1024);
   uint32_t *g[2] = {nullptr, nullptr}, *s[2] = {nullptr, nullptr};
   void* f[2] = {nullptr, nullptr};
   float* sc[2] = {nullptr, nullptr};
   uint8_t* ct = nullptr;
   uint32_t *ss = nullptr, *pm = nullptr, *pc = nullptr;
   uint4 *sr = nullptr, *cr = nullptr;
   float4* bs = nullptr;
   auto release = [&]() {
      void* all[] = {g[0], g[1], s[0], s[1], f[0], f[1], sc[0], sc[1], ct, ss, pm, pc, sr, cr, bs};
      for (void* p : all)
         if (p) (void)hipFree(p);
   };
   hipError_t e = hipSuccess;
   for (int i = 0; i < 2 && e == hipSuccess; ++i) {
      e = hipMalloc(&g[i], ncap * sizeof(uint32_t));
      if (e == hipSuccess) e = hipMalloc(&s[i], ncap * sizeof(uint32_t));
      if (e == hipSuccess) e = hipMalloc(&f[i], ncap * WID3 * elem_bytes(c));
      if (e == hipSuccess && c->opt.storage != VACC_STORE_F32) e = hipMalloc(&sc[i], ncap * sizeof(float));
   }
   if (e == hipSuccess) e = hipMalloc(&ct, ncap);
   if (e == hipSuccess) e = hipMalloc(&ss, ncap * sizeof(uint32_t));
   if (e == hipSuccess) e = hipMalloc(&sr, ncap * sizeof(uint4));
   if (e == hipSuccess) e = hipMalloc(&cr, ncap * sizeof(uint4));
   if (e == hipSuccess) e = hipMalloc(&pm, ncap * sizeof(uint32_t));
   if (e == hipSuccess) e = hipMalloc(&pc, ncap * sizeof(uint32_t));
   if (e == hipSuccess && use_block_sums(c)) e = hipMalloc(&bs, ncap * sizeof(float4));
   if (e == hipSuccess && c->cap && c->nblocks) {
      // carry the current state: list entries and the whole current pool extent
      const size_t n = c->nblocks;
      e = hipMemcpyAsync(g[c->lcur], c->gid[c->lcur], n * sizeof(uint32_t), hipMemcpyDeviceToDevice, c->stream);
      if (e == hipSuccess)
         e = hipMemcpyAsync(s[c->lcur], c->src[c->lcur], n * sizeof(uint32_t), hipMemcpyDeviceToDevice, c->stream);
      if (e == hipSuccess)
         e = hipMemcpyAsync(f[c->fcur], c->f[c->fcur], c->cap * WID3 * elem_bytes(c), hipMemcpyDeviceToDevice, c->stream);
      if (e == hipSuccess && sc[c->fcur])
         e = hipMemcpyAsync(sc[c->fcur], c->scl[c->fcur], c->cap * sizeof(float), hipMemcpyDeviceToDevice, c->stream);
      if (e == hipSuccess) e = hipMemcpyAsync(ct, c->content, n, hipMemcpyDeviceToDevice, c->stream);
      if (e == hipSuccess && bs && c->blk_stats && c->stats_valid)
         e = hipMemcpyAsync(bs, c->blk_stats, c->cap * sizeof(float4), hipMemcpyDeviceToDevice, c->stream);
      const hipError_t e2 = hipStreamSynchronize(c->stream);
      if (e == hipSuccess) e = e2;
   }
   if (e != hipSuccess) {
      release();
      (void)hipGetLastError();
      return set_err(e == hipErrorOutOfMemory ? VACC_E_NOMEM : VACC_E_HIP,
                     std::string("block storage for ") + std::to_string(ncap) + " blocks: " + hipGetErrorString(e));
   }
   const bool keep_stats = bs && c->blk_stats && c->stats_valid && c->nblocks;
   free_blocks(c);
   for (int i = 0; i < 2; ++i) {
      c->gid[i] = g[i];
      c->src[i] = s[i];
      c->f[i] = f[i];
      c->scl[i] = sc[i];
   }
   c->content = ct;
   c->sorted_src = ss;
   c->unit_rec = sr;
   c->col_rec = cr;
   c->perm = pm;
   c->perm_cell = pc;
   c->blk_stats = bs;
   c->stats_valid = keep_stats;   // carried with the pool
   c->cap = ncap;
   return 0;
}

// Read the total block count of the current list (synchronises the stream).
static int refresh_count(vacc_ctx* c) {
   uint32_t tot = 0;
   HIPCHK(hipMemcpyAsync(&tot, c->off[c->lcur] + c->ncells, sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
   HIPCHK(hipStreamSynchronize(c->stream));
   c->nblocks = tot;
   return 0;
}

// Keep head-room for the growth of one step (new target blocks and adjust halo blocks).
static int reserve_for_step(vacc_ctx* c) {
   const size_t want = c->nblocks + c->nblocks / 2 + 64 * (size_t)c->ncells + 4096;
   if (want > c->cap) return ensure_blocks(c, want + want / 4);
   return 0;
}

static float cell_f(const vacc_ctx* c, int d) { return (float)c->mi.cell[d]; }

// One 1-D remap of every cell.
// Column build and remap of one pass from (lcur, fcur) into the other list and pool; the
// caller flips the parity.  grid_mult: persistent remap workgroups per CU (0: the tuned or
// default value).
// grid_mult / wgs: the remap launch shape, or 0 for the context's tuned shape
// stats: the pass also writes the block sums of its output (blk_stats)
static int enqueue_pass(vacc_ctx* c, int dim, int grid_mult, int wgs = 0, bool stats = false) {
   const int li = c->lcur, lo = 1 - c->lcur, fi = c->fcur, fo = 1 - c->fcur;
   ColumnArgs ca{};
   ca.cell_off = c->off[li];
   ca.gid = c->gid[li];
   ca.src = c->src[li];
   ca.isect = c->isect;
   ca.sorted_src = c->sorted_src;
   ca.unit_rec = c->unit_rec;
   ca.col_rec = c->col_rec;
   ca.n_units = c->n_units;
   ca.n_targets = c->n_tgt;
   ca.len_hist = c->len_hist;
   ca.status = c->status;
   ca.ncells = c->ncells;
   for (int d = 0; d < 3; ++d) ca.len[d] = c->mi.len[d];
   ca.gs = make_gid_split(c->mi.len);
   ca.dv = cell_f(c, dim);                 // Realf dv = cellSize (cpu_acc_map.cpp:147)
   ca.vmin = (float)c->mi.vmin[dim];       // Realf v_min (cpu_acc_map.cpp:148)
   ca.wall_margin = c->opt.wall_margin;
   ca.active = c->subcycling ? c->active : nullptr;
   ca.fast = c->opt.fast_arith != 0;
   ca.cell_hist = c->cell_hist;
   {
      ProfScope ps(c, 1);
      launch_fill_u32(c->len_hist, 0u, LEN_BINS, c->stream);
      launch_columns(dim, ca, c->stream);
      ScanArgs sa{};
      sa.a = c->n_tgt;
      sa.b = c->n_units;
      sa.a_off = c->off[lo];
      sa.b_off = c->unit_base;
      sa.n = c->ncells;
      sa.cap_a = (uint32_t)std::min<size_t>(c->cap, 0xFFFFFFFFull);
      sa.flags = c->flags;
      sa.hist = c->len_hist;
      sa.hist_cursor = c->len_hist + LEN_BINS;
      launch_scan(sa, c->stream);
      launch_bin_scan(c->cell_hist, c->ncells, c->stream);
      OrderArgs oa{};
      oa.cell_off = c->off[li];
      oa.unit_base = c->unit_base;
      oa.unit_rec = c->unit_rec;
      oa.hist_cursor = c->len_hist + LEN_BINS;
      oa.cell_hist = c->cell_hist;
      oa.perm = c->perm;
      oa.perm_cell = c->perm_cell;
      oa.flags = c->flags;
      oa.ncells = c->ncells;
      launch_order(oa, c->stream);
   }
   RemapArgs ra{};
   ra.cell_off = c->off[li];
   ra.unit_base = c->unit_base;
   ra.out_off = c->off[lo];
   ra.perm = c->perm;
   ra.perm_cell = c->perm_cell;
   ra.unit_rec = c->unit_rec;
   ra.col_rec = c->col_rec;
   ra.sorted_src = c->sorted_src;
   ra.isect = c->isect;
   ra.in = pool(c, fi);
   ra.out = pool(c, fo);
   ra.gid_out = c->gid[lo];
   ra.src_out = c->src[lo];
   ra.content_out = c->content;
   ra.status = c->status;
   ra.flags = c->flags;
   ra.ncells = c->ncells;
   for (int d = 0; d < 3; ++d) ra.len[d] = c->mi.len[d];
   ra.dv = ca.dv;
   ra.vmin = ca.vmin;
   ra.threshold = (float)c->sp.sparse_min_value;   // Realf threshold argument of compute_*_coeff
   ra.min_value = c->sp.sparse_min_value;
   ra.minval = c->minval;
   ra.len_hist = c->len_hist;
   ra.dv_d = c->mi.cell[dim];
   ra.vmin_d = c->mi.vmin[dim];
   ra.blk_stats = (stats && c->blk_stats) ? c->blk_stats : nullptr;
   {
      ProfScope ps(c, 0);
      launch_remap(dim, c->opt.method, c->opt.fast_arith != 0, ra,
                   grid_mult ? grid_mult : (ra.blk_stats ? c->remap_mult_s[dim] : c->remap_mult[dim]),
                   wgs ? wgs : (ra.blk_stats ? c->remap_wgs_s[dim] : c->remap_wgs[dim]), c->stream);
   }
   HIPCHK(hipGetLastError());
   return 0;
}

// stats: the last pass of an acceleration (the block sums feed the adjust and the moments)
static int do_pass(vacc_ctx* c, int dim, bool stats = false) {
   const int li = c->lcur, lo = 1 - c->lcur, fo = 1 - c->fcur;
   stats = stats && c->blk_stats;
   const int rc = enqueue_pass(c, dim, 0, 0, stats);
   if (rc) return rc;
   if (c->prof && c->nlog < PROF_LOG) {
      HIPCHK(hipMemcpyAsync(c->prof_counts + 2 * c->nlog, c->off[li] + c->ncells, sizeof(uint32_t),
                            hipMemcpyDeviceToDevice, c->stream));
      HIPCHK(hipMemcpyAsync(c->prof_counts + 2 * c->nlog + 1, c->off[lo] + c->ncells, sizeof(uint32_t),
                            hipMemcpyDeviceToDevice, c->stream));
      c->prof_sums[c->nlog] = stats ? 1 : 0;
      c->nlog++;
   }
   c->lcur = lo;
   c->fcur = fo;
   c->content_valid = true;
   c->moments_valid = false;
   c->stats_valid = stats;
   c->sorted = false;   // column order of the pass
   return 0;
}

struct SpatialIn {
   const uint32_t* nbr_off = nullptr;   // device pointers or null
   const uint32_t* nbr = nullptr;
   const uint64_t* ext_off = nullptr;
   const uint32_t* ext = nullptr;
};

// Rebuild the list: add_width >= 0 is the sparsity adjust, add_width < 0 a re-sort by GID.
// With spatial-neighbour inputs the kept count can exceed the reserve: the total is read back
// after the count phase and the block storage grown before the write phase.
static void enqueue_moments(vacc_ctx* c);

static int do_relist(vacc_ctx* c, int add_width, bool want_loss, const SpatialIn* sp = nullptr) {
   ProfScope ps(c, add_width >= 0 ? 2 : 4);
   const int li = c->lcur, lo = 1 - c->lcur;
   const bool dynamic = c->sp.sparse_dynamic_algorithm == 1 || c->sp.sparse_dynamic_algorithm == 2;
   if (add_width >= 0 && dynamic) {
      // update_velocity_block_content_lists (block_adjust_cpu.cpp:46-49): the cell's threshold
      // first (density of the current f, or its block count), then the content flags with it
      if (c->sp.sparse_dynamic_algorithm == 1) enqueue_moments(c);
      SparseArgs spa{c->sp.sparse_dynamic_algorithm, c->sp.sparse_min_value, c->sp.sparse_dynamic_bulk1,
                     c->sp.sparse_dynamic_bulk2, c->sp.sparse_dynamic_min1, c->sp.sparse_dynamic_min2};
      launch_sparse_min(spa, c->moments, c->off[li], c->minval, c->ncells, c->stream);
      c->content_valid = false;
   }
   if (add_width >= 0 && !c->content_valid) {
      launch_content_cells(c->off[li], c->src[li], pool(c, c->fcur), c->minval, c->content, c->ncells, c->stream);
      c->content_valid = true;
   }
   const bool conserve = add_width >= 0 && c->sp.sparse_conserve_mass;
   if (conserve) launch_cell_sums(c->off[li], c->src[li], pool(c, c->fcur), c->csum_pre, c->ncells, c->stream);
   AdjustArgs aa{};
   aa.cell_off = c->off[li];
   aa.gid = c->gid[li];
   aa.src = c->src[li];
   aa.content = c->content;
   aa.pool = pool(c, c->fcur);
   aa.n_keep = c->n_tgt;
   aa.rho_loss = want_loss ? c->rho_loss : nullptr;
   aa.blk_stats = c->stats_valid ? c->blk_stats : nullptr;
   aa.ncells = c->ncells;
   for (int d = 0; d < 3; ++d) aa.len[d] = c->mi.len[d];
   aa.gs = make_gid_split(c->mi.len);
   aa.min_value = c->sp.sparse_min_value;
   aa.dv3 = c->mi.cell[0] * c->mi.cell[1] * c->mi.cell[2];
   aa.add_width = add_width;
   aa.flags = c->flags;
   aa.keep_bm = c->keep_bm;
   aa.keep_box = c->keep_box;
   if (sp) {
      aa.nbr_off = sp->nbr_off;
      aa.nbr = sp->nbr;
      aa.ext_off = sp->ext_off;
      aa.ext_gid = sp->ext;
   }
   launch_adjust(aa, false, c->stream);
   ScanArgs sa{};
   sa.a = c->n_tgt;
   sa.b = nullptr;
   sa.a_off = c->off[lo];
   sa.b_off = nullptr;
   sa.n = c->ncells;
   sa.cap_a = sp ? 0xFFFFFFFFu : (uint32_t)std::min<size_t>(c->cap, 0xFFFFFFFFull);
   sa.flags = c->flags;
   launch_scan(sa, c->stream);
   if (sp) {
      uint32_t total = 0;
      HIPCHK(hipMemcpyAsync(&total, c->off[lo] + c->ncells, sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
      HIPCHK(hipStreamSynchronize(c->stream));
      if (total > c->cap) {
         const int rc = ensure_blocks(c, (size_t)total + total / 4 + 4096);
         if (rc) return rc;
         aa.gid = c->gid[li];
         aa.src = c->src[li];
         aa.content = c->content;
         aa.pool = pool(c, c->fcur);
      }
   }
   aa.out_off = c->off[lo];
   aa.gid_out = c->gid[lo];
   aa.src_out = c->src[lo];
   launch_adjust(aa, true, c->stream);
   if (conserve) {   // block_adjust_cpu.cpp:96-113
      launch_cell_sums(c->off[lo], c->src[lo], pool(c, c->fcur), c->csum_post, c->ncells, c->stream);
      launch_cell_scale(c->off[lo], c->src[lo], pool(c, c->fcur), c->csum_pre, c->csum_post, c->ncells, c->stream);
      c->stats_valid = false;   // f was rescaled
   }
   HIPCHK(hipGetLastError());
   c->lcur = lo;
   c->content_valid = false;
   c->sorted = true;
   if (add_width >= 0) c->moments_valid = false;   // blocks may have been removed
   return 0;
}

// Compact the pool so that src is the identity (block i of the list at pool index i).
static int do_compact(vacc_ctx* c) {
   launch_gather(c->src[c->lcur], pool(c, c->fcur), pool(c, 1 - c->fcur), (uint32_t)c->nblocks, c->stream);
   launch_iota(c->src[c->lcur], (uint32_t)c->nblocks, c->stream);
   HIPCHK(hipGetLastError());
   c->fcur = 1 - c->fcur;
   c->stats_valid = false;   // the pool was gathered
   return 0;
}

// The one synchronisation of a work call: reads the capacity flag, the block total of the
// current list and the OR of the cell status words (kept in status[ncells]) with three small
// copies into pinned memory.  Returns the status OR or an error.
static int finish(vacc_ctx* c) {
   HIPCHK(hipMemcpyAsync(c->h_sum + 0, c->flags, sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
   HIPCHK(hipMemcpyAsync(c->h_sum + 1, c->off[c->lcur] + c->ncells, sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
   HIPCHK(hipMemcpyAsync(c->h_sum + 2, c->status + c->ncells, sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
   HIPCHK(hipStreamSynchronize(c->stream));
   if (c->h_sum[0] & VACC_STATUS_CAPACITY)
      return set_err(VACC_E_STATE, "block pool capacity exceeded; state is undefined");
   c->nblocks = c->h_sum[1];
   return (int)(c->h_sum[2] & 0x7FFFFFFFu);
}

// OR of the per-cell status words (synchronises the stream).
static int status_or(vacc_ctx* c) {
   if (!c->ncells) return 0;
   HIPCHK(hipMemcpyAsync(c->h_sum + 2, c->status + c->ncells, sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
   HIPCHK(hipStreamSynchronize(c->stream));
   return (int)(c->h_sum[2] & 0x7FFFFFFFu);
}

// Block capacity a load of `total` blocks in `ncells` cells allocates: one step's growth (new
// target blocks and adjust halo blocks) on top of the blocks themselves.
static size_t load_capacity(uint64_t total, uint32_t ncells) {
   return (size_t)total + (size_t)total / 2 + 64 * (size_t)ncells + 4096;
}

// Checks of a load's lists, made before any state changes, so that a rejected load keeps the
// context as it was: offsets start at 0 and ascend, every GID lies in the mesh and appears
// once in its cell (VelocityMesh::push_back rejects invalid and repeated GIDs,
// spatial_cells/velocity_mesh_cpu.h:428-444), and the blocks with their head-room fit the
// 32-bit list indices.  Repeats are found with one bit per mesh block, cleared after each cell.
static int check_load(vacc_ctx* c, uint32_t ncells, const uint64_t* offsets, const uint32_t* gids) {
   if (offsets[0] != 0) return set_err(VACC_E_ARG, "offsets[0] must be 0");
   for (uint32_t i = 1; i <= ncells; ++i)
      if (offsets[i] < offsets[i - 1]) return set_err(VACC_E_ARG, "offsets not ascending");
   const uint64_t total = offsets[ncells];
   if (load_capacity(total, ncells) > 0xF0000000ull)
      return set_err(VACC_E_NOMEM, "block count with its growth reserve exceeds 32-bit list indices");
   const uint64_t gmax = (uint64_t)c->mi.len[0] * c->mi.len[1] * c->mi.len[2];
   std::vector<uint64_t> seen((gmax + 63) / 64, 0ull);
   for (uint32_t i = 0; i < ncells; ++i) {
      const uint64_t b = offsets[i], e = offsets[i + 1];
      uint64_t j = b;
      int rc = 0;
      for (; j < e; ++j) {
         const uint32_t g = gids[j];
         if (g >= gmax) {
            rc = set_err(VACC_E_ARG, "GID outside the velocity mesh in cell " + std::to_string(i));
            break;
         }
         uint64_t& w = seen[g >> 6];
         const uint64_t bit = 1ull << (g & 63);
         if (w & bit) {
            rc = set_err(VACC_E_ARG, "GID " + std::to_string(g) + " repeated in cell " + std::to_string(i));
            break;
         }
         w |= bit;
      }
      for (uint64_t q = b; q < j; ++q) seen[gids[q] >> 6] = 0ull;
      if (rc) return rc;
   }
   return 0;
}

// A load that failed after its checks passed leaves an empty context (no cells): nothing
// that a later call could run on refers to the half-written lists.
static void abort_load(vacc_ctx* c) {
   (void)hipStreamSynchronize(c->stream);
   drop_graphs(c);
   c->ncells = 0;
   c->nblocks = 0;
   c->lcur = c->fcur = 0;
   c->sorted = false;
   c->cell_epoch++;
   if (c->off[0]) {
      (void)hipMemsetAsync(c->off[0], 0, sizeof(uint32_t), c->stream);
      (void)hipStreamSynchronize(c->stream);
   }
   c->content_valid = c->moments_valid = c->fields_valid = c->sub_valid = c->stats_valid = false;
}

// First half of a load of new cells: checks the lists, sizes the buffers and copies the
// offsets and GIDs; the caller then fills pool F[0] and calls end_load, or abort_load when
// that fails.
static int begin_load(vacc_ctx* c, uint32_t ncells, const uint64_t* offsets, const uint32_t* gids) {
   int rc = check_load(c, ncells, offsets, gids);
   if (rc) return rc;
   const uint64_t total = offsets[ncells];
   std::vector<uint32_t> off32(ncells + 1);
   for (uint32_t i = 0; i <= ncells; ++i) off32[i] = (uint32_t)offsets[i];
   // from here on a failure empties the context
   c->content_valid = c->moments_valid = c->fields_valid = c->sub_valid = c->stats_valid = false;
   rc = alloc_cells(c, ncells);
   if (!rc) {
      const size_t want = load_capacity(total, ncells);
      if (want > c->cap) {
         c->nblocks = 0;   // nothing of the old state is carried into the new buffers
         rc = ensure_blocks(c, want);
      }
   }
   if (rc) {
      abort_load(c);
      return rc;
   }
   c->ncells = ncells;
   c->nblocks = 0;
   c->lcur = 0;
   c->fcur = 0;
   c->sorted = false;
   c->cell_epoch++;
   hipError_t e = hipMemcpyAsync(c->off[0], off32.data(), (ncells + 1) * sizeof(uint32_t), hipMemcpyHostToDevice,
                                 c->stream);
   if (e == hipSuccess && total)
      e = hipMemcpyAsync(c->gid[0], gids, total * sizeof(uint32_t), hipMemcpyHostToDevice, c->stream);
   if (e == hipSuccess) e = hipStreamSynchronize(c->stream);   // the caller's arrays may go away after return
   if (e != hipSuccess) {
      abort_load(c);
      return set_err(VACC_E_HIP, std::string("load of the block lists: ") + hipGetErrorString(e));
   }
   return 0;
}

static int end_load(vacc_ctx* c, uint64_t total) {
   launch_iota(c->src[0], (uint32_t)total, c->stream);
   // a new cell starts from the species' minValue (spatial_cell_cpu.cpp:73)
   launch_fill_f64(c->minval, c->sp.sparse_min_value, c->ncells, c->stream);
   hipError_t e = hipGetLastError();
   if (e == hipSuccess) e = hipMemsetAsync(c->status, 0, (c->ncells + 1) * sizeof(uint32_t), c->stream);
   if (e == hipSuccess) e = hipMemsetAsync(c->flags, 0, sizeof(uint32_t), c->stream);
   if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
   if (e != hipSuccess) {
      abort_load(c);
      return set_err(VACC_E_HIP, std::string("end of load: ") + hipGetErrorString(e));
   }
   c->nblocks = total;
   c->content_valid = false;
   c->moments_valid = false;
   c->stats_valid = false;
   c->fields_valid = false;
   c->sub_valid = false;
   return 0;
}

// Streamed block transfer: chunk size in blocks (32 MiB of FP32 values per pinned buffer).
static constexpr uint64_t IO_CHUNK = 1u << 17;

static int ensure_io(vacc_ctx* c) {
   for (int i = 0; i < 2; ++i) {
      if (!c->io_pin[i]) HIPCHK(hipHostMalloc(&c->io_pin[i], IO_CHUNK * WID3 * sizeof(float), hipHostMallocDefault));
      if (!c->io_ev[i]) HIPCHK(hipEventCreateWithFlags(&c->io_ev[i], hipEventDisableTiming));
   }
   if (c->opt.storage != VACC_STORE_F32 && !c->io_dev)
      HIPCHK(hipMalloc(&c->io_dev, 2 * IO_CHUNK * WID3 * sizeof(float)));
   return 0;
}

static Pool pool_at(const vacc_ctx* c, int i, uint64_t first) {
   Pool p = pool(c, i);
   p.f = (char*)p.f + first * WID3 * elem_bytes(c);
   if (p.scale) p.scale += first;
   return p;
}

namespace vacc {

int set_error(int code, const std::string& msg) { return set_err(code, msg); }

void ctx_mesh(const vacc_ctx* c, vacc_mesh* mesh, double cell[3]) {
   *mesh = c->mesh;
   for (int d = 0; d < 3; ++d) cell[d] = c->mi.cell[d];
}

int export_lists(vacc_ctx* c, std::vector<uint32_t>& counts, std::vector<uint32_t>& gids) {
   int rc = refresh_count(c);
   if (rc) return rc;
   rc = do_relist(c, -1, false);    // ascending GID order, same blocks
   if (rc) return rc;
   rc = do_compact(c);
   if (rc) return rc;
   std::vector<uint32_t> off(c->ncells + 1);
   HIPCHK(hipMemcpyAsync(off.data(), c->off[c->lcur], (c->ncells + 1) * sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
   HIPCHK(hipStreamSynchronize(c->stream));
   counts.resize(c->ncells);
   for (uint32_t i = 0; i < c->ncells; ++i) counts[i] = off[i + 1] - off[i];
   const size_t n = off[c->ncells];
   gids.resize(n);
   if (n) HIPCHK(hipMemcpy(gids.data(), c->gid[c->lcur], n * sizeof(uint32_t), hipMemcpyDeviceToHost));
   return 0;
}

// Chunk k goes to pinned buffer k % 2: the copy (and decode) of chunk k + 1 is issued before
// the host consumes chunk k.
int export_values(vacc_ctx* c, const BlockSink& sink) {
   const uint64_t n = c->nblocks;
   if (!n) return 0;
   int rc = ensure_io(c);
   if (rc) return rc;
   const uint64_t nchunks = (n + IO_CHUNK - 1) / IO_CHUNK;
   auto issue = [&](uint64_t k) -> int {
      const uint64_t first = k * IO_CHUNK, m = std::min(IO_CHUNK, n - first);
      float* dst = c->io_pin[k & 1];
      if (c->opt.storage == VACC_STORE_F32) {
         HIPCHK(hipMemcpyAsync(dst, (const float*)c->f[c->fcur] + first * WID3, m * WID3 * sizeof(float),
                               hipMemcpyDeviceToHost, c->stream));
      } else {
         float* st = c->io_dev + (k & 1) * IO_CHUNK * WID3;
         launch_decode(pool_at(c, c->fcur, first), st, (uint32_t)m, c->stream);
         HIPCHK(hipGetLastError());
         HIPCHK(hipMemcpyAsync(dst, st, m * WID3 * sizeof(float), hipMemcpyDeviceToHost, c->stream));
      }
      HIPCHK(hipEventRecord(c->io_ev[k & 1], c->stream));
      return 0;
   };
   if ((rc = issue(0))) return rc;
   for (uint64_t k = 0; k < nchunks; ++k) {
      if (k + 1 < nchunks && (rc = issue(k + 1))) return rc;
      HIPCHK(hipEventSynchronize(c->io_ev[k & 1]));
      const uint64_t first = k * IO_CHUNK;
      if ((rc = sink(c->io_pin[k & 1], first, std::min(IO_CHUNK, n - first)))) {
         (void)hipStreamSynchronize(c->stream);
         return rc;
      }
   }
   return 0;
}

// Chunk k is filled into pinned buffer k % 2 once the copy of chunk k - 2 out of it is done.
int import_blocks(vacc_ctx* c, uint32_t ncells, const uint64_t* offsets, const uint32_t* gids,
                  const BlockSource& source) {
   if (!c || !offsets || (ncells && !gids)) return set_err(VACC_E_ARG, "null argument");
   int rc = begin_load(c, ncells, offsets, gids);
   if (rc) return rc;
   const uint64_t n = offsets[ncells];
   if (n) {
      if ((rc = ensure_io(c))) return rc;
      const uint64_t nchunks = (n + IO_CHUNK - 1) / IO_CHUNK;
      for (uint64_t k = 0; k < nchunks; ++k) {
         const uint64_t first = k * IO_CHUNK, m = std::min(IO_CHUNK, n - first);
         float* buf = c->io_pin[k & 1];
         if (k >= 2) HIPCHK(hipEventSynchronize(c->io_ev[k & 1]));
         if ((rc = source(buf, first, m))) {
            abort_load(c);
            return rc;
         }
         if (c->opt.storage == VACC_STORE_F32) {
            HIPCHK(hipMemcpyAsync((float*)c->f[0] + first * WID3, buf, m * WID3 * sizeof(float), hipMemcpyHostToDevice,
                                  c->stream));
         } else {
            float* st = c->io_dev + (k & 1) * IO_CHUNK * WID3;
            HIPCHK(hipMemcpyAsync(st, buf, m * WID3 * sizeof(float), hipMemcpyHostToDevice, c->stream));
            launch_encode(st, pool_at(c, 0, first), (uint32_t)m, c->stream);
            HIPCHK(hipGetLastError());
         }
         HIPCHK(hipEventRecord(c->io_ev[k & 1], c->stream));
      }
   }
   return end_load(c, n);
}

}  // namespace vacc

// Width-only scalars of one h4 face on non-uniform widths, h[0..3] = widths at k-2..k+1,
// with the float and double sub-expressions of cpu_face_estimates.hpp:179-191.
static TransFace trans_face(const float* h) {
   TransFace F{};
   const float h0 = h[0], h1 = h[1], h2 = h[2], h3 = h[3];
   F.D1 = 1.0 / (double)(((h0 + h1) + h2) + h3);
   F.SA = ((h0 + h1) * (h2 + h3)) / (h1 + h2);
   F.DA = 1.0 / (double)((h0 + h1) + h2) + 1.0 / (double)((h1 + h2) + h3);
   F.SB = (h2 * (h2 + h3)) / (((h0 + h1) + h2) * (h0 + h1));
   F.DB = (double)h0 + 2.0 * (double)h1;
   F.SC = (h1 * (h0 + h1)) / (((h1 + h2) + h3) * (h2 + h3));
   F.DC = 2.0 * (double)h2 + (double)h3;
   F.h1 = h1;
   F.h2 = h2;
   return F;
}

// Host plan of one dimension's translation (vacc_trans_set_pencils): the per-cell schedule of
// k_trans and the segments of the pencil walk.
struct TransPlan {
   std::vector<std::vector<TransEntry>> ent;                  // per cell, in accumulation order
   std::vector<std::vector<uint32_t>> tab;                    // per cell: distinct stencil cells
   std::vector<std::vector<std::pair<uint32_t, int>>> where;  // per entry: (pencil, target position)
   size_t total = 0;                                          // entries
   std::vector<uint8_t> wide;                                 // per cell: stencils span > TRANS_MAXN cells
   std::vector<uint8_t> walk;                                 // per cell: written by the walk
   std::vector<WalkSeg> segs;                                 // ordered by level
   std::vector<uint32_t> level_first;                         // [levels + 1] first segment of each level
   std::vector<WalkPos> wpos;
   std::vector<TransFace> wface;
};

// Every contribution each cell receives, in the reference's order: pencils as given (bins,
// then pencils of a bin), then source positions W .. L-W-1 (cpu_trans_map_amr.cpp:59-148).
static int plan_schedule(const vacc_pencils* P, uint32_t nc, TransPlan& T) {
   const int W = VACC_TRANS_STENCIL;
   auto& ent = T.ent;
   auto& tab = T.tab;
   auto& where = T.where;
   ent.assign(nc, {});
   tab.assign(nc, {});
   where.assign(nc, {});
   T.wide.assign(nc, 0);
   size_t total = 0;
   for (uint32_t p = 0; p < P->npencils; ++p) {
      const int L = (int)P->length[p];
      if (L < 2 * W + 1) return set_err(VACC_E_ARG, "pencil shorter than 2*VACC_TRANS_STENCIL+1 positions");
      const uint64_t s0 = P->start[p];
      const uint32_t* pc = P->cell + s0;
      const float* dz = P->dz + s0;
      const float* ratio = P->ratio + s0;
      for (int q = 0; q < L; ++q)
         if (pc[q] != VACC_INVALID_ID && pc[q] >= nc) return set_err(VACC_E_ARG, "pencil cell index outside the context");
      for (int i = W; i < L - W; ++i) {
         const TransFace fl = trans_face(dz + i - 2), fr = trans_face(dz + i - 1);
         // per source the reference adds self, then p1, then m1 (cpu_trans_map_amr.cpp:111-146);
         // the order matters where one cell is the target at two of these positions
         for (int o : {0, 1, -1}) {
            const int q = i + o;
            const uint32_t t = pc[q];
            if (t == VACC_INVALID_ID || ratio[q] == 0.0f) continue;
            TransEntry E{};
            E.fl = fl;
            E.fr = fr;
            E.dz_i = dz[i];
            E.dz_t = dz[q];
            E.ratio = ratio[q];
            E.kind = (o == 0) ? 1 : (o > 0 ? 0 : 2);
            for (int st = 0; st < 5; ++st) {
               const uint32_t sc = pc[i - 2 + st];
               if (sc == VACC_INVALID_ID) {
                  E.st[st] = 0xFF;
                  continue;
               }
               auto& tb = tab[t];
               size_t slot = std::find(tb.begin(), tb.end(), sc) - tb.begin();
               if (slot == tb.size()) {
                  if (tb.size() >= (size_t)TRANS_MAXN) {   // k_trans cannot take this cell
                     T.wide[t] = 1;
                     E.st[st] = 0xFF;
                     continue;
                  }
                  tb.push_back(sc);
               }
               E.st[st] = (uint8_t)slot;
            }
            ent[t].push_back(E);
            where[t].emplace_back(p, q);
            ++total;
         }
      }
   }
   if (total > 0xFFFFFFFFull) return set_err(VACC_E_ARG, "too many contributions");
   T.total = total;
   return 0;
}

// The cells the pencil walk writes, and its segments (windows, flags, faces) in launch levels.
// levels: cells that several pencils write may be walked (FP32 storage; the 16-bit forms would
// encode the partial sums between levels).
static void plan_walk(const vacc_pencils* P, uint32_t nc, bool levels, TransPlan& T) {
   const int W = VACC_TRANS_STENCIL;
   const auto& where = T.where;
   // walk candidates: every entry of the cell from one pencil, at one target position or at two
   // that are at least three apart (no source reaches both), so that walking the sources in
   // order adds the contributions in the reference's order
   struct Cand {
      uint32_t p = ~0u;
      int q1 = -1, q2 = -1;
   };
   std::vector<Cand> cand(nc);
   for (uint32_t t = 0; t < nc; ++t) {
      const auto& wh = where[t];
      if (wh.empty()) continue;
      Cand C;
      C.p = wh[0].first;
      bool ok = true;
      for (const auto& x : wh) {
         if (x.first != C.p) ok = false;
         else if (C.q1 < 0 || x.second == C.q1) C.q1 = x.second;
         else if (C.q2 < 0 || x.second == C.q2) C.q2 = x.second;
         else ok = false;
      }
      if (C.q2 >= 0) {
         if (C.q2 < C.q1) std::swap(C.q1, C.q2);
         ok = ok && C.q2 - C.q1 >= 3;
      }
      if (ok) cand[t] = C;
   }
   // cells written by several pencils, at one run position of each (a coarse cell that the
   // pencils of finer cells pass): each pencil's part is walked in a launch level above those
   // of the earlier pencils that write the cell, starting from the sum they stored
   std::vector<std::vector<std::pair<uint32_t, int>>> multi(nc);   // (pencil, position) per part
   if (levels) {
      for (uint32_t t = 0; t < nc; ++t) {
         const auto& wh = where[t];
         if (wh.empty() || cand[t].p != ~0u) continue;
         std::vector<std::pair<uint32_t, int>> parts;
         bool ok = true;
         for (const auto& x : wh) {
            const int L = (int)P->length[x.first];
            if (!parts.empty() && parts.back().first == x.first) ok = ok && parts.back().second == x.second;
            else if (!parts.empty() && x.first < parts.back().first) ok = false;
            else parts.push_back(x);
            ok = ok && x.second >= W && x.second <= L - W - 1;
         }
         if (!ok) continue;
         for (int k = 1; k < (int)parts.size(); ++k) ok = ok && parts[k].first != parts[k - 1].first;
         if (!ok) continue;
         multi[t] = parts;
      }
   }
   const int span = WALK_MAXN - 6;
   // a cell at two positions of one pencil that no segment holds together (the wrap cells of
   // a periodic pencil longer than a segment): its first part is stored by the segment at
   // the first position, its second starts from that block in a later level
   if (levels) {
      for (uint32_t t = 0; t < nc; ++t) {
         const Cand& C = cand[t];
         if (C.p == ~0u || C.q2 < 0) continue;
         const int L = (int)P->length[C.p];
         bool fits = false;
         for (int qa = W; qa < L - W && !fits; qa += span) {
            const int qb = std::min(qa + span, L - W);
            const int smin = std::max(W, qa - 1), smax = std::min(L - W - 1, qb);
            auto inside = [&](int q) {
               return q >= smin - 1 && q <= smax + 1 && std::max(q - 1, W) >= smin && std::min(q + 1, L - W - 1) <= smax;
            };
            fits = inside(C.q1) && inside(C.q2) && C.q2 >= smin + 2;
         }
         if (fits) continue;
         multi[t] = {{C.p, C.q1}, {C.p, C.q2}};
         cand[t] = Cand{};
      }
   }
   std::vector<uint32_t> last(nc, 0), seglevel;
   uint32_t nlevels = 1;
   // segments: runs [qa, qb) of at most WALK_MAXN - 6 positions with the stencils of their
   // sources smin..smax; a segment walks a candidate whose positions lie in smin-1 .. smax+1
   // and whose contributions all come from sources smin..smax (a one-position cell: its
   // position in [qa, qb)); two-position cells carry the first sum in one of two registers
   auto& walk = T.walk;
   auto& segs = T.segs;
   auto& wpos = T.wpos;
   auto& wface = T.wface;
   walk.assign(nc, 0);
   for (uint32_t t = 0; t < nc; ++t)
      if (!multi[t].empty()) walk[t] = 1;
   segs.clear();
   wpos.clear();
   wface.clear();
   std::vector<uint32_t> flags;
   for (uint32_t p = 0; p < P->npencils; ++p) {
      const int L = (int)P->length[p];
      const uint64_t s0 = P->start[p];
      const uint32_t* pc = P->cell + s0;
      const float* dz = P->dz + s0;
      const float* ratio = P->ratio + s0;
      for (int qa = W; qa < L - W; qa += span) {
         const int qb = std::min(qa + span, L - W);
         const int smin = std::max(W, qa - 1), smax = std::min(L - W - 1, qb);
         const int wa = smin - 2, n = smax - smin + 5;
         // sources reaching position q lie in smin..smax
         auto inside = [&](int q) {
            return q >= smin - 1 && q <= smax + 1 && std::max(q - 1, W) >= smin && std::min(q + 1, L - W - 1) <= smax;
         };
         flags.assign((size_t)n, 0u);
         int slot_end[2] = {-1, -1};
         int nw = 0;
         std::vector<std::pair<uint32_t, int>> parts;   // (cell, part) stored parts of this segment
         for (int q = smin - 1; q <= smax + 1; ++q) {
            const uint32_t t = pc[q];
            if (t == VACC_INVALID_ID || ratio[q] == 0.0f) continue;
            if (!multi[t].empty()) {   // a stored part: its run position, or a pencil end's
               const bool mine = (q >= qa && q < qb) || (q == W - 1 && qa == W) || (q == L - W && qb == L - W);
               if (!mine) continue;
               for (int k = 0; k < (int)multi[t].size(); ++k)
                  if (multi[t][k].first == p && multi[t][k].second == q) {
                     flags[q - wa] = WALK_W | (k > 0 ? WALK_INIT_STORED << WALK_INIT_SHIFT : 0u);
                     parts.emplace_back(t, k);
                     ++nw;
                  }
               continue;
            }
            if (walk[t]) continue;
            const Cand& C = cand[t];
            if (C.p != p || C.q1 != q) continue;
            if (C.q2 < 0) {
               if (q < qa || q >= qb) continue;
               flags[q - wa] = WALK_W;
               walk[t] = 1;
               ++nw;
            } else if (inside(C.q1) && inside(C.q2) && C.q2 >= smin + 2) {
               // the second position's sum starts at source q2 - 1 >= smin + 1, after the
               // first has ended (source q1 + 1); a slot is free once its last pair has ended
               const int k = (slot_end[0] < q) ? 0 : (slot_end[1] < q) ? 1 : -1;
               if (k < 0) continue;
               flags[C.q1 - wa] = WALK_W | ((uint32_t)(k + 1) << WALK_SAVE_SHIFT);
               flags[C.q2 - wa] = WALK_W | ((uint32_t)(k + 1) << WALK_INIT_SHIFT);
               slot_end[k] = C.q2;
               walk[t] = 1;
               nw += 2;
            }
         }
         if (!nw) continue;
         WalkSeg S{};
         S.pos0 = (uint32_t)wpos.size();
         S.n = (uint32_t)n;
         S.js = 2;
         S.je = (uint32_t)(n - 3);
         segs.push_back(S);
         // launch level: above the level of every earlier part of its stored parts' cells
         uint32_t lv = 0;
         for (const auto& x : parts)
            if (x.second > 0) lv = std::max(lv, last[x.first] + 1);
         for (const auto& x : parts) last[x.first] = lv;
         seglevel.push_back(lv);
         nlevels = std::max(nlevels, lv + 1);
         for (int k = 0; k < n; ++k) {
            const int q = wa + k;
            WalkPos X{};
            X.cell = pc[q];
            X.dz = dz[q];
            X.ratio = ratio[q];
            X.flags = flags[k];
            wpos.push_back(X);
            wface.push_back((k >= 2 && k <= n - 2) ? trans_face(dz + q - 2) : TransFace{});
         }
      }
   }
   // segments in level order (stable), and the first segment of each level
   std::vector<uint32_t> order(segs.size());
   for (uint32_t k = 0; k < (uint32_t)order.size(); ++k) order[k] = k;
   std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return seglevel[a] < seglevel[b]; });
   std::vector<WalkSeg> sorted(segs.size());
   for (size_t k = 0; k < order.size(); ++k) sorted[k] = segs[order[k]];
   segs.swap(sorted);
   T.level_first.assign(nlevels + 1, 0);
   for (uint32_t k = 0, lv = 0; lv <= nlevels; ++lv) {
      while (k < order.size() && seglevel[order[k]] < lv) ++k;
      T.level_first[lv] = k;
   }
   T.level_first[nlevels] = (uint32_t)segs.size();
}

// Device array p of at least n elements (contents not kept when it grows).
template <class T>
static int dev_grow(T*& p, size_t& cap, size_t n) {
   if (p && cap >= n) return 0;
   if (p) (void)hipFree(p);
   p = nullptr;
   cap = 0;
   HIPCHK(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)));
   cap = std::max<size_t>(n, 1);
   return 0;
}

// Waves of a pencil walk launch (all segments together), about.
constexpr uint32_t WALK_WAVES = 32768;

// Device temporaries of one call, freed when it returns (on every path).
struct DevTemps {
   std::vector<void*> p;
   template <class T>
   hipError_t alloc(T** out, size_t bytes) {
      void* q = nullptr;
      const hipError_t e = hipMalloc(&q, bytes);
      if (e == hipSuccess) p.push_back(q);
      *out = static_cast<T*>(q);
      return e;
   }
   ~DevTemps() {
      for (void* q : p) (void)hipFree(q);
   }
};

// The load of vacc_init_peaks once its arguments are checked; a failure leaves the caller
// to empty the context.
static int init_peaks_load(vacc_ctx* c, uint32_t ncells, uint32_t max_peaks, const vacc_peak* peaks,
                           const int32_t* npeaks, const int32_t* bbox, const std::vector<uint32_t>& bm_off) {
   int rc = alloc_cells(c, ncells);
   if (rc) return rc;
   HIPCHK(hipStreamSynchronize(c->stream));
   c->ncells = ncells;
   c->lcur = 0;
   c->fcur = 0;
   c->nblocks = 0;
   c->sorted = false;
   c->cell_epoch++;
   c->content_valid = false;
   c->moments_valid = false;
   c->stats_valid = false;
   c->fields_valid = false;
   c->sub_valid = false;
   if (!ncells) return 0;
   DevTemps tmp;
   vacc_peak* d_pk = nullptr;
   int32_t *d_np = nullptr, *d_bb = nullptr;
   uint32_t *d_bmo = nullptr, *d_bm = nullptr;
   float* stage = nullptr;
   HIPCHK(tmp.alloc(&d_pk, (size_t)ncells * max_peaks * sizeof(vacc_peak)));
   HIPCHK(tmp.alloc(&d_np, ncells * sizeof(int32_t)));
   HIPCHK(tmp.alloc(&d_bb, (size_t)ncells * 6 * sizeof(int32_t)));
   HIPCHK(tmp.alloc(&d_bmo, (ncells + 1) * sizeof(uint32_t)));
   HIPCHK(tmp.alloc(&d_bm, std::max<size_t>(bm_off[ncells], 1) * sizeof(uint32_t)));
   HIPCHK(hipMemcpy(d_pk, peaks, (size_t)ncells * max_peaks * sizeof(vacc_peak), hipMemcpyHostToDevice));
   HIPCHK(hipMemcpy(d_np, npeaks, ncells * sizeof(int32_t), hipMemcpyHostToDevice));
   HIPCHK(hipMemcpy(d_bb, bbox, (size_t)ncells * 6 * sizeof(int32_t), hipMemcpyHostToDevice));
   HIPCHK(hipMemcpy(d_bmo, bm_off.data(), (ncells + 1) * sizeof(uint32_t), hipMemcpyHostToDevice));
   InitArgs ia{};
   ia.peaks = d_pk;
   ia.npeaks = d_np;
   ia.bbox = d_bb;
   ia.bm_off = d_bmo;
   ia.content_bm = d_bm;
   ia.n_keep = c->n_tgt;
   ia.ncells = ncells;
   ia.max_peaks = max_peaks;
   for (int d = 0; d < 3; ++d) {
      ia.len[d] = c->mi.len[d];
      ia.vmin[d] = c->mi.vmin[d];
      ia.cell[d] = c->mi.cell[d];
   }
   launch_fill_f64(c->minval, c->sp.sparse_min_value, ncells, c->stream);
   ia.mass = c->sp.mass;
   ia.min_value = c->sp.sparse_min_value;
   ia.add_width = c->sp.sparse_add_width_v;
   launch_init(ia, false, c->stream);
   ScanArgs sa{};
   sa.a = c->n_tgt;
   sa.a_off = c->off[0];
   sa.n = ncells;
   launch_scan(sa, c->stream);
   HIPCHK(hipGetLastError());
   rc = refresh_count(c);
   if (rc) return rc;
   const size_t total = c->nblocks;
   c->nblocks = 0;   // nothing to carry over
   rc = ensure_blocks(c, total + total / 2 + 64 * (size_t)ncells + 4096);
   if (rc) return rc;
   ia.out_off = c->off[0];
   ia.gid_out = c->gid[0];
   ia.src_out = c->src[0];
   if (c->opt.storage == VACC_STORE_F32) {
      ia.f_out = static_cast<float*>(c->f[0]);
   } else {
      HIPCHK(tmp.alloc(&stage, std::max<size_t>(total, 1) * WID3 * sizeof(float)));
      ia.f_out = stage;
   }
   launch_init(ia, true, c->stream);
   if (stage) launch_encode(stage, pool(c, 0), (uint32_t)total, c->stream);
   HIPCHK(hipGetLastError());
   HIPCHK(hipMemsetAsync(c->status, 0, (ncells + 1) * sizeof(uint32_t), c->stream));
   HIPCHK(hipMemsetAsync(c->flags, 0, sizeof(uint32_t), c->stream));
   HIPCHK(hipStreamSynchronize(c->stream));
   c->nblocks = total;
   return 0;
}

extern "C" {

const char* vacc_last_error(void) { return g_err.c_str(); }

// Query staging of vacc_find_blocks_noalloc: cells, GIDs and answers of up to
// VACC_FIND_NOALLOC_MAX queries (768 KiB per device, held until the process ends).  Host code
// never passes the array's name to a kernel or a copy: it uses the device address that
// hipGetSymbolAddress returns, looked up once per device when the first context on that device
// is created.
static __device__ __align__(256) uint32_t scratch[3 * (size_t)VACC_FIND_NOALLOC_MAX];

namespace {
constexpr int SCRATCH_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
uint32_t* g_scratch_ptr[SCRATCH_MAX_DEVICES] = {};
hipError_t g_scratch_err[SCRATCH_MAX_DEVICES] = {};
bool g_scratch_done[SCRATCH_MAX_DEVICES] = {};
}  // namespace

// Address of the current device's `scratch` (dev = that device), looked up on first use.
static void scratch_setup(int dev) {
   if (dev < 0 || dev >= SCRATCH_MAX_DEVICES) return;
   std::lock_guard<std::mutex> lk(g_scratch_mu);
   if (g_scratch_done[dev]) return;
   void* p = nullptr;
   const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
   g_scratch_err[dev] = e;
   g_scratch_ptr[dev] = (e == hipSuccess) ? static_cast<uint32_t*>(p) : nullptr;
   g_scratch_done[dev] = true;
}

int vacc_create(const vacc_mesh* mesh, const vacc_species* species, const vacc_options* opts, vacc_ctx** out) {
   if (!mesh || !species || !opts || !out) return set_err(VACC_E_ARG, "null argument");
   for (int d = 0; d < 3; ++d) {
      if (mesh->grid_length[d] == 0 || mesh->grid_length[d] > 65535)
         return set_err(VACC_E_ARG, "grid_length must be in [1, 65535]");
      if (!(mesh->max_limits[d] > mesh->min_limits[d])) return set_err(VACC_E_ARG, "empty mesh extent");
   }
   if (opts->method < VACC_PLM || opts->method > VACC_PQM) return set_err(VACC_E_ARG, "unknown method");
   if (opts->storage < VACC_STORE_F32 || opts->storage > VACC_STORE_F16) return set_err(VACC_E_ARG, "unknown storage");
   if (species->sparse_dynamic_algorithm < 0 || species->sparse_dynamic_algorithm > 2)
      return set_err(VACC_E_ARG, "sparse_dynamic_algorithm must be 0, 1 or 2");
   if (species->sparse_dynamic_algorithm != 0 && !(species->sparse_dynamic_bulk2 != species->sparse_dynamic_bulk1))
      return set_err(VACC_E_ARG, "sparse_dynamic_bulk1 and _bulk2 must differ");
   if (species->sparse_add_width_v < 0 || species->sparse_add_width_v > 8)
      return set_err(VACC_E_ARG, "sparse_add_width_v must be in [0, 8]");
   int ndev = 0;
   if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return set_err(VACC_E_HIP, "no HIP device");
   int dev = 0;
   if (hipGetDevice(&dev) != hipSuccess) return set_err(VACC_E_HIP, "hipGetDevice failed");
   scratch_setup(dev);
   vacc_ctx* c = new vacc_ctx();
   c->device = dev;
   c->mesh = *mesh;
   c->sp = *species;
   c->opt = *opts;
   for (int d = 0; d < 3; ++d) {
      c->mi.len[d] = mesh->grid_length[d];
      c->mi.vmin[d] = mesh->min_limits[d];
      c->mi.cell[d] = cell_size(mesh->min_limits[d], mesh->max_limits[d], mesh->grid_length[d]);
   }
   if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
       hipMalloc(&c->flags, sizeof(uint32_t)) != hipSuccess || hipMemset(c->flags, 0, sizeof(uint32_t)) != hipSuccess ||
       hipHostMalloc(&c->h_sum, 4 * sizeof(uint32_t), hipHostMallocDefault) != hipSuccess ||
       hipEventCreateWithFlags(&c->mom_ev, hipEventDisableTiming) != hipSuccess) {
      delete c;
      return set_err(VACC_E_HIP, "stream/flag allocation failed");
   }
   *out = c;
   return 0;
}

int vacc_destroy(vacc_ctx* c) {
   if (!c) return 0;
   if (c->stream) (void)hipStreamSynchronize(c->stream);
   drop_graphs(c);
   free_blocks(c);
   free_cells(c);
   for (auto& r : c->recs) {
      (void)hipEventDestroy(r.a);
      (void)hipEventDestroy(r.b);
   }
   for (auto e : c->ev_free) (void)hipEventDestroy(e);
   if (c->prof_counts) (void)hipFree(c->prof_counts);
   if (c->flags) (void)hipFree(c->flags);
   if (c->h_sum) (void)hipHostFree(c->h_sum);
   if (c->mom_ev) (void)hipEventDestroy(c->mom_ev);
   for (int i = 0; i < 2; ++i) {
      if (c->io_pin[i]) (void)hipHostFree(c->io_pin[i]);
      if (c->io_ev[i]) (void)hipEventDestroy(c->io_ev[i]);
   }
   if (c->io_dev) (void)hipFree(c->io_dev);
   for (auto& t : c->trans) {
      void* tp[] = {t.cells, t.entries, t.rest, t.bounds, t.seg, t.pos, t.face};
      for (void* p : tp)
         if (p) (void)hipFree(p);
   }
   void* sp[] = {c->d_nbr_off, c->d_nbr, c->d_ext_off, c->d_ext};
   for (void* p : sp)
      if (p) (void)hipFree(p);
   if (c->stream) (void)hipStreamDestroy(c->stream);
   delete c;
   return 0;
}

int vacc_upload_cells(vacc_ctx* c, uint32_t ncells, const uint64_t* offsets, const uint32_t* gids, const float* f) {
   if (!c || !offsets || (ncells && (!gids || !f))) return set_err(VACC_E_ARG, "null argument");
   int rc = begin_load(c, ncells, offsets, gids);
   if (rc) return rc;
   const uint64_t total = offsets[ncells];
   if (total) {
      hipError_t e = hipSuccess;
      if (c->opt.storage == VACC_STORE_F32) {
         e = hipMemcpyAsync(c->f[0], f, total * WID3 * sizeof(float), hipMemcpyHostToDevice, c->stream);
      } else {
         // 16-bit forms: FP32 chunks go through the two device stages of the streamed
         // transfer and are encoded into the pool (stream order keeps the stages safe)
         if ((rc = ensure_io(c))) {
            abort_load(c);
            return rc;
         }
         for (uint64_t first = 0; first < total && e == hipSuccess; first += IO_CHUNK) {
            const uint64_t m = std::min(IO_CHUNK, total - first);
            float* st = c->io_dev + ((first / IO_CHUNK) & 1) * IO_CHUNK * WID3;
            e = hipMemcpyAsync(st, f + first * WID3, m * WID3 * sizeof(float), hipMemcpyHostToDevice, c->stream);
            if (e == hipSuccess) {
               launch_encode(st, pool_at(c, 0, first), (uint32_t)m, c->stream);
               e = hipGetLastError();
            }
         }
      }
      if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
      if (e != hipSuccess) {
         abort_load(c);
         return set_err(VACC_E_HIP, std::string("upload of f: ") + hipGetErrorString(e));
      }
   }
   return end_load(c, total);
}

int vacc_total_blocks(vacc_ctx* c, uint64_t* total) {
   if (!c || !total) return set_err(VACC_E_ARG, "null argument");
   *total = c->nblocks;
   return 0;
}

int vacc_get_counts(vacc_ctx* c, uint32_t* ncells, uint32_t* counts) {
   if (!c) return set_err(VACC_E_ARG, "null context");
   int rc = refresh_count(c);
   if (rc) return rc;
   if (ncells) *ncells = c->ncells;
   if (counts && c->ncells) {
      std::vector<uint32_t> off(c->ncells + 1);
      HIPCHK(hipMemcpy(off.data(), c->off[c->lcur], (c->ncells + 1) * sizeof(uint32_t), hipMemcpyDeviceToHost));
      for (uint32_t i = 0; i < c->ncells; ++i) counts[i] = off[i + 1] - off[i];
   }
   return 0;
}

int vacc_download_cells(vacc_ctx* c, uint64_t* offsets, uint32_t* gids, float* f) {
   if (!c || !offsets) return set_err(VACC_E_ARG, "null argument");
   int rc = refresh_count(c);
   if (rc) return rc;
   rc = do_relist(c, -1, false);    // ascending GID order, same blocks
   if (rc) return rc;
   rc = do_compact(c);
   if (rc) return rc;
   std::vector<uint32_t> off(c->ncells + 1);
   HIPCHK(hipMemcpyAsync(off.data(), c->off[c->lcur], (c->ncells + 1) * sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
   HIPCHK(hipStreamSynchronize(c->stream));
   for (uint32_t i = 0; i <= c->ncells; ++i) offsets[i] = off[i];
   const size_t n = off[c->ncells];
   if (n && gids) HIPCHK(hipMemcpy(gids, c->gid[c->lcur], n * sizeof(uint32_t), hipMemcpyDeviceToHost));
   if (n && f) {
      if (c->opt.storage == VACC_STORE_F32) {
         HIPCHK(hipMemcpy(f, c->f[c->fcur], n * WID3 * sizeof(float), hipMemcpyDeviceToHost));
      } else {
         // decode chunk by chunk into the two device stages of the streamed transfer
         if ((rc = ensure_io(c))) return rc;
         for (uint64_t first = 0; first < n; first += IO_CHUNK) {
            const uint64_t m = std::min<uint64_t>(IO_CHUNK, n - first);
            float* st = c->io_dev + ((first / IO_CHUNK) & 1) * IO_CHUNK * WID3;
            launch_decode(pool_at(c, c->fcur, first), st, (uint32_t)m, c->stream);
            HIPCHK(hipGetLastError());
            HIPCHK(hipMemcpyAsync(f + first * WID3, st, m * WID3 * sizeof(float), hipMemcpyDeviceToHost, c->stream));
         }
         HIPCHK(hipStreamSynchronize(c->stream));
      }
   }
   return 0;
}

int vacc_set_intersections(vacc_ctx* c, const double* isect) {
   if (!c || (!isect && c->ncells)) return set_err(VACC_E_ARG, "null argument");
   if (c->ncells)
      HIPCHK(hipMemcpyAsync(c->isect, isect, (size_t)c->ncells * 12 * sizeof(double), hipMemcpyHostToDevice, c->stream));
   HIPCHK(hipStreamSynchronize(c->stream));
   return 0;
}

int vacc_get_intersections(vacc_ctx* c, double* isect) {
   if (!c || (!isect && c->ncells)) return set_err(VACC_E_ARG, "null argument");
   if (c->ncells)
      HIPCHK(hipMemcpyAsync(isect, c->isect, (size_t)c->ncells * 12 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
   HIPCHK(hipStreamSynchronize(c->stream));
   return 0;
}

static void to_tin(const vacc_cell_fields& in, TransformIn& t) {
   for (int i = 0; i < 3; ++i) {
      t.B[i] = in.B[i];
      t.V[i] = in.V[i];
      t.EgradPe[i] = in.EgradPe[i];
   }
   for (int i = 0; i < 6; ++i) t.dB[i] = in.dB[i];
   t.rhoq = in.rhoq;
   t.dt = in.dt;
}

// Enqueue the moments of the current list unless they are resident.
static void enqueue_moments(vacc_ctx* c) {
   if (c->moments_valid) return;
   ProfScope ps(c, 3);
   MomentArgs ma{};
   ma.cell_off = c->off[c->lcur];
   ma.gid = c->gid[c->lcur];
   ma.src = c->src[c->lcur];
   ma.pool = pool(c, c->fcur);
   ma.out = c->moments;
   ma.ncells = c->ncells;
   for (int d = 0; d < 3; ++d) {
      ma.len[d] = c->mi.len[d];
      ma.vmin[d] = c->mi.vmin[d];
      ma.cell[d] = c->mi.cell[d];
   }
   ma.gs = make_gid_split(c->mi.len);
   ma.blk_stats = c->stats_valid ? c->blk_stats : nullptr;
   launch_moments(ma, c->stream);
   c->moments_valid = true;
}

// Enqueue the transform + intersections; mode 0: pivot from the fields, 1: from this
// population's moments (computed here when not resident), 2: from c->pivot.
static void enqueue_transform(vacc_ctx* c, uint32_t map_order, int mode) {
   if (mode == 1) enqueue_moments(c);
   ProfScope ps(c, 3);
   const double* mo = (mode == 1) ? c->moments : (mode == 2) ? c->pivot : nullptr;
   launch_transform(c->mi, c->fields, c->subcycling ? c->sub_dt : nullptr, mo, mode, c->sp.mass, c->sp.charge,
                    c->opt.hall_minimum_rhoq,
                    c->opt.ohm_gradpe_term, (int)map_order, c->isect, c->ncells, c->stream);
}

static int upload_fields(vacc_ctx* c, const vacc_cell_fields* fields) {
   std::vector<TransformIn> tin(c->ncells);
   for (uint32_t i = 0; i < c->ncells; ++i) to_tin(fields[i], tin[i]);
   HIPCHK(hipMemcpyAsync(c->fields, tin.data(), tin.size() * sizeof(TransformIn), hipMemcpyHostToDevice, c->stream));
   HIPCHK(hipStreamSynchronize(c->stream));   // tin is released on return
   c->fields_valid = true;
   c->sub_valid = false;
   return 0;
}

int vacc_set_fields(vacc_ctx* c, const vacc_cell_fields* fields) {
   if (!c || (!fields && c->ncells)) return set_err(VACC_E_ARG, "null argument");
   if (!c->ncells) return 0;
   return upload_fields(c, fields);
}

int vacc_compute_cell_intersections(vacc_ctx* c, const vacc_cell_fields* fields, uint32_t map_order, int32_t use_moments) {
   if (!c || map_order > 2) return set_err(VACC_E_ARG, "bad argument");
   if (!c->ncells) return 0;
   if (fields) {
      const int rc = upload_fields(c, fields);
      if (rc) return rc;
   } else if (!c->fields_valid) {
      return set_err(VACC_E_STATE, "no cell fields: pass fields or call vacc_set_fields first");
   }
   enqueue_transform(c, map_order, use_moments ? 1 : 0);
   HIPCHK(hipGetLastError());
   HIPCHK(hipStreamSynchronize(c->stream));
   return 0;
}

int vacc_cell_intersections_host(const vacc_mesh* mesh, const vacc_species* sp, const vacc_options* opt,
                                 const vacc_cell_fields* fields, uint32_t map_order, double out[12]) {
   if (!mesh || !sp || !opt || !fields || !out || map_order > 2) return set_err(VACC_E_ARG, "bad argument");
   MeshInfo mi;
   for (int d = 0; d < 3; ++d) {
      mi.len[d] = mesh->grid_length[d];
      mi.vmin[d] = mesh->min_limits[d];
      mi.cell[d] = cell_size(mesh->min_limits[d], mesh->max_limits[d], mesh->grid_length[d]);
   }
   TransformIn t;
   to_tin(*fields, t);
   const Affine fwd = acceleration_transform(t, sp->mass, sp->charge, opt->hall_minimum_rhoq, opt->ohm_gradpe_term);
   cell_intersections(mi, fwd, (int)map_order, out);
   return 0;
}

double vacc_max_v_dt(const vacc_species* sp, const vacc_options* opt, const double B[3]) {
   const double B_mag = std::sqrt(B[0] * B[0] + B[1] * B[1] + B[2] * B[2]) + 1e-30;
   const double gyro_period = 2 * PI * sp->mass / (sp->charge * B_mag);
   return std::fabs(gyro_period) * (opt->max_rotation_deg / 360.0);
}

int vacc_map_1d(vacc_ctx* c, uint32_t dim) {
   if (!c || dim > 2) return set_err(VACC_E_ARG, "bad argument");
   if (!c->ncells) return 0;
   int rc = reserve_for_step(c);
   if (rc) return rc;
   rc = do_pass(c, (int)dim);
   if (rc) return rc;
   return finish(c);
}

int vacc_accelerate_cells(vacc_ctx* c, uint32_t map_order) {
   if (!c || map_order > 2) return set_err(VACC_E_ARG, "bad argument");
   if (!c->ncells) return 0;
   int rc = reserve_for_step(c);
   if (rc) return rc;
   for (int p = 0; p < 3; ++p) {
      rc = do_pass(c, pass_dim((int)map_order, p), p == 2);
      if (rc) return rc;
   }
   return finish(c);
}

int vacc_adjust_blocks(vacc_ctx* c, double* rho_loss) {
   if (!c) return set_err(VACC_E_ARG, "null context");
   if (!c->ncells) return 0;
   int rc = reserve_for_step(c);
   if (rc) return rc;
   rc = do_relist(c, c->sp.sparse_add_width_v, true);
   if (rc) return rc;
   rc = finish(c);
   if (rc < 0) return rc;
   if (rho_loss) HIPCHK(hipMemcpy(rho_loss, c->rho_loss, c->ncells * sizeof(double), hipMemcpyDeviceToHost));
   return rc;
}

// grow a device buffer to hold n elements of size es
static int grow(void** p, size_t* cap, size_t n, size_t es) {
   if (n <= *cap && *p) return 0;
   if (*p) (void)hipFree(*p);
   *p = nullptr;
   const size_t nc = std::max<size_t>(n + n / 2, 64);
   HIPCHK(hipMalloc(p, nc * es));
   *cap = nc;
   return 0;
}

int vacc_adjust_blocks_spatial(vacc_ctx* c, const uint32_t* nbr_off, const uint32_t* nbr, const uint64_t* extra_off,
                               const uint32_t* extra_gids, double* rho_loss) {
   if (!c || (nbr_off && !nbr && nbr_off[c->ncells]) || (extra_off && !extra_gids && extra_off[c->ncells]))
      return set_err(VACC_E_ARG, "null argument");
   if (!c->ncells) return 0;
   SpatialIn sp;
   const size_t m = (size_t)c->ncells + 1;
   if (nbr_off) {
      for (uint32_t i = 0; i < c->ncells; ++i)
         if (nbr_off[i + 1] < nbr_off[i]) return set_err(VACC_E_ARG, "neighbour offsets not ascending");
      const size_t nn = nbr_off[c->ncells];
      for (size_t k = 0; k < nn; ++k)
         if (nbr[k] >= c->ncells) return set_err(VACC_E_ARG, "neighbour index outside the context");
      int rc = grow((void**)&c->d_nbr_off, &c->off_cap, m, sizeof(uint32_t));
      if (rc) return rc;
      rc = grow((void**)&c->d_nbr, &c->nbr_cap, std::max<size_t>(nn, 1), sizeof(uint32_t));
      if (rc) return rc;
      HIPCHK(hipMemcpyAsync(c->d_nbr_off, nbr_off, m * sizeof(uint32_t), hipMemcpyHostToDevice, c->stream));
      if (nn) HIPCHK(hipMemcpyAsync(c->d_nbr, nbr, nn * sizeof(uint32_t), hipMemcpyHostToDevice, c->stream));
      sp.nbr_off = c->d_nbr_off;
      sp.nbr = c->d_nbr;
   }
   if (extra_off) {
      for (uint32_t i = 0; i < c->ncells; ++i)
         if (extra_off[i + 1] < extra_off[i]) return set_err(VACC_E_ARG, "extra GID offsets not ascending");
      const size_t ne = extra_off[c->ncells];
      int rc = grow((void**)&c->d_ext_off, &c->ext_off_cap, m, sizeof(uint64_t));
      if (rc) return rc;
      rc = grow((void**)&c->d_ext, &c->ext_cap, std::max<size_t>(ne, 1), sizeof(uint32_t));
      if (rc) return rc;
      HIPCHK(hipMemcpyAsync(c->d_ext_off, extra_off, m * sizeof(uint64_t), hipMemcpyHostToDevice, c->stream));
      if (ne) HIPCHK(hipMemcpyAsync(c->d_ext, extra_gids, ne * sizeof(uint32_t), hipMemcpyHostToDevice, c->stream));
      sp.ext_off = c->d_ext_off;
      sp.ext = c->d_ext;
   }
   int rc = reserve_for_step(c);
   if (rc) return rc;
   rc = do_relist(c, c->sp.sparse_add_width_v, true, &sp);
   if (rc) return rc;
   rc = finish(c);
   if (rc < 0) return rc;
   if (rho_loss) HIPCHK(hipMemcpy(rho_loss, c->rho_loss, c->ncells * sizeof(double), hipMemcpyDeviceToHost));
   return rc;
}

int vacc_autotune(vacc_ctx* c, uint32_t* mult_out) {
   if (!c) return set_err(VACC_E_ARG, "null context");
   if (!c->ncells) return 0;
   int rc = reserve_for_step(c);
   if (rc) return rc;
   if (c->fields_valid) enqueue_transform(c, 0, 0);   // intersections of the current fields
   // the trial passes write the other list and pool (scratch) and the status words; keep the
   // status words and leave the state as it was
   uint32_t* saved = nullptr;
   HIPCHK(hipMalloc(&saved, ((size_t)c->ncells + 1) * sizeof(uint32_t)));
   HIPCHK(hipMemcpyAsync(saved, c->status, ((size_t)c->ncells + 1) * sizeof(uint32_t), hipMemcpyDeviceToDevice,
                         c->stream));
   const bool prof = c->prof;
   c->prof = false;
   // workgroups per CU, and (tolerance form) lanes per workgroup: 4, 8 or 16 work units each
   static const int cand[] = {8, 32, 128, 512, 1024, 2048, 4096};
   static const int wg_fast[] = {64, 128, 256};
   static const int wg_exact[] = {256};
   const bool fast = c->opt.fast_arith != 0;
   const int* wgc = fast ? wg_fast : wg_exact;
   const int nwg = fast ? 3 : 1;
   hipEvent_t e0, e1;
   HIPCHK(hipEventCreate(&e0));
   HIPCHK(hipEventCreate(&e1));
   // the tolerance form with block sums tunes the passes that store them separately
   const int nvar = (fast && c->blk_stats) ? 2 : 1;
   for (int v = 0; v < nvar && rc == 0; ++v)
   for (int dim = 0; dim < 3 && rc == 0; ++dim) {
      float best = 1e30f;
      int& mult = v ? c->remap_mult_s[dim] : c->remap_mult[dim];
      int& wgs = v ? c->remap_wgs_s[dim] : c->remap_wgs[dim];
      int arg = mult, argw = wgs;
      for (int rep = 0; rep < 2 && rc == 0; ++rep)
         for (int wi = 0; wi < nwg && rc == 0; ++wi)
            for (int m : cand) {
               (void)hipEventRecord(e0, c->stream);
               rc = enqueue_pass(c, dim, m, wgc[wi], v == 1);
               (void)hipEventRecord(e1, c->stream);
               if (rc) break;
               HIPCHK(hipEventSynchronize(e1));
               float ms = 0.0f;
               HIPCHK(hipEventElapsedTime(&ms, e0, e1));
               if (ms < best) {
                  best = ms;
                  arg = m;
                  argw = wgc[wi];
               }
            }
      mult = arg;
      wgs = argw;
   }
   (void)hipEventDestroy(e0);
   (void)hipEventDestroy(e1);
   c->prof = prof;
   HIPCHK(hipMemcpyAsync(c->status, saved, ((size_t)c->ncells + 1) * sizeof(uint32_t), hipMemcpyDeviceToDevice,
                         c->stream));
   HIPCHK(hipStreamSynchronize(c->stream));
   HIPCHK(hipFree(saved));
   if (rc) return rc;
   c->content_valid = false;   // the trial passes wrote content flags of the scratch list
   if (nvar > 1) c->stats_valid = false;   // and block sums
   drop_graphs(c);             // captured steps hold the old launch shape
   const int fl = finish(c);
   if (fl < 0) return fl;
   if (mult_out)
      for (int d = 0; d < 3; ++d) mult_out[d] = (uint32_t)c->remap_mult[d];
   return 0;
}

int vacc_get_remap_shape(vacc_ctx* c, uint32_t* wg_per_cu, uint32_t* lanes_per_wg) {
   if (!c) return set_err(VACC_E_ARG, "null context");
   for (int d = 0; d < 3; ++d) {
      if (wg_per_cu) wg_per_cu[d] = (uint32_t)c->remap_mult[d];
      if (lanes_per_wg) lanes_per_wg[d] = (uint32_t)(c->opt.fast_arith ? c->remap_wgs[d] : 256);   // the exact kernel: 256
   }
   return 0;
}

// The launches of one step (transform, three passes, adjust) on the context stream.
static int enqueue_step(vacc_ctx* c, uint32_t map_order, int32_t use_moments, int32_t adjust) {
   enqueue_transform(c, map_order, use_moments ? 1 : 0);
   for (int p = 0; p < 3; ++p) {
      const int rc = do_pass(c, pass_dim((int)map_order, p), p == 2);
      if (rc) return rc;
   }
   if (adjust) {
      const int rc = do_relist(c, c->sp.sparse_add_width_v, true);
      if (rc) return rc;
   }
   return 0;
}

// Replay the captured step for the current state, capturing it first when needed.  The host
// state transitions of the step (list/pool parity, validity flags) are recorded at capture
// and applied on replay.
static int graph_step(vacc_ctx* c, uint32_t map_order, int32_t use_moments, int32_t adjust) {
   const int um = use_moments ? 1 : 0, ad = adjust ? 1 : 0;
   for (const auto& g : c->graphs)
      if (g.gen == c->gen && g.map_order == map_order && g.use_moments == um && g.adjust == ad && g.lcur == c->lcur &&
          g.fcur == c->fcur && g.moments_in == c->moments_valid && g.content_in == c->content_valid &&
          g.stats_in == c->stats_valid) {
         HIPCHK(hipGraphLaunch(g.exec, c->stream));
         c->lcur = g.lcur_out;
         c->fcur = g.fcur_out;
         c->moments_valid = g.moments_out;
         c->content_valid = g.content_out;
         c->stats_valid = g.stats_out;
         c->sorted = g.sorted_out;
         return 0;
      }
   vacc_ctx::StepGraph g{};
   g.gen = c->gen;
   g.map_order = map_order;
   g.use_moments = um;
   g.adjust = ad;
   g.lcur = c->lcur;
   g.fcur = c->fcur;
   g.moments_in = c->moments_valid;
   g.content_in = c->content_valid;
   g.stats_in = c->stats_valid;
   HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
   const int rc = enqueue_step(c, map_order, use_moments, adjust);
   hipGraph_t graph = nullptr;
   const hipError_t e = hipStreamEndCapture(c->stream, &graph);
   if (rc) {
      if (graph) (void)hipGraphDestroy(graph);
      return rc;
   }
   if (e != hipSuccess) return set_err(VACC_E_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
   const hipError_t ei = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
   (void)hipGraphDestroy(graph);
   if (ei != hipSuccess) return set_err(VACC_E_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
   g.lcur_out = c->lcur;
   g.fcur_out = c->fcur;
   g.moments_out = c->moments_valid;
   g.content_out = c->content_valid;
   g.sorted_out = c->sorted;
   g.stats_out = c->stats_valid;
   c->graphs.push_back(g);
   HIPCHK(hipGraphLaunch(g.exec, c->stream));
   return 0;
}

int vacc_step(vacc_ctx* c, uint32_t map_order, int32_t use_moments, int32_t adjust, double* rho_loss) {
   if (!c || map_order > 2) return set_err(VACC_E_ARG, "bad argument");
   if (!c->ncells) return 0;
   if (!c->fields_valid) return set_err(VACC_E_STATE, "no cell fields: call vacc_set_fields first");
   int rc = reserve_for_step(c);
   if (rc) return rc;
   rc = (c->opt.use_graph && !c->prof) ? graph_step(c, map_order, use_moments, adjust)
                                       : enqueue_step(c, map_order, use_moments, adjust);
   if (rc) return rc;
   rc = finish(c);
   if (rc < 0) return rc;
   if (adjust && rho_loss) HIPCHK(hipMemcpy(rho_loss, c->rho_loss, c->ncells * sizeof(double), hipMemcpyDeviceToHost));
   return rc;
}

int vacc_subcycle_count(vacc_ctx* c, double dt, uint32_t* nmax) {
   if (!c || !nmax || !std::isfinite(dt)) return set_err(VACC_E_ARG, "bad argument");
   *nmax = 0;
   if (!c->ncells) return 0;
   if (!c->fields_valid) return set_err(VACC_E_STATE, "no cell fields: call vacc_set_fields first");
   if (!(c->opt.max_rotation_deg > 0.0)) return set_err(VACC_E_ARG, "max_rotation_deg must be positive");
   launch_fill_u32(c->sub_nmax, 0u, 1, c->stream);
   launch_subcycle_count(c->fields, c->ncells, c->sp.mass, c->sp.charge, c->opt.max_rotation_deg, dt, c->sub_maxdt,
                         c->sub_n, c->sub_nmax, c->stream);
   HIPCHK(hipGetLastError());
   HIPCHK(hipMemcpyAsync(c->h_sum + 3, c->sub_nmax, sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
   HIPCHK(hipStreamSynchronize(c->stream));
   c->sub_max = c->h_sum[3];
   c->sub_dt_total = dt;
   c->sub_valid = true;
   *nmax = c->sub_max;
   return 0;
}

int vacc_subcycle_step(vacc_ctx* c, uint32_t step, uint32_t map_order, int32_t use_moments, int32_t adjust,
                       double* rho_loss) {
   if (!c || map_order > 2) return set_err(VACC_E_ARG, "bad argument");
   if (!c->ncells) return 0;
   if (!c->fields_valid || !c->sub_valid)
      return set_err(VACC_E_STATE, "no subcycle counts for the current fields: call vacc_subcycle_count first");
   int rc = reserve_for_step(c);
   if (rc) return rc;
   launch_subcycle_step(c->sub_maxdt, c->sub_n, c->ncells, step, c->sub_dt_total, c->sub_dt, c->active, c->stream);
   HIPCHK(hipGetLastError());
   c->subcycling = true;
   rc = enqueue_step(c, map_order, use_moments, adjust);
   c->subcycling = false;
   if (rc) return rc;
   rc = finish(c);
   if (rc < 0) return rc;
   if (adjust && rho_loss) HIPCHK(hipMemcpy(rho_loss, c->rho_loss, c->ncells * sizeof(double), hipMemcpyDeviceToHost));
   return rc;
}

int vacc_accelerate_subcycled(vacc_ctx* c, double dt, uint32_t map_order, int32_t use_moments, uint32_t* nsub,
                              double* rho_loss) {
   if (!c || map_order > 2) return set_err(VACC_E_ARG, "bad argument");
   uint32_t n = 0;
   int rc = vacc_subcycle_count(c, dt, &n);
   if (rc) return rc;
   if (nsub) *nsub = n;
   std::vector<double> loss(rho_loss ? c->ncells : 0, 0.0), part(loss.size());
   if (rho_loss) std::fill(rho_loss, rho_loss + c->ncells, 0.0);
   int st = 0;
   for (uint32_t s = 0; s < n; ++s) {
      rc = vacc_subcycle_step(c, s, map_order, use_moments, 1, rho_loss ? part.data() : nullptr);
      if (rc < 0) return rc;
      st |= rc;
      for (size_t i = 0; i < part.size(); ++i) rho_loss[i] += part[i];
   }
   return st;
}

int vacc_accelerate_populations(vacc_ctx* const* ctxs, uint32_t npop, uint32_t map_order, int32_t use_moments,
                                int32_t adjust, double* const* rho_loss) {
   if (!ctxs || npop == 0 || npop > (uint32_t)MAX_POP || map_order > 2) return set_err(VACC_E_ARG, "bad argument");
   for (uint32_t p = 0; p < npop; ++p) {
      if (!ctxs[p]) return set_err(VACC_E_ARG, "null context");
      if (ctxs[p]->ncells != ctxs[0]->ncells) return set_err(VACC_E_ARG, "populations hold different cell counts");
      if (!ctxs[p]->fields_valid && ctxs[p]->ncells)
         return set_err(VACC_E_STATE, "no cell fields: call vacc_set_fields on every population");
      for (uint32_t q = 0; q < p; ++q)
         if (ctxs[q] == ctxs[p]) return set_err(VACC_E_ARG, "a context appears twice");
   }
   if (!ctxs[0]->ncells) return 0;
   for (uint32_t p = 0; p < npop; ++p) {
      const int rc = reserve_for_step(ctxs[p]);
      if (rc) return rc;
   }
   // populations one after another, as calculateAcceleration's loop over popID
   // (vlasovmover.cpp:503-565): each one's pivot sees the others' current state
   for (uint32_t p = 0; p < npop; ++p) {
      vacc_ctx* c = ctxs[p];
      if (use_moments) {
         PlasmaArgs pa{};
         pa.npop = (int)npop;
         pa.ncells = c->ncells;
         pa.out = c->pivot;
         for (uint32_t q = 0; q < npop; ++q) {
            vacc_ctx* o = ctxs[q];
            enqueue_moments(o);
            if (q != p) {
               HIPCHK(hipEventRecord(o->mom_ev, o->stream));
               HIPCHK(hipStreamWaitEvent(c->stream, o->mom_ev, 0));
            }
            pa.mom[q] = o->moments;
            pa.mass[q] = o->sp.mass;
            pa.charge[q] = o->sp.charge;
         }
         {
            ProfScope ps(c, 3);
            launch_plasma(pa, c->stream);
         }
         enqueue_transform(c, map_order, 2);
      } else {
         enqueue_transform(c, map_order, 0);
      }
      for (int d = 0; d < 3; ++d) {
         const int rc = do_pass(c, pass_dim((int)map_order, d), d == 2);
         if (rc) return rc;
      }
      if (adjust) {
         const int rc = do_relist(c, c->sp.sparse_add_width_v, true);
         if (rc) return rc;
      }
      HIPCHK(hipGetLastError());
   }
   int st = 0;
   for (uint32_t p = 0; p < npop; ++p) {
      const int rc = finish(ctxs[p]);
      if (rc < 0) return rc;
      st |= rc;
   }
   if (adjust && rho_loss)
      for (uint32_t p = 0; p < npop; ++p)
         if (rho_loss[p])
            HIPCHK(hipMemcpy(rho_loss[p], ctxs[p]->rho_loss, ctxs[p]->ncells * sizeof(double), hipMemcpyDeviceToHost));
   return st;
}

int vacc_find_blocks(vacc_ctx* c, uint64_t nq, const uint32_t* cells, const uint32_t* gids, uint32_t* index) {
   if (!c) return set_err(VACC_E_ARG, "null context");
   if (!nq) return 0;
   if (!cells || !gids || !index) return set_err(VACC_E_ARG, "null argument");
   int rc = 0;
   if (c->ncells && !c->sorted && (rc = do_relist(c, -1, false))) return rc;   // ascending GIDs, same blocks
   // queries go through the device in chunks of FIND_CHUNK: cell, GID and the answer
   constexpr uint64_t FIND_CHUNK = 1u << 24;
   const uint64_t chunk = std::min<uint64_t>(nq, FIND_CHUNK);
   DevTemps t;
   uint32_t* d = nullptr;
   HIPCHK(t.alloc(&d, 3 * chunk * sizeof(uint32_t)));
   for (uint64_t first = 0; first < nq; first += chunk) {
      const uint32_t m = (uint32_t)std::min<uint64_t>(chunk, nq - first);
      HIPCHK(hipMemcpyAsync(d, cells + first, m * sizeof(uint32_t), hipMemcpyHostToDevice, c->stream));
      HIPCHK(hipMemcpyAsync(d + chunk, gids + first, m * sizeof(uint32_t), hipMemcpyHostToDevice, c->stream));
      if (c->ncells) launch_find(c->off[c->lcur], c->gid[c->lcur], c->ncells, d, d + chunk, d + 2 * chunk, m, c->stream);
      else launch_fill_u32(d + 2 * chunk, VACC_INVALID_ID, m, c->stream);
      HIPCHK(hipGetLastError());
      HIPCHK(hipMemcpyAsync(index + first, d + 2 * chunk, m * sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
      HIPCHK(hipStreamSynchronize(c->stream));
   }
   return 0;
}

int vacc_find_blocks_noalloc(vacc_ctx* c, uint64_t nq, const uint32_t* cells, const uint32_t* gids, uint32_t* index) {
   if (!c) return set_err(VACC_E_ARG, "null context");
   if (nq > VACC_FIND_NOALLOC_MAX)
      return set_err(VACC_E_ARG, "vacc_find_blocks_noalloc: " + std::to_string(nq) + " queries exceed VACC_FIND_NOALLOC_MAX (" +
                                     std::to_string(VACC_FIND_NOALLOC_MAX) + "); split the call or use vacc_find_blocks");
   if (!nq) return 0;
   if (!cells || !gids || !index) return set_err(VACC_E_ARG, "null argument");
   uint32_t* d = (c->device >= 0 && c->device < SCRATCH_MAX_DEVICES) ? g_scratch_ptr[c->device] : nullptr;
   if (!d) {
      const hipError_t e = (c->device >= 0 && c->device < SCRATCH_MAX_DEVICES) ? g_scratch_err[c->device] : hipErrorInvalidDevice;
      return set_err(VACC_E_HIP, std::string("vacc_find_blocks_noalloc: no scratch address for this device: ") +
                                     hipGetErrorString(e));
   }
   int rc = 0;
   if (c->ncells && !c->sorted && (rc = do_relist(c, -1, false))) return rc;   // ascending GIDs, no allocation
   const uint32_t m = (uint32_t)nq;
   uint32_t* qc = d;
   uint32_t* qg = d + VACC_FIND_NOALLOC_MAX;
   uint32_t* ans = d + 2 * (size_t)VACC_FIND_NOALLOC_MAX;
   HIPCHK(hipMemcpyAsync(qc, cells, m * sizeof(uint32_t), hipMemcpyHostToDevice, c->stream));
   HIPCHK(hipMemcpyAsync(qg, gids, m * sizeof(uint32_t), hipMemcpyHostToDevice, c->stream));
   if (c->ncells) launch_find(c->off[c->lcur], c->gid[c->lcur], c->ncells, qc, qg, ans, m, c->stream);
   else launch_fill_u32(ans, VACC_INVALID_ID, m, c->stream);
   HIPCHK(hipGetLastError());
   HIPCHK(hipMemcpyAsync(index, ans, m * sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
   HIPCHK(hipStreamSynchronize(c->stream));
   return 0;
}

int vacc_moments(vacc_ctx* c, double* out) {
   if (!c || (!out && c->ncells)) return set_err(VACC_E_ARG, "null argument");
   if (!c->ncells) return 0;
   enqueue_moments(c);
   HIPCHK(hipGetLastError());
   HIPCHK(hipMemcpyAsync(out, c->moments, (size_t)c->ncells * 4 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
   HIPCHK(hipStreamSynchronize(c->stream));
   return 0;
}

int vacc_pressure(vacc_ctx* c, const double* pivot, double* out) {
   if (!c || (!out && c->ncells)) return set_err(VACC_E_ARG, "null argument");
   if (!c->ncells) return 0;
   const size_t n = c->ncells;
   std::vector<double> piv(3 * n);
   if (pivot) {
      std::copy(pivot, pivot + 3 * n, piv.begin());
   } else {   // the population's own bulk velocity from its first moments
      std::vector<double> mo(4 * n);
      int rc = vacc_moments(c, mo.data());
      if (rc) return rc;
      for (size_t i = 0; i < n; ++i)
         for (int d = 0; d < 3; ++d) piv[3 * i + d] = (mo[4 * i] != 0.0) ? mo[4 * i + 1 + d] / mo[4 * i] : 0.0;
   }
   DevTemps tmp;
   double *d_piv = nullptr, *d_out = nullptr;
   HIPCHK(tmp.alloc(&d_piv, 3 * n * sizeof(double)));
   HIPCHK(tmp.alloc(&d_out, 6 * n * sizeof(double)));
   HIPCHK(hipMemcpyAsync(d_piv, piv.data(), 3 * n * sizeof(double), hipMemcpyHostToDevice, c->stream));
   MomentArgs ma{};
   ma.cell_off = c->off[c->lcur];
   ma.gid = c->gid[c->lcur];
   ma.src = c->src[c->lcur];
   ma.pool = pool(c, c->fcur);
   ma.ncells = c->ncells;
   for (int d = 0; d < 3; ++d) {
      ma.len[d] = c->mi.len[d];
      ma.vmin[d] = c->mi.vmin[d];
      ma.cell[d] = c->mi.cell[d];
   }
   ma.gs = make_gid_split(c->mi.len);
   launch_pressure(ma, d_piv, c->sp.mass, d_out, c->stream);
   HIPCHK(hipGetLastError());
   HIPCHK(hipMemcpyAsync(out, d_out, 6 * n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
   HIPCHK(hipStreamSynchronize(c->stream));
   return 0;
}

int vacc_init_peaks(vacc_ctx* c, uint32_t ncells, uint32_t max_peaks, const vacc_peak* peaks, const int32_t* npeaks,
                    const int32_t* bbox) {
   if (!c || (ncells && (!peaks || !npeaks || !bbox)) || max_peaks == 0) return set_err(VACC_E_ARG, "bad argument");
   std::vector<uint32_t> bm_off(ncells + 1, 0);
   for (uint32_t i = 0; i < ncells; ++i) {
      const int32_t* b = bbox + 6 * (size_t)i;
      uint32_t words = 0;
      if (npeaks[i] > 0) {
         if (npeaks[i] > (int32_t)max_peaks) return set_err(VACC_E_ARG, "npeaks > max_peaks");
         for (int d = 0; d < 3; ++d)
            if (b[d] < 0 || b[d + 3] < b[d] || b[d + 3] >= (int32_t)c->mi.len[d])
               return set_err(VACC_E_ARG, "candidate box outside the mesh");
         const uint32_t rx = b[3] - b[0] + 1, ry = b[4] - b[1] + 1, rz = b[5] - b[2] + 1;
         words = ry * rz * ((rx + 31) / 32);
         if (words > (uint32_t)init_box_words_max()) return set_err(VACC_E_ARG, "candidate box too large");
      }
      bm_off[i + 1] = bm_off[i] + words;
   }
   c->content_valid = c->moments_valid = c->fields_valid = c->sub_valid = c->stats_valid = false;
   const int rc = init_peaks_load(c, ncells, max_peaks, peaks, npeaks, bbox, bm_off);
   if (rc) abort_load(c);
   return rc;
}

int vacc_profile_begin(vacc_ctx* c) {
   if (!c) return set_err(VACC_E_ARG, "null context");
   HIPCHK(hipStreamSynchronize(c->stream));
   if (!c->prof_counts) HIPCHK(hipMalloc(&c->prof_counts, 2 * PROF_LOG * sizeof(uint32_t)));
   c->prof_sums.assign(PROF_LOG, 0);
   for (auto& r : c->recs) {
      c->ev_free.push_back(r.a);
      c->ev_free.push_back(r.b);
   }
   c->recs.clear();
   c->nlog = 0;
   c->prof = true;
   return 0;
}

int vacc_profile_read(vacc_ctx* c, vacc_profile* out) {
   if (!c || !out) return set_err(VACC_E_ARG, "null argument");
   HIPCHK(hipStreamSynchronize(c->stream));
   std::memset(out, 0, sizeof(*out));
   for (auto& r : c->recs) {
      float ms = 0.0f;
      HIPCHK(hipEventElapsedTime(&ms, r.a, r.b));
      if (r.kind == 0) { out->remap_launches++; out->remap_ms += ms; }
      else if (r.kind == 1) { out->build_launches++; out->build_ms += ms; }
      else if (r.kind == 2) { out->adjust_launches++; out->adjust_ms += ms; }
      else if (r.kind == 3) { out->transform_launches++; out->transform_ms += ms; }
      else if (r.kind == 4) { out->trans_launches++; out->trans_ms += ms; }
   }
   if (c->nlog) {
      std::vector<uint32_t> cnt(2 * c->nlog);
      HIPCHK(hipMemcpy(cnt.data(), c->prof_counts, cnt.size() * sizeof(uint32_t), hipMemcpyDeviceToHost));
      for (uint32_t i = 0; i < c->nlog; ++i) {
         out->remap_src_blocks += cnt[2 * i];
         out->remap_tgt_blocks += cnt[2 * i + 1];
         if (c->prof_sums[i]) out->remap_sum_blocks += cnt[2 * i + 1];
      }
   }
   return 0;
}

int vacc_sync(vacc_ctx* c) {
   if (!c) return set_err(VACC_E_ARG, "null context");
   HIPCHK(hipStreamSynchronize(c->stream));
   return status_or(c);
}

int vacc_get_status(vacc_ctx* c, uint32_t* status) {
   if (!c || (!status && c->ncells)) return set_err(VACC_E_ARG, "null argument");
   HIPCHK(hipStreamSynchronize(c->stream));
   if (c->ncells) {
      HIPCHK(hipMemcpy(status, c->status, c->ncells * sizeof(uint32_t), hipMemcpyDeviceToHost));
      HIPCHK(hipMemset(c->status, 0, (c->ncells + 1) * sizeof(uint32_t)));
   }
   return 0;
}

// ---------------------------------------------------------------------------
// Spatial translation (include/vacc_trans.h)
// ---------------------------------------------------------------------------


int vacc_trans_set_pencils(vacc_ctx* c, uint32_t dim, const vacc_pencils* P) {
   if (!c || !P) return set_err(VACC_E_ARG, "null argument");
   if (dim > 2) return set_err(VACC_E_ARG, "dim must be 0, 1 or 2");
   if (P->npencils && (!P->length || !P->start || !P->cell || !P->dz || !P->ratio))
      return set_err(VACC_E_ARG, "null pencil array");
   const uint32_t nc = c->ncells;
   TransPlan plan;
   int rc = 0;
   if ((rc = plan_schedule(P, nc, plan))) return rc;
   plan_walk(P, nc, c->opt.storage == VACC_STORE_F32, plan);
   for (uint32_t t = 0; t < nc; ++t)   // k_trans reads at most TRANS_MAXN stencil cells per cell
      if (plan.wide[t] && !plan.walk[t]) return set_err(VACC_E_ARG, "a target cell's stencils span more than 12 cells");
   const auto& ent = plan.ent;
   const auto& tab = plan.tab;
   const auto& walk = plan.walk;
   const auto& segs = plan.segs;
   const auto& wpos = plan.wpos;
   const auto& wface = plan.wface;
   const size_t total = plan.total;
   if (wpos.size() > 0xFFFFFFFFull) return set_err(VACC_E_ARG, "too many pencil positions");
   std::vector<TransCell> cells(nc);
   std::vector<TransEntry> flat;
   std::vector<uint32_t> rest;
   flat.reserve(total);
   for (uint32_t t = 0; t < nc; ++t) {
      if (!walk[t]) rest.push_back(t);
      TransCell& C = cells[t];
      C.e0 = (uint32_t)flat.size();
      C.ne = (uint32_t)ent[t].size();
      C.nn = (uint32_t)tab[t].size();
      for (size_t j = 0; j < tab[t].size(); ++j) C.nbr[j] = tab[t][j];
      C.walked = walk[t];
      flat.insert(flat.end(), ent[t].begin(), ent[t].end());
   }
   auto& T = c->trans[dim];
   HIPCHK(hipStreamSynchronize(c->stream));
   if ((rc = dev_grow(T.cells, T.cells_cap, cells.size())) || (rc = dev_grow(T.entries, T.entries_cap, flat.size())) ||
       (rc = dev_grow(T.seg, T.seg_cap, segs.size())) || (rc = dev_grow(T.pos, T.pos_cap, wpos.size())) ||
       (rc = dev_grow(T.face, T.face_cap, wface.size())) || (rc = dev_grow(T.rest, T.rest_cap, rest.size())))
      return rc;
   if (nc) HIPCHK(hipMemcpy(T.cells, cells.data(), nc * sizeof(TransCell), hipMemcpyHostToDevice));
   if (!flat.empty()) HIPCHK(hipMemcpy(T.entries, flat.data(), flat.size() * sizeof(TransEntry), hipMemcpyHostToDevice));
   if (!segs.empty()) {
      HIPCHK(hipMemcpy(T.seg, segs.data(), segs.size() * sizeof(WalkSeg), hipMemcpyHostToDevice));
      HIPCHK(hipMemcpy(T.pos, wpos.data(), wpos.size() * sizeof(WalkPos), hipMemcpyHostToDevice));
      HIPCHK(hipMemcpy(T.face, wface.data(), wface.size() * sizeof(TransFace), hipMemcpyHostToDevice));
   }
   if (!rest.empty()) HIPCHK(hipMemcpy(T.rest, rest.data(), rest.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
   T.nrest = (uint32_t)rest.size();
   T.nseg = (uint32_t)segs.size();
   T.level_first = plan.level_first;
   // walk waves per segment: a power of two, about WALK_WAVES in all, at most 1024 per segment;
   // the share bounds of every wave are computed on the device at each call
   T.wps = 1;
   while (T.wps < 1024 && (uint64_t)T.nseg * T.wps * 2 <= WALK_WAVES) T.wps *= 2;
   if ((rc = dev_grow(T.bounds, T.bounds_cap, (size_t)T.nseg * (T.wps + 1)))) return rc;
   T.epoch = c->cell_epoch;
   return 0;
}

int vacc_trans_map_1d(vacc_ctx* c, uint32_t dim, double dt) {
   if (!c) return set_err(VACC_E_ARG, "null context");
   if (dim > 2) return set_err(VACC_E_ARG, "dim must be 0, 1 or 2");
   const auto& T = c->trans[dim];
   if (T.epoch != c->cell_epoch) return set_err(VACC_E_STATE, "no pencils installed for the current cells and dim");
   if (!c->ncells) return 0;
   // captured steps stay valid: no buffer moves, and the list/pool parity and validity flags
   // that select a captured variant are updated below
   int rc = 0;
   if (!c->sorted && (rc = do_relist(c, -1, false))) return rc;   // ascending GIDs, same blocks
   if ((rc = refresh_count(c))) return rc;
   TransArgs ta{};
   ta.off = c->off[c->lcur];
   ta.gid = c->gid[c->lcur];
   ta.src = c->src[c->lcur];
   ta.in = pool(c, c->fcur);
   ta.out = pool(c, 1 - c->fcur);
   ta.cells = T.cells;
   ta.entries = T.entries;
   ta.rest = T.rest;
   ta.nrest = T.nrest;
   ta.ncells = c->ncells;
   ta.nblocks = (uint32_t)c->nblocks;
   ta.dim = (int)dim;
   for (int d = 0; d < 3; ++d) ta.len[d] = c->mi.len[d];
   ta.dvz = (float)c->mi.cell[dim];
   ta.vz_min = (float)c->mi.vmin[dim];
   ta.dt = (float)dt;
   ta.thr = (float)c->sp.sparse_min_value;
   ta.scale = (float)(1.0 / (double)ta.thr);
   WalkArgs wa{};
   wa.t = ta;
   wa.seg = T.seg;
   wa.pos = T.pos;
   wa.face = T.face;
   wa.nseg = T.nseg;
   wa.nmesh = c->mi.len[0] * c->mi.len[1] * c->mi.len[2];
   wa.wps = T.wps;
   wa.bin_shift = 0;
   while (wa.nmesh && ((wa.nmesh - 1) >> wa.bin_shift) >= 4096u) ++wa.bin_shift;
   {
      ProfScope ps(c, 4);
      for (size_t lv = 0; lv + 1 < T.level_first.size(); ++lv) {   // launch levels in order
         const uint32_t first = T.level_first[lv], count = T.level_first[lv + 1] - first;
         if (!count) continue;
         WalkArgs wl = wa;
         wl.seg = T.seg + first;
         wl.nseg = count;
         uint32_t* bounds = T.bounds + (size_t)first * (T.wps + 1);
         wl.bounds = bounds;
         launch_trans_walk(wl, bounds, c->stream);
      }
      launch_trans(ta, c->stream);
   }
   HIPCHK(hipGetLastError());
   launch_iota(c->src[c->lcur], (uint32_t)c->nblocks, c->stream);
   c->fcur = 1 - c->fcur;
   c->content_valid = false;
   c->moments_valid = false;
   c->stats_valid = false;
   HIPCHK(hipStreamSynchronize(c->stream));
   return 0;
}

}  // extern "C"
