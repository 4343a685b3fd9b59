// vacc_kernels.hip -- gfx950 kernels of the velocity-space acceleration step.
//
// One 1-D remap pass along a dimension is three launches over all cells:
//   k_columns  one workgroup per cell: LDS bitmap counting sort of the cell's
//              blocks into column sets / columns along the pass axis, target
//              extents of every column (cpu_acc_map.cpp:250-353) and the
//              output slot of every target block.
//   k_scan2    one workgroup: exclusive scans of per-cell set and target
//              counts (global work list and output offsets).
//   k_remap    persistent: 16 lanes per column set, one lane per transverse
//              pencil, walks the columns of the set sourcing the blocks from
//              HBM once and writing every target block once, with the PLM,
//              PPM or PQM reconstruction and the piecewise integration of
//              cpu_acc_map.cpp:393-631 in registers and a per-lane target
//              ring in LDS.
// Block adjust (spatial_cell_cpu.cpp:179-273) is k_adjust<false> (count) +
// k_scan2 + k_adjust<true> (write), again with an LDS bitmap per cell.
//
// Floating point follows the reference CPU path built with the portable
// vector class (VEC_FALLBACK_GENERIC): every float operation is rounded to
// float, a float combined with a double literal is evaluated in double and
// rounded.  The file is compiled with -ffp-contract=off.
#include "vacc_kernels.h"
#include "vacc_recon.h"
#include "vacc_lean.h"
#include "vacc_codec.h"

#include <float.h>
#include <type_traits>
#include <limits.h>

namespace vacc {

constexpr int WG = 256;
constexpr int CWG = 512;         // k_columns workgroup: 8 waves share one cell's LDS bitmap
constexpr int AWG = 512;         // k_adjust workgroup, likewise
// LDS bitmap words per chunk of rows, and rows per chunk of k_columns: 37 KiB of LDS per
// k_columns workgroup (four per CU, at 64 VGPRs: 7 % faster than three with 1536-row
// chunks) and at most 40 KiB per k_adjust workgroup (three per CU).  Larger chunks (28 KiB
// bitmaps, two workgroups per CU) made the column build 25 % and the adjust 9 % slower on the
// 8192-cell 75^3 workload; most cells' boxes fit one chunk either way.
constexpr int BM_WORDS = ADJ_BM_WORDS;
constexpr int ROW_MAX = 1024;

// ---------------------------------------------------------------------------
// small helpers
// ---------------------------------------------------------------------------

// DPP row rotation of an int within the 16 lanes of the caller's row (all active)
template <int N>
__device__ inline int row_ror_i(int x) {
   return (int)row_ror_u<N>((uint32_t)x);
}

template <int DIM>
__device__ inline void pass_coords(uint32_t gid, const GidSplit& gs, int& b0, int& b1, int& k) {
   uint32_t x, y, z;
   gid_xyz(gid, gs, x, y, z);
   // frame of map_1d after swapBlockIndices (cpu_acc_map.cpp:85-106)
   if (DIM == 0) { b0 = (int)z; b1 = (int)y; k = (int)x; }
   else if (DIM == 1) { b0 = (int)x; b1 = (int)z; k = (int)y; }
   else { b0 = (int)x; b1 = (int)y; k = (int)z; }
}

template <int DIM>
__device__ inline uint32_t pass_gid(int b0, int b1, int k, const uint32_t* len) {
   // block_indices_to_id of cpu_acc_map.cpp:151-199
   const uint32_t nx = len[0], nxy = len[0] * len[1];
   if (DIM == 0) return (uint32_t)b0 * nxy + (uint32_t)b1 * nx + (uint32_t)k;
   if (DIM == 1) return (uint32_t)b0 + (uint32_t)b1 * nxy + (uint32_t)k * nx;
   return (uint32_t)b0 + (uint32_t)b1 * nx + (uint32_t)k * nxy;
}

template <int DIM>
__device__ inline void pass_isect(const double* is, float& I, float& di, float& dj, float& dk) {
   // (Realf) conversion then the i<->k (dim 0) or j<->k (dim 1) swap, cpu_acc_map.cpp:128-176
   I = (float)is[4 * DIM + 0];
   const float a = (float)is[4 * DIM + 1], b = (float)is[4 * DIM + 2], c = (float)is[4 * DIM + 3];
   if (DIM == 0) { di = c; dj = b; dk = a; }
   else if (DIM == 1) { di = a; dj = c; dk = b; }
   else { di = a; dj = b; dk = c; }
}

// work units are ordered longest first: bin 0 holds units of >= EXACT_BIN-1 source blocks;
// units marked EXACT_FLAG go to EXACT_BIN and units marked WIDE_FLAG to WIDE_BIN, the end of
// the list
__device__ inline uint32_t len_bin(uint32_t nblocks) {
   return (uint32_t)(EXACT_BIN - 1) - min(nblocks, (uint32_t)(EXACT_BIN - 1));
}
__device__ inline uint32_t unit_bin(const uint4& u) {
   return (u.z & WIDE_FLAG) ? (uint32_t)WIDE_BIN : (u.z & EXACT_FLAG) ? (uint32_t)EXACT_BIN : len_bin(u.w);
}

// Per-cell deposit span (dv + 3|di| + 3|dj|)/dk in target cells.  Cells whose Lagrangian
// spacing is not positive, or whose span would not fit even the wide ring, are passed through
// unchanged and flagged VACC_STATUS_DK; cells whose span exceeds the 32-cell ring's bound are
// remapped by the wide instance.
__device__ inline float pass_span(float dv, float di, float dj, float dk) {
   return (dv + 3.0f * fabsf(di) + 3.0f * fabsf(dj)) / dk;
}
__device__ inline bool pass_is_copy(float dv, float di, float dj, float dk) {
   if (!(dk > 0.0f)) return true;
   return !(pass_span(dv, di, dj, dk) <= (float)MAX_SPAN_WIDE);
}
__device__ inline bool pass_is_wide(float dv, float di, float dj, float dk) {
   return !(pass_span(dv, di, dj, dk) <= (float)MAX_SPAN_CELLS);
}

__device__ inline uint32_t wave_incl_scan(uint32_t v) {
   const int lane = threadIdx.x & 63;
#pragma unroll
   for (int o = 1; o < 64; o <<= 1) {
      const uint32_t u = __shfl_up(v, o, 64);
      if (lane >= o) v += u;
   }
   return v;
}

// Exclusive scan of NV values per thread over the workgroup (blockDim.x threads).
template <int NV>
__device__ inline void block_scan(const uint32_t (&v)[NV], uint32_t (&ex)[NV], uint32_t (&tot)[NV], uint32_t* sh) {
   const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
   uint32_t inc[NV];
#pragma unroll
   for (int q = 0; q < NV; ++q) inc[q] = wave_incl_scan(v[q]);
   if (lane == 63) {
#pragma unroll
      for (int q = 0; q < NV; ++q) sh[w * NV + q] = inc[q];
   }
   __syncthreads();
#pragma unroll
   for (int q = 0; q < NV; ++q) {
      uint32_t base = 0, all = 0;
      for (int i = 0; i < nw; ++i) {
         const uint32_t x = sh[i * NV + q];
         if (i < w) base += x;
         all += x;
      }
      ex[q] = base + inc[q] - v[q];
      tot[q] = all;
   }
   __syncthreads();
}

__device__ inline int next_one(const uint32_t* w, int W, int p) {
   int q = p >> 5;
   if (q >= W) return -1;
   uint32_t x = w[q] & (0xFFFFFFFFu << (p & 31));
   for (;;) {
      if (x) return q * 32 + __ffs(x) - 1;
      if (++q >= W) return -1;
      x = w[q];
   }
}
__device__ inline int next_zero(const uint32_t* w, int W, int p) {
   int q = p >> 5;
   if (q >= W) return W * 32;
   uint32_t x = ~w[q] & (0xFFFFFFFFu << (p & 31));
   for (;;) {
      if (x) return q * 32 + __ffs(x) - 1;
      if (++q >= W) return W * 32;
      x = ~w[q];
   }
}
__device__ inline uint32_t rank_in_row(const uint32_t* w, int kk) {
   uint32_t r = 0;
   const int q = kk >> 5;
   for (int i = 0; i < q; ++i) r += __popc(w[i]);
   const int b = kk & 31;
   if (b) r += __popc(w[q] & ((1u << b) - 1u));
   return r;
}

// Target extent of a column set: max/min of the Lagrangian base over the 4 corner cells,
// cpu_acc_map.cpp:258-288 (std::max(a,b) = a<b?b:a, std::min(a,b) = b<a?b:a).
struct SetGeom {
   float maxI, minI;
};
__device__ inline SetGeom set_geom(int b0, int b1, float I, float di, float dj) {
   const float x0 = (float)(uint32_t)(b0 * WID), x3 = (float)(uint32_t)(b0 * WID + WID - 1);
   const float y0 = (float)(uint32_t)(b1 * WID), y3 = (float)(uint32_t)(b1 * WID + WID - 1);
   const float c00 = (I + x0 * di) + y0 * dj;
   const float c03 = (I + x0 * di) + y3 * dj;
   const float c30 = (I + x3 * di) + y0 * dj;
   const float c33 = (I + x3 * di) + y3 * dj;
   SetGeom g;
   float m = c00;
   m = (m < c03) ? c03 : m;
   m = (m < c30) ? c30 : m;
   m = (m < c33) ? c33 : m;
   g.maxI = m;
   m = c00;
   m = (c03 < m) ? c03 : m;
   m = (c30 < m) ? c30 : m;
   m = (c33 < m) ? c33 : m;
   g.minI = m;
   return g;
}

// Target block range of one column, cpu_acc_map.cpp:307-338.
__device__ inline void column_extent(int ks, int ke, const SetGeom& g, float dv, float vmin, float dk,
                                     int maxlen, int wall, int& fk, int& lk, bool& wallhit) {
   const float fbv = (float)(uint32_t)(WID * ks) * dv + vmin;
   const float lbv = (float)(uint32_t)(WID * (ke + 1)) * dv + vmin;
   const int fgk = (int)((fbv - g.maxI) / dk);
   const int lgk = (int)((lbv - g.minI) / dk);
   fk = fgk / WID;
   lk = lgk / WID;
   fk = (fk >= 0) ? fk : 0;
   fk = (fk < maxlen) ? fk : maxlen - 1;
   lk = (lk >= 0) ? lk : 0;
   lk = (lk < maxlen) ? lk : maxlen - 1;
   wallhit = (fk < wall) || (fk >= maxlen - wall) || (lk < wall) || (lk >= maxlen - wall);
}

// The reference's lanes of minimum and maximum Lagrangian index (cpu_acc_map.cpp:515-529)
// start from DBL_MIN, so a group of 8 lanes whose Lagrangian indices at the column start are
// all <= 0 takes its first lane as the maximum.  True when that happens to either group of the
// column's 16 pencils (the expressions are k_remap's, so the signs agree exactly).
__device__ inline bool max_lane_rule_binds(int b0, int b1, int ks, float I, float di, float dj, float dv, float vmin) {
   const float vr = (float)(uint32_t)(WID * ks) * dv + vmin;
   bool none_pos[2] = {true, true};
#pragma unroll
   for (int jp = 0; jp < WID; ++jp)
#pragma unroll
      for (int ip = 0; ip < WID; ++ip) {
         const float imin = (I + (float)(uint32_t)(b0 * WID + ip) * di) + (float)(uint32_t)(b1 * WID + jp) * dj;
         if (vr - imin > 0.0f) none_pos[jp >> 1] = false;
      }
   return none_pos[0] || none_pos[1];
}

// Target ring of the tolerance-form remap, in cells per lane.  Before each source block the
// kernel opens the target blocks up to the highest target any lane of the unit can reach with
// the block's 4 cells (below 4 dv/dk + 2 past its open target) and has flushed those below the
// lowest target any lane still accumulates; the lanes' boundaries lie within
// (3|di| + 3|dj|)/dk + 1 cells of each other, and block alignment adds 7 cells.  So the open
// span stays within (4 dv + 3|di| + 3|dj|)/dk + 10 cells, and a cell whose span bound
// exceeds LEAN_SPAN = LEAN_RING - 10 is left to the exact kernel (32-cell ring).
constexpr int LEAN_RING = 16;
constexpr float LEAN_SPAN = (float)(LEAN_RING - 10);

// ---------------------------------------------------------------------------
// k_columns
// ---------------------------------------------------------------------------

// f(gid[i], i) for i = t, t + STRIDE, ... < n (a workgroup of STRIDE threads walking one cell's
// blocks), with four GID loads issued before the first use
template <int STRIDE, class F>
__device__ inline void for_blocks4(const uint32_t* gid, uint32_t n, int t, F&& f) {
   for (uint32_t i0 = (uint32_t)t; i0 < n; i0 += 4u * STRIDE) {
      uint32_t g[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
         const uint32_t i = i0 + (uint32_t)u * STRIDE;
         g[u] = (i < n) ? gid[i] : 0u;
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
         const uint32_t i = i0 + (uint32_t)u * STRIDE;
         if (i < n) f(g[u], i);
      }
   }
}

template <int DIM>
__global__ __launch_bounds__(CWG, 8) void k_columns(ColumnArgs a) {
   __shared__ uint32_t bm[BM_WORDS];
   __shared__ uint32_t rowv[4][ROW_MAX];
   __shared__ int sbox[6];
   __shared__ uint32_t sh[4 * (CWG / 64)];
   __shared__ uint32_t carry[4];
   __shared__ uint32_t sstat;
   __shared__ uint32_t hist[LEN_BINS];
   const int t = threadIdx.x;
   const uint32_t c = blockIdx.x;
   const uint32_t off = a.cell_off[c];
   const uint32_t n = a.cell_off[c + 1] - off;
   if (n == 0) {
      if (t == 0) {
         a.n_units[c] = 0;
         a.n_targets[c] = 0;
      }
      for (int i = t; i < LEN_BINS; i += CWG) a.cell_hist[(size_t)i * a.ncells + c] = 0u;
      return;
   }
   for (int i = t; i < LEN_BINS; i += CWG) hist[i] = 0;
   if (t < 3) {
      sbox[t] = INT_MAX;
      sbox[t + 3] = INT_MIN;
   }
   if (t < 4) carry[t] = 0;
   if (t == 0) sstat = 0;
   __syncthreads();
   {
      int lo0 = INT_MAX, lo1 = INT_MAX, lok = INT_MAX, hi0 = INT_MIN, hi1 = INT_MIN, hik = INT_MIN;
      for_blocks4<CWG>(a.gid + off, n, t, [&](uint32_t gid, uint32_t) {
         int b0, b1, k;
         pass_coords<DIM>(gid, a.gs, b0, b1, k);
         lo0 = min(lo0, b0); hi0 = max(hi0, b0);
         lo1 = min(lo1, b1); hi1 = max(hi1, b1);
         lok = min(lok, k); hik = max(hik, k);
      });
      atomicMin(&sbox[0], lo0); atomicMin(&sbox[1], lo1); atomicMin(&sbox[2], lok);
      atomicMax(&sbox[3], hi0); atomicMax(&sbox[4], hi1); atomicMax(&sbox[5], hik);
   }
   __syncthreads();
   const int mn0 = sbox[0], mn1 = sbox[1], mnk = sbox[2];
   const int R1 = sbox[4] - mn1 + 1;
   const int R0 = sbox[3] - mn0 + 1;
   const int K = sbox[5] - mnk + 1;
   const int W = (K + 31) >> 5;
   const int rows = R0 * R1;
   const int rpc = min(ROW_MAX, BM_WORDS / W);
   float I, di, dj, dk;
   pass_isect<DIM>(a.isect + 12 * (size_t)c, I, di, dj, dk);
   // a cell outside this subcycle passes through unchanged (no status)
   const bool idle = a.active && !a.active[c];
   const bool copy = idle || pass_is_copy(a.dv, di, dj, dk);
   const bool wide = !copy && pass_is_wide(a.dv, di, dj, dk);
   const uint32_t wf = wide ? WIDE_FLAG : 0u;
   const bool lean_wide = a.fast && !copy && !((4.0f * a.dv + 3.0f * fabsf(di) + 3.0f * fabsf(dj)) / dk <= LEAN_SPAN);
   // the tolerance form writes each target cell once; consecutive columns of a column set
   // (sources at least one block apart) then reach distinct target cells when dk <= 4 dv, here
   // with a margin for the running boundary offsets: wider target cells go to the exact kernel
   const bool lean_shared = a.fast && !copy && !(fabsf(dk) <= 3.0f * a.dv);
   const int maxlen = (int)a.len[DIM];
   uint32_t mystat = (copy && !idle) ? VACC_STATUS_DK : 0u;

   for (int r0 = 0; r0 < rows; r0 += rpc) {
      const int nr = min(rpc, rows - r0);
      for (int i = t; i < nr * W; i += CWG) bm[i] = 0;
      __syncthreads();
      for_blocks4<CWG>(a.gid + off, n, t, [&](uint32_t gid, uint32_t) {
         int b0, b1, k;
         pass_coords<DIM>(gid, a.gs, b0, b1, k);
         const int row = (b0 - mn0) * R1 + (b1 - mn1) - r0;
         if (row >= 0 && row < nr) {
            const int kk = k - mnk;
            atomicOr(&bm[row * W + (kk >> 5)], 1u << (kk & 31));
         }
      });
      __syncthreads();
      // per-row counts: blocks, columns, work units, target blocks.  The columns of a set are
      // separate work units unless the target ranges of two of them share a block, in which
      // case the set is one unit (its columns must add into the shared block in order).
      for (int rr = t; rr < nr; rr += CWG) {
         const uint32_t* w = bm + rr * W;
         uint32_t nb = 0;
         for (int q = 0; q < W; ++q) nb += __popc(w[q]);
         uint32_t ncol = 0, ntgt = 0;
         bool overlap = false;
         if (nb) {
            const int grow = r0 + rr;
            const int b0 = mn0 + grow / R1, b1 = mn1 + grow % R1;
            const SetGeom g = set_geom(b0, b1, I, di, dj);
            int end = -1;
            int p = 0;
            for (;;) {
               const int s = next_one(w, W, p);
               if (s < 0) break;
               const int e = next_zero(w, W, s) - 1;
               int fk, lk;
               if (copy) {
                  fk = mnk + s;
                  lk = mnk + e;
               } else {
                  bool wh;
                  column_extent(mnk + s, mnk + e, g, a.dv, a.vmin, dk, maxlen, a.wall_margin, fk, lk, wh);
                  if (wh) mystat |= VACC_STATUS_WALL;
               }
               if (fk <= end) overlap = true;
               if (lk > end) {
                  ntgt += (uint32_t)(lk - max(fk, end + 1) + 1);
                  end = lk;
               }
               ++ncol;
               p = e + 1;
            }
         }
         rowv[0][rr] = nb;
         rowv[1][rr] = ncol;
         rowv[2][rr] = overlap ? 1u : ncol;
         rowv[3][rr] = ntgt;
      }
      __syncthreads();
      // exclusive scan of the four row quantities (contiguous chunk of rows per thread)
      const int per = (nr + CWG - 1) / CWG;
      const int rb = t * per, re = min(nr, rb + per);
      uint32_t s4[4] = {0, 0, 0, 0}, ex[4], tot[4];
      for (int i = rb; i < re; ++i)
#pragma unroll
         for (int q = 0; q < 4; ++q) s4[q] += rowv[q][i];
      block_scan<4>(s4, ex, tot, sh);
#pragma unroll
      for (int q = 0; q < 4; ++q) {
         uint32_t run = ex[q] + carry[q];
         for (int i = rb; i < re; ++i) {
            const uint32_t x = rowv[q][i];
            rowv[q][i] = run;
            run += x;
         }
      }
      __syncthreads();
      // emit set and column records
      for (int rr = t; rr < nr; rr += CWG) {
         const uint32_t* w = bm + rr * W;
         bool any = false;
         for (int q = 0; q < W; ++q) any |= (w[q] != 0);
         if (!any) continue;
         const int grow = r0 + rr;
         const int b0 = mn0 + grow / R1, b1 = mn1 + grow % R1;
         const SetGeom g = set_geom(b0, b1, I, di, dj);
         const uint32_t ui = rowv[2][rr], cb = rowv[1][rr], spos0 = rowv[0][rr], slot0 = rowv[3][rr];
         int end = -1;
         uint32_t cnt = 0, ncol = 0, bpos = 0;
         bool overlap = false;
         int p = 0;
         for (;;) {
            const int s = next_one(w, W, p);
            if (s < 0) break;
            const int e = next_zero(w, W, s) - 1;
            int fk, lk;
            if (copy) {
               fk = mnk + s;
               lk = mnk + e;
            } else {
               bool wh;
               column_extent(mnk + s, mnk + e, g, a.dv, a.vmin, dk, maxlen, a.wall_margin, fk, lk, wh);
            }
            // slot of block fk in the union of target ranges (ranges ascend in start and end)
            const uint32_t first = (fk <= end) ? (cnt - 1u - (uint32_t)(end - fk)) : cnt;
            if (fk <= end) overlap = true;
            if (lk > end) {
               cnt += (uint32_t)(lk - max(fk, end + 1) + 1);
               end = lk;
            }
            const uint32_t len = (uint32_t)(e - s + 1);
            a.col_rec[off + cb + ncol] = make_uint4((uint32_t)(mnk + s) | (len << 16), off + spos0 + bpos,
                                                    (uint32_t)fk | ((uint32_t)lk << 16), slot0 + first);
            bpos += len;
            ++ncol;
            p = e + 1;
         }
         // work units: {b0 | b1 << 16, first column, columns | COPY_FLAG | EXACT_FLAG, source blocks}
         const uint32_t bb = (uint32_t)b0 | ((uint32_t)b1 << 16), cf = copy ? COPY_FLAG : 0u;
         auto exact_col = [&](uint32_t j) {
            return a.fast && !copy &&
                   max_lane_rule_binds(b0, b1, (int)(a.col_rec[off + cb + j].x & 0xFFFFu), I, di, dj, a.dv, a.vmin);
         };
         if (overlap) {
            bool ex = lean_wide || lean_shared;
            for (uint32_t j = 0; j < ncol && !ex; ++j) ex = exact_col(j);
            a.unit_rec[off + ui] = make_uint4(bb, off + cb, ncol | cf | wf | (ex ? EXACT_FLAG : 0u), bpos);
            atomicAdd(&hist[wide ? (uint32_t)WIDE_BIN : ex ? (uint32_t)EXACT_BIN : len_bin(bpos)], 1u);
         } else {
            for (uint32_t j = 0; j < ncol; ++j) {
               const uint32_t len = a.col_rec[off + cb + j].x >> 16;
               const bool ex = lean_wide || exact_col(j);
               a.unit_rec[off + ui + j] = make_uint4(bb, off + cb + j, 1u | cf | wf | (ex ? EXACT_FLAG : 0u), len);
               atomicAdd(&hist[wide ? (uint32_t)WIDE_BIN : ex ? (uint32_t)EXACT_BIN : len_bin(len)], 1u);
            }
         }
      }
      // source index of every block in sorted (set, k) order
      for_blocks4<CWG>(a.gid + off, n, t, [&](uint32_t gid, uint32_t i) {
         int b0, b1, k;
         pass_coords<DIM>(gid, a.gs, b0, b1, k);
         const int row = (b0 - mn0) * R1 + (b1 - mn1) - r0;
         if (row >= 0 && row < nr) {
            const uint32_t pos = rowv[0][row] + rank_in_row(bm + row * W, k - mnk);
            a.sorted_src[off + pos] = a.src[off + i];
         }
      });
      __syncthreads();
      if (t < 4) carry[t] += tot[t];
      __syncthreads();
   }
   if (mystat) atomicOr(&sstat, mystat);
   __syncthreads();
   if (t == 0) {
      a.n_units[c] = carry[2];
      a.n_targets[c] = carry[3];
      if (sstat) {
         atomicOr(&a.status[c], sstat);
         atomicOr(&a.status[a.ncells], sstat);   // OR over all cells
      }
   }
   for (int i = t; i < LEN_BINS; i += CWG) {
      a.cell_hist[(size_t)i * a.ncells + c] = hist[i];
      if (hist[i]) atomicAdd(&a.len_hist[i], hist[i]);
   }
}

// Per length bin (one workgroup each), the exclusive prefix over cells of the cells' unit
// counts, in place: cell c's units of bin b take list positions from hist_cursor[b] +
// cell_hist[b][c] on.
__global__ __launch_bounds__(1024) void k_bin_scan(uint32_t* cell_hist, uint32_t ncells) {
   constexpr uint32_t K = 8;
   __shared__ uint32_t sh[16];
   uint32_t* row = cell_hist + (size_t)blockIdx.x * ncells;
   uint32_t carry = 0;
   for (uint32_t base = 0; base < ncells; base += 1024 * K) {
      const uint32_t i0 = base + threadIdx.x * K;
      uint32_t v[K];
      uint32_t sum[1] = {0u};
#pragma unroll
      for (uint32_t k = 0; k < K; ++k) {
         v[k] = (i0 + k < ncells) ? row[i0 + k] : 0u;
         sum[0] += v[k];
      }
      uint32_t ex[1], tot[1];
      block_scan<1>(sum, ex, tot, sh);
      uint32_t r = carry + ex[0];
#pragma unroll
      for (uint32_t k = 0; k < K; ++k) {
         if (i0 + k < ncells) row[i0 + k] = r;
         r += v[k];
      }
      carry += tot[0];
   }
}

// ---------------------------------------------------------------------------
// k_scan2: exclusive scans of one or two per-cell count arrays
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(1024) void k_scan2(ScanArgs a) {
   // tiles of 1024 x K entries, K contiguous entries per thread (one workgroup scan per tile)
   constexpr uint32_t K = 8;
   __shared__ uint32_t sh[16 * 2];
   uint32_t carry0 = 0, carry1 = 0;   // running totals, the same in every thread
   for (uint32_t base = 0; base < a.n; base += 1024 * K) {
      const uint32_t i0 = base + threadIdx.x * K;
      uint32_t va[K], vb[K];
      uint32_t v[2] = {0u, 0u};
#pragma unroll
      for (uint32_t k = 0; k < K; ++k) {
         const uint32_t i = i0 + k;
         va[k] = (i < a.n) ? a.a[i] : 0u;
         vb[k] = (a.b && i < a.n) ? a.b[i] : 0u;
         v[0] += va[k];
         v[1] += vb[k];
      }
      uint32_t ex[2], tot[2];
      block_scan<2>(v, ex, tot, sh);
      uint32_t r0 = carry0 + ex[0], r1 = carry1 + ex[1];
#pragma unroll
      for (uint32_t k = 0; k < K; ++k) {
         const uint32_t i = i0 + k;
         if (i < a.n) {
            a.a_off[i] = r0;
            if (a.b) a.b_off[i] = r1;
         }
         r0 += va[k];
         r1 += vb[k];
      }
      carry0 += tot[0];
      carry1 += tot[1];
   }
   if (threadIdx.x == 0) {
      a.a_off[a.n] = carry0;
      if (a.b) a.b_off[a.n] = carry1;
      if (a.cap_a && carry0 > a.cap_a && a.flags) atomicOr(a.flags, VACC_STATUS_CAPACITY);
   }
   if (a.hist) {
      // exclusive scan of the work-unit length histogram -> insertion cursors of k_order
      uint32_t v[2] = {threadIdx.x < LEN_BINS ? a.hist[threadIdx.x] : 0u, 0u};
      uint32_t ex[2], tot[2];
      block_scan<2>(v, ex, tot, sh);
      if (threadIdx.x < LEN_BINS) a.hist_cursor[threadIdx.x] = ex[0];
   }
}

// Global work list of the remap: every unit's global index, ordered by source length
// (longest first; within a length bin by cell, units are independent).
__global__ __launch_bounds__(WG) void k_order(OrderArgs a) {
   // one workgroup per cell: the cell's range in each bin is known (hist_cursor + the bin's
   // prefix over cells); units take positions in it through an LDS counter per bin
   __shared__ uint32_t cnt[LEN_BINS];
   if (*a.flags & VACC_STATUS_CAPACITY) return;
   const uint32_t c = blockIdx.x;
   const uint32_t base = a.unit_base[c];
   const uint32_t nu = a.unit_base[c + 1] - base;
   if (nu == 0) return;
   const uint32_t off = a.cell_off[c];
   for (int i = threadIdx.x; i < LEN_BINS; i += WG) cnt[i] = 0;
   __syncthreads();
   for (uint32_t u = threadIdx.x; u < nu; u += WG) {
      const uint32_t b = unit_bin(a.unit_rec[off + u]);
      const uint32_t pos = a.hist_cursor[b] + a.cell_hist[(size_t)b * a.ncells + c] + atomicAdd(&cnt[b], 1u);
      a.perm[pos] = off + u;   // the unit's record
      a.perm_cell[pos] = c;
   }
}

// ---------------------------------------------------------------------------
// k_remap
// ---------------------------------------------------------------------------

// Value e of block b of a pool (any storage), decoded to FP32.
__device__ inline float pool_value(const Pool& pl, uint32_t b, int e) {
   if (pl.storage == ST_F32) return static_cast<const float*>(pl.f)[(size_t)b * WID3 + e];
   return pl.scale[b] * dec16(pl.storage, static_cast<const uint16_t*>(pl.f)[(size_t)b * WID3 + e]);
}

// The same with the storage form fixed at compile time (ST_F32, ST_BF16, ST_F16).
template <int ST>
__device__ inline float pool_value_t(const Pool& pl, uint32_t b, int e) {
   if constexpr (ST == ST_F32) return static_cast<const float*>(pl.f)[(size_t)b * WID3 + e];
   else return pl.scale[b] * Codec<ST>::dec(static_cast<const uint16_t*>(pl.f)[(size_t)b * WID3 + e]);
}

// Row q (cells 4q..4q+3, i = 0..3 at j = q & 3, k = q >> 2) of block b, decoded to FP32.
template <int ST>
__device__ inline float4 pool_row_t(const Pool& pl, uint32_t b, int q) {
   if (b == INVALID) return make_float4(0.0f, 0.0f, 0.0f, 0.0f);
   if constexpr (ST == ST_F32) {
      return reinterpret_cast<const float4*>(pl.f)[(size_t)b * (WID3 / 4) + q];
   } else {
      const uint2 r = reinterpret_cast<const uint2*>(pl.f)[(size_t)b * (WID3 / 4) + q];
      const float sc = pl.scale[b];
      return make_float4(sc * Pack2<ST>::lo(r.x), sc * Pack2<ST>::hi(r.x), sc * Pack2<ST>::lo(r.y), sc * Pack2<ST>::hi(r.y));
   }
}

__device__ inline float4 pool_row(const Pool& pl, uint32_t b, int q) {
   if (b == INVALID) return make_float4(0.0f, 0.0f, 0.0f, 0.0f);
   if (pl.storage == ST_F32) return reinterpret_cast<const float4*>(pl.f)[(size_t)b * (WID3 / 4) + q];
   const uint2 r = reinterpret_cast<const uint2*>(pl.f)[(size_t)b * (WID3 / 4) + q];
   const float sc = pl.scale[b];
   return make_float4(sc * dec16(pl.storage, r.x & 0xFFFFu), sc * dec16(pl.storage, r.x >> 16),
                      sc * dec16(pl.storage, r.y & 0xFFFFu), sc * dec16(pl.storage, r.y >> 16));
}

template <int DIM>
struct Pencil {
   int o;   // offset of the pencil's first cell in a block
   static constexpr int st = (DIM == 0) ? 1 : (DIM == 1) ? 4 : 16;   // stride along the pass axis
};

// Load the lane's pencil (4 cells along the pass axis) of source block sidx, decoded to FP32.
template <int DIM, int ST>
__device__ inline void load_pencil(const Pool& pl, uint32_t sidx, int o, float (&w)[4]) {
   if (sidx == INVALID) {
      w[0] = w[1] = w[2] = w[3] = 0.0f;
      return;
   }
   constexpr int st = Pencil<DIM>::st;
   if constexpr (ST == ST_F32) {
      const float* p = static_cast<const float*>(pl.f) + (size_t)sidx * WID3 + o;
      if (DIM == 0) {
         const float4 q = *reinterpret_cast<const float4*>(p);
         w[0] = q.x; w[1] = q.y; w[2] = q.z; w[3] = q.w;
      } else {
         w[0] = p[0]; w[1] = p[st]; w[2] = p[2 * st]; w[3] = p[3 * st];
      }
   } else {
      const uint16_t* p = static_cast<const uint16_t*>(pl.f) + (size_t)sidx * WID3 + o;
      const float sc = pl.scale[sidx];
      uint32_t u[4];
      if (DIM == 0) {
         const uint2 q = *reinterpret_cast<const uint2*>(p);
         u[0] = q.x & 0xFFFFu; u[1] = q.x >> 16; u[2] = q.y & 0xFFFFu; u[3] = q.y >> 16;
      } else {
         u[0] = p[0]; u[1] = p[st]; u[2] = p[2 * st]; u[3] = p[3 * st];
      }
#pragma unroll
      for (int e = 0; e < 4; ++e) w[e] = sc * Codec<ST>::dec(u[e]);
   }
}

template <int DIM>
__device__ inline void store_pencil(float* __restrict__ f, size_t slot, int o, const float (&w)[4]) {
   float* p = f + slot * WID3 + o;
   if (DIM == 0) {
      *reinterpret_cast<float4*>(p) = make_float4(w[0], w[1], w[2], w[3]);
   } else {
      constexpr int st = Pencil<DIM>::st;
      p[0] = w[0]; p[st] = w[1]; p[2 * st] = w[2]; p[3 * st] = w[3];
   }
}

// The content test (double)x >= mv of an FP32 value x, as x >= content_threshold(mv): the
// smallest float whose double value is >= mv (x >= it exactly when (double)x >= mv).
__device__ inline float content_threshold(double mv) {
   const float f = (float)mv;   // nearest float: the answer, or the float just below mv
   return ((double)f < mv) ? nextafterf(f, INFINITY) : f;
}

__device__ inline void set_content_flag(uint8_t* content, size_t slot, bool has, int setlane0, int l16) {
   const unsigned long long bal = __ballot(has);
   if (l16 == 0) content[slot] = (uint8_t)(((bal >> setlane0) & 0xFFFFull) != 0ull);
}
__device__ inline void set_content(uint8_t* content, size_t slot, const float (&w)[4], float thc, int setlane0, int l16) {
   set_content_flag(content, slot, (w[0] >= thc) || (w[1] >= thc) || (w[2] >= thc) || (w[3] >= thc), setlane0, l16);
}

// Block sums of a stored block held by the 16 lanes of a DPP row (all active): lane values
// v[e] at in-block index e along axis AX and the lane's fixed indices (cx, cy, cz) on the
// other two axes, times `scale`.  Writes (S, sum i f, sum j f, sum k f) to st[slot], lanes 0..3
// one component each.  FP32 sums of the 64 stored values, for the tolerance form's moments and
// removed mass.
template <int AX>
__device__ inline void put_block_stats(float4* st, size_t slot, int l16, const float (&v)[4], int cx, int cy, int cz,
                                       float scale) {
   const float s = (v[0] + v[1]) + (v[2] + v[3]);
   const float e = fmaf(3.0f, v[3], fmaf(2.0f, v[2], v[1]));
   const float wx = (AX == 0) ? e : (float)cx * s;
   const float wy = (AX == 1) ? e : (float)cy * s;
   const float wz = (AX == 2) ? e : (float)cz * s;
   const float r = row_sum4x16(s, wx, wy, wz, l16);
   if (l16 < 4) reinterpret_cast<float*>(st + slot)[l16] = r * scale;
}

// Block sums of a lane's pencil (the FP32 store layout of pass DIM: lane (ip, jp) = (l16 & 3,
// l16 >> 2) holds i = 0..3 at (j, k) = (jp, ip) for DIM 0, j = 0..3 at (i, k) = (ip, jp) for
// DIM 1, k = 0..3 at (i, j) = (ip, jp) for DIM 2).
template <int DIM>
__device__ inline void put_pencil_stats(float4* st, size_t slot, int l16, const float (&v)[4]) {
   const int ip = l16 & 3, jp = l16 >> 2;
   if (DIM == 0) put_block_stats<0>(st, slot, l16, v, 0, jp, ip, 1.0f);
   else if (DIM == 1) put_block_stats<1>(st, slot, l16, v, ip, 0, jp, 1.0f);
   else put_block_stats<2>(st, slot, l16, v, ip, jp, 0, 1.0f);
}

// Block sums of the canonical rows of a 16-bit block (lane L: codes q of i = 0..3 at
// (j, k) = (L & 3, L >> 2)), the stored values being scale * q.
__device__ inline void put_row_stats(float4* st, size_t slot, int l16, const float (&q)[4], float scale) {
   put_block_stats<0>(st, slot, l16, q, 0, l16 & 3, l16 >> 2, scale);
}

// Store one finished target block from the ring (all 16 lanes of the unit active).  FP32: each
// lane stores its own pencil.  16-bit forms: lane L reads the canonical row (j,k) = (L&3, L>>2)
// of the block from the ring columns of the unit's lanes, encodes the block and stores its 4
// codes (8 contiguous bytes) and, lane 0, the scale.  The content flag is taken on the stored
// (decoded) values.  FAST: the block scale's quotient from the hardware reciprocal (the
// tolerance form).  STATS: also the block sums, when a.blk_stats is set.
template <int DIM, int ST, int RN, int WGS, bool FAST = false, bool STATS = true>
__device__ inline void flush_block(const RemapArgs& a, float (*ring)[WGS], int tid, int l16, int setlane0, int o,
                                   int blk, size_t slot, float thc) {
   if constexpr (ST == ST_F32) {
      float q[4];
#pragma unroll
      for (int e = 0; e < 4; ++e) q[e] = ring[(WID * blk + e) & (RN - 1)][tid];
      store_pencil<DIM>(static_cast<float*>(a.out.f), slot, o, q);
      set_content(a.content_out, slot, q, thc, setlane0, l16);
      if (STATS && a.blk_stats) put_pencil_stats<DIM>(a.blk_stats, slot, l16, q);
   } else {
      const int L = l16, base = tid - l16;
      float x[4];
      __builtin_amdgcn_wave_barrier();
#pragma unroll
      for (int i = 0; i < 4; ++i) {
         // lane (ip, jp) of the unit holds f(i=kk, j=jp, k=ip) (dim 0), f(ip, kk, jp) (dim 1),
         // f(ip, jp, kk) (dim 2) in ring row kk
         const int lane = (DIM == 0) ? ((L >> 2) + 4 * (L & 3)) : (DIM == 1) ? (i + 4 * (L >> 2)) : (i + 4 * (L & 3));
         const int row = (DIM == 0) ? i : (DIM == 1) ? (L & 3) : (L >> 2);
         x[i] = ring[(WID * blk + row) & (RN - 1)][base + lane];
      }
      uint2 pk;
      float qmax;
      const float sc = encode_rows16<ST, FAST>(x, pk, qmax);
      uint16_t* p = static_cast<uint16_t*>(a.out.f) + slot * WID3 + 4 * L;
      *reinterpret_cast<uint2*>(p) = pk;
      if (L == 0) a.out.scale[slot] = sc;
      set_content_flag(a.content_out, slot, sc * qmax >= thc, setlane0, l16);
      if (STATS && a.blk_stats) {
         const float d[4] = {Pack2<ST>::lo(pk.x), Pack2<ST>::hi(pk.x), Pack2<ST>::lo(pk.y), Pack2<ST>::hi(pk.y)};
         put_row_stats(a.blk_stats, slot, l16, d, sc);
      }
   }
}

// Store an all-zero target block (a target block that received no contribution).
template <int DIM, int ST, bool STATS = true>
__device__ inline void zero_block(const RemapArgs& a, int l16, int setlane0, int o, size_t slot, float thc) {
   const float q[4] = {0.0f, 0.0f, 0.0f, 0.0f};
   if constexpr (ST == ST_F32) {
      store_pencil<DIM>(static_cast<float*>(a.out.f), slot, o, q);
   } else {
      uint16_t* p = static_cast<uint16_t*>(a.out.f) + slot * WID3 + 4 * l16;
      *reinterpret_cast<uint2*>(p) = make_uint2(0u, 0u);
      if (l16 == 0) a.out.scale[slot] = 0.0f;
   }
   set_content(a.content_out, slot, q, thc, setlane0, l16);
   if (STATS && a.blk_stats && l16 < 4) reinterpret_cast<float*>(a.blk_stats + slot)[l16] = 0.0f;
}

// Pass-through copy of a source block (cells in copy mode keep their stored representation).
template <int DIM, int ST, bool STATS = true>
__device__ inline void copy_block(const RemapArgs& a, int l16, int setlane0, int o, uint32_t sidx, size_t slot,
                                  float thc) {
   if constexpr (ST == ST_F32) {
      float w[4];
      load_pencil<DIM, ST>(a.in, sidx, o, w);
      store_pencil<DIM>(static_cast<float*>(a.out.f), slot, o, w);
      set_content(a.content_out, slot, w, thc, setlane0, l16);
      if (STATS && a.blk_stats) put_pencil_stats<DIM>(a.blk_stats, slot, l16, w);
   } else {
      uint2 raw = make_uint2(0u, 0u);
      float sc = 0.0f;
      if (sidx != INVALID) {
         raw = *reinterpret_cast<const uint2*>(static_cast<const uint16_t*>(a.in.f) + (size_t)sidx * WID3 + 4 * l16);
         sc = a.in.scale[sidx];
      }
      *reinterpret_cast<uint2*>(static_cast<uint16_t*>(a.out.f) + slot * WID3 + 4 * l16) = raw;
      if (l16 == 0) a.out.scale[slot] = sc;
      set_content_flag(a.content_out, slot, sc * max_code<ST>(raw) >= thc, setlane0, l16);
      if (STATS && a.blk_stats) {
         const float d[4] = {Pack2<ST>::lo(raw.x), Pack2<ST>::hi(raw.x), Pack2<ST>::lo(raw.y), Pack2<ST>::hi(raw.y)};
         put_row_stats(a.blk_stats, slot, l16, d, sc);
      }
   }
}

// RN-cell ring per lane in workgroups of WGS lanes: RING / WG for the work list, RING_WIDE /
// WG_WIDE for the units of wide-span cells (range RANGE_WIDE)
template <int DIM, int METHOD, int ST, int RN, int WGS>
__global__ __launch_bounds__(WGS) void k_remap(RemapArgs a, double i_dv) {
   __shared__ float ring[RN][WGS];
   if (*a.flags & VACC_STATUS_CAPACITY) return;
   const int tid = threadIdx.x;
   const int l16 = tid & 15;
   const int ip = l16 & 3, jp = l16 >> 2;
   const int setlane0 = (tid & 63) & ~15;
   const int o = (DIM == 0) ? (ip * WID2 + jp * WID) : (DIM == 1) ? (jp * WID2 + ip) : (jp * WID + ip);
   // the list ends with the EXACT_BIN units and then the WIDE_BIN units
   const uint32_t all = a.unit_base[a.ncells], nwide = a.len_hist[WIDE_BIN];
   const uint32_t first = (a.range == RANGE_WIDE)    ? all - nwide
                          : (a.range == RANGE_EXACT) ? all - nwide - a.len_hist[EXACT_BIN]
                                                     : 0u;
   const uint32_t total = (a.range == RANGE_WIDE) ? all : all - nwide;
   const uint32_t gstride = gridDim.x * (WGS / 16);
   for (uint32_t gi = first + blockIdx.x * (WGS / 16) + (tid >> 4); gi < total; gi += gstride) {
      const uint32_t c = a.perm_cell[gi];
      const uint4 sr = a.unit_rec[a.perm[gi]];
      const int b0 = (int)(sr.x & 0xFFFFu), b1 = (int)(sr.x >> 16);
      const uint32_t col0 = sr.y;
      // the cell's sparsity threshold: the content test, and the reconstruction threshold as
      // the Realf argument of compute_*_coeff (cpu_acc_map.cpp:539-547)
      const double mv = a.minval[c];
      const float thc = content_threshold(mv);
      const float thr = (float)mv;
      const float scale = (float)(1.0 / (double)thr);
      const uint32_t ncol = sr.z & UNIT_NCOL_MASK;
      const bool copy = (sr.z & COPY_FLAG) != 0u;
      const uint32_t obase = a.out_off[c];
      float I, di, dj, dk;
      pass_isect<DIM>(a.isect + 12 * (size_t)c, I, di, dj, dk);
      // Lagrangian base of this lane's pencil, cpu_acc_map.cpp:493-496
      const float Imin = (I + (float)(uint32_t)(b0 * WID + ip) * di) + (float)(uint32_t)(b1 * WID + jp) * dj;
      // Columns of a unit have non-decreasing target ranges that overlap only where one ends
      // and the next begins.  Blocks from the next column's first target block on stay open
      // in the ring across the column boundary (carry_open), so every target block is
      // accumulated in FP32 registers/LDS and stored exactly once.
      int prev_end = -1;
      int carry_open = -1;
      for (uint32_t ci = 0; ci < ncol; ++ci) {
         const uint4 cr = a.col_rec[col0 + ci];
         const int ks = (int)(cr.x & 0xFFFFu), len = (int)(cr.x >> 16);
         const int fk = (int)(cr.z & 0xFFFFu), lk = (int)(cr.z >> 16);
         const int nfk = (ci + 1 < ncol) ? (int)(a.col_rec[col0 + ci + 1].z & 0xFFFFu) : INT_MAX;
         const uint32_t* srcs = a.sorted_src + cr.y;
         const size_t slotb = (size_t)obase + cr.w;   // output slot of block fk
         for (int tb = max(fk, prev_end + 1) + l16; tb <= lk; tb += 16) {
            const size_t slot = slotb + (size_t)(tb - fk);
            a.gid_out[slot] = pass_gid<DIM>(b0, b1, tb, a.len);
            a.src_out[slot] = (uint32_t)slot;
         }
         if (copy) {
            for (int b = 0; b < len; ++b) copy_block<DIM, ST>(a, l16, setlane0, o, srcs[b], slotb + b, thc);
            prev_end = max(prev_end, lk);
            continue;
         }
         float v_r = (float)(uint32_t)(WID * ks) * a.dv + a.vmin;
         const float lvr = (v_r - Imin) / dk;
         int gkr = (int)lvr;
         const float rdk = 1.0f / dk;
         // lanes of min/max Lagrangian index within the 8-lane vector (cpu_acc_map.cpp:515-529)
         int minL = 0, maxL = 0;
         {
            double maxV = DBL_MIN, minV = DBL_MAX;
#pragma unroll
            for (int L = 0; L < 8; ++L) {
               const float x = __shfl(lvr, L, 8);
               if ((double)x > maxV) { maxV = (double)x; maxL = L; }
               if ((double)x < minV) { minV = (double)x; minL = L; }
            }
         }
         float w0[4] = {0.0f, 0.0f, 0.0f, 0.0f}, w1[4], w2[4];
         // source indices are read one block ahead of the block data they locate, so that no
         // block waits for an index load and then for the data load behind it
         uint32_t s_next = (len > 2) ? srcs[2] : INVALID;   // index of block b + 2
         load_pencil<DIM, ST>(a.in, srcs[0], o, w1);
         load_pencil<DIM, ST>(a.in, (len > 1) ? srcs[1] : INVALID, o, w2);
         int next_flush = fk;
         int next_open = (carry_open >= 0) ? carry_open : fk;
         const int lo_gk = WID * fk, hi_gk = WID * lk + WID - 1;
         Recon<METHOD> rc;
         {
            const float v0[12] = {w0[0], w0[1], w0[2], w0[3], w1[0], w1[1], w1[2], w1[3], w2[0], w2[1], w2[2], w2[3]};
            rc.start(v0 + 4, scale);
         }
         for (int b = 0; b < len; ++b) {
            const uint32_t s_after = (b + 3 < len) ? srcs[b + 3] : INVALID;
            float w3[4];
            load_pencil<DIM, ST>(a.in, s_next, o, w3);   // zeros past the column's end
            s_next = s_after;
            const float v[12] = {w0[0], w0[1], w0[2], w0[3], w1[0], w1[1], w1[2], w1[3], w2[0], w2[1], w2[2], w2[3]};
#pragma unroll
            for (int kk = 0; kk < WID; ++kk) {
               float cf[5];
               rc.coeffs(v + 4 + kk, thr, scale, cf);
               const float v_l = v_r;
               v_r = v_r + a.dv;
               const int gkl = gkr;
               gkr = lagr_index(v_r - Imin, dk, rdk);
               const int mnG = max(__shfl(gkl, minL, 8), lo_gk);
               const int mxG = min(__shfl(gkr, maxL, 8), hi_gk);
               // the two 8-lane groups of the unit (a DPP row): rotating by 8 swaps them
               const int flushLim = min(mnG, row_ror_i<8>(mnG));
               const int openLim = max(max(mxG, row_ror_i<8>(mxG)), flushLim);
               while (next_open <= lk && WID * next_open <= openLim) {
                  if (next_open - next_flush >= RN / WID) {   // ring full: cannot happen for checked spans
                     atomicOr(&a.status[c], VACC_STATUS_DK);
                     atomicOr(&a.status[a.ncells], VACC_STATUS_DK);
                     flush_block<DIM, ST, RN, WGS>(a, ring, tid, l16, setlane0, o, next_flush, slotb + (next_flush - fk), thc);
                     ++next_flush;
                  }
#pragma unroll
                  for (int e = 0; e < 4; ++e) ring[(WID * next_open + e) & (RN - 1)][tid] = 0.0f;
                  ++next_open;
               }
               while (next_flush < next_open && next_flush < nfk && WID * next_flush + WID - 1 < flushLim) {
                  flush_block<DIM, ST, RN, WGS>(a, ring, tid, l16, setlane0, o, next_flush, slotb + (next_flush - fk), thc);
                  ++next_flush;
               }
               float tdr = 0.0f;
               for (int gk = mnG; gk <= mxG; ++gk) {
                  const float y = (float)(gk + 1) * dk + Imin;
                  // (y > v_l ? y : v_l), then (mx < v_r ? mx : v_r) (cpu_acc_map.cpp:580-584); no
                  // operand is NaN or a negative zero here, so max/min give the same value
                  const float mx = fmaxf(y, v_l);
                  const float mn = fminf(mx, v_r);
                  const float tdl = tdr;
                  const float vn = (float)((double)(mn - v_l) * i_dv);
                  tdr = integral<METHOD>(vn, cf);
                  float& r = ring[gk & (RN - 1)][tid];
                  r = r + (tdr - tdl);
               }
            }
#pragma unroll
            for (int e = 0; e < 4; ++e) {
               w0[e] = w1[e];
               w1[e] = w2[e];
               w2[e] = w3[e];
            }
         }
         while (next_flush < next_open && next_flush < nfk) {
            flush_block<DIM, ST, RN, WGS>(a, ring, tid, l16, setlane0, o, next_flush, slotb + (next_flush - fk), thc);
            ++next_flush;
         }
         for (int tb = next_open; tb <= lk && tb < nfk; ++tb) zero_block<DIM, ST>(a, l16, setlane0, o, slotb + (tb - fk), thc);
         carry_open = (next_open > nfk) ? next_open : -1;
         prev_end = max(prev_end, lk);
      }
   }
}

// ---------------------------------------------------------------------------
// k_remap_lean: the tolerance form of the remap (options.fast_arith = 1)
// ---------------------------------------------------------------------------
// The work units, column records, target ring and block flushes of k_remap, with the
// arithmetic of vacc_lean.h per source cell: reconstructions in deviations from the cell
// average and the flux-form deposit (one integral per Lagrangian boundary, the cell average
// for the rest of the cell), from the FP64 intersections.  Ring opening and block flushing
// run once per source block on unit-wide bounds (DPP row reductions: a unit is one 16-lane
// row).  Units marked EXACT_FLAG are left to k_remap (range RANGE_EXACT), and those marked
// WIDE_FLAG to its wide instance.

template <int DIM>
__device__ inline void pass_isect_d(const double* is, double& I, double& di, double& dj, double& dk) {
   I = is[4 * DIM + 0];
   const double a = is[4 * DIM + 1], b = is[4 * DIM + 2], c = is[4 * DIM + 3];
   if (DIM == 0) { di = c; dj = b; dk = a; }
   else if (DIM == 1) { di = a; dj = c; dk = b; }
   else { di = a; dj = b; dk = c; }
}

// maximum / minimum over the 16 lanes of the caller's DPP row (all active), in every lane of
// the row (rotations folded into v_max_i32_dpp / v_min_i32_dpp)
__device__ inline int row_max16(int x) {
   x = max(x, row_ror_i<8>(x));
   x = max(x, row_ror_i<4>(x));
   x = max(x, row_ror_i<2>(x));
   return max(x, row_ror_i<1>(x));
}
__device__ inline int row_min16(int x) {
   x = min(x, row_ror_i<8>(x));
   x = min(x, row_ror_i<4>(x));
   x = min(x, row_ror_i<2>(x));
   return min(x, row_ror_i<1>(x));
}

// At least five waves per SIMD (at most 96 VGPRs).  The values that no longer fit are kept in
// scratch across the per-unit setup and the column ends, not in the block loop: 12 to 20
// bytes per lane for FP32 storage (36 to 44 with block sums), 44 to 64 for the 16-bit forms
// (their codec alone would need 111 to 116 VGPRs).  Five waves issue about 7 % more per pass
// than four on the PQM workloads for FP32 and make the 16-bit forms' step 3 % faster; six and
// eight are slower.
template <int DIM, int METHOD, int ST, int WGS, bool STATS>
__global__ __launch_bounds__(WGS, 5) void k_remap_lean(RemapArgs a) {
   using R = ReconLean<METHOD>;
   constexpr int NC = R::NC;
   // rows 0 .. LEAN_RING - 1: the target ring; row LEAN_RING: sink of deposits outside the
   // column's target range
   __shared__ float ring[LEAN_RING + 1][WGS];
   if (*a.flags & VACC_STATUS_CAPACITY) return;
   const int tid = threadIdx.x;
   const int l16 = tid & 15;
   const int ip = l16 & 3, jp = l16 >> 2;
   const int setlane0 = (tid & 63) & ~15;
   const int o = (DIM == 0) ? (ip * WID2 + jp * WID) : (DIM == 1) ? (jp * WID2 + ip) : (jp * WID + ip);
   // the units before EXACT_BIN's; k_remap takes the EXACT_BIN and WIDE_BIN units
   const uint32_t total = a.unit_base[a.ncells] - a.len_hist[EXACT_BIN] - a.len_hist[WIDE_BIN];
   const uint32_t gstride = gridDim.x * (WGS / 16);
   for (uint32_t gi = blockIdx.x * (WGS / 16) + (tid >> 4); gi < total; gi += gstride) {
      const uint32_t c = a.perm_cell[gi];
      const uint4 sr = a.unit_rec[a.perm[gi]];
      const float thc = content_threshold(a.minval[c]);
      const int b0 = (int)(sr.x & 0xFFFFu), b1 = (int)(sr.x >> 16);
      const uint32_t col0 = sr.y;
      const uint32_t ncol = sr.z & UNIT_NCOL_MASK;
      const bool copy = (sr.z & COPY_FLAG) != 0u;
      const uint32_t obase = a.out_off[c];
      double I, di, dj, dk;
      pass_isect_d<DIM>(a.isect + 12 * (size_t)c, I, di, dj, dk);
      // Lagrangian base of this lane's pencil and the boundary spacing in cells
      const double imin = I + (double)(b0 * WID + ip) * di + (double)(b1 * WID + jp) * dj;
      const double rdk = 1.0 / dk;
      const float r = (float)(dk / a.dv_d), inv_r = (float)(a.dv_d * rdk);
      int prev_end = -1;
      int carry_open = -1;
      for (uint32_t ci = 0; ci < ncol; ++ci) {
         const uint4 cr = a.col_rec[col0 + ci];
         const int ks = (int)(cr.x & 0xFFFFu), len = (int)(cr.x >> 16);
         const int fk = (int)(cr.z & 0xFFFFu), lk = (int)(cr.z >> 16);
         const int nfk = (ci + 1 < ncol) ? (int)(a.col_rec[col0 + ci + 1].z & 0xFFFFu) : INT_MAX;
         const uint32_t* srcs = a.sorted_src + cr.y;
         const size_t slotb = (size_t)obase + cr.w;   // output slot of block fk
         for (int tb = max(fk, prev_end + 1) + l16; tb <= lk; tb += 16) {
            const size_t slot = slotb + (size_t)(tb - fk);
            a.gid_out[slot] = pass_gid<DIM>(b0, b1, tb, a.len);
            a.src_out[slot] = (uint32_t)slot;
         }
         if (copy) {
            for (int b = 0; b < len; ++b) copy_block<DIM, ST, STATS>(a, l16, setlane0, o, srcs[b], slotb + b, thc);
            prev_end = max(prev_end, lk);
            continue;
         }
         const int lo = WID * fk, hi = WID * lk + WID - 1;
         // Each target cell ends with its whole mass in one store: the deposit closes a target
         // once (the shares of earlier source cells ride in the running sum), the partial sum a
         // cell without a boundary stores into the open target is overwritten when it closes,
         // and the columns of a unit reach distinct target cells (dk <= 3 dv, checked in the
         // column build).  Ring cells that no deposit reaches keep the zero of their opening.  A target
         // outside the column's range goes to the sink row (no branch).
         auto emit = [&](int gt, float m) {
            const bool in = (uint32_t)(gt - lo) <= (uint32_t)(hi - lo);
            ring[in ? (gt & (LEAN_RING - 1)) : LEAN_RING][tid] = m;
         };
         Flux<NC> fx;
         fx.begin((a.vmin_d + (double)(WID * ks) * a.dv_d - imin) * rdk, r);
         float w0[4] = {0.0f, 0.0f, 0.0f, 0.0f}, w1[4], w2[4];
         // source indices are read one block ahead of the block data they locate, so that no
         // block waits for an index load and then for the data load behind it
         uint32_t s_next = (len > 2) ? srcs[2] : INVALID;   // index of block b + 2
         load_pencil<DIM, ST>(a.in, srcs[0], o, w1);
         load_pencil<DIM, ST>(a.in, (len > 1) ? srcs[1] : INVALID, o, w2);
         int next_flush = fk;
         int next_open = (carry_open >= 0) ? carry_open : fk;
         R rc;
         {
            const float v0[12] = {w0[0], w0[1], w0[2], w0[3], w1[0], w1[1], w1[2], w1[3], w2[0], w2[1], w2[2], w2[3]};
            rc.start(v0 + 4);
         }
         for (int b = 0; b < len; ++b) {
            // targets the block's 4 cells can reach (one cell of margin for rounding), and the
            // lowest target any lane still accumulates
            const int reach = fx.o + 1 + ((fx.t < 4.0f) ? (int)((4.0f - fx.t) * inv_r) + 1 : 0);
            const int openLim = row_max16(min(reach, hi));
            const int flushLim = row_min16(fx.o);
            while (next_open <= lk && WID * next_open <= openLim) {
               if (next_open - next_flush >= LEAN_RING / WID) {   // ring full: not for LEAN_SPAN cells
                  atomicOr(&a.status[c], VACC_STATUS_DK);
                  atomicOr(&a.status[a.ncells], VACC_STATUS_DK);
                  flush_block<DIM, ST, LEAN_RING, WGS, true, STATS>(a, ring, tid, l16, setlane0, o, next_flush, slotb + (next_flush - fk), thc);
                  ++next_flush;
               }
#pragma unroll
               for (int e = 0; e < 4; ++e) ring[(WID * next_open + e) & (LEAN_RING - 1)][tid] = 0.0f;
               ++next_open;
            }
            // Memory order of the iteration: the first cell uses block b + 1, whose load (and the
            // previous iteration's stores) the compiler waits for in full, since the stores vary
            // in number; only then are block b + 2 read and the finished target blocks (all below
            // flushLim, which no deposit of this block reaches) stored, so that cells 1..3 cover
            // their latency.  The scheduling barriers keep that order.
            const float v[12] = {w0[0], w0[1], w0[2], w0[3], w1[0], w1[1], w1[2], w1[3], w2[0], w2[1], w2[2], w2[3]};
            {
               float cf[NC];
               rc.coeffs(v + 4, cf);
               fx.cell(cf, v[4], emit);
            }
            __builtin_amdgcn_sched_barrier(0);
            const uint32_t s_after = (b + 3 < len) ? srcs[b + 3] : INVALID;
            float w3[4];
            load_pencil<DIM, ST>(a.in, s_next, o, w3);   // zeros past the column's end
            s_next = s_after;
            while (next_flush < next_open && next_flush < nfk && WID * next_flush + WID - 1 < flushLim) {
               flush_block<DIM, ST, LEAN_RING, WGS, true, STATS>(a, ring, tid, l16, setlane0, o, next_flush, slotb + (next_flush - fk), thc);
               ++next_flush;
            }
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int kk = 1; kk < WID; ++kk) {
               float cf[NC];
               rc.coeffs(v + 4 + kk, cf);
               fx.cell(cf, v[4 + kk], emit);
            }
#pragma unroll
            for (int e = 0; e < 4; ++e) {
               w0[e] = w1[e];
               w1[e] = w2[e];
               w2[e] = w3[e];
            }
         }
         fx.end(emit);
         while (next_flush < next_open && next_flush < nfk) {
            flush_block<DIM, ST, LEAN_RING, WGS, true, STATS>(a, ring, tid, l16, setlane0, o, next_flush, slotb + (next_flush - fk), thc);
            ++next_flush;
         }
         for (int tb = next_open; tb <= lk && tb < nfk; ++tb) zero_block<DIM, ST, STATS>(a, l16, setlane0, o, slotb + (tb - fk), thc);
         carry_open = (next_open > nfk) ? next_open : -1;
         prev_end = max(prev_end, lk);
      }
   }
}

// ---------------------------------------------------------------------------
// k_adjust: keep = content blocks and their (2w+1)^3 velocity neighbours
// (spatial_cell_cpu.cpp:190-210), or, for w < 0, the existing blocks.  The
// count phase writes the kept count and the mass of removed blocks
// (:227-267); the write phase writes the kept list in ascending GID order.
// ---------------------------------------------------------------------------

// Dilate an LDS bitmap (rows of W words over an Rx x Ry x Rz box, row = z*Ry + y) by one
// block along x, y and z in place: the union of the 3x3x3 neighbourhoods of the set bits,
// clipped to the box.  Applied w times this is the (2w+1)^3 neighbourhood.
__device__ inline void dilate_box1(uint32_t* bm, int W, int Rx, int Ry, int Rz) {
   const int t = threadIdx.x;
   const int rows = Ry * Rz;
   const uint32_t last = ((Rx & 31) == 0) ? 0xFFFFFFFFu : ((1u << (Rx & 31)) - 1u);
   for (int r = t; r < rows; r += (int)blockDim.x) {            // x: one thread per row, words in order
      uint32_t* row = bm + r * W;
      uint32_t prev = 0u, cur = row[0];
      for (int q = 0; q < W; ++q) {
         const uint32_t next = (q + 1 < W) ? row[q + 1] : 0u;
         uint32_t d = cur | (cur << 1) | (cur >> 1) | (prev >> 31) | (next << 31);
         if (q == W - 1) d &= last;
         row[q] = d;
         prev = cur;
         cur = next;
      }
   }
   __syncthreads();
   for (int i = t; i < Rz * W; i += (int)blockDim.x) {           // y: one thread per (z, word) column
      const int z = i / W, q = i - z * W;
      uint32_t* col = bm + (size_t)z * Ry * W + q;
      uint32_t prev = 0u, cur = col[0];
      for (int y = 0; y < Ry; ++y) {
         const uint32_t next = (y + 1 < Ry) ? col[(y + 1) * W] : 0u;
         col[y * W] = prev | cur | next;
         prev = cur;
         cur = next;
      }
   }
   __syncthreads();
   for (int i = t; i < Ry * W; i += (int)blockDim.x) {           // z: one thread per (y, word) column
      uint32_t* col = bm + i;
      const int stride = Ry * W;
      uint32_t prev = 0u, cur = col[0];
      for (int z = 0; z < Rz; ++z) {
         const uint32_t next = (z + 1 < Rz) ? col[(z + 1) * stride] : 0u;
         col[z * stride] = prev | cur | next;
         prev = cur;
         cur = next;
      }
   }
   __syncthreads();
}

template <bool WRITE, int ST>
__global__ __launch_bounds__(AWG) void k_adjust(AdjustArgs a) {
   __shared__ uint32_t bm[BM_WORDS];
   __shared__ uint32_t wpre[WRITE ? BM_WORDS : 1];
   __shared__ int sbox[6];
   __shared__ uint32_t sh[AWG / 64];
   __shared__ uint32_t carry;
   __shared__ double wloss[AWG / 64];
   if (WRITE && (*a.flags & VACC_STATUS_CAPACITY)) return;
   const int t = threadIdx.x;
   const uint32_t c = blockIdx.x;
   const uint32_t off = a.cell_off[c];
   const uint32_t n = a.cell_off[c + 1] - off;
   const int w = a.add_width;
   const int ww = (w > 0) ? w : 0;
   const uint32_t nx = a.len[0], ny = a.len[1], nz = a.len[2];
   const uint32_t ntot = nx * ny * nz;
   // content GIDs of the spatial neighbours (spatial_cell_cpu.cpp:212-222): content blocks of
   // neighbour cells of this context, and caller-supplied GIDs; kept as they are
   const bool spatial = (w >= 0) && (a.nbr_off || a.ext_off);
   auto for_spatial = [&](auto&& fn) {
      if (!spatial) return;
      if (a.nbr_off)
         for (uint32_t k = a.nbr_off[c]; k < a.nbr_off[c + 1]; ++k) {
            const uint32_t o = a.cell_off[a.nbr[k]], m = a.cell_off[a.nbr[k] + 1] - o;
            for (uint32_t i = t; i < m; i += AWG)
               if (a.content[o + i]) fn(a.gid[o + i]);
         }
      if (a.ext_off)
         for (uint64_t i = a.ext_off[c] + t; i < a.ext_off[c + 1]; i += AWG) {
            const uint32_t g = a.ext_gid[i];
            if (g < ntot) fn(g);
         }
   };
   bool any_spatial = false;
   if (spatial) {
      if (a.ext_off && a.ext_off[c + 1] > a.ext_off[c]) any_spatial = true;
      if (a.nbr_off)
         for (uint32_t k = a.nbr_off[c]; k < a.nbr_off[c + 1]; ++k)
            if (a.cell_off[a.nbr[k] + 1] > a.cell_off[a.nbr[k]]) any_spatial = true;
   }
   if (n == 0 && !any_spatial) {
      if (!WRITE && t == 0) {
         a.n_keep[c] = 0;
         if (a.rho_loss) a.rho_loss[c] = 0.0;
         if (a.keep_box) a.keep_box[8 * (size_t)c + 6] = 0;
      }
      return;
   }
   // the count phase leaves the keep bitmap and box of a cell that it built in one chunk for
   // the write phase (keep_box[8c + 6] = 1), which then reads them instead of building them again
   int* const kbox = a.keep_box ? a.keep_box + 8 * (size_t)c : nullptr;
   const bool reuse = WRITE && kbox && kbox[6] != 0;
   if (t < 3) {
      sbox[t] = reuse ? kbox[t] : INT_MAX;
      sbox[t + 3] = reuse ? kbox[t + 3] : INT_MIN;
   }
   if (t == 0) carry = 0;
   __syncthreads();
   if (!reuse) {
      int lx = INT_MAX, ly = INT_MAX, lz = INT_MAX, hx = INT_MIN, hy = INT_MIN, hz = INT_MIN;
      auto grow = [&](uint32_t g) {
         uint32_t ux, uy, uz;
         gid_xyz(g, a.gs, ux, uy, uz);
         const int x = (int)ux, y = (int)uy, z = (int)uz;
         lx = min(lx, x); hx = max(hx, x);
         ly = min(ly, y); hy = max(hy, y);
         lz = min(lz, z); hz = max(hz, z);
      };
      for_blocks4<AWG>(a.gid + off, n, t, [&](uint32_t g, uint32_t) { grow(g); });
      for_spatial(grow);
      atomicMin(&sbox[0], lx); atomicMin(&sbox[1], ly); atomicMin(&sbox[2], lz);
      atomicMax(&sbox[3], hx); atomicMax(&sbox[4], hy); atomicMax(&sbox[5], hz);
   }
   __syncthreads();
   if (sbox[0] == INT_MAX) {   // no block of its own and no neighbour content: nothing to keep
      if (!WRITE && t == 0) {
         a.n_keep[c] = 0;
         if (a.rho_loss) a.rho_loss[c] = 0.0;
         if (kbox) kbox[6] = 0;
      }
      return;
   }
   const int mnx = max(sbox[0] - ww, 0), mny = max(sbox[1] - ww, 0), mnz = max(sbox[2] - ww, 0);
   const int mxx = min(sbox[3] + ww, (int)nx - 1), mxy = min(sbox[4] + ww, (int)ny - 1), mxz = min(sbox[5] + ww, (int)nz - 1);
   const int Rx = mxx - mnx + 1, Ry = mxy - mny + 1, Rz = mxz - mnz + 1;
   const int W = (Rx + 31) >> 5;
   const int rows = Ry * Rz;
   const int rpc = BM_WORDS / W;
   // one chunk holds the whole box: mark content blocks, then dilate separably in LDS
   const bool separable = (w >= 0) && (rows <= rpc);
   double lsum = 0.0;      // removed f, summed per lane (count phase)
   for (int r0 = 0; r0 < rows; r0 += rpc) {
      const int nr = min(rpc, rows - r0);
      if (reuse) {   // one chunk (the count phase found the box separable)
         const uint32_t* src = a.keep_bm + (size_t)c * BM_WORDS;
         for (int i = t; i < nr * W; i += AWG) bm[i] = src[i];
         __syncthreads();
      } else {
      for (int i = t; i < nr * W; i += AWG) bm[i] = 0;
      __syncthreads();
      if (separable) {
         for_blocks4<AWG>(a.gid + off, n, t, [&](uint32_t g, uint32_t i) {
            if (!a.content[off + i]) return;
            uint32_t ux, uy, uz;
            gid_xyz(g, a.gs, ux, uy, uz);
            const int x = (int)ux, y = (int)uy, z = (int)uz;
            const int kx = x - mnx;
            atomicOr(&bm[((z - mnz) * Ry + (y - mny)) * W + (kx >> 5)], 1u << (kx & 31));
         });
         __syncthreads();
         for (int k = 0; k < w; ++k) dilate_box1(bm, W, Rx, Ry, Rz);
         for_spatial([&](uint32_t g) {
            uint32_t ux, uy, uz;
            gid_xyz(g, a.gs, ux, uy, uz);
            const int x = (int)ux, y = (int)uy, z = (int)uz;
            const int kx = x - mnx;
            atomicOr(&bm[((z - mnz) * Ry + (y - mny)) * W + (kx >> 5)], 1u << (kx & 31));
         });
         __syncthreads();
         if (!WRITE && kbox) {   // the finished keep bitmap of the whole box, for the write phase
            uint32_t* dst = a.keep_bm + (size_t)c * BM_WORDS;
            for (int i = t; i < nr * W; i += AWG) dst[i] = bm[i];
         }
      } else {
      for_spatial([&](uint32_t g) {
         uint32_t ux, uy, uz;
         gid_xyz(g, a.gs, ux, uy, uz);
         const int x = (int)ux, y = (int)uy, z = (int)uz;
         const int row = (z - mnz) * Ry + (y - mny) - r0;
         if (row >= 0 && row < nr) {
            const int kx = x - mnx;
            atomicOr(&bm[row * W + (kx >> 5)], 1u << (kx & 31));
         }
      });
      for (uint32_t i = t; i < n; i += AWG) {
         const uint32_t g = a.gid[off + i];
         uint32_t ux, uy, uz;
         gid_xyz(g, a.gs, ux, uy, uz);
         const int x = (int)ux, y = (int)uy, z = (int)uz;
         if (w < 0) {
            const int row = (z - mnz) * Ry + (y - mny) - r0;
            if (row >= 0 && row < nr) {
               const int kx = x - mnx;
               atomicOr(&bm[row * W + (kx >> 5)], 1u << (kx & 31));
            }
         } else if (a.content[off + i]) {
            const int x0 = max(x - w, mnx), x1 = min(x + w, mxx);
            for (int dz = -w; dz <= w; ++dz) {
               const int zz = z + dz;
               if (zz < mnz || zz > mxz) continue;
               for (int dy = -w; dy <= w; ++dy) {
                  const int yy = y + dy;
                  if (yy < mny || yy > mxy) continue;
                  const int row = (zz - mnz) * Ry + (yy - mny) - r0;
                  if (row < 0 || row >= nr) continue;
                  for (int xx = x0; xx <= x1; ++xx) {
                     const int kx = xx - mnx;
                     atomicOr(&bm[row * W + (kx >> 5)], 1u << (kx & 31));
                  }
               }
            }
         }
      }
      }
      }
      __syncthreads();
      // word prefix counts over the chunk (contiguous words per thread)
      const int nwords = nr * W;
      const int per = (nwords + AWG - 1) / AWG;
      const int wb = t * per, we = min(nwords, wb + per);
      uint32_t s1[1] = {0}, ex1[1], tot1[1];
      for (int q = wb; q < we; ++q) s1[0] += __popc(bm[q]);
      block_scan<1>(s1, ex1, tot1, sh);
      if (WRITE) {
         uint32_t run = ex1[0] + carry;
         for (int q = wb; q < we; ++q) {
            wpre[q] = run;
            run += __popc(bm[q]);
         }
         __syncthreads();
         const uint32_t obase = a.out_off[c];
         for (int q = t; q < nwords; q += AWG) {
            uint32_t x = bm[q];
            uint32_t pos = wpre[q];
            const int row = q / W + r0;
            const int zz = mnz + row / Ry, yy = mny + row % Ry;
            const int xb = mnx + (q % W) * 32;
            while (x) {
               const int b = __ffs(x) - 1;
               x &= x - 1;
               a.gid_out[obase + pos] = (uint32_t)(xb + b) + (uint32_t)yy * nx + (uint32_t)zz * nx * ny;
               a.src_out[obase + pos] = INVALID;
               ++pos;
            }
         }
         __syncthreads();
         for_blocks4<AWG>(a.gid + off, n, t, [&](uint32_t g, uint32_t i) {
            uint32_t ux, uy, uz;
            gid_xyz(g, a.gs, ux, uy, uz);
            const int x = (int)ux, y = (int)uy, z = (int)uz;
            const int row = (z - mnz) * Ry + (y - mny) - r0;
            if (row < 0 || row >= nr) return;
            const int kx = x - mnx;
            const int q = row * W + (kx >> 5);
            const uint32_t bit = 1u << (kx & 31);
            if (bm[q] & bit) {
               const uint32_t pos = wpre[q] + __popc(bm[q] & (bit - 1u));
               a.src_out[obase + pos] = a.src[off + i];
            }
         });
      } else {
         // removed blocks: no content and not kept; their mass goes to RHOLOSSADJUST
         // (spatial_cell_cpu.cpp:227-267).  Each wave takes 64 consecutive list entries, one
         // per lane, ballots the removed ones and then sums them block by block in lane order,
         // all 64 lanes reading one block (coalesced) into per-lane FP64 sums: the same blocks
         // reach the same lanes in the same order on every run.
         if (w >= 0) {
            const int lane = t & 63;
            for (uint32_t i0 = (uint32_t)(t - lane); i0 < n; i0 += AWG) {
               const uint32_t i = i0 + (uint32_t)lane;
               uint32_t sidx = INVALID;
               if (i < n && !a.content[off + i]) {
                  uint32_t ux, uy, uz;
                  gid_xyz(a.gid[off + i], a.gs, ux, uy, uz);
                  const int row = ((int)uz - mnz) * Ry + ((int)uy - mny) - r0;
                  const int kx = (int)ux - mnx;
                  if (row >= 0 && row < nr && !(bm[row * W + (kx >> 5)] & (1u << (kx & 31)))) sidx = a.src[off + i];
               }
               if (a.blk_stats) {   // the block's sum from the remap's store
                  if (sidx != INVALID) lsum += (double)a.blk_stats[sidx].x;
                  continue;
               }
               unsigned long long bal = __ballot(sidx != INVALID);
               while (bal) {   // up to 8 removed blocks per batch, their loads in flight together
                  constexpr int B = 8;
                  uint32_t sb[B];
#pragma unroll
                  for (int u = 0; u < B; ++u) {
                     sb[u] = INVALID;
                     if (bal) {
                        sb[u] = (uint32_t)__builtin_amdgcn_readlane((int)sidx, __ffsll((long long)bal) - 1);
                        bal &= bal - 1ull;
                     }
                  }
                  // unconditional loads (an empty slot reads pool entry 0, which exists when a
                  // removed block does, and is discarded) so that all of them are in flight
                  float v[B];
#pragma unroll
                  for (int u = 0; u < B; ++u) v[u] = pool_value_t<ST>(a.pool, (sb[u] == INVALID) ? 0u : sb[u], lane);
#pragma unroll
                  for (int u = 0; u < B; ++u) v[u] = (sb[u] == INVALID) ? 0.0f : v[u];
#pragma unroll
                  for (int u = 0; u < B; ++u) lsum += (double)v[u];
               }
            }
         }
      }
      __syncthreads();
      if (t == 0) carry += tot1[0];
      __syncthreads();
   }
   if (!WRITE) {
      // fixed-order reduction (wave tree, then waves in order): the loss is reproducible
      double v = a.dv3 * lsum;
      for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
      if ((t & 63) == 0) wloss[t >> 6] = v;
      __syncthreads();
      if (t == 0) {
         double sum = 0.0;
         for (int q = 0; q < AWG / 64; ++q) sum += wloss[q];
         a.n_keep[c] = carry;
         if (a.rho_loss) a.rho_loss[c] = sum;
         if (kbox) {
            for (int q = 0; q < 6; ++q) kbox[q] = sbox[q];
            kbox[6] = separable ? 1 : 0;
         }
      }
   }
}

// content flag of each block of a list, from f (spatial_cell_cpu.cpp:290-316)
// One wave per block, grid-stride: a dispatch holds fewer than 2^32 work-items, so the grid is
// capped (launch_waves) and each wave walks blocks wave, wave + nwaves, ...
__global__ __launch_bounds__(WG) void k_content(const uint32_t* src, Pool pl, uint8_t* content,
                                                uint32_t n, double mv) {
   const int lane = threadIdx.x & 63;
   const uint32_t nw = gridDim.x * (WG / 64);
   for (uint32_t wave = blockIdx.x * (WG / 64) + (threadIdx.x >> 6); wave < n; wave += nw) {
      const uint32_t s = src[wave];
      const float x = (s == INVALID) ? 0.0f : pool_value(pl, s, lane);
      const unsigned long long b = __ballot((double)x >= mv);
      if (lane == 0) content[wave] = (uint8_t)(b != 0ull);
   }
}

// Second moments of one population about a pivot (arch_moments.h:123-150): one workgroup per
// cell, one lane per velocity cell of a block, FP64 sums; DV3 and the mass applied once.
__global__ __launch_bounds__(WG) void k_pressure(MomentArgs a, const double* pivot, double mass, double* out) {
   __shared__ double red[6][WG / 64];
   const int t = threadIdx.x, lane = t & 63, w = t >> 6;
   const uint32_t c = blockIdx.x;
   const uint32_t off = a.cell_off[c], n = a.cell_off[c + 1] - off;
   const int i = lane & 3, j = (lane >> 2) & 3, k = lane >> 4;
   const double px = pivot[3 * (size_t)c], py = pivot[3 * (size_t)c + 1], pz = pivot[3 * (size_t)c + 2];
   const uint32_t nx = a.len[0], nxy = a.len[0] * a.len[1];
   double s[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
   for (uint32_t b = w; b < n; b += WG / 64) {
      const uint32_t sidx = a.src[off + b];
      if (sidx == INVALID) continue;
      const double f = pool_value(a.pool, sidx, lane);
      const uint32_t g = a.gid[off + b];
      uint32_t x, y, z;
      gid_xyz(g, a.gs, x, y, z);
      const double dx = (a.vmin[0] + (x * WID + i + 0.5) * a.cell[0]) - px;
      const double dy = (a.vmin[1] + (y * WID + j + 0.5) * a.cell[1]) - py;
      const double dz = (a.vmin[2] + (z * WID + k + 0.5) * a.cell[2]) - pz;
      s[0] += f * dx * dx;
      s[1] += f * dy * dy;
      s[2] += f * dz * dz;
      s[3] += f * dy * dz;
      s[4] += f * dx * dz;
      s[5] += f * dx * dy;
   }
#pragma unroll
   for (int q = 0; q < 6; ++q) {
      for (int o = 32; o > 0; o >>= 1) s[q] += __shfl_down(s[q], o, 64);
      if (lane == 0) red[q][w] = s[q];
   }
   __syncthreads();
   if (t < 6) {
      const double DV3 = a.cell[0] * a.cell[1] * a.cell[2];
      double sum = 0.0;
      for (int q = 0; q < WG / 64; ++q) sum += red[t][q];
      out[6 * (size_t)c + t] = mass * (sum * DV3);
   }
}

// spatial_cell_cpu.cpp:699-727
__global__ void k_sparse_min(SparseArgs p, const double* moments, const uint32_t* off, double* minval,
                             uint32_t ncells) {
   const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
   if (c >= ncells) return;
   double v = p.min_value;
   if (p.algorithm == 1 || p.algorithm == 2) {
      const double k = (p.min2 - p.min1) / (p.bulk2 - p.bulk1);
      const double b = p.min1 - k * p.bulk1;
      const double x = (p.algorithm == 1) ? moments[4 * (size_t)c] : (double)(off[c + 1] - off[c]);
      const double y = k * x + b;
      v = (y < p.min1) ? p.min1 : (y > p.min2) ? p.min2 : y;
   }
   minval[c] = v;
}

// content flags with the cell's threshold: one workgroup per cell, one wave per block
__global__ __launch_bounds__(WG) void k_content_cells(const uint32_t* off, const uint32_t* src, Pool pl,
                                                      const double* minval, uint8_t* content) {
   const uint32_t c = blockIdx.x;
   const int lane = threadIdx.x & 63;
   const double mv = minval[c];
   for (uint32_t b = off[c] + (threadIdx.x >> 6); b < off[c + 1]; b += WG / 64) {
      const uint32_t s = src[b];
      const float x = (s == INVALID) ? 0.0f : pool_value(pl, s, lane);
      const unsigned long long bal = __ballot((double)x >= mv);
      if (lane == 0) content[b] = (uint8_t)(bal != 0ull);
   }
}

// FP64 sum of each cell's f: one workgroup per cell, one wave per block
__global__ __launch_bounds__(WG) void k_cell_sums(const uint32_t* off, const uint32_t* src, Pool pl, double* out) {
   __shared__ double part[WG / 64];
   const uint32_t c = blockIdx.x;
   const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
   double acc = 0.0;
   for (uint32_t b = off[c] + w; b < off[c + 1]; b += WG / 64) {
      const uint32_t s = src[b];
      acc += (s == INVALID) ? 0.0 : (double)pool_value(pl, s, lane);
   }
   for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
   if (lane == 0) part[w] = acc;
   __syncthreads();
   if (threadIdx.x == 0) {
      double t = 0.0;
      for (int q = 0; q < WG / 64; ++q) t += part[q];
      out[c] = t;
   }
}

// f *= pre / post per cell (block_adjust_cpu.cpp:108-112); one workgroup per cell
__global__ __launch_bounds__(WG) void k_cell_scale(const uint32_t* off, const uint32_t* src, Pool pl,
                                                   const double* pre, const double* post) {
   const uint32_t c = blockIdx.x;
   if (post[c] == 0.0) return;
   const double r = pre[c] / post[c];
   const int lane = threadIdx.x & 63;
   for (uint32_t b = off[c] + (threadIdx.x >> 6); b < off[c + 1]; b += WG / 64) {
      const uint32_t s = src[b];
      if (s == INVALID) continue;
      if (pl.storage == ST_F32) {
         float* f = static_cast<float*>(pl.f) + (size_t)s * WID3 + lane;
         *f = (float)((double)*f * r);
      } else if (lane == 0) {
         pl.scale[s] = (float)((double)pl.scale[s] * r);
      }
   }
}

// ---------------------------------------------------------------------------
// moments, transform, gather
// ---------------------------------------------------------------------------

template <int ST>
__global__ __launch_bounds__(WG) void k_moments(MomentArgs a) {
   // One workgroup per cell; 16 lanes per block, each lane one float4 (i = 0..3 at fixed j, k),
   // UNR blocks per 16-lane group in flight.  Per value the reference adds f*DV3, f*VX*DV3, ...
   // (arch_moments.h:98-108); here DV3 is applied once to the cell's FP64 sums.  A lane keeps
   // per row the sum S and the index-weighted sum W = f1 + 2 f2 + 3 f3, and accumulates S, W and
   // S times the block's corner velocity; its fixed offsets within the block (x: 0.5 dv and W dv,
   // y, z: (j, k) + 0.5 cells) are applied once at the end.
   __shared__ double red[4][WG / 64];
   const int t = threadIdx.x;
   const uint32_t c = blockIdx.x;
   const uint32_t off = a.cell_off[c];
   const uint32_t n = a.cell_off[c + 1] - off;
   const int q = t & 15;
   const int j = q & 3, k = q >> 2;
   const double bx = a.cell[0] * WID, by = a.cell[1] * WID, bz = a.cell[2] * WID;
   double s0 = 0.0, sw = 0.0, sx = 0.0, sy = 0.0, sz = 0.0;
   constexpr uint32_t GROUPS = WG / 16;
   constexpr int UNR = 4;
   for (uint32_t b = t >> 4; b < n; b += UNR * GROUPS) {
      uint32_t sv[UNR], gv[UNR];
      float4 v[UNR];
#pragma unroll
      for (int h = 0; h < UNR; ++h) {
         const uint32_t bb = b + h * GROUPS;
         sv[h] = (bb < n) ? a.src[off + bb] : INVALID;
         gv[h] = (bb < n) ? a.gid[off + bb] : 0u;
      }
#pragma unroll
      for (int h = 0; h < UNR; ++h) v[h] = pool_row_t<ST>(a.pool, sv[h], q);
#pragma unroll
      for (int h = 0; h < UNR; ++h) {
         uint32_t x, y, z;
         gid_xyz(gv[h], a.gs, x, y, z);
         const double v0 = v[h].x, v1 = v[h].y, v2 = v[h].z, v3 = v[h].w;
         const double S = (v0 + v1) + (v2 + v3);
         const double Wt = fma(3.0, v3, fma(2.0, v2, v1));
         s0 += S;
         sw += Wt;
         sx = fma(S, fma((double)x, bx, a.vmin[0]), sx);
         sy = fma(S, fma((double)y, by, a.vmin[1]), sy);
         sz = fma(S, fma((double)z, bz, a.vmin[2]), sz);
      }
   }
   double s1 = sx + (0.5 * a.cell[0]) * s0 + a.cell[0] * sw;
   double s2 = sy + ((j + 0.5) * a.cell[1]) * s0;
   double s3 = sz + ((k + 0.5) * a.cell[2]) * s0;
   for (int o = 32; o > 0; o >>= 1) {
      s0 += __shfl_down(s0, o, 64);
      s1 += __shfl_down(s1, o, 64);
      s2 += __shfl_down(s2, o, 64);
      s3 += __shfl_down(s3, o, 64);
   }
   if ((t & 63) == 0) {
      red[0][t >> 6] = s0;
      red[1][t >> 6] = s1;
      red[2][t >> 6] = s2;
      red[3][t >> 6] = s3;
   }
   __syncthreads();
   if (t < 4) {
      const double DV3 = a.cell[0] * a.cell[1] * a.cell[2];
      double sum = 0.0;
      for (int w = 0; w < WG / 64; ++w) sum += red[t][w];
      a.out[4 * (size_t)c + t] = sum * DV3;
   }
}

// The same moments from the block sums of the last remap pass (MomentArgs::blk_stats): one
// workgroup per cell, one lane per block, FP64 sums of S, S times the block's corner velocity
// and the in-block index sums; the 0.5-cell offsets and DV3 applied once per cell.
__global__ __launch_bounds__(WG) void k_moments_sums(MomentArgs a) {
   __shared__ double red[4][WG / 64];
   const int t = threadIdx.x;
   const uint32_t c = blockIdx.x;
   const uint32_t off = a.cell_off[c];
   const uint32_t n = a.cell_off[c + 1] - off;
   const double bx = a.cell[0] * WID, by = a.cell[1] * WID, bz = a.cell[2] * WID;
   double s0 = 0.0, sx = 0.0, sy = 0.0, sz = 0.0, wx = 0.0, wy = 0.0, wz = 0.0;
   constexpr int UNR = 4;   // blocks per lane in flight
   for (uint32_t b0 = t; b0 < n; b0 += UNR * WG) {
      uint32_t sidx[UNR], g[UNR];
#pragma unroll
      for (int h = 0; h < UNR; ++h) {
         const uint32_t b = b0 + h * WG;
         sidx[h] = (b < n) ? a.src[off + b] : INVALID;
         g[h] = (b < n) ? a.gid[off + b] : 0u;
      }
      float4 st[UNR];
#pragma unroll
      for (int h = 0; h < UNR; ++h)   // a block the adjust created (INVALID) is all zero
         st[h] = (sidx[h] != INVALID) ? a.blk_stats[sidx[h]] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
#pragma unroll
      for (int h = 0; h < UNR; ++h) {
         uint32_t x, y, z;
         gid_xyz(g[h], a.gs, x, y, z);
         const double S = st[h].x;
         s0 += S;
         sx = fma(S, fma((double)x, bx, a.vmin[0]), sx);
         sy = fma(S, fma((double)y, by, a.vmin[1]), sy);
         sz = fma(S, fma((double)z, bz, a.vmin[2]), sz);
         wx += (double)st[h].y;
         wy += (double)st[h].z;
         wz += (double)st[h].w;
      }
   }
   double s1 = sx + (0.5 * a.cell[0]) * s0 + a.cell[0] * wx;
   double s2 = sy + (0.5 * a.cell[1]) * s0 + a.cell[1] * wy;
   double s3 = sz + (0.5 * a.cell[2]) * s0 + a.cell[2] * wz;
   for (int o = 32; o > 0; o >>= 1) {
      s0 += __shfl_down(s0, o, 64);
      s1 += __shfl_down(s1, o, 64);
      s2 += __shfl_down(s2, o, 64);
      s3 += __shfl_down(s3, o, 64);
   }
   if ((t & 63) == 0) {
      red[0][t >> 6] = s0;
      red[1][t >> 6] = s1;
      red[2][t >> 6] = s2;
      red[3][t >> 6] = s3;
   }
   __syncthreads();
   if (t < 4) {
      const double DV3 = a.cell[0] * a.cell[1] * a.cell[2];
      double sum = 0.0;
      for (int w = 0; w < WG / 64; ++w) sum += red[t][w];
      a.out[4 * (size_t)c + t] = sum * DV3;
   }
}

__global__ void k_transform(MeshInfo m, const TransformIn* fields, const double* dt_override, const double* moments,
                            int use_moments, double mass, double charge, double hall_min_rhoq, int ohm_gradpe,
                            int map_order, double* isect_out, uint32_t ncells) {
   const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
   if (c >= ncells) return;
   TransformIn in = fields[c];
   if (dt_override) in.dt = dt_override[c];
   if (use_moments == 1) {
      // one population: VX_V = (nVx * m) / (n * m), RHOQ_V = n * q (arch_moments.cpp:419-440)
      const double* mo = moments + 4 * (size_t)c;
      const double rhom = mo[0] * mass;
      if (rhom != 0.0) {
         in.V[0] = (mo[1] * mass) / rhom;
         in.V[1] = (mo[2] * mass) / rhom;
         in.V[2] = (mo[3] * mass) / rhom;
      }
      in.rhoq = mo[0] * charge;
   } else if (use_moments == 2) {
      // plasma pivot of several populations (k_plasma): {VX_V, VY_V, VZ_V, RHOQ_V}
      const double* pv = moments + 4 * (size_t)c;
      in.V[0] = pv[0];
      in.V[1] = pv[1];
      in.V[2] = pv[2];
      in.rhoq = pv[3];
   }
   const Affine fwd = acceleration_transform(in, mass, charge, hall_min_rhoq, ohm_gradpe);
   cell_intersections(m, fwd, map_order, isect_out + 12 * (size_t)c);
}

// Plasma moments of several populations (arch_moments.cpp:406-440): RHOM_V = sum n*m,
// V_V = sum (nV*m) / RHOM_V (0 where RHOM_V = 0), RHOQ_V = sum n*q, summed in population order.
__global__ void k_plasma(PlasmaArgs a) {
   const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
   if (c >= a.ncells) return;
   double rhom = 0.0, vx = 0.0, vy = 0.0, vz = 0.0, rhoq = 0.0;
   for (int p = 0; p < a.npop; ++p) {
      const double* mo = a.mom[p] + 4 * (size_t)c;
      rhom += mo[0] * a.mass[p];
      vx += mo[1] * a.mass[p];
      vy += mo[2] * a.mass[p];
      vz += mo[3] * a.mass[p];
      rhoq += mo[0] * a.charge[p];
   }
   double* o = a.out + 4 * (size_t)c;
   o[0] = (rhom != 0.0) ? vx / rhom : 0.0;
   o[1] = (rhom != 0.0) ? vy / rhom : 0.0;
   o[2] = (rhom != 0.0) ? vz / rhom : 0.0;
   o[3] = rhoq;
}

void launch_plasma(const PlasmaArgs& a, hipStream_t s) {
   if (a.ncells == 0) return;
   hipLaunchKernelGGL(k_plasma, dim3((a.ncells + 255) / 256), dim3(256), 0, s, a);
}

// 16 lanes per block, grid-stride (capped grid, see launch_gather)
__global__ void k_gather(const uint32_t* src, const float4* f_in, float4* f_out, uint32_t n) {
   const size_t stride = (size_t)gridDim.x * blockDim.x;
   for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < (size_t)n * 16; q += stride) {
      const uint32_t b = (uint32_t)(q >> 4);
      const uint32_t s = src[b];
      f_out[q] = (s == INVALID) ? make_float4(0.f, 0.f, 0.f, 0.f) : f_in[(size_t)s * 16 + (q & 15)];
   }
}

// 16-bit pools: 16 lanes per block, 8 bytes of codes each, lane 0 the scale.
__global__ void k_gather16(const uint32_t* src, const uint2* q_in, const float* s_in, uint2* q_out, float* s_out,
                           uint32_t n) {
   const size_t stride = (size_t)gridDim.x * blockDim.x;
   for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < (size_t)n * 16; q += stride) {
      const uint32_t b = (uint32_t)(q >> 4);
      const uint32_t s = src[b];
      q_out[q] = (s == INVALID) ? make_uint2(0u, 0u) : q_in[(size_t)s * 16 + (q & 15)];
      if ((q & 15) == 0) s_out[b] = (s == INVALID) ? 0.0f : s_in[s];
   }
}

// FP32 blocks -> 16-bit pool, one wave per block.
template <int ST>
__global__ __launch_bounds__(WG) void k_encode(const float* f, uint16_t* q, float* scale, uint32_t n) {
   const int e = threadIdx.x & 63;
   const uint32_t nw = gridDim.x * (WG / 64);
   // whole waves take the same blocks, so the wave-wide shuffles see 64 active lanes
   for (uint32_t b = blockIdx.x * (WG / 64) + (threadIdx.x >> 6); b < n; b += nw) {
      uint32_t code;
      const float sc = encode_wave64<ST>(f[(size_t)b * WID3 + e], code);
      q[(size_t)b * WID3 + e] = (uint16_t)code;
      if (e == 0) scale[b] = sc;
   }
}

// 16-bit pool -> FP32 blocks, one wave per block, grid-stride.
__global__ __launch_bounds__(WG) void k_decode(Pool pl, float* f, uint32_t n) {
   const int e = threadIdx.x & 63;
   const uint32_t nw = gridDim.x * (WG / 64);
   for (uint32_t b = blockIdx.x * (WG / 64) + (threadIdx.x >> 6); b < n; b += nw)
      f[(size_t)b * WID3 + e] = pool_value(pl, b, e);
}

__global__ void k_iota(uint32_t* dst, uint32_t n) {
   const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
   if (i < n) dst[i] = i;
}

__global__ void k_fill_u32(uint32_t* dst, uint32_t v, uint32_t n) {
   const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
   if (i < n) dst[i] = v;
}

// ---------------------------------------------------------------------------
// launch wrappers
// ---------------------------------------------------------------------------

void launch_columns(int dim, const ColumnArgs& a, hipStream_t s) {
   if (a.ncells == 0) return;
   if (dim == 0) hipLaunchKernelGGL(k_columns<0>, dim3(a.ncells), dim3(CWG), 0, s, a);
   else if (dim == 1) hipLaunchKernelGGL(k_columns<1>, dim3(a.ncells), dim3(CWG), 0, s, a);
   else hipLaunchKernelGGL(k_columns<2>, dim3(a.ncells), dim3(CWG), 0, s, a);
}

void launch_order(const OrderArgs& a, hipStream_t s) {
   if (a.ncells == 0) return;
   hipLaunchKernelGGL(k_order, dim3(a.ncells), dim3(WG), 0, s, a);
}

void launch_bin_scan(uint32_t* cell_hist, uint32_t ncells, hipStream_t s) {
   if (ncells == 0) return;
   hipLaunchKernelGGL(k_bin_scan, dim3(LEN_BINS), dim3(1024), 0, s, cell_hist, ncells);
}

void launch_scan(const ScanArgs& a, hipStream_t s) {
   hipLaunchKernelGGL(k_scan2, dim3(1), dim3(1024), 0, s, a);
}

static int g_cus = 0;
int device_cus() {
   if (g_cus == 0) {
      int dev = 0, cus = 0;
      (void)hipGetDevice(&dev);
      (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
      g_cus = (cus > 0) ? cus : 256;
   }
   return g_cus;
}

template <int DIM, int METHOD, int ST>
static void remap_dms(RemapArgs a, uint32_t grid, double i_dv, bool fast, int wgs, hipStream_t s) {
   if (!fast) {
      a.range = RANGE_ALL;
      hipLaunchKernelGGL((k_remap<DIM, METHOD, ST, RING, WG>), dim3(grid), dim3(WG), 0, s, a, i_dv);
   } else {
      // tolerance form: the lean kernel (workgroups of wgs lanes, wgs/16 work units each), then
      // the exact kernel over the few units it leaves
      // (with block sums: their own instance, whose extra registers the other passes do not carry)
      if (a.blk_stats) {
         if (wgs == 64) hipLaunchKernelGGL((k_remap_lean<DIM, METHOD, ST, 64, true>), dim3(grid), dim3(64), 0, s, a);
         else if (wgs == 128) hipLaunchKernelGGL((k_remap_lean<DIM, METHOD, ST, 128, true>), dim3(grid), dim3(128), 0, s, a);
         else hipLaunchKernelGGL((k_remap_lean<DIM, METHOD, ST, WG, true>), dim3(grid), dim3(WG), 0, s, a);
      } else {
         if (wgs == 64) hipLaunchKernelGGL((k_remap_lean<DIM, METHOD, ST, 64, false>), dim3(grid), dim3(64), 0, s, a);
         else if (wgs == 128) hipLaunchKernelGGL((k_remap_lean<DIM, METHOD, ST, 128, false>), dim3(grid), dim3(128), 0, s, a);
         else hipLaunchKernelGGL((k_remap_lean<DIM, METHOD, ST, WG, false>), dim3(grid), dim3(WG), 0, s, a);
      }
      a.range = RANGE_EXACT;
      hipLaunchKernelGGL((k_remap<DIM, METHOD, ST, RING, WG>), dim3((uint32_t)device_cus()), dim3(WG), 0, s, a, i_dv);
   }
   // the units of wide-span cells (none in most passes: the launch then exits at once)
   a.range = RANGE_WIDE;
   hipLaunchKernelGGL((k_remap<DIM, METHOD, ST, RING_WIDE, WG_WIDE>), dim3((uint32_t)device_cus()), dim3(WG_WIDE), 0, s,
                      a, i_dv);
}

template <int DIM, int METHOD>
static void remap_dm(const RemapArgs& a, uint32_t grid, double i_dv, bool fast, int wgs, hipStream_t s) {
   // the source and target pools share one storage form
   const int st = a.in.storage;
   if (st == ST_F32) remap_dms<DIM, METHOD, ST_F32>(a, grid, i_dv, fast, wgs, s);
   else if (st == ST_BF16) remap_dms<DIM, METHOD, ST_BF16>(a, grid, i_dv, fast, wgs, s);
   else remap_dms<DIM, METHOD, ST_F16>(a, grid, i_dv, fast, wgs, s);
}

template <int DIM>
static void remap_d(int method, const RemapArgs& a, uint32_t grid, double i_dv, bool fast, int wgs, hipStream_t s) {
   if (method == VACC_PLM) remap_dm<DIM, VACC_PLM>(a, grid, i_dv, fast, wgs, s);
   else if (method == VACC_PPM) remap_dm<DIM, VACC_PPM>(a, grid, i_dv, fast, wgs, s);
   else remap_dm<DIM, VACC_PQM>(a, grid, i_dv, fast, wgs, s);
}

void launch_remap(int dim, int method, bool fast, RemapArgs a, int grid_mult, int wgs, hipStream_t s) {
   if (a.ncells == 0) return;
   const uint32_t grid = (uint32_t)device_cus() * (uint32_t)std::max(grid_mult, 1);
   const double i_dv = 1.0 / (double)a.dv;   // cpu_acc_map.cpp:201, Real i_dv = 1.0/dv
   if (dim == 0) remap_d<0>(method, a, grid, i_dv, fast, wgs, s);
   else if (dim == 1) remap_d<1>(method, a, grid, i_dv, fast, wgs, s);
   else remap_d<2>(method, a, grid, i_dv, fast, wgs, s);
}

void launch_adjust(const AdjustArgs& a, bool write_phase, hipStream_t s) {
   if (a.ncells == 0) return;
   const int st = a.pool.storage;
   if (write_phase) {
      // the write phase reads no f
      hipLaunchKernelGGL((k_adjust<true, ST_F32>), dim3(a.ncells), dim3(AWG), 0, s, a);
   } else if (st == ST_F32) {
      hipLaunchKernelGGL((k_adjust<false, ST_F32>), dim3(a.ncells), dim3(AWG), 0, s, a);
   } else if (st == ST_BF16) {
      hipLaunchKernelGGL((k_adjust<false, ST_BF16>), dim3(a.ncells), dim3(AWG), 0, s, a);
   } else {
      hipLaunchKernelGGL((k_adjust<false, ST_F16>), dim3(a.ncells), dim3(AWG), 0, s, a);
   }
}

// Workgroups for one wave per block, capped so that a dispatch stays far below 2^32 work-items.
static uint32_t launch_waves(uint32_t n) {
   const uint64_t want = ((uint64_t)n + (WG / 64) - 1) / (WG / 64);
   return (uint32_t)std::min<uint64_t>(want, 1u << 20);
}

void launch_content(const uint32_t* src, const Pool& pool, uint8_t* content, uint32_t n, double mv, hipStream_t s) {
   if (n == 0) return;
   const uint32_t blocks = launch_waves(n);
   hipLaunchKernelGGL(k_content, dim3(blocks), dim3(WG), 0, s, src, pool, content, n, mv);
}

void launch_moments(const MomentArgs& a, hipStream_t s) {
   if (a.ncells == 0) return;
   if (a.blk_stats) hipLaunchKernelGGL(k_moments_sums, dim3(a.ncells), dim3(WG), 0, s, a);
   else if (a.pool.storage == ST_F32) hipLaunchKernelGGL(k_moments<ST_F32>, dim3(a.ncells), dim3(WG), 0, s, a);
   else if (a.pool.storage == ST_BF16) hipLaunchKernelGGL(k_moments<ST_BF16>, dim3(a.ncells), dim3(WG), 0, s, a);
   else hipLaunchKernelGGL(k_moments<ST_F16>, dim3(a.ncells), dim3(WG), 0, s, a);
}

void launch_transform(const MeshInfo& m, const TransformIn* fields, const double* dt_override, const double* moments,
                      int use_moments, double mass, double charge, double hall_min_rhoq, int ohm_gradpe, int map_order,
                      double* isect_out, uint32_t ncells, hipStream_t s) {
   if (ncells == 0) return;
   hipLaunchKernelGGL(k_transform, dim3((ncells + 127) / 128), dim3(128), 0, s, m, fields, dt_override, moments,
                      use_moments, mass, charge, hall_min_rhoq, ohm_gradpe, map_order, isect_out, ncells);
}

// max_v_dt per cell (cpu_acc_transform.cpp:41-55) and subcycles (vlasovmover.cpp:398-401)
__global__ void k_subcycle_count(const TransformIn* fields, uint32_t ncells, double mass, double charge,
                                 double max_rotation_deg, double dt, double* maxdt, uint32_t* nsub, uint32_t* nmax) {
   const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
   if (c >= ncells) return;
   const double* B = fields[c].B;
   const double B_mag = sqrt(B[0] * B[0] + B[1] * B[1] + B[2] * B[2]) + 1e-30;
   const double gyro_period = 2 * PI * mass / (charge * B_mag);
   const double m = fabs(gyro_period) * (max_rotation_deg / 360.0);
   const double q = ceil(fabs(dt) / m);
   const uint32_t n = (q >= 1.0) ? (q < 4294967295.0 ? (uint32_t)q : 0xFFFFFFFFu) : 1u;
   maxdt[c] = m;
   nsub[c] = n;
   atomicMax(nmax, n);
}

// subcycle dt of step `step` (vlasovmover.cpp:429-456)
__global__ void k_subcycle_step(const double* maxdt, const uint32_t* nsub, uint32_t ncells, uint32_t step, double dt,
                                double* subdt, uint8_t* active) {
   const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
   if (c >= ncells) return;
   const double m = maxdt[c];
   double d = ((double)(step + 1) * m > fabs(dt)) ? fmax(fabs(dt) - (double)step * m, 0.0) : m;
   if (dt < 0) d = -d;
   const bool on = step < nsub[c];
   active[c] = on ? 1 : 0;
   subdt[c] = on ? d : 0.0;
}

void launch_subcycle_count(const TransformIn* fields, uint32_t ncells, double mass, double charge,
                           double max_rotation_deg, double dt, double* maxdt, uint32_t* nsub, uint32_t* nmax,
                           hipStream_t s) {
   if (ncells == 0) return;
   hipLaunchKernelGGL(k_subcycle_count, dim3((ncells + 127) / 128), dim3(128), 0, s, fields, ncells, mass, charge,
                      max_rotation_deg, dt, maxdt, nsub, nmax);
}

void launch_subcycle_step(const double* maxdt, const uint32_t* nsub, uint32_t ncells, uint32_t step, double dt,
                          double* subdt, uint8_t* active, hipStream_t s) {
   if (ncells == 0) return;
   hipLaunchKernelGGL(k_subcycle_step, dim3((ncells + 127) / 128), dim3(128), 0, s, maxdt, nsub, ncells, step, dt,
                      subdt, active);
}

void launch_gather(const uint32_t* src, const Pool& in, const Pool& out, uint32_t n, hipStream_t s) {
   if (n == 0) return;
   const uint32_t grid = (uint32_t)std::min<uint64_t>(((uint64_t)n * 16 + 255) / 256, 1u << 20);
   if (in.storage == ST_F32)
      hipLaunchKernelGGL(k_gather, dim3(grid), dim3(256), 0, s, src,
                         static_cast<const float4*>(in.f), static_cast<float4*>(out.f), n);
   else
      hipLaunchKernelGGL(k_gather16, dim3(grid), dim3(256), 0, s, src,
                         static_cast<const uint2*>(in.f), in.scale, static_cast<uint2*>(out.f), out.scale, n);
}

void launch_encode(const float* f, const Pool& out, uint32_t n, hipStream_t s) {
   if (n == 0) return;
   const uint32_t blocks = launch_waves(n);
   if (out.storage == ST_BF16)
      hipLaunchKernelGGL(k_encode<ST_BF16>, dim3(blocks), dim3(WG), 0, s, f, static_cast<uint16_t*>(out.f), out.scale, n);
   else
      hipLaunchKernelGGL(k_encode<ST_F16>, dim3(blocks), dim3(WG), 0, s, f, static_cast<uint16_t*>(out.f), out.scale, n);
}

void launch_decode(const Pool& in, float* f, uint32_t n, hipStream_t s) {
   if (n == 0) return;
   const uint32_t blocks = launch_waves(n);
   hipLaunchKernelGGL(k_decode, dim3(blocks), dim3(WG), 0, s, in, f, n);
}

void launch_iota(uint32_t* dst, uint32_t n, hipStream_t s) {
   if (n == 0) return;
   hipLaunchKernelGGL(k_iota, dim3((n + 255) / 256), dim3(256), 0, s, dst, n);
}

__global__ void k_fill_f64(double* dst, double v, uint32_t n) {
   const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
   if (i < n) dst[i] = v;
}

void launch_fill_f64(double* dst, double v, uint32_t n, hipStream_t s) {
   if (n == 0) return;
   hipLaunchKernelGGL(k_fill_f64, dim3((n + 255) / 256), dim3(256), 0, s, dst, v, n);
}

void launch_pressure(const MomentArgs& a, const double* pivot, double mass, double* out, hipStream_t s) {
   if (a.ncells == 0) return;
   hipLaunchKernelGGL(k_pressure, dim3(a.ncells), dim3(WG), 0, s, a, pivot, mass, out);
}

void launch_sparse_min(const SparseArgs& p, const double* moments, const uint32_t* off, double* minval,
                       uint32_t ncells, hipStream_t s) {
   if (ncells == 0) return;
   hipLaunchKernelGGL(k_sparse_min, dim3((ncells + 127) / 128), dim3(128), 0, s, p, moments, off, minval, ncells);
}

void launch_content_cells(const uint32_t* off, const uint32_t* src, const Pool& pool, const double* minval,
                          uint8_t* content, uint32_t ncells, hipStream_t s) {
   if (ncells == 0) return;
   hipLaunchKernelGGL(k_content_cells, dim3(ncells), dim3(WG), 0, s, off, src, pool, minval, content);
}

void launch_cell_sums(const uint32_t* off, const uint32_t* src, const Pool& pool, double* out, uint32_t ncells,
                      hipStream_t s) {
   if (ncells == 0) return;
   hipLaunchKernelGGL(k_cell_sums, dim3(ncells), dim3(WG), 0, s, off, src, pool, out);
}

void launch_cell_scale(const uint32_t* off, const uint32_t* src, const Pool& pool, const double* pre,
                       const double* post, uint32_t ncells, hipStream_t s) {
   if (ncells == 0) return;
   hipLaunchKernelGGL(k_cell_scale, dim3(ncells), dim3(WG), 0, s, off, src, pool, pre, post);
}

void launch_fill_u32(uint32_t* dst, uint32_t v, uint32_t n, hipStream_t s) {
   if (n == 0) return;
   hipLaunchKernelGGL(k_fill_u32, dim3((n + 255) / 256), dim3(256), 0, s, dst, v, n);
}

// GID -> local index (VelocityMesh::getLocalID, velocity_mesh_cpu.h:371-377): lane i finds
// qgid[i] in the ascending GID list of cell qcell[i] by bisection.  The index is the block's
// position within its cell (the order vacc_download_cells returns); absent blocks and cells
// outside the context give VACC_INVALID_ID.
__global__ void k_find(const uint32_t* off, const uint32_t* gid, uint32_t ncells, const uint32_t* qcell,
                       const uint32_t* qgid, uint32_t* out, uint32_t nq) {
   const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
   if (i >= nq) return;
   const uint32_t c = qcell[i], g = qgid[i];
   uint32_t r = 0xFFFFFFFFu;
   if (c < ncells) {
      const uint32_t b = off[c], e = off[c + 1];
      uint32_t lo = b, hi = e;
      while (lo < hi) {
         const uint32_t mid = lo + ((hi - lo) >> 1);
         if (gid[mid] < g) lo = mid + 1;
         else hi = mid;
      }
      if (lo < e && gid[lo] == g) r = lo - b;
   }
   out[i] = r;
}

void launch_find(const uint32_t* off, const uint32_t* gid, uint32_t ncells, const uint32_t* qcell, const uint32_t* qgid,
                 uint32_t* out, uint32_t nq, hipStream_t s) {
   if (nq == 0) return;
   hipLaunchKernelGGL(k_find, dim3((nq + 255) / 256), dim3(256), 0, s, off, gid, ncells, qcell, qgid, out, nq);
}

}  // namespace vacc
