// vacc_trans.hip -- spatial translation kernels (trans_map_1d_amr, TRANS_SEMILAG_PPM).
//
// Three kernels per call; the last two write disjoint blocks of a compact output pool.
//   k_walk_bounds  splits each pencil segment's GID range into shares of equal block counts.
//   k_trans_walk   walks pencil segments, one wave per velocity block, for the cells that
//                  receive all their contributions from one pencil at one position, or at two
//                  positions at least three apart (the run cells of a pencil, including a
//                  periodic pencil's wrap cells); with FP32 storage also, in launch levels,
//                  for cells that several pencils write at one run position each and for the
//                  wrap cells of pencils longer than a walk segment.
//   k_trans        takes every other cell: cells written at adjacent positions (walls where
//                  the stencil repeats a cell), cells of several pencils and the wrap cells
//                  of long pencils in the 16-bit forms, and cells that are not written
//                  (copied).
//
// In both writing kernels lane l holds the block's velocity cell of transposed index
// l = i + 4j + 16k, k along the translation dimension (cpu_trans_map_amr.cpp:224-285), so
// the reference's 8-lane groups of check_skip_remapping (VECL = 8, planes of 16 cells) are
// aligned groups of 8 lanes; values are accumulated in FP32 in the reference's order.
//
// k_trans: for each of its cells the host lists every contribution the cell receives (source
// position of a pencil, kind, widths, ratio) in the reference's accumulation order, and the
// distinct cells of their stencils.  A wave takes TCHUNK consecutive blocks of one cell (GIDs
// ascending within each cell).  For each stencil cell j, a window of its sorted block list
// (GIDs and pool indices) is staged in LDS from the first entry not below the current GID; a
// block's index in cell j is then one compare per lane and a ballot, and the window moves on
// only when the GIDs pass its end, so the one global load that depends on the lookup is the
// value itself.
#include <hip/hip_runtime.h>

#include "vacc_codec.h"
#include "vacc_kernels.h"
#include "vacc_recon.h"

namespace vacc {

constexpr int TWG = 256;     // four independent waves per workgroup
constexpr int TCHUNK = 64;   // consecutive list blocks per wave visit
constexpr int WIN = 64;      // entries of a stencil cell's list staged in LDS
constexpr int TSPLIT = 4;    // workgroups per cell of k_trans

// h4 left face of cell u0 on non-uniform widths (cpu_face_estimates.hpp:179-191); the
// width-only scalars come from the host, the double-scalar products round into the lane.
__device__ inline float face_nu(const TransFace& F, float um2, float um1, float u0, float up1) {
   const float V1 = um1 * F.h2 + u0 * F.h1;
   const float A = (float)(F.DA * (double)(F.SA * V1));
   const float V2 = (float)(F.DB * (double)um1) - um2 * F.h1;
   const float B = F.SB * V2;
   const float V3 = (float)(F.DC * (double)u0) - up1 * F.h2;
   const float C = F.SC * V3;
   return (float)(F.D1 * (double)((A + B) + C));
}

// Filter and monotonicity of PPM on non-uniform widths (cpu_face_estimates.hpp:327-372,
// cpu_1d_ppm_nonuniform.hpp:40-62) from the raw h4 faces m, p of the source cell u0 with
// neighbours um1, up1.
__device__ inline void ppm_nu_coeffs(float m, float p, float um1, float u0, float up1, float thr, float scale,
                                     float& c0, float& c1, float& c2) {
   float sl, sabs;
   if (thr > 0.0f) {
      const float sm = um1 * scale, s0 = u0 * scale, sp = up1 * scale;
      sl = superbee_ab(sp - s0, s0 - sm);
      sabs = fabsf(sl) * thr;
   } else {
      sl = superbee_ab(up1 - u0, u0 - um1);
      sabs = fabsf(sl);
   }
   const float ssgn = (sl > 0.0f) ? 1.0f : -1.0f;
   if (sabs == 0.0f) {
      p = u0;
      m = u0;
   }
   if ((um1 - m) * (m - u0) < 0.0f) m = u0 - (0.5f * ssgn) * sabs;
   if ((up1 - p) * (p - u0) < 0.0f) p = u0 + (0.5f * ssgn) * sabs;
   const float one_sixth = (float)(1.0 / 6.0);
   {
      const float d = p - m;
      const float mid = u0 - 0.5f * (m + p);
      if (d * mid > (d * d) * one_sixth) m = 3.0f * u0 - 2.0f * p;
   }
   {
      const float d = p - m;
      const float mid = u0 - 0.5f * (m + p);
      if (((-d) * d) * one_sixth > d * mid) p = 3.0f * u0 - 2.0f * m;
   }
   c0 = m;
   c1 = (3.0f * u0 - 2.0f * m) - p;
   c2 = (m + p) - 2.0f * u0;
}

// PPM of the source with stencil u[0..4] = values at k-2..k+2.
__device__ inline void ppm_nu(const TransEntry& E, const float (&u)[5], float thr, float scale, float& c0, float& c1,
                              float& c2) {
   const float m = face_nu(E.fl, u[0], u[1], u[2], u[3]);
   const float p = face_nu(E.fr, u[1], u[2], u[3], u[4]);
   ppm_nu_coeffs(m, p, u[1], u[2], u[3], thr, scale, c0, c1, c2);
}

// Integral of the reconstruction over the part of the cell that leaves it
// (cpu_trans_map_amr.cpp:75-106): ng for a Courant number zt.
__device__ inline float leaving(float zt, float c0, float c1, float c2) {
   const bool pos = zt > 0.0f;
   const float z1 = pos ? (float)(1.0 - (double)zt) : 0.0f;
   const float z2 = pos ? 1.0f : -zt;
   return z2 * (c0 + z2 * (c1 + z2 * c2)) - z1 * (c0 + z1 * (c1 + z1 * c2));
}

// Value e of pool block idx: FP32, or scale * decoded 16-bit code (vacc_codec.h).
template <int ST>
__device__ inline float pool_val(const Pool& p, uint32_t idx, int e) {
   if constexpr (ST == ST_F32) {
      return static_cast<const float*>(p.f)[(size_t)idx * 64 + e];
   } else {
      return p.scale[idx] * Codec<ST>::dec(static_cast<const uint16_t*>(p.f)[(size_t)idx * 64 + e]);
   }
}

// Writes block b: lane l holds the value of in-block index ab.  The 16-bit forms encode the
// block with the canonical lane order (lane e = in-block index e), so the values are first
// moved there: lane e reads the lane whose transposed index maps to e.
template <int ST>
__device__ inline void store_block(const Pool& p, uint32_t b, float x, int ab, int dim, int lane) {
   if constexpr (ST == ST_F32) {
      static_cast<float*>(p.f)[(size_t)b * 64 + ab] = x;
   } else {
      const int e = lane;
      const int from = (dim == 0) ? (e >> 4) + 4 * ((e >> 2) & 3) + 16 * (e & 3)
                       : (dim == 1) ? (e & 3) + 4 * (e >> 4) + 16 * ((e >> 2) & 3)
                                    : e;
      const float xe = __shfl(x, from, 64);
      uint32_t code;
      const float sc = encode_wave64<ST>(xe, code);
      static_cast<uint16_t*>(p.f)[(size_t)b * 64 + e] = (uint16_t)code;
      if (e == 0) p.scale[b] = sc;
   }
}

// Copies block sb of the input pool (INVALID: a zero block) to block b of the output.
template <int ST>
__device__ inline void copy_block(const Pool& in, const Pool& out, uint32_t sb, uint32_t b, int lane) {
   const bool ok = sb != VACC_INVALID_ID;
   if constexpr (ST == ST_F32) {
      static_cast<float*>(out.f)[(size_t)b * 64 + lane] = ok ? static_cast<const float*>(in.f)[(size_t)sb * 64 + lane] : 0.0f;
   } else {
      static_cast<uint16_t*>(out.f)[(size_t)b * 64 + lane] =
          ok ? static_cast<const uint16_t*>(in.f)[(size_t)sb * 64 + lane] : (uint16_t)0;
      if (lane == 0) out.scale[b] = ok ? in.scale[sb] : 0.0f;
   }
}

// v dt of the lane's velocity cell along dim for block g (cpu_trans_map_amr.cpp:74-75): the
// cell-centre velocity in double, rounded to Realf, times the Realf dt.
__device__ inline float cell_vdt(const TransArgs& a, uint32_t g, int tk) {
   const uint32_t bz = (a.dim == 0) ? g % a.len[0] : (a.dim == 1) ? (g / a.len[0]) % a.len[1] : g / (a.len[0] * a.len[1]);
   const float cell_vz = (float)(((double)(bz * 4u + (uint32_t)tk) + 0.5) * (double)a.dvz + (double)a.vz_min);
   return cell_vz * a.dt;
}

// Lane q's value of x (uniform q).
__device__ inline float lanef(float x, int q) {
   return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), q));
}

template <int ST>
__global__ __launch_bounds__(TWG) void k_trans(TransArgs a) {
   // per wave: a window of WIN consecutive entries (GID, pool index) of each stencil cell's
   // sorted block list, starting at the first entry not below the chunk's first GID
   __shared__ uint32_t s_gid[TWG / 64][TRANS_MAXN][WIN];
   __shared__ uint32_t s_src[TWG / 64][TRANS_MAXN][WIN];
   static_assert(sizeof(TransFace) == 64 && sizeof(TransEntry) == 160 && sizeof(TransCell) == 64, "layouts");
   const int lane = threadIdx.x & 63;
   const int wv = threadIdx.x >> 6;
   const int ti = lane & 3, tj = (lane >> 2) & 3, tk = lane >> 4;
   // vcell_transpose: in-block index of transposed cell (ti, tj, tk)
   const int ab = (a.dim == 0) ? ti * 16 + tj * 4 + tk : (a.dim == 1) ? ti + tj * 16 + tk * 4 : lane;
   // work item = (cell of the rest list, one of TSPLIT workgroups); its waves take the
   // cell's chunks of TCHUNK blocks in turn
   const uint32_t items = a.nrest * (uint32_t)TSPLIT;
   for (uint32_t item = blockIdx.x; item < items; item += gridDim.x) {
      const uint32_t t = a.rest[item / TSPLIT];
      const uint32_t cb0 = a.off[t], cb1 = a.off[t + 1];
      const uint32_t nch = (cb1 - cb0 + TCHUNK - 1) / TCHUNK;
      // cell header: e0, ne, nn, nbr[12], walked in lanes 0..15
      const uint32_t* cw = (const uint32_t*)(a.cells + t);
      const uint32_t hw = (lane < (int)(sizeof(TransCell) / 4)) ? cw[lane] : 0u;
      const uint32_t e0 = (uint32_t)__builtin_amdgcn_readlane((int)hw, 0);
      const uint32_t ne = (uint32_t)__builtin_amdgcn_readlane((int)hw, 1);
      const uint32_t nn = (uint32_t)__builtin_amdgcn_readlane((int)hw, 2);
      for (uint32_t k = (item % TSPLIT) * (TWG / 64) + (uint32_t)wv; k < nch; k += TSPLIT * (TWG / 64)) {
         const uint32_t b0 = cb0 + k * TCHUNK;
         const uint32_t b1 = min(b0 + (uint32_t)TCHUNK, cb1);
         if (ne == 0) {   // not written: the blocks keep their values
            for (uint32_t b = b0; b < b1; ++b) copy_block<ST>(a.in, a.out, a.src[b], b, lane);
            continue;
         }
         // the chunk's GIDs, lane k holding block b0 + k
         const uint32_t gch = (b0 + (uint32_t)lane < b1) ? a.gid[b0 + lane] : 0u;
         // lane j < nn: window of stencil cell j = list entries [wlo, wlo + WIN) clipped to wend
         uint32_t wlo = 0, wend = 0;
         for (uint32_t b = b0; b < b1; ++b) {
            const bool fresh = b == b0;
            const uint32_t g = (uint32_t)__builtin_amdgcn_readlane((int)gch, (int)(b - b0));
            // lanes j < nn: does window j need (re)loading for g?
            bool reload = false;
            if ((uint32_t)lane < nn) {
               uint32_t lo, hi;
               if (fresh) {
                  const uint32_t c = a.cells[t].nbr[lane];
                  lo = a.off[c];
                  hi = a.off[c + 1];
                  wend = hi;
                  reload = true;
               } else {
                  lo = wlo + WIN;
                  hi = wend;
                  // all WIN window entries below g while the list goes on: move past them
                  reload = lo < hi && s_gid[wv][lane][WIN - 1] < g;
               }
               if (reload) {
                  while (lo < hi) {
                     const uint32_t mid = (lo + hi) >> 1;
                     if (a.gid[mid] < g) lo = mid + 1;
                     else hi = mid;
                  }
                  wlo = lo;
               }
            }
            const unsigned long long rmask = __ballot(reload);
            if (rmask) {
   #pragma unroll
               for (int j = 0; j < TRANS_MAXN; ++j) {
                  if ((rmask >> j) & 1ull) {
                     const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)wlo, j);
                     const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)wend, j);
                     const uint32_t k = lo + (uint32_t)lane;
                     s_gid[wv][j][lane] = (k < hi) ? a.gid[k] : VACC_INVALID_ID;
                     s_src[wv][j][lane] = (k < hi) ? a.src[k] : VACC_INVALID_ID;
                  }
               }
               __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
               __builtin_amdgcn_wave_barrier();
               __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            }
            float v[TRANS_MAXN];
   #pragma unroll
            for (int j = 0; j < TRANS_MAXN; ++j) {
               v[j] = 0.0f;
               if ((uint32_t)j < nn) {
                  const unsigned long long hit = __ballot(s_gid[wv][j][lane] == g);
                  if (hit) {
                     const uint32_t idx = s_src[wv][j][__builtin_ctzll(hit)];
                     if (idx != VACC_INVALID_ID) v[j] = pool_val<ST>(a.in, idx, ab);
                  }
               }
            }
            const float vdt = cell_vdt(a, g, tk);
            float acc = 0.0f;
            for (uint32_t e = 0; e < ne; ++e) {
               const TransEntry& E = a.entries[e0 + e];
               float u[5];
   #pragma unroll
               for (int s = 0; s < 5; ++s) {
                  const uint32_t sl = E.st[s];
                  float x = 0.0f;
   #pragma unroll
                  for (int j = 0; j < TRANS_MAXN; ++j) x = (sl == (uint32_t)j) ? v[j] : x;
                  u[s] = x;
               }
               // check_skip_remapping (cpu_trans_map_amr.cpp:18-25): a group of 8 lanes is left out
               // when none of its stencil values is positive
               const bool any = (u[0] > 0.0f) || (u[1] > 0.0f) || (u[2] > 0.0f) || (u[3] > 0.0f) || (u[4] > 0.0f);
               const unsigned long long mask = __ballot(any);
               if (((mask >> (lane & 56)) & 0xFFull) == 0ull) continue;
               const float zt = vdt / E.dz_i;
               const bool pos = zt > 0.0f;
               const bool take = (E.kind == 0) ? pos : !pos;
               // a contribution that is +-0 in every lane leaves the sum unchanged: the sum starts
               // at +0 and x + (+-0) = x for x != 0, +0 + (+-0) = +0
               if (E.kind != 1 && __ballot(take) == 0ull) continue;
               float c0, c1, c2;
               ppm_nu(E, u, a.thr, a.scale, c0, c1, c2);
               const float ng = leaving(zt, c0, c1, c2);
               float contrib;
               if (E.kind == 1) {
                  contrib = (u[2] - ng) * E.ratio;
               } else {
                  contrib = (take ? (ng * E.dz_i) / E.dz_t : 0.0f) * E.ratio;
               }
               acc = acc + contrib;
            }
            store_block<ST>(a.out, b, acc, ab, a.dim, lane);
         }
      }
   }
}

// Minimum and maximum over the wave (uniform results).
__device__ inline uint32_t wave_min(uint32_t x) {
#pragma unroll
   for (int o = 32; o > 0; o >>= 1) x = min(x, (uint32_t)__shfl_xor((int)x, o, 64));
   return (uint32_t)__builtin_amdgcn_readfirstlane((int)x);
}
__device__ inline uint32_t wave_max(uint32_t x) {
#pragma unroll
   for (int o = 32; o > 0; o >>= 1) x = max(x, (uint32_t)__shfl_xor((int)x, o, 64));
   return (uint32_t)__builtin_amdgcn_readfirstlane((int)x);
}

// Work split of the pencil walk: for each segment, the walked cells' blocks (pairs of a walked
// position and a block it holds) counted in a histogram of WB_BINS GID bins in LDS, then
// bounds[s][0..wps] = GID bounds of wps shares holding equal numbers of pairs, to a bin:
// bounds[k] = start of the first bin whose exclusive prefix is >= k * total / wps.
constexpr int WB_BINS = 4096;

__global__ __launch_bounds__(TWG) void k_walk_bounds(WalkArgs w, uint32_t* bounds) {
   __shared__ uint32_t hist[WB_BINS + 1];
   __shared__ uint32_t part[TWG];
   const TransArgs& a = w.t;
   const int t = threadIdx.x;
   constexpr int PER = WB_BINS / TWG;   // bins per thread
   for (uint32_t s = blockIdx.x; s < w.nseg; s += gridDim.x) {
      const WalkSeg S = w.seg[s];
      for (int b = t; b <= WB_BINS; b += TWG) hist[b] = 0u;
      __syncthreads();
      for (uint32_t q = 0; q < S.n; ++q) {
         const WalkPos P = w.pos[S.pos0 + q];
         if (!(P.flags & WALK_W) || P.cell == VACC_INVALID_ID) continue;
         const uint32_t lo = a.off[P.cell], hi = a.off[P.cell + 1];
         for (uint32_t k = lo + (uint32_t)t; k < hi; k += TWG) atomicAdd(&hist[min(a.gid[k] >> w.bin_shift, (uint32_t)WB_BINS - 1u)], 1u);
      }
      __syncthreads();
      // exclusive prefix of the bins: per-thread sums, then a scan over the threads
      uint32_t loc = 0;
#pragma unroll
      for (int i = 0; i < PER; ++i) loc += hist[t * PER + i];
      part[t] = loc;
      __syncthreads();
      if (t == 0) {
         uint32_t run = 0;
         for (int i = 0; i < TWG; ++i) {
            const uint32_t x = part[i];
            part[i] = run;
            run += x;
         }
         hist[WB_BINS] = run;   // total
      }
      __syncthreads();
      const uint64_t total = hist[WB_BINS];
      uint32_t* out = bounds + (size_t)s * (w.wps + 1);
      uint64_t pb = part[t];   // exclusive prefix of this thread's first bin
      for (int i = 0; i < PER; ++i) {
         const int b = t * PER + i;
         const uint64_t pe = pb + hist[b];   // exclusive prefix of bin b + 1
         // shares k whose first bin is b + 1: pb * wps < k * total <= pe * wps
         if (total) {
            const uint64_t k0 = pb * w.wps / total + 1, k1 = pe * w.wps / total;
            for (uint64_t k = k0; k <= k1 && k < w.wps; ++k) out[k] = min((uint32_t)(b + 1) << w.bin_shift, w.nmesh);
         }
         pb = pe;
      }
      if (t == 0) {
         out[0] = 0u;
         out[w.wps] = w.nmesh;
         if (!total)
            for (uint32_t k = 1; k < w.wps; ++k) out[k] = w.nmesh;
      }
      __syncthreads();
   }
}

// Pencil walk (cpu_trans_map_amr.cpp:59-148 for one pencil and one velocity block, over the
// walked positions of one segment).  Each segment has wps waves; wave k takes the GID range
// [bounds[k], bounds[k+1]) of k_walk_bounds.  Lane q < n holds a cursor into the sorted block
// list of window position q's cell, and the wave visits, in ascending order, every GID of
// its range that a walked position's cell holds (the minimum over the walked lanes'
// cursors).  For such a GID it visits, in order, the sources next to a walked position that
// holds the block (the others contribute to no walked block), a run of consecutive sources
// sharing faces and values: one new face, one reconstruction and one leaving integral per
// source, whose self, p1 and m1 contributions go into three rolling sums (positions j, j+1,
// j-1); position j-1 is complete after source j.  A walked position's sum starts from zero,
// or from a carry register holding the finished sum of an earlier position of the same cell
// (at least three positions before it, so that no source reaches both), and ends in the
// block or in a carry register.  Each cell thus receives its contributions in the order of
// the reference's source loop: p1 from j-1, self from j, m1 from j+1, source after source.
// A cell that several pencils write (one run position each), or whose two positions lie in
// different segments of one pencil, is walked in launch levels in the reference's order; a
// later level's sum starts from the block the earlier one stored (FP32).
// Waves per SIMD: seven for FP32 (72 VGPRs with 12 bytes per lane spilled, 5 % faster than
// six waves at 80 VGPRs without a spill) and six for the 16-bit forms (80 VGPRs instead of 85,
// no spill).
template <int ST>
__global__ __launch_bounds__(TWG) __attribute__((amdgpu_waves_per_eu(ST == ST_F32 ? 7 : 6))) void k_trans_walk(
    WalkArgs w) {
   // per wave: the current segment's faces (window positions 0..n-1)
   __shared__ TransFace s_face[TWG / 64][WALK_MAXN];
   const TransArgs& a = w.t;
   const int lane = threadIdx.x & 63;
   const int wv = threadIdx.x >> 6;
   const uint32_t wave = (uint32_t)__builtin_amdgcn_readfirstlane((int)((blockIdx.x * (uint32_t)TWG + threadIdx.x) >> 6));
   const uint32_t nwaves = (gridDim.x * (uint32_t)TWG) >> 6;
   const uint32_t nunits = w.nseg * w.wps;
   const int ti = lane & 3, tj = (lane >> 2) & 3, tk = lane >> 4;
   const int ab = (a.dim == 0) ? ti * 16 + tj * 4 + tk : (a.dim == 1) ? ti + tj * 16 + tk * 4 : lane;
   for (uint32_t u = wave; u < nunits; u += nwaves) {
      const uint32_t s = u / w.wps;
      const uint32_t glo = w.bounds[u + s], ghi = w.bounds[u + s + 1];   // row s has wps + 1 bounds
      if (glo >= ghi) continue;
      const WalkSeg S = w.seg[s];
      const TransFace* F = s_face[wv];
      bool faces = false;   // F holds this segment's faces
      // lane q < n: window position q, its list [cur, hi) from the first GID >= glo
      uint32_t cur = 0, hi = 0, cg = VACC_INVALID_ID, pf = 0;
      float dzl = 0.0f, rl = 0.0f;
      if ((uint32_t)lane < S.n) {
         const WalkPos P = w.pos[S.pos0 + lane];
         dzl = P.dz;
         rl = P.ratio;
         pf = P.flags;
         if (P.cell != VACC_INVALID_ID) {
            uint32_t lo = a.off[P.cell];
            hi = a.off[P.cell + 1];
            uint32_t h = hi;
            while (lo < h) {
               const uint32_t mid = (lo + h) >> 1;
               if (a.gid[mid] < glo) lo = mid + 1;
               else h = mid;
            }
            cur = lo;
            cg = (cur < hi) ? a.gid[cur] : VACC_INVALID_ID;
         }
      }
      const bool walked = (pf & WALK_W) != 0u;
      const unsigned long long wmask = __ballot(walked);
      // carry slots: positions whose finished sum goes to carry 1 / 2, whose sum starts there
      const unsigned long long save1 = __ballot(((pf >> WALK_SAVE_SHIFT) & 3u) == 1u);
      const unsigned long long save2 = __ballot(((pf >> WALK_SAVE_SHIFT) & 3u) == 2u);
      const unsigned long long init1 = __ballot(((pf >> WALK_INIT_SHIFT) & 3u) == 1u);
      const unsigned long long init2 = __ballot(((pf >> WALK_INIT_SHIFT) & 3u) == 2u);
      const unsigned long long inits = __ballot(((pf >> WALK_INIT_SHIFT) & 3u) == WALK_INIT_STORED);
      const int js = (int)S.js, je = (int)S.je, n = (int)S.n;
      const unsigned long long smask = ((je >= 63) ? ~0ull : ((1ull << (je + 1)) - 1ull)) & ~((1ull << js) - 1ull);
      for (;;) {
         const uint32_t g = wave_min(walked ? cg : VACC_INVALID_ID);
         if (g >= ghi) break;
         // every cursor to its first GID >= g
         while (__ballot(cg < g)) {
            if (cg < g) {
               ++cur;
               cg = (cur < hi) ? a.gid[cur] : VACC_INVALID_ID;
            }
         }
         const bool have = cg == g;
         const unsigned long long tmask = wmask & __ballot(have);   // positions written for g
         const uint32_t pidx = have ? a.src[cur] : VACC_INVALID_ID;
         const uint32_t ob = cur;                                    // output block where have
         // the cursors that hold g move on now, so that the next GID's lookup is in flight
         // during the walk
         if (have) {
            ++cur;
            cg = (cur < hi) ? a.gid[cur] : VACC_INVALID_ID;
         }
         if (!faces) {
            const uint32_t* src = (const uint32_t*)(w.face + S.pos0);
            uint32_t* dst = (uint32_t*)s_face[wv];
            constexpr uint32_t FW = sizeof(TransFace) / 4;
            for (uint32_t k = (uint32_t)lane; k < S.n * FW; k += 64) dst[k] = src[k];
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            faces = true;
         }
         const float vdt = cell_vdt(a, g, tk);
         // value at window position min(q, n-1) (uniform q); the load is not conditional (block
         // 0 stands in for a missing block, whose value is then 0), so that it can be in
         // flight while earlier work runs
         auto val = [&](int q) -> float {
            const uint32_t idx = (uint32_t)__builtin_amdgcn_readlane((int)pidx, min(q, n - 1));
            const float x = pool_val<ST>(a.in, (idx != VACC_INVALID_ID) ? idx : 0u, ab);
            return (idx != VACC_INVALID_ID) ? x : 0.0f;
         };
         float c1 = 0.0f, c2 = 0.0f;   // carry registers
         // start of window position q's sum: zero or a carry register
         const unsigned long long inited = init1 | init2 | inits;
         auto start = [&](int q) -> float {
            if (!((inited >> q) & 1ull)) return 0.0f;   // the common case, a uniform branch
            if constexpr (ST == ST_F32) {
               // the sum an earlier launch level stored into this block
               if (((inits & tmask) >> q) & 1ull)
                  return static_cast<const float*>(a.out.f)[(size_t)__builtin_amdgcn_readlane((int)ob, q) * 64 + ab];
            }
            return ((init1 >> q) & 1ull) ? c1 : ((init2 >> q) & 1ull) ? c2 : 0.0f;
         };
         // end of window position q's sum x: into a carry register or the block
         auto retire = [&](int q, float x) {
            if (!((tmask >> q) & 1ull)) return;
            if ((save1 >> q) & 1ull) c1 = x;
            else if ((save2 >> q) & 1ull) c2 = x;
            else store_block<ST>(a.out, (uint32_t)__builtin_amdgcn_readlane((int)ob, q), x, ab, a.dim, lane);
         };
         // sources next to a written position.  Where every lane moves up (v dt > 0) each m1 is
         // +-0 (zt > 0 or, if zt underflows, a leaving integral of +-0), and where none does
         // each p1 is +0, so the sources only reached that way are left out
         const unsigned long long up = __ballot(vdt > 0.0f);
         const unsigned long long reach = (up == ~0ull) ? (tmask | (tmask >> 1))
                                          : (up == 0ull) ? (tmask | (tmask << 1))
                                                         : (tmask | (tmask << 1) | (tmask >> 1));
         // positions js-1 and je+1 (pencil ends: the parts of wrap cells) are only reached from
         // js and je, which stay in so that their sums are always stored or carried, also when
         // they receive nothing
         unsigned long long rel = (reach & smask) | (((tmask >> (je + 1)) & 1ull) << je) |
                                  (((tmask >> (js - 1)) & 1ull) << js);
         float u0 = 0.0f, u1 = 0.0f, u2 = 0.0f, u3 = 0.0f, u4 = 0.0f, nx = 0.0f, fl = 0.0f;
         float am = 0.0f, a0 = 0.0f;   // sums of positions j-1, j
         float dzc = __builtin_nanf(""), zt = 0.0f;   // zt = vdt / dzc
         int prev = -2;
         while (rel) {
            const int j = __builtin_ctzll(rel);
            rel &= rel - 1ull;
            if (j != prev + 1) {   // a new run: the previous one ends, the stencil is loaded
               if (prev >= 0) {
                  retire(prev, am);
                  retire(prev + 1, a0);
               }
               u0 = val(j - 2);
               u1 = val(j - 1);
               u2 = val(j);
               u3 = val(j + 1);
               u4 = val(j + 2);
               nx = val(j + 3);
               fl = face_nu(F[j], u0, u1, u2, u3);
               am = start(j - 1);
               a0 = start(j);
            } else {
               u0 = u1;
               u1 = u2;
               u2 = u3;
               u3 = u4;
               u4 = nx;
               nx = val(j + 3);
            }
            prev = j;
            float ap = start(j + 1);   // sum of position j+1
            const float fr = face_nu(F[j + 1], u1, u2, u3, u4);
            const bool wm = (tmask >> (j - 1)) & 1ull, ws = (tmask >> j) & 1ull, wp = (tmask >> (j + 1)) & 1ull;
            // check_skip_remapping (cpu_trans_map_amr.cpp:18-25) per group of 8 lanes
            const bool any = (u0 > 0.0f) || (u1 > 0.0f) || (u2 > 0.0f) || (u3 > 0.0f) || (u4 > 0.0f);
            const unsigned long long amask = __ballot(any);
            const bool act = ((amask >> (lane & 56)) & 0xFFull) != 0ull;
            if (amask) {
               float k0, k1, k2;
               ppm_nu_coeffs(fl, fr, u1, u2, u3, a.thr, a.scale, k0, k1, k2);
               const float dzj = lanef(dzl, j);
               if (!(dzj == dzc)) {
                  dzc = dzj;
                  zt = vdt / dzj;
               }
               const float ng = leaving(zt, k0, k1, k2);
               const bool pos = zt > 0.0f;
               if (ws && act) a0 = a0 + (u2 - ng) * lanef(rl, j);
               // a contribution that is +-0 in every lane leaves a sum unchanged (sums start
               // at +0 or at a sum, and x + (+-0) = x for x != 0, +0 + (+-0) = +0)
               if (wp && __ballot(pos && act)) {
                  const float x = (ng * dzj) / lanef(dzl, j + 1);
                  if (act) ap = ap + (pos ? x : 0.0f) * lanef(rl, j + 1);
               }
               if (wm && __ballot(!pos && act)) {
                  const float x = (ng * dzj) / lanef(dzl, j - 1);
                  if (act) am = am + (!pos ? x : 0.0f) * lanef(rl, j - 1);
               }
            }
            retire(j - 1, am);
            am = a0;
            a0 = ap;
            fl = fr;
         }
         if (prev >= 0) {
            retire(prev, am);
            retire(prev + 1, a0);
         }
      }
   }
}

void launch_trans(const TransArgs& a, hipStream_t s) {
   if (a.nblocks == 0 || a.nrest == 0) return;
   const uint64_t wgs = std::min<uint64_t>((uint64_t)a.nrest * TSPLIT, 1u << 20);
   if (a.in.storage == ST_BF16) hipLaunchKernelGGL(k_trans<ST_BF16>, dim3((uint32_t)wgs), dim3(TWG), 0, s, a);
   else if (a.in.storage == ST_F16) hipLaunchKernelGGL(k_trans<ST_F16>, dim3((uint32_t)wgs), dim3(TWG), 0, s, a);
   else hipLaunchKernelGGL(k_trans<ST_F32>, dim3((uint32_t)wgs), dim3(TWG), 0, s, a);
}

void launch_trans_walk(const WalkArgs& w, uint32_t* bounds, hipStream_t s) {
   const uint64_t units = (uint64_t)w.nseg * w.wps;
   if (units == 0 || w.t.nblocks == 0) return;
   hipLaunchKernelGGL(k_walk_bounds, dim3(std::min<uint32_t>(w.nseg, 8192u)), dim3(TWG), 0, s, w, bounds);
   const uint64_t wgs = std::min<uint64_t>((units + (TWG / 64) - 1) / (TWG / 64), 1u << 20);
   if (w.t.in.storage == ST_BF16) hipLaunchKernelGGL(k_trans_walk<ST_BF16>, dim3((uint32_t)wgs), dim3(TWG), 0, s, w);
   else if (w.t.in.storage == ST_F16) hipLaunchKernelGGL(k_trans_walk<ST_F16>, dim3((uint32_t)wgs), dim3(TWG), 0, s, w);
   else hipLaunchKernelGGL(k_trans_walk<ST_F32>, dim3((uint32_t)wgs), dim3(TWG), 0, s, w);
}

}  // namespace vacc
