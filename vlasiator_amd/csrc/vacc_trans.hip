// vacc_trans.hip -- spatial translation kernel (trans_map_1d_amr, TRANS_SEMILAG_PPM).
//
// One wave per target block, one lane per velocity cell: lane l holds the block's cell of
// transposed index l = i + 4j + 16k, k along the translation dimension
// (cpu_trans_map_amr.cpp:224-285), so the reference's 8-lane groups of check_skip_remapping
// (VECL = 8, planes of 16 cells) are aligned groups of 8 lanes.  For target cell t the
// host lists every contribution t receives (source position of a pencil, kind, widths,
// ratio) in the reference's accumulation order, and the distinct cells of their stencils.
// A wave walks TCHUNK consecutive blocks of the list (GIDs ascending within each cell).  For
// each stencil cell j, a window of its sorted block list (GIDs and pool indices) is staged in
// LDS from the first entry not below the current GID; a block's index in cell j is then one
// compare per lane and a ballot, and the window moves on only when the GIDs pass its end, so
// the one global load that depends on the lookup is the value itself.  The new values are accumulated in FP32
// in the reference's order and written to a compact output pool; cells that are not
// written keep their values.
#include <hip/hip_runtime.h>

#include "vacc_codec.h"
#include "vacc_kernels.h"
#include "vacc_recon.h"

namespace vacc {

constexpr int TWG = 256;     // four independent waves per workgroup
constexpr int TCHUNK = 64;   // consecutive list blocks per wave visit
constexpr int WIN = 64;      // entries of a stencil cell's list staged in LDS

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

// Largest t with off[t] <= b (the non-empty cell holding list block b).
__device__ inline uint32_t cell_of(const uint32_t* off, uint32_t ncells, uint32_t b) {
   uint32_t lo = 0, hi = ncells - 1;
   while (lo < hi) {
      const uint32_t mid = (lo + hi + 1) >> 1;
      if (off[mid] <= b) lo = mid;
      else hi = mid - 1;
   }
   return lo;
}

template <int ST>
__global__ __launch_bounds__(TWG) void k_trans(TransArgs a) {
   // per wave: a window of WIN consecutive entries (GID, pool index) of each stencil cell's
   // sorted block list, starting at the first entry not below the chunk's first GID
   __shared__ uint32_t s_gid[TWG / 64][TRANS_MAXN][WIN];
   __shared__ uint32_t s_src[TWG / 64][TRANS_MAXN][WIN];
   // per wave: the current cell's three entries when it is canonical
   static_assert(sizeof(TransFace) == 64 && sizeof(TransEntry) == 160, "entry layout");
   constexpr int EW = (int)(sizeof(TransEntry) / 4);   // dwords per entry
   __shared__ TransEntry s_ent[TWG / 64][3];
   const int lane = threadIdx.x & 63;
   const int wv = threadIdx.x >> 6;
   const uint32_t wave = (blockIdx.x * (uint32_t)TWG + threadIdx.x) >> 6;
   const uint32_t nwaves = (gridDim.x * (uint32_t)TWG) >> 6;
   const uint32_t nchunks = (a.nblocks + TCHUNK - 1) / TCHUNK;
   const int ti = lane & 3, tj = (lane >> 2) & 3, tk = lane >> 4;
   // vcell_transpose: in-block index of transposed cell (ti, tj, tk)
   const int ab = (a.dim == 0) ? ti * 16 + tj * 4 + tk : (a.dim == 1) ? ti + tj * 16 + tk * 4 : lane;
   for (uint32_t ch = wave; ch < nchunks; ch += nwaves) {
      const uint32_t b0 = ch * TCHUNK;
      const uint32_t b1 = min(b0 + (uint32_t)TCHUNK, a.nblocks);
      // the chunk's GIDs, lane k holding block b0 + k
      const uint32_t gch = (b0 + (uint32_t)lane < b1) ? a.gid[b0 + lane] : 0u;
      uint32_t t = cell_of(a.off, a.ncells, b0);
      uint32_t tend = a.off[t + 1];
      bool newcell = true;
      // lane j < nn: window of stencil cell j = list entries [wlo, wlo + WIN) clipped to wend
      uint32_t wlo = 0, wend = 0;
      // the current cell's header (uniform)
      uint32_t ne = 0, nn = 0, e0 = 0, canon = 0;
      for (uint32_t b = b0; b < b1; ++b) {
         while (b >= tend) {
            ++t;
            tend = a.off[t + 1];
            newcell = true;
         }
         const bool fresh = newcell;
         newcell = false;
         if (fresh) {
            // cell header: e0, ne, nn, nbr[12], canonical in lanes 0..15
            const uint32_t* cw = (const uint32_t*)(a.cells + t);
            const uint32_t hw = (lane < (int)(sizeof(TransCell) / 4)) ? cw[lane] : 0u;
            e0 = (uint32_t)__builtin_amdgcn_readlane((int)hw, 0);
            ne = (uint32_t)__builtin_amdgcn_readlane((int)hw, 1);
            nn = (uint32_t)__builtin_amdgcn_readlane((int)hw, 2);
            canon = (uint32_t)__builtin_amdgcn_readlane((int)hw, 3 + TRANS_MAXN);
            if (canon) {
               // the three entries (120 dwords) into this wave's LDS slot, two lane-parallel loads
               const uint32_t* ew = (const uint32_t*)(a.entries + e0);
               uint32_t* sw = (uint32_t*)&s_ent[wv][0];
               sw[lane] = ew[lane];
               if (lane < 3 * EW - 64) sw[64 + lane] = ew[64 + lane];
               __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
               __builtin_amdgcn_wave_barrier();
               __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            }
         }
         if (ne == 0) {   // not written: the block keeps its values
            copy_block<ST>(a.in, a.out, a.src[b], b, lane);
            continue;
         }
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
         // cell-centre velocity along dim (cpu_trans_map_amr.cpp:74): double, then Realf
         const uint32_t bz = (a.dim == 0) ? g % a.len[0] : (a.dim == 1) ? (g / a.len[0]) % a.len[1] : g / (a.len[0] * a.len[1]);
         const float cell_vz = (float)(((double)(bz * 4u + (uint32_t)tk) + 0.5) * (double)a.dvz + (double)a.vz_min);
         const float vdt = cell_vz * a.dt;
         float acc = 0.0f;
         if (canon) {
            // target position q: p1 from source q-1 (cells 0..4), self from q (1..5), m1 from
            // q+1 (2..6); the raw faces k = q-1 .. q+2 are shared between the sources
            const bool p0 = v[0] > 0.0f, p1 = v[1] > 0.0f, p2 = v[2] > 0.0f, p3 = v[3] > 0.0f, p4 = v[4] > 0.0f,
                       p5 = v[5] > 0.0f, p6 = v[6] > 0.0f;
            const int grp = lane & 56;
            const bool do0 = ((__ballot(p0 || p1 || p2 || p3 || p4) >> grp) & 0xFFull) != 0ull;
            const bool do1 = ((__ballot(p1 || p2 || p3 || p4 || p5) >> grp) & 0xFFull) != 0ull;
            const bool do2 = ((__ballot(p2 || p3 || p4 || p5 || p6) >> grp) & 0xFFull) != 0ull;
            const TransEntry& E0 = s_ent[wv][0];
            const TransEntry& E1 = s_ent[wv][1];
            const TransEntry& E2 = s_ent[wv][2];
            const float f1 = face_nu(E0.fr, v[1], v[2], v[3], v[4]);
            const float f2 = face_nu(E1.fr, v[2], v[3], v[4], v[5]);
            const float dzi0 = E0.dz_i, dzi1 = E1.dz_i, dzi2 = E2.dz_i;
            const float zt0 = vdt / dzi0;
            const float zt1 = (dzi1 == dzi0) ? zt0 : vdt / dzi1;
            const float zt2 = (dzi2 == dzi1) ? zt1 : vdt / dzi2;
            // p1 is select(zt0 > 0, .., 0) * ratio and m1 select(zt2 <= 0, .., 0) * ratio: where
            // no lane of the block moves that way (a block's velocities along dim mostly share
            // one sign) the contribution is +0 in every lane, so the source's reconstruction is
            // skipped and the +0 is still added
            const bool up0 = __ballot(zt0 > 0.0f) != 0ull;
            const bool dn2 = __ballot(!(zt2 > 0.0f)) != 0ull;
            float c0, c1, c2;
            if (up0) {
               const float f0 = face_nu(E0.fl, v[0], v[1], v[2], v[3]);
               if (do0) {
                  ppm_nu_coeffs(f0, f1, v[1], v[2], v[3], a.thr, a.scale, c0, c1, c2);
                  const float ng = leaving(zt0, c0, c1, c2);
                  acc = acc + ((zt0 > 0.0f) ? (ng * dzi0) / E0.dz_t : 0.0f) * E0.ratio;
               }
            } else if (do0) {
               acc = acc + 0.0f * E0.ratio;
            }
            if (do1) {
               ppm_nu_coeffs(f1, f2, v[2], v[3], v[4], a.thr, a.scale, c0, c1, c2);
               const float ng = leaving(zt1, c0, c1, c2);
               acc = acc + (v[3] - ng) * E1.ratio;
            }
            if (dn2) {
               const float f3 = face_nu(E2.fr, v[3], v[4], v[5], v[6]);
               if (do2) {
                  ppm_nu_coeffs(f2, f3, v[3], v[4], v[5], a.thr, a.scale, c0, c1, c2);
                  const float ng = leaving(zt2, c0, c1, c2);
                  acc = acc + (!(zt2 > 0.0f) ? (ng * dzi2) / E2.dz_t : 0.0f) * E2.ratio;
               }
            } else if (do2) {
               acc = acc + 0.0f * E2.ratio;
            }
            store_block<ST>(a.out, b, acc, ab, a.dim, lane);
            continue;
         }
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
            if (E.kind != 1 && __ballot(take) == 0ull) {   // +0 in every lane (see the canonical path)
               acc = acc + 0.0f * E.ratio;
               continue;
            }
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

void launch_trans(const TransArgs& a, hipStream_t s) {
   if (a.nblocks == 0 || a.ncells == 0) return;
   const uint64_t chunks = ((uint64_t)a.nblocks + TCHUNK - 1) / TCHUNK;
   const uint64_t wgs = std::min<uint64_t>((chunks + (TWG / 64) - 1) / (TWG / 64), 1u << 20);
   if (a.in.storage == ST_BF16) hipLaunchKernelGGL(k_trans<ST_BF16>, dim3((uint32_t)wgs), dim3(TWG), 0, s, a);
   else if (a.in.storage == ST_F16) hipLaunchKernelGGL(k_trans<ST_F16>, dim3((uint32_t)wgs), dim3(TWG), 0, s, a);
   else hipLaunchKernelGGL(k_trans<ST_F32>, dim3((uint32_t)wgs), dim3(TWG), 0, s, a);
}

}  // namespace vacc
