// vacc_trans.hip -- spatial translation kernel (trans_map_1d_amr, TRANS_SEMILAG_PPM).
//
// One wave per target block, one lane per velocity cell: lane l holds the block's cell of
// transposed index l = i + 4j + 16k, k along the translation dimension
// (cpu_trans_map_amr.cpp:224-285), so the reference's 8-lane groups of check_skip_remapping
// (VECL = 8, planes of 16 cells) are aligned groups of 8 lanes.  For target cell t the
// host lists every contribution t receives (source position of a pencil, kind, widths,
// ratio) in the reference's accumulation order, and the distinct cells of their stencils.
// A wave walks TCHUNK consecutive blocks of the list (GIDs ascending within each cell): lane
// j keeps a cursor into the sorted block list of stencil cell j, so after one binary search
// per cell each further block costs a few compares.  The new values are accumulated in FP32
// in the reference's order and written to a compact output pool; cells that are not
// written keep their values.
#include <hip/hip_runtime.h>

#include "vacc_kernels.h"
#include "vacc_recon.h"

namespace vacc {

constexpr int TWG = 256;     // four independent waves per workgroup
constexpr int TCHUNK = 8;    // consecutive list blocks per wave visit

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

// PPM on non-uniform widths, cpu_1d_ppm_nonuniform.hpp:40-62 with the filter of
// cpu_face_estimates.hpp:327-372; u[0..4] = values at k-2..k+2.
__device__ inline void ppm_nu(const TransEntry& E, const float (&u)[5], float thr, float scale, float& c0, float& c1,
                              float& c2) {
   float m = face_nu(E.fl, u[0], u[1], u[2], u[3]);
   float p = face_nu(E.fr, u[1], u[2], u[3], u[4]);
   float sl, sabs;
   if (thr > 0.0f) {
      const float sm = u[1] * scale, s0 = u[2] * scale, sp = u[3] * scale;
      sl = superbee_ab(sp - s0, s0 - sm);
      sabs = fabsf(sl) * thr;
   } else {
      sl = superbee_ab(u[3] - u[2], u[2] - u[1]);
      sabs = fabsf(sl);
   }
   const float ssgn = (sl > 0.0f) ? 1.0f : -1.0f;
   if (sabs == 0.0f) {
      p = u[2];
      m = u[2];
   }
   if ((u[1] - m) * (m - u[2]) < 0.0f) m = u[2] - (0.5f * ssgn) * sabs;
   if ((u[3] - p) * (p - u[2]) < 0.0f) p = u[2] + (0.5f * ssgn) * sabs;
   const float one_sixth = (float)(1.0 / 6.0);
   {
      const float d = p - m;
      const float mid = u[2] - 0.5f * (m + p);
      if (d * mid > (d * d) * one_sixth) m = 3.0f * u[2] - 2.0f * p;
   }
   {
      const float d = p - m;
      const float mid = u[2] - 0.5f * (m + p);
      if (((-d) * d) * one_sixth > d * mid) p = 3.0f * u[2] - 2.0f * m;
   }
   c0 = m;
   c1 = (3.0f * u[2] - 2.0f * m) - p;
   c2 = (m + p) - 2.0f * u[2];
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

__global__ __launch_bounds__(TWG) void k_trans(TransArgs a) {
   const int lane = threadIdx.x & 63;
   const uint32_t wave = (blockIdx.x * (uint32_t)TWG + threadIdx.x) >> 6;
   const uint32_t nwaves = (gridDim.x * (uint32_t)TWG) >> 6;
   const uint32_t nchunks = (a.nblocks + TCHUNK - 1) / TCHUNK;
   const int ti = lane & 3, tj = (lane >> 2) & 3, tk = lane >> 4;
   // vcell_transpose: in-block index of transposed cell (ti, tj, tk)
   const int ab = (a.dim == 0) ? ti * 16 + tj * 4 + tk : (a.dim == 1) ? ti + tj * 16 + tk * 4 : lane;
   for (uint32_t ch = wave; ch < nchunks; ch += nwaves) {
      const uint32_t b0 = ch * TCHUNK;
      const uint32_t b1 = min(b0 + (uint32_t)TCHUNK, a.nblocks);
      uint32_t t = cell_of(a.off, a.ncells, b0);
      uint32_t tend = a.off[t + 1];
      bool fresh = true;
      uint32_t cur = 0, cend = 0;
      for (uint32_t b = b0; b < b1; ++b) {
         while (b >= tend) {
            ++t;
            tend = a.off[t + 1];
            fresh = true;
         }
         const TransCell& C = a.cells[t];
         const uint32_t ne = C.ne;
         if (ne == 0) {
            const uint32_t sb = a.src[b];
            a.fout[(size_t)b * 64 + lane] = (sb != VACC_INVALID_ID) ? a.fin[(size_t)sb * 64 + lane] : 0.0f;
            continue;
         }
         const uint32_t g = a.gid[b];
         const uint32_t nn = C.nn;
         if ((uint32_t)lane < nn) {
            if (fresh) {
               const uint32_t c = C.nbr[lane];
               uint32_t lo = a.off[c], hi = a.off[c + 1];
               cend = hi;
               while (lo < hi) {
                  const uint32_t mid = (lo + hi) >> 1;
                  if (a.gid[mid] < g) lo = mid + 1;
                  else hi = mid;
               }
               cur = lo;
            } else {
               while (cur < cend && a.gid[cur] < g) ++cur;
            }
         }
         fresh = false;
         const uint32_t my = ((uint32_t)lane < nn && cur < cend && a.gid[cur] == g) ? a.src[cur] : VACC_INVALID_ID;
         float v[TRANS_MAXN];
#pragma unroll
         for (int j = 0; j < TRANS_MAXN; ++j) {
            v[j] = 0.0f;
            if ((uint32_t)j < nn) {
               const uint32_t idx = (uint32_t)__builtin_amdgcn_readlane((int)my, j);
               if (idx != VACC_INVALID_ID) v[j] = a.fin[(size_t)idx * 64 + ab];
            }
         }
         // cell-centre velocity along dim (cpu_trans_map_amr.cpp:74): double, then Realf
         const uint32_t bz = (a.dim == 0) ? g % a.len[0] : (a.dim == 1) ? (g / a.len[0]) % a.len[1] : g / (a.len[0] * a.len[1]);
         const float cell_vz = (float)(((double)(bz * 4u + (uint32_t)tk) + 0.5) * (double)a.dvz + (double)a.vz_min);
         float acc = 0.0f;
         for (uint32_t e = 0; e < ne; ++e) {
            const TransEntry& E = a.entries[C.e0 + e];
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
            const float zt = (cell_vz * a.dt) / E.dz_i;
            const bool pos = zt > 0.0f;
            const float z1 = pos ? (float)(1.0 - (double)zt) : 0.0f;
            const float z2 = pos ? 1.0f : -zt;
            float c0, c1, c2;
            ppm_nu(E, u, a.thr, a.scale, c0, c1, c2);
            const float ng = z2 * (c0 + z2 * (c1 + z2 * c2)) - z1 * (c0 + z1 * (c1 + z1 * c2));
            float contrib;
            if (E.kind == 1) {
               contrib = (u[2] - ng) * E.ratio;
            } else {
               const bool take = (E.kind == 0) ? pos : !pos;
               contrib = (take ? (ng * E.dz_i) / E.dz_t : 0.0f) * E.ratio;
            }
            acc = acc + contrib;
         }
         a.fout[(size_t)b * 64 + ab] = acc;
      }
   }
}

void launch_trans(const TransArgs& a, hipStream_t s) {
   if (a.nblocks == 0 || a.ncells == 0) return;
   const uint64_t chunks = ((uint64_t)a.nblocks + TCHUNK - 1) / TCHUNK;
   const uint64_t wgs = std::min<uint64_t>((chunks + (TWG / 64) - 1) / (TWG / 64), 1u << 20);
   hipLaunchKernelGGL(k_trans, dim3((uint32_t)wgs), dim3(TWG), 0, s, a);
}

}  // namespace vacc
