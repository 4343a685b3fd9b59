// vacc_init.hip -- synthetic initial state on the device (project helper).
//
// Mirrors Project::setVelocitySpace (projects/project.cpp:239-263): every block of a
// candidate box is sampled at cell centres as the float sum of tri-Maxwellian peaks
// (projects/project.h:66-72, MultiPeak.cpp:205-214) or gyrotropic rings, then the initial
// single-cell adjust (grid.cpp:289) keeps the blocks with content and their neighbours.
#include "vacc_kernels.h"

#include <limits.h>

namespace vacc {

constexpr int IWG = 256;
constexpr int IB_WORDS = 4096;   // LDS words per bitmap (candidate box <= 131072 blocks)
constexpr double K_B = 1.3806503e-23;


__device__ inline double peak_value(const vacc_peak& p, double mass, double vx, double vy, double vz) {
   const double x = vx - p.V[0], y = vy - p.V[1], z = vz - p.V[2];
   if (p.kind == 0) {
      const double T = p.T;
      return p.n * pow(mass / (2.0 * PI * K_B), 1.5) * exp(-mass * (x * x / T + y * y / T + z * z / T) / (2.0 * K_B)) /
             sqrt(T * T * T);
   }
   const double par = x * p.axis[0] + y * p.axis[1] + z * p.axis[2];
   const double px = x - par * p.axis[0], py = y - par * p.axis[1], pz = z - par * p.axis[2];
   const double perp = sqrt(px * px + py * py + pz * pz);
   const double s2 = p.vth * p.vth;
   const double norm = 1.0 / ((2.0 * PI) * (2.0 * PI) * p.vperp * s2);
   const double w = perp - p.vperp;
   return p.n * norm * exp(-(w * w + par * par) / (2.0 * s2));
}

// value of cell e (0..63) of block (bx,by,bz)
__device__ inline float sample(const InitArgs& a, const vacc_peak* pk, int np, int bx, int by, int bz, int e) {
   const int i = e & 3, j = (e >> 2) & 3, k = e >> 4;
   const double x0 = a.vmin[0] + bx * (a.cell[0] * WID);
   const double y0 = a.vmin[1] + by * (a.cell[1] * WID);
   const double z0 = a.vmin[2] + bz * (a.cell[2] * WID);
   const double vx = x0 + (i + 0.5) * a.cell[0], vy = y0 + (j + 0.5) * a.cell[1], vz = z0 + (k + 0.5) * a.cell[2];
   float value = 0.0f;
   for (int q = 0; q < np; ++q) value += (float)peak_value(pk[q], a.mass, vx, vy, vz);
   return value;
}

__device__ inline void dilate(const uint32_t* cbm, uint32_t* kbm, int Rx, int Ry, int Rz, int W, int w) {
   const int nwords = Ry * Rz * W;
   for (int q = threadIdx.x; q < nwords; q += IWG) {
      uint32_t x = cbm[q];
      const int row = q / W;
      const int y = row % Ry, z = row / Ry;
      const int xb = (q % W) * 32;
      while (x) {
         const int b = __ffs(x) - 1;
         x &= x - 1;
         const int xx = xb + b;
         const int x0 = max(xx - w, 0), x1 = min(xx + w, Rx - 1);
         for (int dz = -w; dz <= w; ++dz) {
            const int zz = z + dz;
            if (zz < 0 || zz >= Rz) continue;
            for (int dy = -w; dy <= w; ++dy) {
               const int yy = y + dy;
               if (yy < 0 || yy >= Ry) continue;
               uint32_t* r = kbm + (zz * Ry + yy) * W;
               for (int xk = x0; xk <= x1; ++xk) atomicOr(&r[xk >> 5], 1u << (xk & 31));
            }
         }
      }
   }
}

template <bool WRITE>
__global__ __launch_bounds__(IWG) void k_init(InitArgs a) {
   __shared__ uint32_t cbm[IB_WORDS];
   __shared__ uint32_t kbm[IB_WORDS];
   __shared__ uint32_t wpre[WRITE ? IB_WORDS : 1];
   __shared__ uint32_t sh[4];
   const uint32_t c = blockIdx.x;
   const int t = threadIdx.x;
   const int np = a.npeaks[c];
   const int* bb = a.bbox + 6 * (size_t)c;
   const int Rx = bb[3] - bb[0] + 1, Ry = bb[4] - bb[1] + 1, Rz = bb[5] - bb[2] + 1;
   if (np <= 0 || Rx <= 0 || Ry <= 0 || Rz <= 0) {
      if (!WRITE && t == 0) a.n_keep[c] = 0;
      return;
   }
   const int W = (Rx + 31) >> 5;
   const int nwords = Ry * Rz * W;
   const vacc_peak* pk = a.peaks + (size_t)c * a.max_peaks;
   uint32_t* gbm = a.content_bm + a.bm_off[c];
   for (int q = t; q < nwords; q += IWG) kbm[q] = 0;
   if (!WRITE) {
      for (int q = t; q < nwords; q += IWG) cbm[q] = 0;
      __syncthreads();
      // content of every candidate block: one wave per block, one lane per cell
      const int nb = Rx * Ry * Rz;
      const int lane = t & 63;
      for (int b = t >> 6; b < nb; b += IWG / 64) {
         const int x = b % Rx, y = (b / Rx) % Ry, z = b / (Rx * Ry);
         const float v = sample(a, pk, np, bb[0] + x, bb[1] + y, bb[2] + z, lane);
         const unsigned long long bal = __ballot((double)v >= a.min_value);
         if (lane == 0 && bal) atomicOr(&cbm[(z * Ry + y) * W + (x >> 5)], 1u << (x & 31));
      }
      __syncthreads();
      for (int q = t; q < nwords; q += IWG) gbm[q] = cbm[q];
   } else {
      for (int q = t; q < nwords; q += IWG) cbm[q] = gbm[q];
   }
   __syncthreads();
   dilate(cbm, kbm, Rx, Ry, Rz, W, a.add_width);
   __syncthreads();
   const int per = (nwords + IWG - 1) / IWG;
   const int wb = t * per, we = min(nwords, wb + per);
   uint32_t s1[1] = {0}, ex1[1], tot1[1];
   for (int q = wb; q < we; ++q) s1[0] += __popc(kbm[q]);
   {
      // workgroup exclusive scan of s1 (4 waves)
      const int lane = t & 63, wv = t >> 6;
      uint32_t v = s1[0];
      for (int o = 1; o < 64; o <<= 1) {
         const uint32_t u = __shfl_up(v, o, 64);
         if (lane >= o) v += u;
      }
      if (lane == 63) sh[wv] = v;
      __syncthreads();
      uint32_t base = 0, all = 0;
      for (int i = 0; i < IWG / 64; ++i) {
         if (i < wv) base += sh[i];
         all += sh[i];
      }
      ex1[0] = base + v - s1[0];
      tot1[0] = all;
      __syncthreads();
   }
   if (!WRITE) {
      if (t == 0) a.n_keep[c] = tot1[0];
      return;
   }
   uint32_t run = ex1[0];
   for (int q = wb; q < we; ++q) {
      wpre[q] = run;
      run += __popc(kbm[q]);
   }
   __syncthreads();
   const uint32_t obase = a.out_off[c];
   const uint32_t nx = a.len[0], ny = a.len[1];
   for (int q = t; q < nwords; q += IWG) {
      uint32_t x = kbm[q];
      uint32_t pos = wpre[q];
      const int row = q / W;
      const int yy = bb[1] + row % Ry, zz = bb[2] + row / Ry;
      const int xb = bb[0] + (q % W) * 32;
      while (x) {
         const int b = __ffs(x) - 1;
         x &= x - 1;
         a.gid_out[obase + pos] = (uint32_t)(xb + b) + (uint32_t)yy * nx + (uint32_t)zz * nx * ny;
         a.src_out[obase + pos] = obase + pos;
         ++pos;
      }
   }
   __syncthreads();
   const uint32_t nk = tot1[0];
   const int lane = t & 63;
   for (uint32_t b = t >> 6; b < nk; b += IWG / 64) {
      const uint32_t g = a.gid_out[obase + b];
      const int x = (int)(g % nx), y = (int)((g / nx) % ny), z = (int)(g / (nx * ny));
      a.f_out[(size_t)(obase + b) * WID3 + lane] = sample(a, pk, np, x, y, z, lane);
   }
}

void launch_init(const InitArgs& a, bool write_phase, hipStream_t s) {
   if (a.ncells == 0) return;
   if (write_phase) hipLaunchKernelGGL(k_init<true>, dim3(a.ncells), dim3(IWG), 0, s, a);
   else hipLaunchKernelGGL(k_init<false>, dim3(a.ncells), dim3(IWG), 0, s, a);
}

int init_box_words_max() { return IB_WORDS; }

}  // namespace vacc
