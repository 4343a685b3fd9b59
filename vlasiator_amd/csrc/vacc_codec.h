// vacc_codec.h -- 16-bit block storage of the distribution function.
//
// A block of the BF16 or FP16 form is 64 16-bit codes plus one FP32 scale per block;
// the value of cell e is  f_e = scale * decode(code_e)  (one FP32 multiply).  Encoding
// of a block x[64]:
//   m      = max_e |x_e| = mant * 2^p, mant in [0.5, 1)   (m = 0 -> all codes 0, scale 0)
//   code_e = round-to-nearest-even(x_e * 2^-p) in BF16 or FP16   (exact normalisation)
//   Sf     = csum(x), Sq = csum(decode(code))               (canonical FP32 sums, below)
//   scale  = Sf/Sq if Sf > 0, Sq > 0 and 2^(p-1) <= Sf/Sq <= 2^(p+1), else 2^p
// so that the decoded block carries the block's FP32 mass (sum_e f_e = Sf up to the
// rounding of the 64 products) while each value keeps the code's relative precision.
// csum(x): s_L = (x[4L] + x[4L+1]) + (x[4L+2] + x[4L+3]) for L = 0..15 (rows of fixed j,k),
// then s_L <- s_L + s_{L^m} for m = 8, 4, 2, 1.  The order is fixed so that any lane layout
// that holds the block produces the same bits.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace vacc {

enum { ST_F32 = 0, ST_BF16 = 1, ST_F16 = 2 };

template <int ST>
struct Codec;

template <>
struct Codec<ST_BF16> {
   __device__ static inline float dec(uint32_t u) { return __uint_as_float(u << 16); }
   __device__ static inline uint32_t enc(float x) {
      const uint32_t u = __float_as_uint(x);
      return (u + 0x7FFFu + ((u >> 16) & 1u)) >> 16;   // round to nearest even (finite x)
   }
};

template <>
struct Codec<ST_F16> {
   __device__ static inline float dec(uint32_t u) { return (float)__builtin_bit_cast(_Float16, (uint16_t)u); }
   __device__ static inline uint32_t enc(float x) { return __builtin_bit_cast(uint16_t, (_Float16)x); }
};

__device__ inline float dec16(int st, uint32_t u) {
   return (st == ST_BF16) ? Codec<ST_BF16>::dec(u) : Codec<ST_F16>::dec(u);
}

__device__ inline float pick_scale(float p, float sf, float sq) {
   const float r = (sq > 0.0f) ? sf / sq : 0.0f;
   return (sf > 0.0f && r >= 0.5f * p && r <= 2.0f * p) ? r : p;
}

// Encode one block held by 16 lanes, lane L holding the canonical row x[4] = f(i=0..3, j=L&3,
// k=L>>2).  All 16 lanes must be active.  Writes the lane's 4 codes and returns the block
// scale (equal in all 16 lanes).
template <int ST>
__device__ inline float encode_rows16(const float (&x)[4], uint32_t (&code)[4]) {
   float mx = fmaxf(fmaxf(fabsf(x[0]), fabsf(x[1])), fmaxf(fabsf(x[2]), fabsf(x[3])));
#pragma unroll
   for (int m = 8; m > 0; m >>= 1) mx = fmaxf(mx, __shfl_xor(mx, m, 16));
   if (!(mx > 0.0f)) {
      code[0] = code[1] = code[2] = code[3] = 0u;
      return 0.0f;
   }
   const int ex = __builtin_amdgcn_frexp_expf(mx);
   float d[4];
#pragma unroll
   for (int i = 0; i < 4; ++i) {
      code[i] = Codec<ST>::enc(__builtin_amdgcn_ldexpf(x[i], -ex));
      d[i] = Codec<ST>::dec(code[i]);
   }
   float sf = (x[0] + x[1]) + (x[2] + x[3]);
   float sq = (d[0] + d[1]) + (d[2] + d[3]);
#pragma unroll
   for (int m = 8; m > 0; m >>= 1) {
      sf = sf + __shfl_xor(sf, m, 16);
      sq = sq + __shfl_xor(sq, m, 16);
   }
   return pick_scale(__builtin_amdgcn_ldexpf(1.0f, ex), sf, sq);
}

// Encode one block held by a whole wave, lane e holding x = f_e (e = i + 4j + 16k).
template <int ST>
__device__ inline float encode_wave64(float x, uint32_t& code) {
   float mx = fabsf(x);
#pragma unroll
   for (int m = 32; m > 0; m >>= 1) mx = fmaxf(mx, __shfl_xor(mx, m, 64));
   if (!(mx > 0.0f)) {
      code = 0u;
      return 0.0f;
   }
   const int ex = __builtin_amdgcn_frexp_expf(mx);
   code = Codec<ST>::enc(__builtin_amdgcn_ldexpf(x, -ex));
   const float d = Codec<ST>::dec(code);
   float sf = x, sq = d;
   // rows: ((x0 + x1) + (x2 + x3)), then rows L ^ 8, 4, 2, 1 (lanes ^ 32, 16, 8, 4)
   sf = sf + __shfl_xor(sf, 1, 64);
   sq = sq + __shfl_xor(sq, 1, 64);
   sf = sf + __shfl_xor(sf, 2, 64);
   sq = sq + __shfl_xor(sq, 2, 64);
#pragma unroll
   for (int m = 32; m >= 4; m >>= 1) {
      sf = sf + __shfl_xor(sf, m, 64);
      sq = sq + __shfl_xor(sq, m, 64);
   }
   return pick_scale(__builtin_amdgcn_ldexpf(1.0f, ex), sf, sq);
}

}  // namespace vacc
