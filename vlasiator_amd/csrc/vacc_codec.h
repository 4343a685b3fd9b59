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
// The tolerance form's BF16 encoder takes p = 0 (codes RN(x_e), scale Sf/Sq near 1, else 1):
// BF16 has the exponent range of FP32 and scaling by 2^-p commutes with the rounding, so
// the decoded values are those of the normalised encoding, without the block maximum.
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

// Two codes packed in one word (first value in the low half), as the block is stored.
template <int ST>
struct Pack2;

template <>
struct Pack2<ST_BF16> {
   typedef float f2 __attribute__((ext_vector_type(2)));
   typedef __bf16 b2 __attribute__((ext_vector_type(2)));
   // v_cvt_pk_bf16_f32: round to nearest even, the rounding of Codec<ST_BF16>::enc for every
   // finite FP32 value
   __device__ static inline uint32_t enc(float a, float b) {
      const f2 in = {a, b};
      return __builtin_bit_cast(uint32_t, __builtin_convertvector(in, b2));
   }
   __device__ static inline float lo(uint32_t p) { return __uint_as_float(p << 16); }
   __device__ static inline float hi(uint32_t p) { return __uint_as_float(p & 0xFFFF0000u); }
};

template <>
struct Pack2<ST_F16> {
   __device__ static inline uint32_t enc(float a, float b) {
      return Codec<ST_F16>::enc(a) | (Codec<ST_F16>::enc(b) << 16);
   }
   __device__ static inline float lo(uint32_t p) { return Codec<ST_F16>::dec(p & 0xFFFFu); }
   __device__ static inline float hi(uint32_t p) { return Codec<ST_F16>::dec(p >> 16); }
};

// DPP row rotation (row_ror:N, N = 1..15) within the 16 lanes of the caller's row.  A rotation
// reads no lane outside the row, and callers have all 16 lanes of the row active, so the
// bound_ctrl setting changes no value; with it set the compiler folds the rotation into the
// instruction that uses it (v_add_f32_dpp, v_max_u32_dpp, ...).
template <int N>
__device__ inline uint32_t row_ror_u(uint32_t x) {
   return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x120 + N, 0xF, 0xF, true);
}
template <int N>
__device__ inline float row_ror(float x) {
   return __builtin_bit_cast(float, row_ror_u<N>(__builtin_bit_cast(uint32_t, x)));
}

// Reductions over the 16 lanes of a DPP row (lanes 16r .. 16r+15, all active).  Rotating by
// 8, 4, 2, 1 gives lane L the value of lane L^m up to bits already combined, so the sum has
// the canonical butterfly order s_L + s_{L^m}, m = 8, 4, 2, 1, bit for bit (a + b = b + a).
__device__ inline float row_sum16(float s) {
   s = s + row_ror<8>(s);
   s = s + row_ror<4>(s);
   s = s + row_ror<2>(s);
   return s + row_ror<1>(s);
}
// DPP quad permutation within each group of 4 lanes (quad_perm [1,0,3,2]: lane ^ 1,
// [2,3,0,1]: lane ^ 2); as with the rotations, all lanes of the row are active.
template <int CTRL>
__device__ inline float quad_perm(float x) {
   return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, x), CTRL, 0xF, 0xF, true));
}

// Totals of four quantities a, b, c, d over the 16 lanes of a DPP row: lane L returns the
// total of quantity L & 3 (a, b, c, d).  Lanes L and L ^ 2 first exchange the pair the other
// keeps ((a, b) for bit 1 clear), lanes L and L ^ 1 then the value the other keeps, and the
// four quads are summed by rotations of 8 and 4: 5 additions instead of the 16 of four
// separate butterflies (summation order not canonical).
__device__ inline float row_sum4x16(float a, float b, float c, float d, int l16) {
   const bool b1 = (l16 & 2) != 0, b0 = (l16 & 1) != 0;
   float k0 = b1 ? c : a, k1 = b1 ? d : b;
   const float s0 = b1 ? a : c, s1 = b1 ? b : d;
   k0 = k0 + quad_perm<0x4E>(s0);
   k1 = k1 + quad_perm<0x4E>(s1);
   float m = b0 ? k1 : k0;
   const float s = b0 ? k0 : k1;
   m = m + quad_perm<0xB1>(s);
   m = m + row_ror<8>(m);
   return m + row_ror<4>(m);
}

// maximum of non-negative finite values: their bit patterns order as unsigned integers
__device__ inline float row_maxf16(float m) {
   uint32_t u = __builtin_bit_cast(uint32_t, m);
   u = max(u, row_ror_u<8>(u));
   u = max(u, row_ror_u<4>(u));
   u = max(u, row_ror_u<2>(u));
   return __builtin_bit_cast(float, max(u, row_ror_u<1>(u)));
}

__device__ inline float dec16(int st, uint32_t u) {
   return (st == ST_BF16) ? Codec<ST_BF16>::dec(u) : Codec<ST_F16>::dec(u);
}

// FAST: the quotient Sf/Sq from the hardware reciprocal (within 2 ulp of the IEEE quotient),
// for the tolerance form of the remap; otherwise the IEEE division.
template <bool FAST = false>
__device__ inline float pick_scale(float p, float sf, float sq) {
   const float q = FAST ? sf * __builtin_amdgcn_rcpf(sq) : sf / sq;
   const float r = (sq > 0.0f) ? q : 0.0f;
   return (sf > 0.0f && r >= 0.5f * p && r <= 2.0f * p) ? r : p;
}

// Encode one block held by the 16 lanes of a DPP row, lane L holding the canonical row
// x[4] = f(i=0..3, j=L&3, k=L>>2).  All 16 lanes of the row must be active.  Writes the lane's
// 4 codes packed as stored (x[0] in the low half of p.x) and returns the block scale (equal in
// all 16 lanes).
template <int ST, bool FAST = false>
// qmax: the largest of the lane's 4 decoded codes (0 for an all-zero block).  With the block
// scale sc >= 0, the largest stored value max_e RN(sc * d_e) is RN(sc * qmax), rounding being
// monotone, so one product gives the block's content test.
__device__ inline float encode_rows16(const float (&x)[4], uint2& p, float& qmax) {
   if constexpr (FAST && ST == ST_BF16) {   // p = 0 (above)
      p.x = Pack2<ST>::enc(x[0], x[1]);
      p.y = Pack2<ST>::enc(x[2], x[3]);
      const float q0 = Pack2<ST>::lo(p.x), q1 = Pack2<ST>::hi(p.x), q2 = Pack2<ST>::lo(p.y), q3 = Pack2<ST>::hi(p.y);
      qmax = fmaxf(fmaxf(q0, q1), fmaxf(q2, q3));
      const float sf = row_sum16((x[0] + x[1]) + (x[2] + x[3]));
      const float sq = row_sum16((q0 + q1) + (q2 + q3));
      return pick_scale<FAST>(1.0f, sf, sq);
   }
   float mx = fmaxf(fmaxf(fabsf(x[0]), fabsf(x[1])), fmaxf(fabsf(x[2]), fabsf(x[3])));
   mx = row_maxf16(mx);
   if (!(mx > 0.0f)) {   // the same in all lanes of the row
      p = make_uint2(0u, 0u);
      qmax = 0.0f;
      return 0.0f;
   }
   const int ex = __builtin_amdgcn_frexp_expf(mx);
   p.x = Pack2<ST>::enc(__builtin_amdgcn_ldexpf(x[0], -ex), __builtin_amdgcn_ldexpf(x[1], -ex));
   p.y = Pack2<ST>::enc(__builtin_amdgcn_ldexpf(x[2], -ex), __builtin_amdgcn_ldexpf(x[3], -ex));
   const float q0 = Pack2<ST>::lo(p.x), q1 = Pack2<ST>::hi(p.x), q2 = Pack2<ST>::lo(p.y), q3 = Pack2<ST>::hi(p.y);
   qmax = fmaxf(fmaxf(q0, q1), fmaxf(q2, q3));
   const float sf = row_sum16((x[0] + x[1]) + (x[2] + x[3]));
   const float sq = row_sum16((q0 + q1) + (q2 + q3));
   return pick_scale<FAST>(__builtin_amdgcn_ldexpf(1.0f, ex), sf, sq);
}

// Largest decoded code of a lane's 4 packed codes (the content test of a stored block, as above).
template <int ST>
__device__ inline float max_code(const uint2& p) {
   return fmaxf(fmaxf(Pack2<ST>::lo(p.x), Pack2<ST>::hi(p.x)), fmaxf(Pack2<ST>::lo(p.y), Pack2<ST>::hi(p.y)));
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
