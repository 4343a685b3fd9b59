// vacc_recon.h -- PLM / PPM / PQM reconstruction of one pencil, streamed source cell by
// source cell (host + device: the device remap and the host-side checks compile the
// same code).
#pragma once

#include <math.h>

#include "vacc_common.h"

namespace vacc {

// ---------------------------------------------------------------------------
// Reconstructions along a pencil, streamed source cell by source cell.
// cpu_face_estimates.hpp, cpu_slope_limiters.hpp, cpu_1d_plm/ppm/pqm.hpp
//
// Every operation rounds exactly as the reference's: a float times a float-representable
// double literal is written as a float multiply (the double product of two floats is
// exact, so rounding it to float equals the float product); x/3.0 in double rounded to
// float equals the float quotient (53 >= 2*24+2 bits, so double rounding is innocuous);
// literals such as 1/12, 1/840, 1/5040, 1/3 and 10/3 keep the double multiply.
// Face values and derivatives of the face between source cells k and k+1 are computed
// once and reused as the left face of k+1; scaled limiter inputs are reused likewise.
// ---------------------------------------------------------------------------

VACC_HD float face_h4(const float* v) {
   // left face of cell v[0], cpu_face_estimates.hpp:160-167
   float s = -v[-2];
   s = s + 7.0f * v[-1];
   s = s + 7.0f * v[0];
   s = s - v[1];
   return (float)((1.0 / 12.0) * (double)s);
}
VACC_HD float face_h8(const float* v) {
   // cpu_face_estimates.hpp:45-56
   float s = -3.0f * v[-4];
   s = s + 29.0f * v[-3];
   s = s - 139.0f * v[-2];
   s = s + 533.0f * v[-1];
   s = s + 533.0f * v[0];
   s = s - 139.0f * v[1];
   s = s + 29.0f * v[2];
   s = s - 3.0f * v[3];
   return (float)((1.0 / 840.0) * (double)s);
}
VACC_HD float deriv_h7(const float* v) {
   // cpu_face_estimates.hpp:68-78
   float s = 9.0f * v[-4];
   s = s - 119.0f * v[-3];
   s = s + 889.0f * v[-2];
   s = s - 7175.0f * v[-1];
   s = s + 7175.0f * v[0];
   s = s - 889.0f * v[1];
   s = s + 119.0f * v[2];
   s = s - 9.0f * v[3];
   return (float)((1.0 / 5040.0) * (double)s);
}
VACC_HD float mod_min(float s1, float s2) {
   // cpu_slope_limiters.hpp:30-35
   const float slope = (fabsf(s1) < fabsf(s2)) ? s1 : s2;
   const float p = s1 * s2;
   return (p <= 0.0f) ? 0.0f : slope;
}
VACC_HD float mod_max(float s1, float s2) {
   // cpu_slope_limiters.hpp:36-41
   const float slope = (fabsf(s1) > fabsf(s2)) ? s1 : s2;
   const float p = s1 * s2;
   return (p <= 0.0f) ? 0.0f : slope;
}
// superbee on a = r - m, b = m - l, cpu_slope_limiters.hpp:47-54
VACC_HD float superbee_ab(float a, float b) {
   return mod_max(mod_min(a, 2.0f * b), mod_min(2.0f * a, b));
}

// x / 3 with IEEE rounding from a multiply by RN(1/3) and one FMA residual correction
// (checked against the division for every finite FP32 x; a zero residual returns the
// product, which also keeps the sign of a zero x).
VACC_HD float div3(float x) {
   const float r3 = 1.0f / 3.0f;
   const float q0 = x * r3;
   const float e = fmaf(-q0, 3.0f, x);
   return (e == 0.0f) ? q0 : fmaf(e, r3, q0);
}

// (int)(x / dk), the truncated IEEE quotient (Lagrangian cell index, cpu_acc_map.cpp:505-508,
// :564), from q = x * rdk with rdk = RN(1/dk).  q is within 2^-22.4 |q| of the IEEE quotient,
// so when no integer lies within 2^-21 |q| of q both truncate alike; otherwise divide.
VACC_HD int lagr_index(float x, float dk, float rdk) {
   const float q = x * rdk;
   const float t = truncf(q);
   const float fr = fabsf(q - t);
   const float d = fminf(fr, 1.0f - fr);
   if (d > fabsf(q) * 0x1p-21f) return (int)t;
   return (int)(x / dk);
}

template <int METHOD>
struct Recon;

// PLM, cpu_1d_plm.hpp:38-49
template <>
struct Recon<VACC_PLM> {
   static constexpr int NC = 2;
   float sm, s0, b;   // scaled v[k-1], v[k]; b = s0 - sm
   VACC_HD void start(const float* v, float scale) {
      sm = v[-1] * scale;
      s0 = v[0] * scale;
      b = s0 - sm;
   }
   VACC_HD void coeffs(const float* v, float thr, float scale, float* c) {
      const float s1 = v[1] * scale;
      const float a = s1 - s0;
      const float d = superbee_ab(a, b) * thr;
      c[0] = v[0] - 0.5f * d;
      c[1] = 0.5f * d;
      sm = s0;
      s0 = s1;
      b = a;
   }
};

// PPM with h4 faces, cpu_1d_ppm.hpp:35-54 and cpu_face_estimates.hpp:278-323
template <>
struct Recon<VACC_PPM> {
   static constexpr int NC = 3;
   float fl_raw;      // h4 estimate of the left face of the current cell
   float s0, b;       // scaled v[k]; b = s0 - scaled v[k-1]
   bool out;          // (v[-1] - fl_raw)(fl_raw - v[0]) < 0
   VACC_HD void start(const float* v, float scale) {
      fl_raw = face_h4(v);
      s0 = v[0] * scale;
      b = s0 - v[-1] * scale;
      out = (v[-1] - fl_raw) * (fl_raw - v[0]) < 0.0f;
   }
   VACC_HD void coeffs(const float* v, float thr, float scale, float* c) {
      const float fr_raw = face_h4(v + 1);
      const float s1 = v[1] * scale;
      const float a = s1 - s0;
      const float sl = superbee_ab(a, b);
      const float sabs = fabsf(sl) * thr;
      const float ssgn = (sl > 0.0f) ? 1.0f : -1.0f;
      float m = fl_raw, p = fr_raw;
      // The face tests (v[-1] - m)(m - v[0]) < 0 and (v[1] - p)(p - v[0]) < 0 of a face shared
      // by two cells are the same product (both factors negated, exactly), so the right face's
      // test is carried to the next cell's left face.  A flat cell tests v[0] instead (product
      // +-0, never below 0).
      const bool flat = (sabs == 0.0f);
      const bool lout = !flat && out;
      out = (v[1] - fr_raw) * (fr_raw - v[0]) < 0.0f;
      if (flat) {
         p = v[0];
         m = v[0];
      }
      if (lout) m = v[0] - (0.5f * ssgn) * sabs;
      if (!flat && out) p = v[0] + (0.5f * ssgn) * sabs;
      const float one_sixth = (float)(1.0 / 6.0);
      {
         const float d = p - m;
         const float mid = v[0] - 0.5f * (m + p);
         if (d * mid > (d * d) * one_sixth) m = 3.0f * v[0] - 2.0f * p;
      }
      {
         const float d = p - m;
         const float mid = v[0] - 0.5f * (m + p);
         if (((-d) * d) * one_sixth > d * mid) p = 3.0f * v[0] - 2.0f * m;
      }
      c[0] = m;
      c[1] = (3.0f * v[0] - 2.0f * m) - p;
      c[2] = (m + p) - 2.0f * v[0];
      fl_raw = fr_raw;
      s0 = s1;
      b = a;
   }
};

// PQM with h8 faces and h7 derivatives, cpu_1d_pqm.hpp:31-173 and cpu_face_estimates.hpp:215-271
template <>
struct Recon<VACC_PQM> {
   static constexpr int NC = 5;
   float fvl_raw, fdl_raw;   // h8 / h7 estimates at the left face of the current cell
   float s0, b;
   bool out;                 // (v[-1] - fvl_raw)(fvl_raw - v[0]) < 0
   VACC_HD void start(const float* v, float scale) {
      fvl_raw = face_h8(v);
      fdl_raw = deriv_h7(v);
      s0 = v[0] * scale;
      b = s0 - v[-1] * scale;
      out = (v[-1] - fvl_raw) * (fvl_raw - v[0]) < 0.0f;
   }
   VACC_HD void coeffs(const float* v, float thr, float scale, float* c) {
      const float fvr_raw = face_h8(v + 1), fdr_raw = deriv_h7(v + 1);
      const float s1 = v[1] * scale;
      const float a = s1 - s0;
      const float sl = superbee_ab(a, b);
      const float sabs = fabsf(sl) * thr;
      const float ssgn = (sl > 0.0f) ? 1.0f : -1.0f;
      const float vk = v[0];
      float fvl = fvl_raw, fvr = fvr_raw, fdl = fdl_raw, fdr = fdr_raw;
      fvl_raw = fvr_raw;
      fdl_raw = fdr_raw;
      s0 = s1;
      b = a;
      // the face tests are carried from cell to cell as in Recon<VACC_PPM>
      const bool flat = (sabs == 0.0f);
      const bool lout = !flat && out;
      out = (v[1] - fvr_raw) * (fvr_raw - vk) < 0.0f;
      if (flat) {
         fvr = vk;
         fvl = vk;
         fdl = 0.0f;
         fdr = 0.0f;
      }
      if (lout || ssgn * fdl < 0.0f) {
         fvl = vk - (0.5f * ssgn) * sabs;
         fdl = ssgn * sabs;
      }
      if ((!flat && out) || ssgn * fdr < 0.0f) {
         fvr = vk + (0.5f * ssgn) * sabs;
         fdr = ssgn * sabs;
      }
      // monotonicity filter (White & Adcroft 2008 eqs 21, 23, 28, 29)
      const float b0 = ((60.0f * vk - 24.0f * fvr) - 36.0f * fvl) + 3.0f * (fdr - 3.0f * fdl);
      const float b1 = (((-360.0f * vk + 36.0f * fdl) - 24.0f * fdr) + 168.0f * fvr) + 192.0f * fvl;
      const float b2 = (360.0f * vk + 30.0f * (fdr - fdl)) - 180.0f * (fvl + fvr);
      const float disc = (b1 * b1) - (4.0f * b0) * b2;
      const float sq = (disc < 0.0f) ? (b1 + 200.0f * b2) : sqrtf(disc);
      const float tb2 = 2.0f * b2;
      const float root1 = ((-b1) + sq) / tb2;
      const float root2 = ((-b1) - sq) / tb2;
      const float psl = 2.0f * (vk - v[-1]);
      const float psr = 2.0f * (v[1] - vk);
      const float ssg = psl + psr;
      const float c1 = b0;
      const float c2 = 0.5f * b1;
      const float c3 = div3(b2);
      const float r1s = (root1 >= 0.0f && root1 <= 1.0f) ? (fdl + root1 * (c1 + root1 * (c2 + root1 * c3))) : ssg;
      const float r2s = (root2 >= 0.0f && root2 <= 1.0f) ? (fdl + root2 * (c1 + root2 * (c2 + root2 * c3))) : ssg;
      if (r1s * ssg < 0.0f || r2s * ssg < 0.0f) {
         if (fabsf(psl) <= fabsf(psr)) {
            fdl = (float)((1.0 / 3.0) * (((double)(10.0f * vk) - 2.0 * (double)fvr) - 8.0 * (double)fvl));
            fdr = (float)((-10.0 * (double)vk + 6.0 * (double)fvr) + 4.0 * (double)fvl);
            if (ssg * fdl < 0.0f) {
               fdl = 0.0f;
               fvr = 5.0f * vk - 4.0f * fvl;
               fdr = 20.0f * (vk - fvl);
            } else if (ssg * fdr < 0.0f) {
               fdr = 0.0f;
               fvl = 0.5f * (5.0f * vk - 3.0f * fvr);
               fdl = (float)((10.0 / 3.0) * (double)(-vk + fvr));
            }
         } else {
            fdl = (float)((10.0 * (double)vk - 6.0 * (double)fvl) - 4.0 * (double)fvr);
            fdr = (float)((1.0 / 3.0) * ((-10.0 * (double)vk + (double)(2.0f * fvl)) + (double)(8.0f * fvr)));
            if (ssg * fdl < 0.0f) {
               fdl = 0.0f;
               fvr = 0.5f * (5.0f * vk - 3.0f * fvl);
               fdr = (float)((10.0 / 3.0) * (double)(vk - fvl));
            } else if (ssg * fdr < 0.0f) {
               fdr = 0.0f;
               fvl = 5.0f * vk - 4.0f * fvr;
               fdl = 20.0f * (-vk + fvr);
            }
         }
      }
      c[0] = fvl;
      c[1] = 0.5f * fdl;
      c[2] = ((10.0f * vk - 4.0f * fvr) - 6.0f * fvl) + 0.5f * (fdr - 3.0f * fdl);
      c[3] = (((-15.0f * vk + 1.5f * fdl) - fdr) + 7.0f * fvr) + 8.0f * fvl;
      c[4] = (6.0f * vk + 0.5f * (fdr - fdl)) - 3.0f * (fvl + fvr);
   }
};

template <int METHOD>
VACC_HD float integral(float x, const float* c) {
   // cpu_acc_map.cpp:590-602
   if (METHOD == VACC_PLM) return x * (c[0] + x * c[1]);
   if (METHOD == VACC_PPM) return x * (c[0] + x * (c[1] + x * c[2]));
   return x * (c[0] + x * (c[1] + x * (c[2] + x * (c[3] + x * c[4]))));
}


}  // namespace vacc
