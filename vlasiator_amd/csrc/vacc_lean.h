// vacc_lean.h -- tolerance form of the remap arithmetic (options.fast_arith = 1), host +
// device: the PLM / PPM / PQM reconstructions rewritten for few VALU instructions, and the
// flux-form deposit of one pencil.  The device remap (k_remap_lean) and the CPU checks
// compile the same code.
//
// Reconstructions (cpu_1d_plm.hpp:38-49, cpu_1d_ppm.hpp:35-54, cpu_1d_pqm.hpp:31-173,
// cpu_face_estimates.hpp, cpu_slope_limiters.hpp) in deviations from the cell average v:
// u_l = (left face) - v, u_r = (right face) - v.  The reference's tests keep their meaning:
//   (v[-1] - f_l)(f_l - v) < 0   <=>  (b + u_l) u_l > 0      with b = v - v[-1]
//   (v[+1] - f_r)(f_r - v) < 0   <=>  (a - u_r) u_r < 0      with a = v[+1] - v
// and the PQM quartic F(x) = x (c0 + x (c1 + x (c2 + x (c3 + x c4)))) has
//   c0 = v + u_l, c1 = d_l / 2, c2 = -6 u_l - 4 u_r - (3 d_l - d_r) / 2,
//   c3 = 8 u_l + 7 u_r + (3 d_l - 2 d_r) / 2, c4 = -3 (u_l + u_r) + (d_r - d_l) / 2,
// whose monotonicity coefficients are b0 = 6 c2, b1 = 24 c3, b2 = 60 c4 (White & Adcroft
// 2008).  The critical points of F' are then (-2 c3 +- sqrt(4 c3^2 - 10 c2 c4)) / (10 c4) and
// F' = d_l + x (6 c2 + x (12 c3 + x 20 c4)).  The reference tests F' at both critical points
// (with a fallback root for a negative discriminant, sqrt -> b1 + 200 b2); here a Bernstein
// bound passes most cells without either, and the others test the one critical point where
// F' can first take the wrong sign (the same decision: see ReconLean<PQM>::coeffs).
// Multiplies are fused, 1/x and sqrt are the hardware's; results agree with the reference's
// FP32 rounding to a few ulp of the faces.
//
// Deposit (cpu_acc_map.cpp:533-628) in flux form: Lagrangian boundary g of the pencil lies at
// Imin + g dk, t is the next boundary's offset from the current source cell's left edge in
// units of dv, r = dk / dv.  A source cell closes every target whose right boundary it holds
// with F(t) - F(previous boundary in the cell) (plus what earlier cells gave that target) and
// leaves v - F(last boundary) in the open target: every source cell's mass is deposited once,
// with F(1) taken as the cell average v.  The reference evaluates F at both ends of every
// overlap instead, with the Lagrangian coordinates in FP32 (|v| up to 4e6 m/s, ulp 0.25 m/s,
// about 1e-5 of a cell).
#pragma once

#include <math.h>

#include "vacc_common.h"

namespace vacc {

VACC_HD float lean_rcp(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
   return __builtin_amdgcn_rcpf(x);
#else
   return 1.0f / x;
#endif
}
VACC_HD float lean_sqrt(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
   return __builtin_amdgcn_sqrtf(x);
#else
   return sqrtf(x);
#endif
}

// Face estimates with pre-divided weights (cpu_face_estimates.hpp:45-56, :68-78, :160-167).
VACC_HD float lean_h4(const float* v) {   // left face of v[0]
   return fmaf(7.0f / 12.0f, v[-1] + v[0], (-1.0f / 12.0f) * (v[-2] + v[1]));
}
VACC_HD float lean_h8(const float* v) {
   float s = (533.0f / 840.0f) * (v[-1] + v[0]);
   s = fmaf(-139.0f / 840.0f, v[-2] + v[1], s);
   s = fmaf(29.0f / 840.0f, v[-3] + v[2], s);
   return fmaf(-3.0f / 840.0f, v[-4] + v[3], s);
}
VACC_HD float lean_h7(const float* v) {
   float s = (7175.0f / 5040.0f) * (v[0] - v[-1]);
   s = fmaf(-889.0f / 5040.0f, v[1] - v[-2], s);
   s = fmaf(119.0f / 5040.0f, v[2] - v[-3], s);
   return fmaf(-9.0f / 5040.0f, v[3] - v[-4], s);
}

// Superbee (cpu_slope_limiters.hpp:30-54) of a = v[1] - v[0], b = v[0] - v[-1], with 2|b|
// carried from the previous cell: zero unless a and b share a sign, else
// max(min(|a|, 2|b|), min(2|a|, |b|)) with that sign.
VACC_HD float lean_superbee(float a, float b, float a2, float b2) {
   const float m = fmaxf(fminf(fabsf(a), b2), fminf(a2, fabsf(b)));
   return (a * b > 0.0f) ? copysignf(m, a) : 0.0f;
}

template <int METHOD>
struct ReconLean;

template <>
struct ReconLean<VACC_PLM> {
   static constexpr int NC = 2;
   float b, b2;
   VACC_HD void start(const float* v) {
      b = v[0] - v[-1];
      b2 = 2.0f * fabsf(b);
   }
   VACC_HD void coeffs(const float* v, float* c) {
      const float a = v[1] - v[0], a2 = 2.0f * fabsf(a);
      const float hd = 0.5f * lean_superbee(a, b, a2, b2);
      b = a;
      b2 = a2;
      c[0] = v[0] - hd;
      c[1] = hd;
   }
};

template <>
struct ReconLean<VACC_PPM> {
   static constexpr int NC = 3;
   float fl, b, b2;   // h4 left face of the current cell; v[0] - v[-1]; 2|v[0] - v[-1]|
   bool out;          // the left face lies outside [v[-1], v[0]] (the previous cell's right test)
   VACC_HD void start(const float* v) {
      fl = lean_h4(v);
      b = v[0] - v[-1];
      b2 = 2.0f * fabsf(b);
      out = (fl - v[-1]) * (fl - v[0]) > 0.0f;
   }
   VACC_HD void coeffs(const float* v, float* c) {
      const float vk = v[0];
      const float fr = lean_h4(v + 1);
      const float a = v[1] - vk, a2 = 2.0f * fabsf(a);
      const float hs = 0.5f * lean_superbee(a, b, a2, b2);
      float ul = fl - vk, ur = fr - vk;
      // flat cell (zero slope) or a face outside its neighbours: the face moves to v -+ slope/2.
      // The test of a face is the same product for both cells that share it, so the right
      // face's result serves the next cell's left face.
      const bool flat = (hs == 0.0f);
      const bool lout = out;
      out = (a - ur) * ur < 0.0f;
      if (flat || lout) ul = -hs;
      if (flat || out) ur = hs;
      // Colella-Woodward: (p - m)(v - (m + p)/2) > (p - m)^2 / 6  <=>  (u_r - u_l)(u_l + 2 u_r) < 0
      if ((ur - ul) * (ul + 2.0f * ur) < 0.0f) ul = -2.0f * ur;
      if ((ur - ul) * (2.0f * ul + ur) > 0.0f) ur = -2.0f * ul;
      fl = fr;
      b = a;
      b2 = a2;
      c[0] = vk + ul;
      c[1] = fmaf(-2.0f, ul, -ur);
      c[2] = ul + ur;
   }
};

template <>
struct ReconLean<VACC_PQM> {
   static constexpr int NC = 5;
   float fv, fd, b, b2;   // h8 value and h7 derivative of the left face; v[0] - v[-1]; 2|.|
   bool out;              // the left face lies outside [v[-1], v[0]] (the previous cell's right test)
   VACC_HD void start(const float* v) {
      fv = lean_h8(v);
      fd = lean_h7(v);
      b = v[0] - v[-1];
      b2 = 2.0f * fabsf(b);
      out = (fv - v[-1]) * (fv - v[0]) > 0.0f;
   }
   VACC_HD void coeffs(const float* v, float* c) {
      const float vk = v[0];
      const float fvr = lean_h8(v + 1), fdr = lean_h7(v + 1);
      const float a = v[1] - vk, a2 = 2.0f * fabsf(a);
      const float ds = lean_superbee(a, b, a2, b2);   // signed limited slope
      const float hs = 0.5f * ds;
      const float bo = b;
      float ul = fv - vk, ur = fvr - vk, dl = fd, dr = fdr;
      fv = fvr;
      fd = fdr;
      b = a;
      b2 = a2;
      // flat cell, a face outside its neighbours or a face derivative against the slope: PLM
      // face.  The outside test of a face is the same product for both cells that share it, so
      // the right face's result serves the next cell's left face.
      const bool flat = (ds == 0.0f);
      const bool lout = out;
      out = (a - ur) * ur < 0.0f;
      if (flat || lout || ds * dl < 0.0f) {
         ul = -hs;
         dl = ds;
      }
      if (flat || out || ds * dr < 0.0f) {
         ur = hs;
         dr = ds;
      }
      float c2 = fmaf(-6.0f, ul, fmaf(-4.0f, ur, 0.5f * fmaf(-3.0f, dl, dr)));
      float c3 = fmaf(8.0f, ul, fmaf(7.0f, ur, fmaf(1.5f, dl, -dr)));
      float c4 = fmaf(-3.0f, ul + ur, 0.5f * (dr - dl));
      // monotonicity: the quartic's derivative at its critical points inside [0, 1] must
      // have the sign of psl + psr = 2 (a + b).  F' = dl + x (6 c2 + x (12 c3 + x 20 c4)) has
      // the Bernstein coefficients dl, dl + 2 c2, dl + 4 (c2 + c3), dr on [0, 1]; the end ones
      // carry the sign of ssg (or are 0) after the face checks above.  When the inner two do
      // as well, so does F' on all of [0, 1] (it lies in their convex hull) and the cell
      // passes; otherwise the critical points are tested.
      const float ssg = a + bo;
      const float bn1 = fmaf(2.0f, c2, dl) * ssg, bn2 = fmaf(4.0f, c2 + c3, dl) * ssg;
      bool bad = false;
      if (!(fminf(bn1, bn2) >= 0.0f)) {
         // The critical points are x = (-2 c3 +- sq) / (10 c4), sq = sqrt(4 c3^2 - 10 c2 c4).
         // The third derivative there is +-12 sq, so the + root is the local minimum of F' and
         // the - root its local maximum.  With both ends of F' on the side of ssg, F' crosses
         // to the other side at a critical point only if it does at the local minimum of
         // ssg F': the + root for ssg > 0, the - root for ssg < 0 (a wrong-signed local maximum
         // of ssg F' inside [0, 1] has a lower local minimum between it and an end).  A
         // negative discriminant leaves F' monotone between its two well-signed ends, and the
         // reference's fallback root then finds nothing either.  A point outside [0, 1] is
         // moved to the nearer end, where F' has the side of ssg, in place of the range test.
         const float disc = fmaf(4.0f * c3, c3, -10.0f * (c2 * c4));
         const float x0 = (copysignf(lean_sqrt(fmaxf(disc, 0.0f)), ssg) - 2.0f * c3) * lean_rcp(10.0f * c4);
         const float x = fminf(fmaxf(x0, 0.0f), 1.0f);
         const float d = fmaf(x, fmaf(x, fmaf(x, 20.0f * c4, 12.0f * c3), 6.0f * c2), dl);
         bad = disc >= 0.0f && d * ssg < 0.0f;
      }
      if (bad) {
         // repair (cpu_1d_pqm.hpp:79-145) on the side of the smaller one-sided slope
         const bool lft = fabsf(bo) <= fabsf(a);
         const float dlx = lft ? (-2.0f / 3.0f) * fmaf(4.0f, ul, ur) : fmaf(-6.0f, ul, -4.0f * ur);
         const float drx = lft ? fmaf(4.0f, ul, 6.0f * ur) : (2.0f / 3.0f) * fmaf(4.0f, ur, ul);
         if (ssg * dlx < 0.0f) {          // left derivative flattened
            dl = 0.0f;
            // products first, then the choice: a choice between two constants would hold
            // both in registers for the whole kernel
            ur = lft ? -4.0f * ul : -1.5f * ul;
            dr = lft ? -20.0f * ul : (-10.0f / 3.0f) * ul;
         } else if (ssg * drx < 0.0f) {   // right derivative flattened
            dr = 0.0f;
            ul = lft ? -1.5f * ur : -4.0f * ur;
            dl = lft ? (10.0f / 3.0f) * ur : 20.0f * ur;
         } else {
            dl = dlx;
            dr = drx;
         }
         c2 = fmaf(-6.0f, ul, fmaf(-4.0f, ur, 0.5f * fmaf(-3.0f, dl, dr)));
         c3 = fmaf(8.0f, ul, fmaf(7.0f, ur, fmaf(1.5f, dl, -dr)));
         c4 = fmaf(-3.0f, ul + ur, 0.5f * (dr - dl));
      }
      c[0] = vk + ul;
      c[1] = 0.5f * dl;
      c[2] = c2;
      c[3] = c3;
      c[4] = c4;
   }
};

// F(x) = x (c0 + x (c1 + ...)), the integral of the reconstruction over [0, x] (units of dv).
template <int NC>
VACC_HD float lean_integral(float x, const float* c) {
   float s = c[NC - 1];
#pragma unroll
   for (int i = NC - 2; i >= 0; --i) s = fmaf(x, s, c[i]);
   return x * s;
}

// Flux-form deposit along one pencil (see the file comment).
template <int NC>
struct Flux {
   float t;     // next boundary, from the current source cell's left edge, in units of dv
   float r;     // dk / dv
   float acc;   // mass of the open target o from earlier source cells
   int o;       // the open target: its right edge is the next boundary

   // x0 = (v0 - Imin) / dk: the Lagrangian coordinate of the first source cell's left edge
   VACC_HD void begin(double x0, float r_) {
      const double fl = floor(x0);
      o = (int)fl;
      t = (float)((fl + 1.0 - x0) * (double)r_);
      r = r_;
      acc = 0.0f;
   }
   // One source cell with integral coefficients c and average v; emit(target, mass) receives
   // each target this cell closes, with the target's whole mass from this pencil.  The receiver
   // stores the emits: the first boundary is handled without a branch, and a cell that holds
   // none emits the open target's mass so far, which the target's closing emit supersedes.
   template <class Emit>
   VACC_HD void cell(const float* c, float v, Emit&& emit) {
      const bool hit = t < 1.0f;
      float Fp = lean_integral<NC>(hit ? t : 0.0f, c);   // F(0) = 0
      emit(o, acc + Fp);
      acc = hit ? 0.0f : acc;
      o += hit ? 1 : 0;
      t += hit ? r : 0.0f;
      while (t < 1.0f) {   // more boundaries in this cell (dk < dv)
         const float Ft = lean_integral<NC>(t, c);
         emit(o, Ft - Fp);
         Fp = Ft;
         ++o;
         t += r;
      }
      acc += v - Fp;
      t -= 1.0f;
   }
   // the open target after the pencil's last source cell
   template <class Emit>
   VACC_HD void end(Emit&& emit) {
      emit(o, acc);
   }
};

}  // namespace vacc
