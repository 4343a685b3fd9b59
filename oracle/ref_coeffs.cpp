// ref_coeffs.cpp -- TEST INFRASTRUCTURE ONLY.
//
// Thin C-ABI driver around the reference's own reconstruction headers
// (vlasovsolver/cpu_1d_plm.hpp, cpu_1d_ppm.hpp, cpu_1d_pqm.hpp,
// cpu_face_estimates.hpp, cpu_slope_limiters.hpp, vec.h with
// VEC_FALLBACK_GENERIC and VECL=8).  The headers are compiled where they lie
// under the reference checkout (-I<reference root>); nothing of them is
// copied here.  The resulting oracle/_ref/libref_coeffs.so is used by the
// tests to pin the oracle's restatement of these functions bit for bit, and
// to write the golden vectors under tests/golden/.
#include "vlasovsolver/vec.h"
#include "vlasovsolver/cpu_1d_plm.hpp"
#include "vlasovsolver/cpu_1d_ppm.hpp"
#include "vlasovsolver/cpu_1d_pqm.hpp"
#include "vlasovsolver/cpu_1d_ppm_nonuniform.hpp"

extern "C" {

// stencils: n x 9 floats, the source cell at index 4.  out: n x (2|3|5).
// method 0 = PLM, 1 = PPM (h4), 2 = PQM (h8), as selected in cpu_acc_map.cpp:537-548.
int ref_coeff_batch(int method, const float* stencils, int n, float threshold, float* out) {
   const int nc = (method == 0) ? 2 : (method == 1) ? 3 : 5;
   for (int base = 0; base < n; base += VECL) {
      Vec values[9];
      for (int t = 0; t < 9; ++t) {
         for (int l = 0; l < VECL; ++l) {
            const int s = (base + l < n) ? base + l : base;
            values[t].insert(l, stencils[9 * s + t]);
         }
      }
      Vec a[5];
      if (method == 0) compute_plm_coeff(values, 4, a, threshold);
      else if (method == 1) compute_ppm_coeff(values, h4, 4, a, threshold);
      else compute_pqm_coeff(values, h8, 4, a, threshold);
      for (int l = 0; l < VECL && base + l < n; ++l)
         for (int c = 0; c < nc; ++c) out[nc * (base + l) + c] = a[c][l];
   }
   return VECL;
}

// PPM on non-uniform cell widths, as the spatial translation calls it
// (cpu_trans_map_amr.cpp:100-103: compute_ppm_coeff_nonuniform(dz, values, h4, 2, a, thr)).
// h5: the 5 widths of the stencil (shared by every lane of a Vec, as in the reference);
// stencils: n x 5 values, the source cell at index 2.  out: n x 3.
int ref_ppm_nu_batch(const float* h5, const float* stencils, int n, float threshold, float* out) {
   Realf h[5];
   for (int t = 0; t < 5; ++t) h[t] = h5[t];
   for (int base = 0; base < n; base += VECL) {
      Vec values[5];
      for (int t = 0; t < 5; ++t) {
         for (int l = 0; l < VECL; ++l) {
            const int s = (base + l < n) ? base + l : base;
            values[t].insert(l, stencils[5 * s + t]);
         }
      }
      Vec a[3];
      compute_ppm_coeff_nonuniform(h, values, h4, 2, a, threshold);
      for (int l = 0; l < VECL && base + l < n; ++l)
         for (int c = 0; c < 3; ++c) out[3 * (base + l) + c] = a[c][l];
   }
   return VECL;
}

}
