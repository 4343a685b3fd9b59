// Host build of the tolerance-form remap arithmetic (vlasiator_amd/csrc/vacc_lean.h) for the
// CPU tests: the lean reconstructions of single stencils, and one column of one pencil
// mapped with the flux-form deposit exactly as k_remap_lean does.
#include "../vlasiator_amd/csrc/vacc_lean.h"

template <int M>
static void coeffs(const float* st, int n, float* out) {
   using R = vacc::ReconLean<M>;
   for (int i = 0; i < n; ++i) {
      R rc;
      rc.start(st + 9 * i + 4);
      rc.coeffs(st + 9 * i + 4, out + R::NC * i);
   }
}

// stencils of 9 values (the cell at 4) -> NC coefficients each
extern "C" int lean_coeff_batch(int method, const float* st, int n, float* out) {
   if (method == VACC_PLM) coeffs<VACC_PLM>(st, n, out);
   else if (method == VACC_PPM) coeffs<VACC_PPM>(st, n, out);
   else coeffs<VACC_PQM>(st, n, out);
   return 0;
}

template <int M>
static void column(const float* col, int n, double x0, float r, int lo, int hi, float* out, int* span) {
   using R = vacc::ReconLean<M>;
   R rc;
   vacc::Flux<R::NC> fx;
   fx.begin(x0, r);
   // stores, as k_remap_lean does: each target's closing emit carries its whole mass; span
   // receives the lowest and highest target written
   span[0] = hi + 1;
   span[1] = lo - 1;
   auto emit = [&](int g, float m) {
      if (g >= lo && g <= hi) {
         out[g - lo] = m;
         span[0] = g < span[0] ? g : span[0];
         span[1] = g > span[1] ? g : span[1];
      }
   };
   rc.start(col + 4);
   float c[R::NC];
   for (int k = 0; k < n; ++k) {
      rc.coeffs(col + 4 + k, c);
      fx.cell(c, col[4 + k], emit);
   }
   fx.end(emit);
}

// col: n + 8 values, the column's n source cells at col[4 .. n+3] with 4 zero cells each side;
// x0: Lagrangian coordinate of the first source cell's left edge; r = dk / dv.  The masses of
// targets lo .. hi are stored into out[0 .. hi-lo]; span[0 .. 1] = lowest and highest target
// written (span[0] > span[1] if none).
extern "C" int lean_column(int method, const float* col, int n, double x0, float r, int lo, int hi, float* out,
                           int* span) {
   if (method == VACC_PLM) column<VACC_PLM>(col, n, x0, r, lo, hi, out, span);
   else if (method == VACC_PPM) column<VACC_PPM>(col, n, x0, r, lo, hi, out, span);
   else column<VACC_PQM>(col, n, x0, r, lo, hi, out, span);
   return 0;
}
