// Host build of the device reconstruction code (vlasiator_amd/csrc/vacc_recon.h) for the
// CPU tests: streams a padded column through Recon<METHOD> exactly as the remap kernel does
// (start on the first source cell, then one coeffs() call per cell) and returns the
// coefficients of every source cell.
#include "../vlasiator_amd/csrc/vacc_recon.h"

template <class R>
static void run(const float* col, int n, float thr, float* out) {
   const float scale = (float)(1.0 / (double)thr);
   R rc;
   rc.start(col + 4, scale);
   for (int k = 0; k < n; ++k) rc.coeffs(col + 4 + k, thr, scale, out + R::NC * k);
}

// col: n + 8 values, the n source cells at col[4 .. n+3].
extern "C" int recon_stream(int method, const float* col, int n, float thr, float* out) {
   using namespace vacc;
   if (method == VACC_PLM) run<Recon<VACC_PLM>>(col, n, thr, out);
   else if (method == VACC_PPM) run<Recon<VACC_PPM>>(col, n, thr, out);
   else run<Recon<VACC_PQM>>(col, n, thr, out);
   return 0;
}

// Mismatches of vacc::div3 against the IEEE quotient x / 3 over the FP32 bit patterns
// [lo, hi) (non-finite patterns skipped); compares bits, so the sign of zero counts.
extern "C" long long div3_mismatches(unsigned long long lo, unsigned long long hi) {
   long long bad = 0;
   for (unsigned long long i = lo; i < hi; ++i) {
      const uint32_t u = (uint32_t)i;
      float x;
      __builtin_memcpy(&x, &u, 4);
      if (!isfinite(x)) continue;
      const float a = vacc::div3(x), b = x / 3.0f;
      uint32_t ua, ub;
      __builtin_memcpy(&ua, &a, 4);
      __builtin_memcpy(&ub, &b, 4);
      bad += (ua != ub);
   }
   return bad;
}

// vacc::lagr_index(x, dk, 1/dk) against (int)(x / dk) for n pairs; returns the mismatches.
extern "C" long long lagr_mismatches(const float* x, const float* dk, long long n) {
   long long bad = 0;
   for (long long i = 0; i < n; ++i) {
      const float rdk = 1.0f / dk[i];
      bad += (vacc::lagr_index(x[i], dk[i], rdk) != (int)(x[i] / dk[i]));
   }
   return bad;
}
