// Host build of the device reconstruction code (vlasiator_amd/csrc/vacc_recon.h) for the
// CPU tests: streams a padded column through Recon<METHOD> exactly as the remap kernel does
// (start on the first source cell, then one coeffs() call per cell) and returns the
// coefficients of every source cell.
#include "../vlasiator_amd/csrc/vacc_recon.h"

template <int M>
static void run(const float* col, int n, float thr, float* out) {
   const float scale = (float)(1.0 / (double)thr);
   vacc::Recon<M> rc;
   rc.start(col + 4, scale);
   for (int k = 0; k < n; ++k) rc.coeffs(col + 4 + k, thr, scale, out + vacc::Recon<M>::NC * k);
}

// col: n + 8 values, the n source cells at col[4 .. n+3].
extern "C" int recon_stream(int method, const float* col, int n, float thr, float* out) {
   if (method == VACC_PLM) run<VACC_PLM>(col, n, thr, out);
   else if (method == VACC_PPM) run<VACC_PPM>(col, n, thr, out);
   else run<VACC_PQM>(col, n, thr, out);
   return 0;
}
