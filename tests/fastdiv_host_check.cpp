// Host build of the GID splitting of vlasiator_amd/csrc/vacc_common.h (FastDiv, gid_xyz), the
// same code the device kernels compile, for the CPU tests.
#include "../vlasiator_amd/csrc/vacc_common.h"

extern "C" void fastdiv_batch(uint32_t d, const uint32_t* g, int n, uint32_t* q) {
   const vacc::FastDiv f = vacc::make_fastdiv(d);
   for (int i = 0; i < n; ++i) q[i] = vacc::fdiv(g[i], f);
}

extern "C" void gid_xyz_batch(const uint32_t* len, const uint32_t* g, int n, uint32_t* xyz) {
   const vacc::GidSplit s = vacc::make_gid_split(len);
   for (int i = 0; i < n; ++i) vacc::gid_xyz(g[i], s, xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
}
