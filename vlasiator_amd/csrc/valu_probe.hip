// valu_probe -- VALU issue rate of the GPU for the arithmetic the kernels are made of.
//
// Each thread runs long chains of independent operations (eight per kind, no loads, no
// fused multiply-adds, no packed FP32) over a grid of 16 workgroups of 256 lanes per CU:
//   fp32     v_mul_f32 and v_add_f32
//   fp64mul  v_mul_f64
//   cvt+mul  (float)(double constant * (double)x): v_cvt_f64_f32, v_mul_f64, v_cvt_f32_f64,
//            the pattern of the reference's float-times-double expressions
// and prints, per kind, wave64 VALU instructions per second and lane operations per second.
// Divided by the clock (rocprofv3 GRBM_GUI_ACTIVE per XCD) and the 1024 SIMDs this is the
// issue rate per SIMD and cycle that SQ_INSTS_VALU of a kernel is compared with.
//
//     vlasiator_amd/lib/vacc_valu_probe
#include <hip/hip_runtime.h>

#include <cstdio>

namespace {

constexpr int PROBE_WG = 256;

template <int KIND>
__global__ __launch_bounds__(PROBE_WG) void k_valu_probe(float* out, int iters, float x, float y) {
   float a[8];
#pragma unroll
   for (int e = 0; e < 8; ++e) a[e] = (float)threadIdx.x * 1e-3f + (float)e;
   double d[4];
#pragma unroll
   for (int e = 0; e < 4; ++e) d[e] = (double)a[e];
   const double D = 0.9990000000000001;
   for (int i = 0; i < iters; ++i) {
      if (KIND == 0) {   // 16 instructions
#pragma unroll
         for (int e = 0; e < 8; ++e) a[e] = a[e] * x;
#pragma unroll
         for (int e = 0; e < 8; ++e) a[e] = a[e] + y;
      } else if (KIND == 1) {   // 8 instructions
#pragma unroll
         for (int e = 0; e < 4; ++e) d[e] = d[e] * (double)x;
#pragma unroll
         for (int e = 0; e < 4; ++e) d[e] = d[e] * (double)y;
      } else {   // 24 instructions
#pragma unroll
         for (int e = 0; e < 8; ++e) a[e] = (float)(D * (double)a[e]);
      }
   }
   float s = 0.0f;
#pragma unroll
   for (int e = 0; e < 8; ++e) s = s + a[e];
   out[blockIdx.x * PROBE_WG + threadIdx.x] = s + (float)((d[0] + d[1]) + (d[2] + d[3]));
}

template <int KIND>
bool run(const char* name, int per_iter, float* out, int blocks, int iters) {
   hipEvent_t e0, e1;
   if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return false;
   hipLaunchKernelGGL(k_valu_probe<KIND>, dim3(blocks), dim3(PROBE_WG), 0, 0, out, iters, 0.999f, 1e-7f);   // warm-up
   if (hipDeviceSynchronize() != hipSuccess) return false;
   (void)hipEventRecord(e0, 0);
   hipLaunchKernelGGL(k_valu_probe<KIND>, dim3(blocks), dim3(PROBE_WG), 0, 0, out, iters, 0.999f, 1e-7f);
   (void)hipEventRecord(e1, 0);
   if (hipEventSynchronize(e1) != hipSuccess) return false;
   float ms = 0.0f;
   (void)hipEventElapsedTime(&ms, e0, e1);
   const double instr = (double)blocks * (PROBE_WG / 64) * iters * per_iter;
   std::printf("{\"kind\": \"%s\", \"ms\": %.4f, \"wave64_instr\": %.6e, \"wave64_instr_per_s\": %.4e, "
               "\"lane_ops_per_s\": %.4e}\n",
               name, ms, instr, instr / (ms * 1e-3), instr * 64 / (ms * 1e-3));
   (void)hipEventDestroy(e0);
   (void)hipEventDestroy(e1);
   return true;
}

}  // namespace

int main() {
   int dev = 0, cus = 0;
   if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
      return 1;
   const int blocks = cus * 16;
   const int iters = 20000;
   float* out = nullptr;
   if (hipMalloc(&out, (size_t)blocks * PROBE_WG * sizeof(float)) != hipSuccess) return 1;
   const bool ok = run<0>("fp32", 16, out, blocks, iters) && run<1>("fp64mul", 8, out, blocks, iters) &&
                   run<2>("cvt+mul", 24, out, blocks, iters);
   (void)hipFree(out);
   return ok ? 0 : 1;
}
