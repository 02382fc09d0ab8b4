// Matrix-core throughput probe: the sustained rate of the FP16 MFMA shapes
// the mixed-precision kernels could use, measured on the device at hand, so
// that a kernel's issued MFMA rate can be set beside the rate its
// instruction actually sustains (the published dense peak assumes the
// fastest shape).  Each wave runs independent accumulator chains of one
// instruction; the grid puts two, four or eight waves on every SIMD and the best
// rate is reported.  The 16x16x32 form runs eight chains a wave, as the
// back-projection does (two site tiles by four kinds).
#include "mgpu_internal.h"

#include <algorithm>

namespace {

typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef float floatx16 __attribute__((ext_vector_type(16)));

// v_mfma_f32_16x16x32_f16, eight chains per wave
__global__ __launch_bounds__(256) void k_rate16(float *out, int iters) {
   const int lane = threadIdx.x & 63;
   half8 a, b;
   for (int i = 0; i < 8; i++) {
      a[i] = (_Float16)(0.001f * (lane + i));
      b[i] = (_Float16)(0.002f * (i - lane));
   }
   floatx4 acc[8];
   for (int q = 0; q < 8; q++) acc[q] = floatx4{0, 0, 0, 0};
   for (int it = 0; it < iters; it++) {
#pragma unroll
      for (int q = 0; q < 8; q++) acc[q] = __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, acc[q], 0, 0, 0);
   }
   float t = 0;
   for (int q = 0; q < 8; q++) t += acc[q][0] + acc[q][1] + acc[q][2] + acc[q][3];
   out[blockIdx.x * blockDim.x + threadIdx.x] = t;
}

// v_mfma_f32_32x32x16_f16, four chains per wave
__global__ __launch_bounds__(256) void k_rate32(float *out, int iters) {
   const int lane = threadIdx.x & 63;
   half8 a, b;
   for (int i = 0; i < 8; i++) {
      a[i] = (_Float16)(0.001f * (lane + i));
      b[i] = (_Float16)(0.002f * (i - lane));
   }
   floatx16 acc[4];
   for (int q = 0; q < 4; q++)
      for (int r = 0; r < 16; r++) acc[q][r] = 0;
   for (int it = 0; it < iters; it++) {
#pragma unroll
      for (int q = 0; q < 4; q++) acc[q] = __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, acc[q], 0, 0, 0);
   }
   float t = 0;
   for (int q = 0; q < 4; q++)
      for (int r = 0; r < 16; r++) t += acc[q][r];
   out[blockIdx.x * blockDim.x + threadIdx.x] = t;
}

}  // namespace

extern "C" int mgpu_mfma_rate(int device, int shape, double *tflops) {
   if (!tflops || (shape != 16 && shape != 32)) return MGPU_E_ARG;
   try {
      HIPCHK(hipSetDevice(device));
      hipDeviceProp_t prop;
      HIPCHK(hipGetDeviceProperties(&prop, device));
      const int ncu = prop.multiProcessorCount, iters = 2000;
      const int maxblk = 8 * ncu;
      float *d = nullptr;
      HIPCHK(hipMalloc(&d, sizeof(float) * 256 * (size_t)maxblk));
      hipEvent_t e0, e1;
      HIPCHK(hipEventCreate(&e0));
      HIPCHK(hipEventCreate(&e1));
      double best = 0;
      for (int wps = 2; wps <= 8; wps *= 2) {   // blocks of four waves: wps waves on every SIMD
         const int blocks = wps * ncu;
         for (int rep = 0; rep < 2; rep++) {   // the first launch warms the clocks
            HIPCHK(hipEventRecord(e0, nullptr));
            if (shape == 16) k_rate16<<<blocks, 256>>>(d, iters);
            else k_rate32<<<blocks, 256>>>(d, iters);
            HIPCHK(hipGetLastError());
            HIPCHK(hipEventRecord(e1, nullptr));
            HIPCHK(hipEventSynchronize(e1));
            float ms = 0;
            HIPCHK(hipEventElapsedTime(&ms, e0, e1));
            // flops per instruction (2 M N K, both shapes 16384 or 32768)
            const double per = shape == 16 ? 2.0 * 16 * 16 * 32 * 8 : 2.0 * 32 * 32 * 16 * 4;
            const double fl = per * iters * (double)blocks * 4;
            if (rep == 1) best = std::max(best, fl / (ms * 1e-3) / 1e12);
         }
      }
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      HIPCHK(hipFree(d));
      *tflops = best;
      return 0;
   } catch (const mgpu::HipError &) {
      return MGPU_E_HIP;
   }
}
