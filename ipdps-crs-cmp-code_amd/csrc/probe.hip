// FP32 vector FMA rate of the device (sb_device_fp32_fma_rate): the ceiling
// the search kernels are held to.  They run no packed FP32 (the library is
// built without it), so their FP32 arithmetic issues one FMA per lane per
// instruction; this kernel does nothing else -- 16 independent FMA chains
// per lane, no memory traffic inside the loop -- over a grid that fills
// every SIMD eight waves deep.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "sb_internal.hpp"

namespace {

constexpr int kChains = 16;
constexpr int kIters = 4096;

__global__ __launch_bounds__(256) void k_fma_rate(float* __restrict__ out, float x, float y) {
  float acc[kChains];
#pragma unroll
  for (int i = 0; i < kChains; i++) acc[i] = (float)threadIdx.x * 1e-3f + (float)i;
  for (int it = 0; it < kIters; it++) {
#pragma unroll
    for (int i = 0; i < kChains; i++) acc[i] = fmaf(acc[i], x, y);
  }
  float s = 0.f;
#pragma unroll
  for (int i = 0; i < kChains; i++) s += acc[i];
  out[(size_t)blockIdx.x * blockDim.x + threadIdx.x] = s;  // keeps the chains live
}

}  // namespace

extern "C" int sb_device_fp32_fma_rate(int device, double* tflops) {
  using sb::fail;
  if (!tflops) return fail(SB_ERR_INVALID, "sb_device_fp32_fma_rate: null output");
  *tflops = 0;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return fail(SB_ERR_HIP, "no HIP device available");
  if (device < 0 || device >= n) return fail(SB_ERR_INVALID, "bad device index");
  if (hipSetDevice(device) != hipSuccess) return fail(SB_ERR_HIP, "hipSetDevice");
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) return fail(SB_ERR_HIP, "device properties");
  const int blocks = std::max(1, prop.multiProcessorCount) * 32;  // 8 waves per SIMD, 4 times over
  float* out = nullptr;
  hipStream_t st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = SB_OK;
  if (hipMalloc(&out, (size_t)blocks * 256 * sizeof(float)) != hipSuccess ||
      hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
      hipEventCreate(&e1) != hipSuccess) {
    rc = fail(SB_ERR_HIP, "probe set-up");
  }
  double best_ms = 0;
  for (int rep = 0; rep < 3 && rc == SB_OK; rep++) {  // the first launch warms the clocks up
    hipEventRecord(e0, st);
    hipLaunchKernelGGL(k_fma_rate, dim3(blocks), dim3(256), 0, st, out, 0.999f, 1e-3f);
    hipEventRecord(e1, st);
    float ms = 0;
    if (hipGetLastError() != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
        hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
      rc = fail(SB_ERR_HIP, "probe launch");
      break;
    }
    if (rep == 0 || ms < best_ms) best_ms = ms;
  }
  if (rc == SB_OK) *tflops = 2.0 * kChains * kIters * (double)blocks * 256 / (best_ms * 1e-3) / 1e12;
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (st) (void)hipStreamDestroy(st);
  if (out) (void)hipFree(out);
  return rc;
}
