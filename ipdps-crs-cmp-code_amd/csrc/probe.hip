// Device measurements for the roofline.  FP32 vector FMA rate of the device
// (sb_device_fp32_fma_rate): the ceiling the search kernels are held to.  They run no packed FP32 (the library is
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

// Interpolated window samples per second for four forms of the inner step of
// the search (one window sample: interpolate, add to the window sum, add its
// square), each as 8 independent chains per lane with the instructions
// written out, so the compiler cannot fold or reorder them:
//   0  FP32 samples s, s':  d = s' - s; v = fma(d, x, s); n += v; q = fma(v, v, q)
//   1  FP16 pair (s, d):    v = fma_mix(x, d, s); n = fma(c, v, n); q = fma(v, v, q)
//   2  FP16 pair (s, s'):   v = dot2((1 - x, x), (s, s')); n = fma(c, v, n); q = fma(v, v, q)
//   3  FP32 pair (s, d):    v = fma(x, d, s); n = fma(c, v, n); q = fma(v, v, q)
// Forms 1 and 2 are what a 16-bit store could do instead of restoring its
// samples to FP32 (form 0); form 3 is the floor of one instruction per
// interpolation.
constexpr int kWinChains = 8;
constexpr int kWinIters = 2048;

template <int FORM>
__global__ __launch_bounds__(256) void k_window_rate(float* __restrict__ out, float x, float c, unsigned pair) {
  float a[kWinChains], b[kWinChains], n[kWinChains], q[kWinChains];
  unsigned p[kWinChains];
#pragma unroll
  for (int i = 0; i < kWinChains; i++) {
    a[i] = (float)threadIdx.x * 1e-3f + (float)i;
    b[i] = a[i] + 0.5f;
    n[i] = 0.f;
    q[i] = 0.f;
    p[i] = pair + (unsigned)i;
  }
  const unsigned w = 0x38003800u;  // (0.5, 0.5) in FP16
  for (int it = 0; it < kWinIters; it++) {
#pragma unroll
    for (int i = 0; i < kWinChains; i++) {
      float v, d;
      if constexpr (FORM == 0) {
        asm volatile("v_sub_f32 %0, %1, %2" : "=v"(d) : "v"(b[i]), "v"(a[i]));
        asm volatile("v_fma_f32 %0, %1, %2, %3" : "=v"(v) : "v"(d), "v"(x), "v"(a[i]));
        asm volatile("v_add_f32 %0, %0, %1" : "+v"(n[i]) : "v"(v));
      } else if constexpr (FORM == 1) {
        asm volatile("v_fma_mix_f32 %0, %1, %2, %2 op_sel:[0,1,0] op_sel_hi:[0,1,1]" : "=v"(v) : "v"(x), "v"(p[i]));
        asm volatile("v_fmac_f32 %0, %1, %2" : "+v"(n[i]) : "v"(c), "v"(v));
      } else if constexpr (FORM == 2) {
        asm volatile("v_dot2_f32_f16 %0, %1, %2, 0" : "=v"(v) : "v"(w), "v"(p[i]));
        asm volatile("v_fmac_f32 %0, %1, %2" : "+v"(n[i]) : "v"(c), "v"(v));
      } else {
        asm volatile("v_fma_f32 %0, %1, %2, %3" : "=v"(v) : "v"(x), "v"(b[i]), "v"(a[i]));
        asm volatile("v_fmac_f32 %0, %1, %2" : "+v"(n[i]) : "v"(c), "v"(v));
      }
      asm volatile("v_fmac_f32 %0, %1, %1" : "+v"(q[i]) : "v"(v));
    }
  }
  float s = 0.f;
#pragma unroll
  for (int i = 0; i < kWinChains; i++) s += n[i] + q[i];
  out[(size_t)blockIdx.x * blockDim.x + threadIdx.x] = s;  // keeps the chains live
}

// Shared set-up of the probes: a grid 8 waves deep on every SIMD, an output
// buffer, a stream and two events; `run` launches one kernel and returns its
// best time over three launches (the first warms the clocks up).
struct Probe {
  int blocks = 0;
  float* out = nullptr;
  hipStream_t st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int open(int device) {
    using sb::fail;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return fail(SB_ERR_HIP, "no HIP device available");
    if (device < 0 || device >= n) return fail(SB_ERR_INVALID, "bad device index");
    if (hipSetDevice(device) != hipSuccess) return fail(SB_ERR_HIP, "hipSetDevice");
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return fail(SB_ERR_HIP, "device properties");
    blocks = std::max(1, prop.multiProcessorCount) * 32;  // 8 waves per SIMD, 4 times over
    if (hipMalloc(&out, (size_t)blocks * 256 * sizeof(float)) != hipSuccess ||
        hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
        hipEventCreate(&e1) != hipSuccess)
      return fail(SB_ERR_HIP, "probe set-up");
    return SB_OK;
  }
  template <class F>
  int run(F launch, double* best_ms) {
    for (int rep = 0; rep < 3; rep++) {
      hipEventRecord(e0, st);
      launch();
      hipEventRecord(e1, st);
      float ms = 0;
      if (hipGetLastError() != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
          hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
        return sb::fail(SB_ERR_HIP, "probe launch");
      if (rep == 0 || ms < *best_ms) *best_ms = ms;
    }
    return SB_OK;
  }
  ~Probe() {
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st) (void)hipStreamDestroy(st);
    if (out) (void)hipFree(out);
  }
};

}  // namespace

extern "C" int sb_device_window_rates(int device, double* samples_per_s) {
  using sb::fail;
  if (!samples_per_s) return fail(SB_ERR_INVALID, "sb_device_window_rates: null output");
  for (int f = 0; f < 4; f++) samples_per_s[f] = 0;
  Probe pr;
  int rc = pr.open(device);
  for (int f = 0; f < 4 && rc == SB_OK; f++) {
    double ms = 0;
    auto launch = [&] {
      void (*k)(float*, float, float, unsigned) = f == 0   ? k_window_rate<0>
                                                  : f == 1 ? k_window_rate<1>
                                                  : f == 2 ? k_window_rate<2>
                                                           : k_window_rate<3>;
      hipLaunchKernelGGL(k, dim3(pr.blocks), dim3(256), 0, pr.st, pr.out, 0.375f, 0.5f, 0x3c003800u);
    };
    rc = pr.run(launch, &ms);
    if (rc == SB_OK) samples_per_s[f] = (double)kWinChains * kWinIters * pr.blocks * 256 / (ms * 1e-3);
  }
  return rc;
}

extern "C" int sb_device_fp32_fma_rate(int device, double* tflops) {
  using sb::fail;
  if (!tflops) return fail(SB_ERR_INVALID, "sb_device_fp32_fma_rate: null output");
  *tflops = 0;
  Probe pr;
  int rc = pr.open(device);
  double ms = 0;
  if (rc == SB_OK)
    rc = pr.run([&] { hipLaunchKernelGGL(k_fma_rate, dim3(pr.blocks), dim3(256), 0, pr.st, pr.out, 0.999f, 1e-3f); },
                &ms);
  if (rc == SB_OK) *tflops = 2.0 * kChains * kIters * (double)pr.blocks * 256 / (ms * 1e-3) / 1e12;
  return rc;
}
