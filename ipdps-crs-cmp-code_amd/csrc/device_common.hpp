// Device helpers shared by kernels.hip and de.hip.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <hip/hip_bf16.h>

#include "device_types.hpp"
#include "semblance.h"

namespace sb {

// ------------------------------------------------------------------ stores
// Samples of row r, element i, restored to FP32.
struct LoadF32 {
  static constexpr bool kQuantized = false;
  const float* p;
  int64_t stride;
  __device__ float operator()(int64_t row, int i) const { return p[row * stride + i]; }
};
struct LoadF16 {
  static constexpr bool kQuantized = true;
  const __half* p;
  const float* scale;
  int64_t stride;
  __device__ float operator()(int64_t row, int i) const {
    return __half2float(p[row * stride + i]) * scale[row];
  }
};
struct LoadBF16 {
  static constexpr bool kQuantized = true;
  const uint16_t* p;
  const float* scale;
  int64_t stride;
  __device__ float operator()(int64_t row, int i) const {
    return __uint_as_float((uint32_t)p[row * stride + i] << 16) * scale[row];
  }
};

// Raw buffer resources over one row: an offset at or beyond num_records
// reads zeros, so lanes whose trace falls outside [tau, ns-tau-1) load from
// kOob and contribute nothing, with no branch around the loads.
constexpr unsigned kOob = 0x7ffffff0u;
constexpr int kRsrcWord3 = 0x00020000;

template <int N>
struct Win {
  float v[N];
};

// N consecutive 16-bit samples as loaded (no conversion), before load_win
// restores them to FP32.
template <int N>
struct WinH {
  _Float16 h[N];
};

// N consecutive FP16 samples starting at element e of a row: dword-aligned
// loads (a 2-byte-misaligned multi-dword load costs about twice as much in
// the address path) and a funnel shift by 16 bits when e is odd.  `off` is
// the byte offset of element (e & ~1); `sh` is 16 * (e & 1).
template <int N>
__device__ __forceinline__ WinH<N> load_win_h(__amdgpu_buffer_rsrc_t rs, unsigned off, unsigned sh) {
  typedef unsigned u2 __attribute__((ext_vector_type(2)));
  typedef unsigned u3 __attribute__((ext_vector_type(3)));
  typedef unsigned u4 __attribute__((ext_vector_type(4)));
  typedef _Float16 h2 __attribute__((ext_vector_type(2)));
  constexpr int ND = N / 2 + 1;  // dwords covering N halves at either parity
  unsigned d[ND];
  if constexpr (ND == 2) {
    u2 q = __builtin_amdgcn_raw_buffer_load_b64(rs, off, 0, 0);
    d[0] = q.x;
    d[1] = q.y;
  } else if constexpr (ND == 3) {
    u3 q = __builtin_amdgcn_raw_buffer_load_b96(rs, off, 0, 0);
    d[0] = q.x;
    d[1] = q.y;
    d[2] = q.z;
  } else if constexpr (ND == 4) {
    u4 q = __builtin_amdgcn_raw_buffer_load_b128(rs, off, 0, 0);
    d[0] = q.x;
    d[1] = q.y;
    d[2] = q.z;
    d[3] = q.w;
  } else {
    u4 q = __builtin_amdgcn_raw_buffer_load_b128(rs, off, 0, 0);
    d[0] = q.x;
    d[1] = q.y;
    d[2] = q.z;
    d[3] = q.w;
#pragma unroll
    for (int k = 4; k < ND; k++) d[k] = __builtin_amdgcn_raw_buffer_load_b32(rs, off + 4u * k, 0, 0);
  }
  WinH<N> r;
#pragma unroll
  for (int k = 0; k < (N + 1) / 2; k++) {
    const h2 p = __builtin_bit_cast(h2, __builtin_amdgcn_alignbit(d[k + 1], d[k], sh));
    r.h[2 * k] = p.x;
    if (2 * k + 1 < N) r.h[2 * k + 1] = p.y;
  }
  return r;
}

// W+1 consecutive samples restored to FP32 in the row's stored units (the
// per-trace scale is applied by the caller).  F32: byte offset `off` of the
// first sample.  16-bit stores: as load_win_h (dword-aligned, shift `sh`).
template <int N, int STORE>
__device__ __forceinline__ Win<N> load_win(__amdgpu_buffer_rsrc_t rs, unsigned off, unsigned sh) {
  typedef float f4 __attribute__((ext_vector_type(4)));
  typedef float f2 __attribute__((ext_vector_type(2)));
  Win<N> r;
  if constexpr (STORE == SB_STORE_F32) {
    int i = 0;
#pragma unroll
    for (; i + 4 <= N; i += 4) {
      f4 q = __builtin_amdgcn_raw_buffer_load_b128(rs, off + 4u * i, 0, 0);
      r.v[i] = q.x;
      r.v[i + 1] = q.y;
      r.v[i + 2] = q.z;
      r.v[i + 3] = q.w;
    }
    if constexpr (N % 4 >= 2) {
      f2 q = __builtin_amdgcn_raw_buffer_load_b64(rs, off + 4u * i, 0, 0);
      r.v[i] = q.x;
      r.v[i + 1] = q.y;
      i += 2;
    }
    if constexpr (N % 2 == 1) r.v[i] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rs, off + 4u * i, 0, 0));
  } else {
    const WinH<N> h = load_win_h<N>(rs, off, sh);
#pragma unroll
    for (int k = 0; k < N; k++) {
      const unsigned short u = __builtin_bit_cast(unsigned short, h.h[k]);
      if constexpr (STORE == SB_STORE_F16)
        r.v[k] = (float)h.h[k];
      else
        r.v[k] = __uint_as_float((unsigned)u << 16);
    }
  }
  return r;
}

}  // namespace sb
