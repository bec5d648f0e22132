// LDS-staged CMP search (k_search_lds).
//
// Same work items and the same arithmetic as k_search_fast (lanes = 64
// consecutive t0 of one CDP, every lane walks the CDP's traces in order,
// packed 64-bit arg-max keys), but the sample windows come from LDS instead
// of the vector L1, whose address path bounds the global kernel (~17.5
// CU-cycles per 64-lane window load).  A block works through its parameter
// range in rounds of 4*CC consecutive parameters (CC per wave).  Per round it
// computes, for every trace, the range of window starts its t0 span and the
// round's c range can reach (the CMP traveltime sqrt(t0^2 + c h) is monotone
// in t0 and in c >= 0), packs the traces into chunks that fit the LDS, and
// for each chunk stages exactly those samples and runs the trace loop from
// LDS; the windowed sums stay in registers across chunks.
//
// Slab layout ("pair layout"): pair i of a trace holds samples (lo+i,
// lo+i+1), so the W+1 samples of any window start sit in pairs e-lo, e-lo+2,
// ... -- aligned ds_read2 reads at any start parity.  FP32: 8-byte pairs;
// FP16/BF16: 4-byte pairs (two halves, interpolated with mixed-precision
// FMAs).  Pairs 0..W-1 are zero: invalid windows (outside [tau, ns-tau-1))
// read them, as the global kernel's out-of-range buffer loads read zeros.
//
// The values staged are the stored bits and every rounding step matches
// k_search_fast, in the same trace order, so both kernels produce identical
// keys.
#include <hip/hip_runtime.h>

#include "device_common.hpp"
#include "device_types.hpp"
#include "launch.hpp"
#include "semblance.h"

namespace sb {

namespace {
typedef float f2v __attribute__((ext_vector_type(2)));
typedef _Float16 h2v __attribute__((ext_vector_type(2)));

// One window's raw words: W+1 FP32 samples, or (W+1)/2 dwords of two halves.
template <int W, int STORE>
struct RawWin {
  static constexpr int NP = (W + 1) / 2;
  float f[STORE == SB_STORE_F32 ? W + 1 : 1];
  uint32_t d[STORE == SB_STORE_F32 ? 1 : NP];
};

// Reads the window whose first pair is `idx` (pair layout: pairs idx,
// idx+2, ... hold samples 0-1, 2-3, ...).
template <int W, int STORE>
__device__ __forceinline__ void read_win(const unsigned char* pairs, unsigned idx, RawWin<W, STORE>& w) {
  constexpr int NP = (W + 1) / 2;
  if constexpr (STORE == SB_STORE_F32) {
    const f2v* pp = (const f2v*)(pairs + (size_t)idx * 8);
#pragma unroll
    for (int j = 0; j < NP; j++) {
      const f2v v = pp[2 * j];
      w.f[2 * j] = v.x;
      w.f[2 * j + 1] = v.y;
    }
  } else {
    const uint32_t* pp = (const uint32_t*)(pairs + (size_t)idx * 4);
#pragma unroll
    for (int j = 0; j < NP; j++) w.d[j] = pp[2 * j];
  }
}

// k_search_fast's window arithmetic, operation for operation.
template <int W, int STORE>
__device__ __forceinline__ void acc_win(const RawWin<W, STORE>& w, float x, float scale, float scale2, float* num,
                                        float& den) {
  constexpr int NP = (W + 1) / 2;
  if constexpr (STORE == SB_STORE_F32) {
#pragma unroll
    for (int j = 0; j < W; j++) {
      const float v = fmaf(__fsub_rn(w.f[j + 1], w.f[j]), x, w.f[j]);
      num[j] = __fadd_rn(num[j], v);
      den = fmaf(v, v, den);
    }
  } else if constexpr (STORE == SB_STORE_F16) {
    _Float16 hv[W + 1];
#pragma unroll
    for (int j = 0; j < NP; j++) {
      const h2v t = __builtin_bit_cast(h2v, w.d[j]);
      hv[2 * j] = t.x;
      hv[2 * j + 1] = t.y;
    }
    float q = 0.f;
#pragma unroll
    for (int j = 0; j < W; j++) {
      const float hj = (float)hv[j], hj1 = (float)hv[j + 1];
      const float v = fmaf(-x, hj, fmaf(x, hj1, hj));
      num[j] = fmaf(scale, v, num[j]);
      q = fmaf(v, v, q);
    }
    den = fmaf(scale2, q, den);
  } else {
    float f[W + 1];
#pragma unroll
    for (int j = 0; j < NP; j++) {
      f[2 * j] = __uint_as_float(w.d[j] << 16);
      f[2 * j + 1] = __uint_as_float(w.d[j] & 0xffff0000u);
    }
    float q = 0.f;
#pragma unroll
    for (int j = 0; j < W; j++) {
      const float v = fmaf(__fsub_rn(f[j + 1], f[j]), x, f[j]);
      num[j] = fmaf(scale, v, num[j]);
      q = fmaf(v, v, q);
    }
    den = fmaf(scale2, q, den);
  }
}
}  // namespace

template <int W, int R, int CC, int STORE, int NW>
__global__ __launch_bounds__(64 * NW) void k_search_lds(SearchArgs a) {
  constexpr int NT = 64 * NW;                        // threads per block
  constexpr int PB = STORE == SB_STORE_F32 ? 8 : 4;  // bytes per pair
  constexpr int ZP = W;                              // zero pairs (a window at 0 reads pairs 0..W-1)
  constexpr int RP = NW * CC;                        // parameters per round (CC per wave)
  extern __shared__ __align__(16) unsigned char dyn[];
  __shared__ unsigned long long s_best[NW][64 * R];
  __shared__ int s_wsum[NW], s_wmax[NW];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const WorkItem wi = a.work[blockIdx.x];
  if (wi.cdp < 0) return;
  const int cdp = wi.cdp;
  const int k0 = a.slot_off[cdp];
  const int S = a.slot_off[cdp + 1] - k0;
  if (S <= 0) return;
  const SlotF* __restrict__ sl = a.slots + k0;

  // Dynamic LDS: s_pre[S+1] (exclusive prefix of the round's slab lengths),
  // s_off[S] (s_pre[k] - first sample of slab k), s_first[S+2] (first trace
  // of each chunk), s_h[S] (each trace's h), then the pairs.
  int* s_pre = (int*)dyn;
  int* s_off = s_pre + (S + 1);
  int* s_first = s_off + S;
  float* s_h = (float*)(s_first + S + 2);
  const int hdr = ((4 * S + 3) * 4 + 15) & ~15;
  unsigned char* pairs = dyn + hdr;
  const int capP = (a.lds_bytes - hdr) / PB - ZP;  // pairs available to one chunk

  for (int i = tid; i < ZP * PB / 4; i += NT) ((uint32_t*)pairs)[i] = 0u;
  for (int k = tid; k < S; k += NT) s_h[k] = sl[k].h;  // read back below by the same thread

  const int nwin = max(a.ns - a.w, 0);
  const int t0min = wi.tile * 64 * R;
  const int t0max = min(t0min + 64 * R - 1, a.ns - 1);
  const float sqlo = __fmul_rn((float)t0min, (float)t0min), sqhi = __fmul_rn((float)t0max, (float)t0max);

  float T0sq[R];
#pragma unroll
  for (int r = 0; r < R; r++) {
    const float T0 = (float)(wi.tile * 64 * R + r * 64 + lane);
    T0sq[r] = __fmul_rn(T0, T0);
  }
  unsigned long long best[R];
#pragma unroll
  for (int r = 0; r < R; r++) best[r] = 0ull;

  for (int q0 = wi.p0; q0 < wi.p1; q0 += RP) {
    const int q1 = min(q0 + RP, wi.p1);
    // ---- 1. slab of every trace for this round: window starts
    // e = floor(sqrt(T0^2 + c h)) - tau are monotone in T0 and c >= 0, so the
    // block's extreme t0 and the round's extreme c bound them; widened by 2
    // for the approximate square root and clipped to valid starts.
    const float cA = __fadd_rn(a.fc0, __fmul_rn(a.finc_c, (float)q0));
    const float cB = __fadd_rn(a.fc0, __fmul_rn(a.finc_c, (float)(q1 - 1)));
    const float cmin = fminf(cA, cB), cmax = fmaxf(cA, cB);
    int carry = 0, lmax = 0;
    for (int kb = 0; kb < S; kb += NT) {
      const int k = kb + tid;
      int lo = 0, len = 0;
      if (k < S) {
        const float h = s_h[k];
        const int elo = (int)__builtin_amdgcn_sqrtf(fmaf(cmin, h, sqlo)) - a.tau - 2;
        const int ehi = (int)__builtin_amdgcn_sqrtf(fmaf(cmax, h, sqhi)) - a.tau + 2;
        lo = max(elo, 0);
        const int hi = min(ehi, nwin - 1);
        len = hi >= lo ? hi - lo + W : 0;
      }
      int x = len, m = len;
#pragma unroll
      for (int d = 1; d < 64; d <<= 1) {
        const int y = __shfl_up(x, d);
        if (lane >= d) x += y;
        m = max(m, __shfl_xor(m, d));
      }
      if (lane == 63) s_wsum[wave] = x;
      if (lane == 0) s_wmax[wave] = m;
      __syncthreads();
      int off = carry, tot = 0, mx = 0;
#pragma unroll
      for (int q = 0; q < NW; q++) {
        off += q < wave ? s_wsum[q] : 0;
        tot += s_wsum[q];
        mx = max(mx, s_wmax[q]);
      }
      if (k < S) {
        s_pre[k] = off + x - len;
        s_off[k] = off + x - len - lo;
      }
      carry += tot;
      lmax = max(lmax, mx);
      __syncthreads();
    }
    // ---- 2. chunks: chunk j holds the traces whose slab starts in
    // [j*H, (j+1)*H); with H = capP - lmax a chunk never exceeds capP, and
    // lmax <= capP / 2 keeps H >= lmax, so at most S + 1 chunks (s_first's
    // size).  The host only launches shapes whose slabs are that narrow.
    const int H = capP - lmax;
    if (2 * lmax > capP) {
      if (tid == 0) atomicOr(a.status, 1u);
      return;
    }
    const int nch = carry > 0 ? (carry + H - 1) / H : 1;
    if (tid == 0) {
      s_pre[S] = carry;
      s_first[0] = 0;
      s_first[nch] = S;
    }
    __syncthreads();
    for (int k = 1 + tid; k <= S; k += NT) {
      const int jlo = s_pre[k - 1] / H, jhi = min(s_pre[k] / H, nch - 1);
      for (int j = jlo + 1; j <= jhi; j++) s_first[j] = k;
    }

    // ---- 3. this wave's CC parameters of the round, over all chunks
    const int p = q0 + wave * CC;
    float pc[CC];
    bool act[CC];
#pragma unroll
    for (int ci = 0; ci < CC; ci++) {
      act[ci] = p + ci < q1;
      pc[ci] = __fadd_rn(a.fc0, __fmul_rn(a.finc_c, (float)(act[ci] ? p + ci : q0)));
    }
    float num[R][CC][W], den[R][CC];
    int cnt[R][CC];
#pragma unroll
    for (int r = 0; r < R; r++)
#pragma unroll
      for (int ci = 0; ci < CC; ci++) {
        den[r][ci] = 0.f;
        cnt[r][ci] = 0;
#pragma unroll
        for (int j = 0; j < W; j++) num[r][ci][j] = 0.f;
      }
    __syncthreads();  // s_first, and the previous round's last chunk is consumed

    for (int j = 0; j < nch; j++) {
      const int ks = s_first[j], ke = s_first[j + 1];
      if (ks >= ke) continue;  // uniform
      const int pb0 = s_pre[ks] - ZP;  // pair index of trace k: s_pre[k] - pb0
      // stage: one wave per trace, lanes over pairs, two loads in flight
      for (int k = ks + wave; k < ke; k += NW) {
        const int base = s_pre[k] - pb0, len = s_pre[k + 1] - s_pre[k], lo = s_pre[k] - s_off[k];
        const int64_t row = sl[k].row;
        if constexpr (STORE == SB_STORE_F32) {
          const float* __restrict__ src = (const float*)a.samples + row * a.stride + lo;
          for (int i = lane; i < len; i += 128) {
            const bool two = i + 64 < len;
            f2v u, v;
            u.x = src[i];
            u.y = src[i + 1];
            if (two) {
              v.x = src[i + 64];
              v.y = src[i + 65];
            }
            *(f2v*)(pairs + (size_t)(base + i) * 8) = u;
            if (two) *(f2v*)(pairs + (size_t)(base + i + 64) * 8) = v;
          }
        } else {
          const uint16_t* __restrict__ src = (const uint16_t*)a.samples + row * a.stride + lo;
          for (int i = lane; i < len; i += 128) {
            const bool two = i + 64 < len;
            const uint32_t u = (uint32_t)src[i] | ((uint32_t)src[i + 1] << 16);
            uint32_t v = 0;
            if (two) v = (uint32_t)src[i + 64] | ((uint32_t)src[i + 65] << 16);
            *(uint32_t*)(pairs + (size_t)(base + i) * 4) = u;
            if (two) *(uint32_t*)(pairs + (size_t)(base + i + 64) * 4) = v;
          }
        }
      }
      __syncthreads();

      // Trace loop, software-pipelined by one trace: the windows of trace
      // k+1 are read from LDS while trace k accumulates.
      RawWin<W, STORE> wa[R][CC], wb[R][CC];
      float xa[R][CC], xb[R][CC];
      bool oa[R][CC], ob[R][CC];
      auto issue = [&](int k, RawWin<W, STORE>(&w)[R][CC], float(&xx)[R][CC], bool(&okk)[R][CC]) {
        const float h = s_h[k];
        const int sbk = s_off[k] - pb0;  // pair index of window start e is sbk + e
#pragma unroll
        for (int r = 0; r < R; r++) {
#pragma unroll
          for (int ci = 0; ci < CC; ci++) {
            const float tt = fmaf(pc[ci], h, T0sq[r]);
            const float ts = __builtin_amdgcn_sqrtf(tt);  // tt >= 0: the host checks c0, c1 >= 0
            const int e = (int)ts - a.tau;
            const bool ok = (unsigned)e < (unsigned)nwin && tt < a.hi2;  // the global kernel's rule
            okk[r][ci] = ok;
            xx[r][ci] = __builtin_amdgcn_fractf(ts);
            read_win<W, STORE>(pairs, ok ? (unsigned)(sbk + e) : 0u, w[r][ci]);
          }
        }
      };
      auto consume = [&](int k, const RawWin<W, STORE>(&w)[R][CC], const float(&xx)[R][CC],
                         const bool(&okk)[R][CC]) {
        float scale = 1.f, scale2 = 1.f;
        if constexpr (STORE != SB_STORE_F32) {
          const SlotF s = sl[k];
          scale = s.scale;
          scale2 = s.scale2;
        }
#pragma unroll
        for (int r = 0; r < R; r++) {
#pragma unroll
          for (int ci = 0; ci < CC; ci++) {
            cnt[r][ci] += okk[r][ci] ? 1 : 0;
            acc_win<W, STORE>(w[r][ci], xx[r][ci], scale, scale2, num[r][ci], den[r][ci]);
          }
        }
      };
      issue(ks, wa, xa, oa);
      for (int k = ks; k < ke; k += 2) {
        issue(k + 1 < ke ? k + 1 : k, wb, xb, ob);
        consume(k, wa, xa, oa);
        if (k + 1 >= ke) break;
        issue(k + 2 < ke ? k + 2 : k + 1, wa, xa, oa);
        consume(k + 1, wb, xb, ob);
      }
      __syncthreads();  // chunk consumed before the next one is staged
    }

#pragma unroll
    for (int r = 0; r < R; r++) {
#pragma unroll
      for (int ci = 0; ci < CC; ci++) {
        float acsq = 0.f;
#pragma unroll
        for (int jj = 0; jj < W; jj++) acsq = fmaf(num[r][ci][jj], num[r][ci][jj], acsq);
        const float m = (float)cnt[r][ci];
        const bool valid = act[ci] && den[r][ci] > 1e-13f && cnt[r][ci] > 0 && (S - cnt[r][ci]) < 2;
        const float sem = __fmul_rn(acsq, __builtin_amdgcn_rcpf(__fmul_rn(den[r][ci], m)));
        if (valid && sem > 0.f) {
          const unsigned long long key =
              ((unsigned long long)__float_as_uint(sem) << 32) | (unsigned long long)(~(uint32_t)(p + ci));
          best[r] = key > best[r] ? key : best[r];
        }
      }
    }
  }

#pragma unroll
  for (int r = 0; r < R; r++) s_best[wave][r * 64 + lane] = best[r];
  __syncthreads();
  if (wave == 0) {
#pragma unroll
    for (int r = 0; r < R; r++) {
      unsigned long long m = s_best[0][r * 64 + lane];
#pragma unroll
      for (int q = 1; q < NW; q++) {
        const unsigned long long v = s_best[q][r * 64 + lane];
        m = v > m ? v : m;
      }
      const int t0 = wi.tile * 64 * R + r * 64 + lane;
      if (t0 < a.ns && m) atomicMax(&a.keys[(int64_t)(cdp - a.cdp_base) * a.ns + t0], m);
    }
  }
}

// ------------------------------------------------------------- launchers
template <int W, int R, int CC, int NW>
static hipError_t lds_store(const SearchArgs& a, int store, int nblocks, hipStream_t st) {
  const size_t sh = (size_t)a.lds_bytes;
  if (store == SB_STORE_F32)
    hipLaunchKernelGGL((k_search_lds<W, R, CC, SB_STORE_F32, NW>), dim3(nblocks), dim3(64 * NW), sh, st, a);
  else if (store == SB_STORE_F16)
    hipLaunchKernelGGL((k_search_lds<W, R, CC, SB_STORE_F16, NW>), dim3(nblocks), dim3(64 * NW), sh, st, a);
  else
    hipLaunchKernelGGL((k_search_lds<W, R, CC, SB_STORE_BF16, NW>), dim3(nblocks), dim3(64 * NW), sh, st, a);
  return hipGetLastError();
}

template <int R, int CC, int NW>
static hipError_t lds_w(const SearchArgs& a, int store, int nblocks, hipStream_t st) {
  switch (a.w) {
    case 1: return lds_store<1, R, CC, NW>(a, store, nblocks, st);
    case 3: return lds_store<3, R, CC, NW>(a, store, nblocks, st);
    case 5: return lds_store<5, R, CC, NW>(a, store, nblocks, st);
    case 7: return lds_store<7, R, CC, NW>(a, store, nblocks, st);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_search_lds(const SearchArgs& a, int store, const FastVariant& v, int nblocks, hipStream_t st) {
  if (nblocks <= 0) return hipSuccess;
  if (a.method != SB_METHOD_CMP || a.lds_bytes <= 0 || a.lds_bytes > kLdsMaxBytes) return hipErrorInvalidValue;
  if (v.R == 1 && v.CC == 4 && v.waves == 8) return lds_w<1, 4, 8>(a, store, nblocks, st);
  if (v.R == 2 && v.CC == 2 && v.waves == 8) return lds_w<2, 2, 8>(a, store, nblocks, st);
  if (v.R == 1 && v.CC == 2 && v.waves == 8) return lds_w<1, 2, 8>(a, store, nblocks, st);
  if (v.R == 1 && v.CC == 4 && v.waves == 4) return lds_w<1, 4, 4>(a, store, nblocks, st);
  return hipErrorInvalidValue;
}

}  // namespace sb
