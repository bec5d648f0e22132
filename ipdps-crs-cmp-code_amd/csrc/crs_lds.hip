// LDS-staged CRS grid search (k_search_crs_lds).
//
// Same evaluations, sums and 64-bit arg-max keys as k_search_fast<CRS>
// (lanes = 64 consecutive t0 of one CDP, every (t0, parameter) summed over
// the CDP's supergather in slot order), with every window read from LDS.
// The global kernel loads one 16-byte window per evaluation and keeps the
// texture data path ~99 % busy; a supergather (hundreds to thousands of
// traces) cannot be staged at once, so the workgroup streams it:
//
// * The workgroup covers one (CDP, 64-t0 tile) and NW (a, b) x c-chunk
//   combinations, one per wave: a wave holds the sums of NC consecutive c
//   values of one (a, b) for its 64 t0 in registers while the whole
//   supergather passes.  Every evaluation of the workgroup lies in the box of
//   its (a, b) pairs and c values, so each trace's window starts are bounded
//   by (T0 + a m)^2 + b m^2 + c h at its extremes over that box and the
//   tile's t0 span, square-rooted, widened by 2 samples and clipped to the
//   valid starts; the workgroup plans batches of consecutive traces (at most
//   64) whose ranges fit the slab once, in an LDS table.
// * Each batch is staged by NW threads per trace as (sample, difference)
//   pairs (or samples), then every wave runs the batch's traces in order:
//   Q = (a^2 + b) m^2 + 2 T0 a m + T0^2 once per trace, then for each of its
//   c values t^2 = c h + Q, the window at LDS byte base + 8 floor(t).  The
//   lanes' windows are near-contiguous (consecutive t0), so the reads do not
//   conflict on LDS banks.  Batches whose traces have no gated window for
//   any evaluation of the box run without the window test, the clamp and the
//   per-trace count.
//
// Every rounding is the global kernel's (the grid values as its fused
// multiply-adds, the traveltime as c h + Q, the difference a paired slab
// holds, zeros for gated windows), in the same trace order, so both kernels
// produce the same keys.  FP32 rows (16-bit stores run on their restored
// FP32 copy, or on k_search_fast).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "device_common.hpp"
#include "device_types.hpp"
#include "launch.hpp"
#include "semblance.h"
#include "staged.hpp"

namespace sb {

template <int W, int NC, int NW, bool PAIRS>
__global__ __launch_bounds__(64 * NW) void k_search_crs_lds(SearchArgs a) {
  // every rounding below is explicit (fmaf; plain operators under this pragma)
#pragma clang fp contract(off)
  constexpr int NT = 64 * NW;
  const WorkItem wi = a.work[blockIdx.x];
  if (wi.cdp < 0) return;  // the whole workgroup
  const int cdp = wi.cdp;
  const int k0 = a.slot_off[cdp];
  const int S = a.slot_off[cdp + 1] - k0;
  if (S <= 0) return;  // no traces: nothing beats 0, keys stay empty
  const SlotF* __restrict__ sl = a.slots + k0;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);

  // ---- this wave's (a, b) pair and c values; the workgroup's box
  const int nchunk = (a.nc + NC - 1) / NC;
  const int pair0 = wi.p0 / a.nc, pair1 = (wi.p1 + a.nc - 1) / a.nc;
  const int pair = pair0 + wave / nchunk;
  const int cb = (wave % nchunk) * NC;
  const bool wact = pair < pair1;
  const int ncw = wact ? min(NC, a.nc - cb) : 0;  // c values of this wave (uniform)
  const int ia = pair / a.nb, ib = pair % a.nb;
  // grid values as k_search_fast computes them (its a0 + inc i contracts to one FMA)
  const float pa = fmaf(a.finc_a, (float)ia, a.fa0);
  const float pb = fmaf(a.finc_b, (float)ib, a.fb0);
  float pc[NC];
#pragma unroll
  for (int ci = 0; ci < NC; ci++) pc[ci] = fmaf(a.finc_c, (float)(cb + ci), a.fc0);
  const int ia0 = pair0 / a.nb, ia1 = (pair1 - 1) / a.nb;
  const int ib0 = ia0 == ia1 ? pair0 % a.nb : 0, ib1 = ia0 == ia1 ? (pair1 - 1) % a.nb : a.nb - 1;
  const float ea0 = fmaf(a.finc_a, (float)ia0, a.fa0), ea1 = fmaf(a.finc_a, (float)ia1, a.fa0);
  const float eb0 = fmaf(a.finc_b, (float)ib0, a.fb0), eb1 = fmaf(a.finc_b, (float)ib1, a.fb0);
  const float ec0 = fmaf(a.finc_c, 0.f, a.fc0), ec1 = fmaf(a.finc_c, (float)(a.nc - 1), a.fc0);
  const staged::Box box{{fminf(ea0, ea1), fminf(eb0, eb1), fminf(ec0, ec1)},
                        {fmaxf(ea0, ea1), fmaxf(eb0, eb1), fmaxf(ec0, ec1)}};
  // the whole 64-t0 span, past ns for the last tile, so that every lane's
  // windows lie in the staged ranges
  const float T0lo = (float)(wi.tile * 64), T0hi = (float)(wi.tile * 64 + 63);
  const float T0 = (float)(wi.tile * 64 + lane);
  const float T0sq = T0 * T0;
  const float K1 = (2.f * T0) * pa, K2 = fmaf(pa, pa, pb);
  const unsigned nwin = (unsigned)max(a.ns - a.w, 0);

  // ---- plan (once): the batches and window bases
  const int nbatch = staged::plan<W, NT, PAIRS>(a, sl, S, T0lo, T0hi, box);
  auto stage = [&](int kb, int ke) {
    staged::stage<NW, PAIRS>(a, S, kb, ke, [&](int k) {
      return (const float*)((const char*)a.samples + (int64_t)sl[k].row * a.stride * 4);
    });
  };

  float num[NC][W], den[NC];
  int cnt[NC];
#pragma unroll
  for (int ci = 0; ci < NC; ci++) {
    den[ci] = 0.f;
    cnt[ci] = 0;
#pragma unroll
    for (int j = 0; j < W; j++) num[ci][j] = 0.f;
  }
  // The window at LDS byte addr into the sums of c value ci.
  auto accumulate = [&](int ci, unsigned addr, float x) {
    if constexpr (PAIRS) {
      typedef float f2 __attribute__((ext_vector_type(2)));
      f2 pr[W];
#pragma unroll
      for (int q = 0; q < W; q++) pr[q] = staged::at<f2>(addr + 8u * q);
#pragma unroll
      for (int q = 0; q < W; q++) {
        const float v = fmaf(pr[q].y, x, pr[q].x);
        num[ci][q] = num[ci][q] + v;
        den[ci] = fmaf(v, v, den[ci]);
      }
    } else {
      float s[W + 1];
#pragma unroll
      for (int q = 0; q <= W; q++) s[q] = staged::at<float>(addr + 4u * q);
#pragma unroll
      for (int q = 0; q < W; q++) {
        const float v = fmaf(s[q + 1] - s[q], x, s[q]);
        num[ci][q] = num[ci][q] + v;
        den[ci] = fmaf(v, v, den[ci]);
      }
    }
  };
  // One batch's traces for this wave's evaluations, in order; lane j holds
  // trace kb + j's window base.
  auto run_batch = [&](int kb, int ke, bool ungated) {
    typedef const __attribute__((address_space(4))) SlotF CSlot;
    const int K = ke - kb;
    const unsigned mybase = lane < K ? staged::at<unsigned>(staged::kTable + 8u * (unsigned)(kb + lane)) : 0u;
    CSlot* sk = (CSlot*)(size_t)(sl + kb);
    for (int j = 0; j < K; j++) {
      const float m = sk[j].m, m2 = sk[j].m2, h = sk[j].h;
      const unsigned base = __builtin_amdgcn_readlane(mybase, j);
      const float Q = fmaf(K2, m2, fmaf(K1, m, T0sq));
#pragma unroll
      for (int ci = 0; ci < NC; ci++) {
        if (ci < ncw) {
          const float tt = fmaf(pc[ci], h, Q);
          if (ungated) {
            const float ts = __builtin_amdgcn_sqrtf(tt);
            accumulate(ci, base + ((unsigned)(int)ts << (PAIRS ? 3 : 2)), __builtin_amdgcn_fractf(ts));
          } else {
            const float ts = __builtin_amdgcn_sqrtf(fmaxf(tt, 0.f));
            const int it = (int)ts;
            bool ok = (unsigned)(it - a.tau) < nwin;
            if constexpr (W == 1) ok = ok && tt >= 0.f;
            cnt[ci] += ok ? 1 : 0;
            accumulate(ci, ok ? base + ((unsigned)it << (PAIRS ? 3 : 2)) : 0u, __builtin_amdgcn_fractf(ts));
          }
        }
      }
    }
    if (ungated) {
#pragma unroll
      for (int ci = 0; ci < NC; ci++) cnt[ci] += K;
    }
  };

  if (nbatch == 1) {
    stage(0, S);
    __syncthreads();
    if (wact) run_batch(0, S, (staged::batch_at(S, 0) & 0x8000) != 0);
  } else {
    for (int b = 0; b < nbatch; b++) {
      const int vb = staged::batch_at(S, b), kb = vb & 0x7fff, ke = staged::batch_at(S, b + 1) & 0x7fff;
      __syncthreads();  // every wave is done with the previous batch's slab
      stage(kb, ke);
      __syncthreads();
      if (wact) run_batch(kb, ke, (vb & 0x8000) != 0);
    }
  }

  // ---- keys: k_search_fast's validity, semblance and packing
  unsigned long long best = 0ull;
#pragma unroll
  for (int ci = 0; ci < NC; ci++) {
    if (ci < ncw) {
      float acsq = 0.f;
#pragma unroll
      for (int j = 0; j < W; j++) acsq = fmaf(num[ci][j], num[ci][j], acsq);
      const float mm = (float)cnt[ci];
      const bool valid = den[ci] > 1e-13f && cnt[ci] > 0 && (S - cnt[ci]) < 2;
      const float sem = acsq * __builtin_amdgcn_rcpf(den[ci] * mm);
      if (valid && sem > 0.f) {
        const int p = pair * a.nc + cb + ci;
        const unsigned long long key =
            ((unsigned long long)__float_as_uint(sem) << 32) | (unsigned long long)(~(uint32_t)p);
        best = key > best ? key : best;
      }
    }
  }
  __syncthreads();  // every wave is done with the tables and slab
  staged::put<unsigned long long>(64u + 8u * (unsigned)(wave * 64 + lane), best);
  __syncthreads();
  if (wave == 0) {
    unsigned long long mx = staged::at<unsigned long long>(64u + 8u * (unsigned)lane);
#pragma unroll
    for (int q = 1; q < NW; q++) {
      const unsigned long long v = staged::at<unsigned long long>(64u + 8u * (unsigned)(q * 64 + lane));
      mx = v > mx ? v : mx;
    }
    const int t0 = wi.tile * 64 + lane;
    if (t0 < a.ns && mx) atomicMax(&a.keys[(int64_t)(cdp - a.cdp_base) * a.ns + t0], mx);
  }
}

int crs_lds_min_bytes(int ns, int smax, int pairs, int waves) {
  // the plan of the longest supergather and the largest single-trace range
  // (every valid start plus the window, as quads, one more for the
  // alignment); the waves' keys reuse the space from byte 64
  return std::max(staged::min_bytes(ns, smax, pairs != 0), 64 + 8 * 64 * waves);
}

int crs_lds_c_per_wave(int nc, int cpw) {
  if (cpw == 6 || cpw == 12) return cpw;
  return 6;
}

bool crs_lds_supported(int w, int nc, int cpw, int waves, int pairs, int lds_bytes, int ns, int smax) {
  const int NC = crs_lds_c_per_wave(nc, cpw);
  return (w == 1 || w == 3 || w == 5 || w == 7) && waves == 8 && nc >= 1 && (nc + NC - 1) / NC <= waves &&
         smax < 32768 && lds_bytes >= crs_lds_min_bytes(ns, smax, pairs, waves) && lds_bytes <= 64 * 1024;
}

int crs_lds_pairs_per_block(int nc, int cpw, int waves) {
  const int NC = crs_lds_c_per_wave(nc, cpw);
  return waves / ((nc + NC - 1) / NC);
}

template <int W, int NC, bool PAIRS>
static hipError_t crs_lds_go(const SearchArgs& a, int nblocks, hipStream_t st) {
  hipLaunchKernelGGL((k_search_crs_lds<W, NC, 8, PAIRS>), dim3(nblocks), dim3(64 * 8), (size_t)a.lds_bytes, st, a);
  return hipGetLastError();
}

template <int NC, bool PAIRS>
static hipError_t crs_lds_w(const SearchArgs& a, int nblocks, hipStream_t st) {
  switch (a.w) {
    case 1: return crs_lds_go<1, NC, PAIRS>(a, nblocks, st);
    case 3: return crs_lds_go<3, NC, PAIRS>(a, nblocks, st);
    case 5: return crs_lds_go<5, NC, PAIRS>(a, nblocks, st);
    case 7: return crs_lds_go<7, NC, PAIRS>(a, nblocks, st);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_search_crs_lds(const SearchArgs& a, int cpw, int waves, int pairs, int smax, int nblocks,
                                 hipStream_t st) {
  if (nblocks <= 0) return hipSuccess;
  if (a.method != SB_METHOD_CRS || !crs_lds_supported(a.w, a.nc, cpw, waves, pairs, a.lds_bytes, a.ns, smax))
    return hipErrorInvalidValue;
  const int NC = crs_lds_c_per_wave(a.nc, cpw);
  if (NC == 6) return pairs ? crs_lds_w<6, true>(a, nblocks, st) : crs_lds_w<6, false>(a, nblocks, st);
  return pairs ? crs_lds_w<12, true>(a, nblocks, st) : crs_lds_w<12, false>(a, nblocks, st);
}

}  // namespace sb
