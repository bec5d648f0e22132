// Batch-staged windows in LDS, shared by the CRS grid search
// (k_search_crs_lds, crs_lds.hip) and the CRS-DE evaluator (k_crs_de_lds,
// de.hip).
//
// A workgroup's evaluations all lie in a box of traveltime parameters
// (a, b, c) -- the CRS grid's pairs and c values it covers, or the DE box --
// and its t0 span.  For each trace of the supergather the window starts
// floor(sqrt((T0 + a m)^2 + b m^2 + c h)) - tau of every such evaluation lie
// between the square roots of that expression's smallest and largest values
// over the box, widened by 2 samples (the FP32 traveltime's rounding moves
// floor(t) by far less) and clipped to the valid starts [0, ns - w).  The
// workgroup plans once how the supergather splits into batches of
// consecutive traces (at most 64) whose ranges fit the slab, then each
// batch is staged -- as (sample, forward difference) pairs, 8 bytes a sample,
// or as samples, 4 -- and its traces are evaluated in supergather order.
//
// LDS layout (dynamic LDS from address 0) for a supergather of S traces:
//   [0, 64)            zeros: the window of a gated evaluation reads them
//   [64, 64 + 8 S)     trace table: window base, then qlo | nq << 16 |
//                      ungated << 31 (the trace's whole unclipped range is
//                      valid, so no evaluation of the box is gated)
//   batches(S)         u16 x (S + 2): batch starts with bit 15 set when every
//                      trace of the batch is ungated, the end S, and the
//                      batch count at index S + 1
//   slab(S) ..         the staged quads
// The window of traveltime t of trace k starts at byte base_k + SB floor(t)
// (SB = 8 for pairs, 4 for samples).
#pragma once

#include <hip/hip_runtime.h>

#include "device_types.hpp"

namespace sb {
namespace staged {

constexpr unsigned kTable = 64u;
__host__ __device__ constexpr int batches(int S) { return 64 + 8 * S; }
__host__ __device__ constexpr int slab(int S) { return (64 + 8 * S + 2 * (S + 2) + 15) & ~15; }
// LDS bytes for the plan of the longest supergather (smax traces) plus the
// largest single-trace range: every valid start plus the window, as quads,
// one more for the alignment.
__host__ __device__ constexpr int min_bytes(int ns, int smax, bool pairs) {
  return slab(smax) + (pairs ? 32 : 16) * ((ns + 3) / 4 + 2);
}

template <class T>
__device__ __forceinline__ T at(unsigned addr) {
  return *(__attribute__((address_space(3))) T*)(size_t)addr;
}
template <class T>
__device__ __forceinline__ void put(unsigned addr, T v) {
  *(__attribute__((address_space(3))) T*)(size_t)addr = v;
}

// The parameter box: lo/hi of a, b and c (c on the scale of the slots' h).
struct Box {
  float lo[3], hi[3];
};

// Quads of trace s's window range over T0 in [T0lo, T0hi] and the box:
// pairs of samples elo .. ehi + W - 1 (samples: .. ehi + W) as the row's
// quads qlo .. qlo + nq - 1 (nq = 0: no window of the span and box is valid).
template <int W, bool PAIRS>
__device__ __forceinline__ void range(const SearchArgs& a, const SlotF& s, float T0lo, float T0hi, const Box& b,
                                      int& qlo, int& nq, bool& full) {
  const int nwin = max(a.ns - a.w, 0);
  const float am0 = b.lo[0] * s.m, am1 = b.hi[0] * s.m;
  const float ulo = T0lo + fminf(am0, am1), uhi = T0hi + fmaxf(am0, am1);
  const float u2lo = ulo * ulo, u2hi = uhi * uhi;
  const float usqmin = (ulo <= 0.f && uhi >= 0.f) ? 0.f : fminf(u2lo, u2hi);
  const float usqmax = fmaxf(u2lo, u2hi);
  const float bm0 = b.lo[1] * s.m2, bm1 = b.hi[1] * s.m2;
  const float ch0 = b.lo[2] * s.h, ch1 = b.hi[2] * s.h;
  const float ttlo = usqmin + fminf(bm0, bm1) + fminf(ch0, ch1);
  const float tthi = usqmax + fmaxf(bm0, bm1) + fmaxf(ch0, ch1);
  const int elo_b = (int)sqrtf(fmaxf(ttlo, 0.f)) - a.tau - 2;
  const int ehi_b = (int)sqrtf(fmaxf(tthi, 0.f)) - a.tau + 2;
  // every window of the span and box valid: no evaluation is gated, and
  // t^2 >= (tau + 2)^2 > 0 for all of them
  full = elo_b >= 0 && ehi_b <= nwin - 1;
  const int elo = max(elo_b, 0), ehi = min(ehi_b, nwin - 1);
  qlo = 0;
  nq = 0;
  if (ehi >= elo) {
    qlo = elo >> 2;
    nq = ((ehi + W - (PAIRS ? 1 : 0)) >> 2) - qlo + 1;
  }
}

// The plan, by the whole workgroup of NT threads: each trace's quads (all
// threads), then the batches and window bases (one wave, sequentially: the
// next up to 64 traces whose quads fit the slab -- the inclusive sums are
// nondecreasing, so those that fit are a prefix).  Returns the batch count
// (uniform).  Writes the 64 zero bytes too.
template <int W, int NT, bool PAIRS>
__device__ int plan(const SearchArgs& a, const SlotF* __restrict__ sl, int S, float T0lo, float T0hi, const Box& box) {
  constexpr unsigned SB = PAIRS ? 8u : 4u, QB = 4u * SB;
  constexpr int QSH = PAIRS ? 5 : 4;
  const int tid = threadIdx.x, lane = tid & 63;
  const unsigned bat = (unsigned)batches(S), sl0 = (unsigned)slab(S);
  const int capq = (a.lds_bytes - (int)sl0) >> QSH;
  if (tid < 16) put<unsigned>(4u * tid, 0u);
  for (int k = tid; k < S; k += NT) {
    int qlo, nq;
    bool full;
    range<W, PAIRS>(a, sl[k], T0lo, T0hi, box, qlo, nq, full);
    put<unsigned>(kTable + 8u * k + 4u, (unsigned)qlo | ((unsigned)nq << 16) | (full ? 0x80000000u : 0u));
  }
  __syncthreads();
  if (tid < 64) {
    int nb = 0;
    for (int kb = 0; kb < S;) {
      const int k = kb + lane;
      const unsigned e = k < S ? at<unsigned>(kTable + 8u * k + 4u) : 0u;
      const int nq = (int)((e >> 16) & 0x7fffu), qlo = (int)(e & 0xffffu);
      int x = nq;
#pragma unroll
      for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(x, o);
        if (lane >= o) x += y;
      }
      int K = __builtin_amdgcn_readfirstlane(__popcll(__builtin_amdgcn_ballot_w64(k < S && x <= capq)));
      K = K > 0 ? K : 1;  // the host sizes the slab for any single trace (min_bytes)
      if (lane < K) put<unsigned>(kTable + 8u * k, sl0 + QB * (unsigned)(x - nq - qlo) - SB * (unsigned)a.tau);
      const bool gated = __builtin_amdgcn_ballot_w64(lane < K && !(e >> 31)) != 0ull;
      if (lane == 0) put<unsigned short>(bat + 2u * nb, (unsigned short)(kb | (gated ? 0 : 0x8000)));
      nb++;
      kb += K;
    }
    if (lane == 0) {
      put<unsigned short>(bat + 2u * nb, (unsigned short)S);
      put<unsigned short>(bat + 2u * (S + 1), (unsigned short)nb);
    }
  }
  __syncthreads();
  return __builtin_amdgcn_readfirstlane((int)at<unsigned short>(bat + 2u * (S + 1)));
}

// Batch b's first trace (bit 15: every trace of the batch is ungated).
__device__ __forceinline__ int batch_at(int S, int b) {
  return __builtin_amdgcn_readfirstlane((int)at<unsigned short>((unsigned)batches(S) + 2u * b));
}

// Stage the batch [kb, ke) of FP32 rows (row_of(k): trace k's row): trace
// kb + j by the NW threads j NW .. j NW + NW - 1, thread r taking the
// trace's quads r, r + NW, ...; global loads through per-lane row pointers
// (a buffer resource would have to be uniform), quads up to the one holding
// sample ns - 1 and one sample more: inside the row's padding (stride >=
// ns + 8).
template <int NW, bool PAIRS, class RowOf>
__device__ __forceinline__ void stage(const SearchArgs& a, int S, int kb, int ke, const RowOf& row_of) {
  typedef float f4 __attribute__((ext_vector_type(4)));
  typedef const __attribute__((address_space(1))) float GF;
  typedef const __attribute__((address_space(1))) f4 GF4;
  constexpr unsigned SB = PAIRS ? 8u : 4u, QB = 4u * SB;
  constexpr int QSH = PAIRS ? 5 : 4;
  const unsigned sl0 = (unsigned)slab(S);
  const int j = (int)threadIdx.x / NW, r = (int)threadIdx.x % NW;
  if (j < ke - kb) {
    const int k = kb + j;
    const unsigned base = at<unsigned>(kTable + 8u * k), e = at<unsigned>(kTable + 8u * k + 4u);
    const int nq = (int)((e >> 16) & 0x7fffu), qlo = (int)(e & 0xffffu);
    // base = slab + QB (qoff - qlo) - SB tau
    const int qoff = ((int)(base - sl0 + SB * (unsigned)a.tau) >> QSH) + qlo;
    GF* row = (GF*)row_of(k);
    for (int q = r; q < nq; q += NW) {
      GF* p = row + 4 * (qlo + q);
      const f4 v = *(GF4*)p;
      const unsigned dst = sl0 + QB * (unsigned)(qoff + q);
      if constexpr (PAIRS) {
        const float nx = p[4];
        put<f4>(dst, f4{v.x, v.y - v.x, v.y, v.z - v.y});
        put<f4>(dst + 16u, f4{v.z, v.w - v.z, v.w, nx - v.w});
      } else {
        put<f4>(dst, v);
      }
    }
  }
}

}  // namespace staged
}  // namespace sb
