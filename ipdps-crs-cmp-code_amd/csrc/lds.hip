// LDS-staged CMP search (k_search_lds).
//
// Same work items, evaluations and 64-bit arg-max keys as k_search_fast
// (lanes = 64 consecutive t0 of one CDP, every lane walks the CDP's traces in
// slot order by increasing h), but each window is read from LDS: the global
// kernel spends ~16.6 CU-cycles of the vector-memory address path per 64
// window loads, two aligned ds_read2_b32 cost about half of that and leave
// the loop bound by its VALU work.
//
// A workgroup first bounds, for every trace, the window starts that any of
// its lanes can reach over its whole parameter range: e = floor(sqrt(T0^2 +
// c h)) - tau is monotone in T0 and in c >= 0, so the first t0 with the
// smallest c and the last t0 (t0min + 64R - 1, past ns for the last tile, so
// that every lane's reads stay inside the staged range) with the largest c
// bound it, widened by 2 for the approximate square root.  The ranges of the
// first K traces in slot order (all of them when they fit; the nearest
// offsets have the shortest ranges) are staged once as FP32 in 16-byte quads
// aligned as in the row (16-bit stores restored exactly: the stored value
// times the trace's power-of-two scale), and every wave then takes
// groups of CC parameters in turn, reading the windows of those traces at
// byte address base_k + 4 floor(t) (one v_lshl_add_u32; gated windows read
// the zeros at address 0) and the windows of the other traces from global
// memory as k_search_fast does.
//
// Paired slabs hold every sample with its forward difference (8 bytes a
// sample, the difference the trace loop would form), so a window sample is
// one FMA; they hold half the samples per byte.  PM selects the slabs:
// 0 samples only; 1 pairs only (the first K traces that fit); 2 pairs for
// the first Kp traces and samples for traces Kp .. K-1 -- every trace as
// pairs when that fits, else every trace staged with as many pairs as the
// rest of the slab holds when the samples fit, else pairs in at most
// pair_share/16 of the slab bytes and samples after them.  A window read
// from a paired slab costs 24 LDS bytes at w = 3 against 16 from a sample
// slab and 3 fewer VALU: mixing the two balances the loop between the LDS
// data path and VALU issue.
//
// Every rounding step matches k_search_fast, in the same trace order, with
// the same general/prefix split, last-trace gate and second-to-last-trace
// validity, so both kernels produce identical keys.  Every store runs the
// FP32 evaluation on its restored samples, so a trace's contribution is the
// same whether its window comes from LDS or from global memory.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "device_common.hpp"
#include "device_types.hpp"
#include "launch.hpp"
#include "semblance.h"

namespace sb {

namespace {
// (x << 2) + y in one v_lshl_add_u32, y in a VGPR.
__device__ __forceinline__ unsigned lshl2_add(int x, unsigned y) {
  unsigned d;
  asm("v_lshl_add_u32 %0, %1, 2, %2" : "=v"(d) : "v"(x), "v"(y));
  return d;
}
// (x << 1) + y and (x << 4) + y, y uniform (the compiler would reassociate
// them into an add and a shift).
__device__ __forceinline__ unsigned lshl1_add(int x, unsigned y) {
  unsigned d;
  asm("v_lshl_add_u32 %0, %1, 1, %2" : "=v"(d) : "v"(x), "s"(y));
  return d;
}
__device__ __forceinline__ unsigned lshl3_add(int x, unsigned y) {
  unsigned d;
  asm("v_lshl_add_u32 %0, %1, 3, %2" : "=v"(d) : "v"(x), "v"(y));
  return d;
}
__device__ __forceinline__ unsigned lshl4_add(int x, unsigned y) {
  unsigned d;
  asm("v_lshl_add_u32 %0, %1, 4, %2" : "=v"(d) : "v"(x), "s"(y));
  return d;
}

// LDS by byte address (the dynamic LDS starts at address 0): the window
// reads take the computed address as is.
template <class T>
__device__ __forceinline__ T lds_at(unsigned addr) {
  return *(const __attribute__((address_space(3))) T*)(size_t)addr;
}

// One trace of the CDP as the trace loop uses it: h and the 16-bit stores'
// scale (restores windows read from global memory) from the slot table
// (scalar loads), the window base from LDS.
struct TraceRec {
  float h;
  unsigned base;  // the window of traveltime t starts at LDS byte base + 4 floor(t)
  float scale;
};

// A trace's entry in the LDS table (h for the range scan, then the base).
struct TraceLds {
  float h;
  unsigned base;
};

// One evaluation in flight: its window (W + 1 samples, or W (sample,
// difference) pairs from a paired slab) and interpolation fraction.
template <int W>
struct Ev {
  float v[2 * W > W + 1 ? 2 * W : W + 1];
  float x;
  bool ok;
};
}  // namespace

// Dynamic LDS, from address 0 (the kernel has no other LDS): 64 zero bytes,
// the scan scratch, then per-CDP tables and the slabs.  At the end the
// waves' keys reuse the space from byte 64.
template <int NW>
struct LdsLayout {
  static constexpr int kScan = 64;
  static constexpr int kTr = kScan + ((4 * NW + 15) & ~15);
  // First slab byte for a CDP of S traces: TraceLds[S], then s_pre[S+1],
  // s_lo[S].
  static __host__ __device__ constexpr int slab(int S) { return kTr + ((8 * S + 4 * (2 * S + 1) + 15) & ~15); }
};

template <int W, int R, int CC, int STORE, int NW, int PM>
__global__ __launch_bounds__(64 * NW) void k_search_lds(SearchArgs a) {
  using L = LdsLayout<NW>;
  constexpr int NT = 64 * NW;
  constexpr int RP = NW * CC;  // parameters per pass of the block's waves
  constexpr int ESZ = STORE == SB_STORE_F32 ? 4 : 2;
  static_assert(PM >= 0 && PM <= 2, "slab layout");
  extern __shared__ __align__(16) unsigned char dyn[];
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);  // uniform: scalar trace indices
  const WorkItem wi = a.work[blockIdx.x];
  if (wi.cdp < 0) return;
  const int cdp = wi.cdp;
  const int k0 = a.slot_off[cdp];
  const int S = a.slot_off[cdp + 1] - k0;
  if (S <= 0) return;
  const SlotF* __restrict__ sl = a.slots + k0;

  int* s_wsum = (int*)(dyn + L::kScan);
  TraceLds* s_tr = (TraceLds*)(dyn + L::kTr);
  int* s_pre = (int*)(dyn + L::kTr + 8 * S);
  int* s_lo = s_pre + (S + 1);
  const int smp = L::slab(S);
  // Slab capacity in 16-byte units: a quad of samples takes one, a quad of
  // (sample, difference) pairs two.
  const int cap16 = (a.lds_bytes - smp) >> 4;

  if (tid < 16) ((uint32_t*)dyn)[tid] = 0u;
  for (int k = tid; k < S; k += NT) s_tr[k] = TraceLds{sl[k].h, 0u};  // h read back below by the same thread

  const int nwin = max(a.ns - a.w, 0);
  const int t0min = wi.tile * 64 * R;
  const int t0max = t0min + 64 * R - 1;
  const float sqlo = __fmul_rn((float)t0min, (float)t0min), sqhi = __fmul_rn((float)t0max, (float)t0max);
  float T0sq[R];
#pragma unroll
  for (int r = 0; r < R; r++) {
    const float T0 = (float)(t0min + r * 64 + lane);
    T0sq[r] = __fmul_rn(T0, T0);
  }
  // k_search_fast's prefix path: traces before the last run ungated, the last
  // one gated by the t^2 bound, validity from the second-to-last trace.
  const bool prefix = t0min > a.tau;
  const float h2 = S >= 2 ? sl[S - 2].h : 0.f;
  // Global-memory window addressing, as k_search_fast's: FP32 rows from tau
  // samples before their start, 16-bit rows from tau rounded up to even.
  const int tau_even = (a.tau + 1) & ~1, even_adj = tau_even - a.tau;
  const unsigned adj2 = 2u * (unsigned)even_adj, adj16 = 16u * (unsigned)even_adj;
  unsigned long long best[R];
#pragma unroll
  for (int r = 0; r < R; r++) best[r] = 0ull;

  // ---- one wave's group of CC parameters [p, p + CC) (those below qe)
  float pc[CC];
  bool act[CC];
  bool head_ok[R][CC];
  bool wlive = true;
  float num[R][CC][W], den[R][CC];
  int cnt[R][CC];
  auto begin = [&](int p, int qe) {
#pragma unroll
    for (int ci = 0; ci < CC; ci++) {
      act[ci] = p + ci < qe;
      pc[ci] = __fadd_rn(a.fc0, __fmul_rn(a.finc_c, (float)(act[ci] ? p + ci : p)));
    }
    wlive = true;
    if (prefix) {
      bool live = false;
#pragma unroll
      for (int r = 0; r < R; r++)
#pragma unroll
        for (int ci = 0; ci < CC; ci++) {
          head_ok[r][ci] = S < 2 || fmaf(pc[ci], h2, T0sq[r]) < a.hi2;
          live = live || (head_ok[r][ci] && act[ci]);
        }
      // no result possible for any lane and parameter of the group: the
      // wave skips it (same keys as k_search_fast's skipped wave)
      wlive = __builtin_amdgcn_ballot_w64(live) != 0ull;
    }
#pragma unroll
    for (int r = 0; r < R; r++)
#pragma unroll
      for (int ci = 0; ci < CC; ci++) {
        den[r][ci] = 0.f;
        cnt[r][ci] = 0;
#pragma unroll
        for (int j = 0; j < W; j++) num[r][ci][j] = 0.f;
      }
  };

  // Window sources of a trace: global memory, a sample slab, a paired slab.
  constexpr int kGlobal = 0, kSamples = 1, kPairs = 2;
  // Reads the windows of trace `t` (slabs: `t.base`; global memory: `row`)
  // for every (row, parameter).  GATE 0: window and t^2 test (general path);
  // 1: none; 2: the t^2 bound (last trace).  Every source gives the same
  // values: the stored samples converted exactly to FP32 (a paired slab also
  // the differences the consumer would form), zeros for gated windows.
  auto issue = [&](const TraceRec& t, int row, Ev<W>(&ev)[R][CC], auto gate_tag, auto src_tag) {
    constexpr int GATE = decltype(gate_tag)::value;
    constexpr int SRC = decltype(src_tag)::value;
    constexpr bool FROM_LDS = SRC != kGlobal;
    __amdgpu_buffer_rsrc_t rs;
    if constexpr (!FROM_LDS) {
      const int lead = STORE == SB_STORE_F32 ? 4 * a.tau : 2 * tau_even;
      rs = __builtin_amdgcn_make_buffer_rsrc(
          (void*)((const char*)a.samples + (int64_t)row * a.stride * ESZ - lead), (short)0,
          (int)(a.stride * ESZ) + lead, kRsrcWord3);
    }
#pragma unroll
    for (int r = 0; r < R; r++) {
#pragma unroll
      for (int ci = 0; ci < CC; ci++) {
        const float tt = fmaf(pc[ci], t.h, T0sq[r]);
        const float ts = __builtin_amdgcn_sqrtf(tt);  // tt >= 0: the host checks c0, c1 >= 0
        const int it = (int)ts;
        bool ok = true;
        if constexpr (GATE == 0) ok = (unsigned)(it - a.tau) < (unsigned)nwin && tt < a.hi2;
        if constexpr (GATE == 2) ok = tt < a.hi2;
        ev[r][ci].ok = ok;
        ev[r][ci].x = __builtin_amdgcn_fractf(ts);
        if constexpr (FROM_LDS) {
          unsigned addr = SRC == kPairs ? lshl3_add(it, t.base) : lshl2_add(it, t.base);
          if constexpr (GATE != 1) addr = ok ? addr : 0u;
          if constexpr (SRC == kPairs) {  // 8-byte aligned (sample, difference) pairs
            typedef float f2 __attribute__((ext_vector_type(2)));
#pragma unroll
            for (int j = 0; j < W; j++) {
              const f2 q = lds_at<f2>(addr + 8u * j);
              ev[r][ci].v[2 * j] = q.x;
              ev[r][ci].v[2 * j + 1] = q.y;
            }
          } else {
#pragma unroll
            for (int j = 0; j <= W; j++) ev[r][ci].v[j] = lds_at<float>(addr + 4u * j);
          }
        } else {
          unsigned off, shf = 0u;
          if constexpr (STORE == SB_STORE_F32) {
            off = (unsigned)it * 4u;
          } else {
            off = lshl1_add(it, adj2) & ~3u;
            shf = lshl4_add(it, adj16);
          }
          if constexpr (GATE != 1) off = ok ? off : kOob;
          const Win<W + 1> wv = load_win<W + 1, STORE>(rs, off, shf);
#pragma unroll
          for (int j = 0; j <= W; j++) ev[r][ci].v[j] = STORE == SB_STORE_F32 ? wv.v[j] : wv.v[j] * t.scale;
        }
      }
    }
  };
  // k_search_fast's window arithmetic, operation for operation (a paired
  // slab holds the difference the other windows form: the same bits).
  auto consume = [&](const Ev<W>(&ev)[R][CC], auto gate_tag, auto src_tag) {
    constexpr int GATE = decltype(gate_tag)::value;
    constexpr bool PAIRED = decltype(src_tag)::value == kPairs;
#pragma unroll
    for (int r = 0; r < R; r++) {
#pragma unroll
      for (int ci = 0; ci < CC; ci++) {
        const Ev<W>& e = ev[r][ci];
        if constexpr (GATE != 1) cnt[r][ci] += e.ok ? 1 : 0;
        const float x = e.x;
#pragma unroll
        for (int j = 0; j < W; j++) {
          const float v = PAIRED ? fmaf(e.v[2 * j + 1], x, e.v[2 * j])
                                 : fmaf(__fsub_rn(e.v[j + 1], e.v[j]), x, e.v[j]);
          num[r][ci][j] = __fadd_rn(num[r][ci][j], v);
          den[r][ci] = fmaf(v, v, den[r][ci]);
        }
      }
    }
  };
  // Trace k's record: h and scales by scalar loads from the slot table (the
  // kernel only reads it), the base by a 4-byte LDS read (a 16-byte record
  // read per trace and wave returned 1 KiB over the LDS data path the window
  // reads use).
  typedef const __attribute__((address_space(4))) SlotF CSlot;
  CSlot* slc = (CSlot*)(size_t)sl;
  auto tr = [&](int k, bool with_base) {
    TraceRec t;
    t.h = slc[k].h;
    t.base = with_base ? lds_at<unsigned>((unsigned)(L::kTr + 8 * k + 4)) : 0u;
    t.scale = STORE == SB_STORE_F32 || with_base ? 1.f : slc[k].scale;
    return t;
  };
  // Traces [kb, ke): each trace's windows are read and accumulated in turn
  // (the waves of the CU hide the latency); global-memory windows take the
  // row from the slot table (scalar: uniform row addressing).
  auto run = [&](int kb, int ke, auto gate_tag, auto src_tag) {
    constexpr bool FROM_LDS = decltype(src_tag)::value != kGlobal;
    for (int k = kb; k < ke; k++) {
      const TraceRec t = tr(k, FROM_LDS);
      Ev<W> A[R][CC];
      issue(t, FROM_LDS ? 0 : slc[k].row, A, gate_tag, src_tag);
      consume(A, gate_tag, src_tag);
    }
  };
  // All traces of the group, in slot order: traces [0, Kp) from paired
  // slabs, [Kp, K) from sample slabs, [K, S) from global memory.
  auto accumulate = [&](int Kp, int K) {
    using G0 = std::integral_constant<int, 0>;
    using G1 = std::integral_constant<int, 1>;
    using G2 = std::integral_constant<int, 2>;
    using SG = std::integral_constant<int, kGlobal>;
    using SS = std::integral_constant<int, kSamples>;
    using SP = std::integral_constant<int, kPairs>;
    if (prefix) {
      const int kp = min(Kp, S - 1), ks = min(K, S - 1);
      if constexpr (PM != 0)
        if (kp > 0) run(0, kp, G1{}, SP{});
      if constexpr (PM != 1)
        if (ks > kp) run(kp, ks, G1{}, SS{});
      if (S - 1 > ks) run(ks, S - 1, G1{}, SG{});
      if (PM != 0 && Kp == S)
        run(S - 1, S, G2{}, SP{});
      else if (PM != 1 && K == S)
        run(S - 1, S, G2{}, SS{});
      else
        run(S - 1, S, G2{}, SG{});
    } else {
      if constexpr (PM != 0)
        if (Kp > 0) run(0, Kp, G0{}, SP{});
      if constexpr (PM != 1)
        if (K > Kp) run(Kp, K, G0{}, SS{});
      if (S > K) run(K, S, G0{}, SG{});
    }
  };
  // The group's results into the running keys.
  auto finish = [&](int p) {
    if (prefix) {
#pragma unroll
      for (int r = 0; r < R; r++)
#pragma unroll
        for (int ci = 0; ci < CC; ci++) cnt[r][ci] += S - 1;
    } else {
#pragma unroll
      for (int r = 0; r < R; r++)
#pragma unroll
        for (int ci = 0; ci < CC; ci++) head_ok[r][ci] = (S - cnt[r][ci]) < 2;
    }
#pragma unroll
    for (int r = 0; r < R; r++) {
#pragma unroll
      for (int ci = 0; ci < CC; ci++) {
        float acsq = 0.f;
#pragma unroll
        for (int jj = 0; jj < W; jj++) acsq = fmaf(num[r][ci][jj], num[r][ci][jj], acsq);
        const float m = (float)cnt[r][ci];
        const bool valid = act[ci] && den[r][ci] > 1e-13f && cnt[r][ci] > 0 && head_ok[r][ci];
        const float sem = __fmul_rn(acsq, __builtin_amdgcn_rcpf(__fmul_rn(den[r][ci], m)));
        if (valid && sem > 0.f) {
          const unsigned long long key =
              ((unsigned long long)__float_as_uint(sem) << 32) | (unsigned long long)(~(uint32_t)(p + ci));
          best[r] = key > best[r] ? key : best[r];
        }
      }
    }
  };

  // ---- the work item's parameter range [p0, p1): c in [cmin, cmax]
  const float cA = __fadd_rn(a.fc0, __fmul_rn(a.finc_c, (float)wi.p0));
  const float cB = __fadd_rn(a.fc0, __fmul_rn(a.finc_c, (float)(wi.p1 - 1)));
  const float cmin = fminf(cA, cB), cmax = fmaxf(cA, cB);
  // No lane of the block has a result when the smallest c gives an
  // out-of-range second-to-last t^2 at the first t0 (the same float values
  // as that lane's); block-uniform.
  const bool dead = prefix && S >= 2 && !(fmaf(cmin, h2, sqlo) < a.hi2);
  int K = 0;   // traces staged in LDS (the first K in slot order)
  int Kp = 0;  // of which the first Kp as (sample, difference) pairs
  if (!dead) {
    // ---- each trace's range of window starts: e = floor(sqrt(T0^2 + c h))
    // - tau is monotone in T0 and in c >= 0, so (first t0, cmin) and (last
    // t0, cmax) bound every lane's, widened by 2 for the approximate square
    // root.  As 16-byte quads of samples aligned as in the row (a staged quad
    // is one 16-byte load and one aligned LDS store), prefix-summed.
    int carry = 0;
    for (int kb = 0; kb < S; kb += NT) {
      const int k = kb + tid;
      int lo = 0, len = 0;
      if (k < S) {
        const float h = s_tr[k].h;
        const int elo = max((int)__builtin_amdgcn_sqrtf(fmaf(cmin, h, sqlo)) - a.tau - 2, 0);
        const int ehi = (int)__builtin_amdgcn_sqrtf(fmaf(cmax, h, sqhi)) - a.tau + 2;
        lo = elo >> 2;
        len = ehi >= elo ? ((ehi + W) >> 2) - lo + 1 : 0;  // quads of samples elo .. ehi + W
      }
      int x = len;
#pragma unroll
      for (int d = 1; d < 64; d <<= 1) {
        const int y = __shfl_up(x, d);
        if (lane >= d) x += y;
      }
      if (lane == 63) s_wsum[wave] = x;
      __syncthreads();
      int off = carry, tot = 0;
#pragma unroll
      for (int q = 0; q < NW; q++) {
        off += q < wave ? s_wsum[q] : 0;
        tot += s_wsum[q];
      }
      if (k < S) {
        s_pre[k] = off + x - len;
        s_lo[k] = lo;
      }
      carry += tot;
      __syncthreads();
    }
    // Stage the first K traces, as many as the slabs hold (all when every
    // range fits; block-uniform): the nearest offsets, whose ranges are the
    // shortest.  The other traces' windows are read from global memory as
    // k_search_fast reads them.
    if (tid == 0) s_pre[S] = carry;
    __syncthreads();  // s_pre[S]; every read of the scan scratch is done
    const int total = __builtin_amdgcn_readfirstlane(carry);  // s_pre[S] (carry is uniform)
    // The largest k in [0, S] with s_pre[k] <= x (x >= 0, block-uniform):
    // s_pre is nondecreasing from s_pre[0] = 0, so when s_pre[S] > x exactly
    // one thread sees the boundary s_pre[k] <= x < s_pre[k + 1].
    auto largest_le = [&](int x) {
      if (total <= x) return S;
      for (int k = tid; k < S; k += NT)
        if (s_pre[k] <= x && s_pre[k + 1] > x) s_wsum[0] = k;
      __syncthreads();
      const int r = __builtin_amdgcn_readfirstlane(s_wsum[0]);
      __syncthreads();  // read by every thread before the next search writes it
      return r;
    };
    if constexpr (PM == 0) {
      K = largest_le(cap16);
    } else if constexpr (PM == 1) {
      K = Kp = largest_le(cap16 >> 1);
    } else {
      if (2 * total <= cap16) {
        K = Kp = S;
      } else if (total <= cap16) {
        K = S;
        Kp = largest_le(cap16 - total);
      } else {
        Kp = largest_le((a.pair_share * cap16) >> 5);
        K = largest_le(cap16 - __builtin_amdgcn_readfirstlane(s_pre[Kp]));
      }
    }
    // The paired slabs take 32 bytes a quad from smp, the sample slabs 16
    // from sbase, after them.
    const int pre_kp = PM == 2 ? __builtin_amdgcn_readfirstlane(s_pre[Kp]) : 0;
    const unsigned sbase = (unsigned)(smp + 32 * pre_kp);
    if (K > 0) {
      // A wave takes G consecutive traces at a time, lanes over their
      // quads, every load of the group in flight before the stores; past the
      // row the loads read zeros.  Sets each trace's window address base.
      constexpr int G = 8;
      typedef float f4 __attribute__((ext_vector_type(4)));
      typedef unsigned u2 __attribute__((ext_vector_type(2)));
      for (int kg = wave * G; kg < K; kg += NW * G) {
        __amdgpu_buffer_rsrc_t rs[G];
        int qlo[G], nq[G];
        unsigned dq[G];
        bool pg[G];   // the trace's slab is paired (uniform)
        float sc[G];  // 16-bit stores: the trace's scale (restored exactly: a power of two)
        int nqm = 0;
#pragma unroll
        for (int g = 0; g < G; g++) {
          const int k = min(kg + g, K - 1);
          const int qb = __builtin_amdgcn_readfirstlane(s_pre[k]);
          qlo[g] = __builtin_amdgcn_readfirstlane(s_lo[k]);
          nq[g] = kg + g < K ? __builtin_amdgcn_readfirstlane(s_pre[k + 1]) - qb : 0;
          nqm = max(nqm, nq[g]);
          pg[g] = PM == 1 || (PM == 2 && k < Kp);
          // paired: 32 bytes a quad from smp; samples: 16 from sbase
          const int qs = pg[g] ? 32 : 16;
          dq[g] = pg[g] ? (unsigned)(smp + 32 * qb) : sbase + 16u * (unsigned)(qb - pre_kp);
          if (lane == 0 && kg + g < K) s_tr[k].base = dq[g] - (unsigned)(qs * qlo[g] + (qs / 4) * a.tau);
          const int row = slc[k].row;
          sc[g] = STORE == SB_STORE_F32 ? 1.f : slc[k].scale;
          rs[g] = __builtin_amdgcn_make_buffer_rsrc((void*)((const char*)a.samples + (int64_t)row * a.stride * ESZ),
                                                    (short)0, (int)(a.stride * ESZ), kRsrcWord3);
        }
        for (int u = lane; u < nqm; u += 64) {
          f4 v[G];
          float nx[G];  // paired slabs: the sample after the quad (its last difference)
#pragma unroll
          for (int g = 0; g < G; g++) {
            const unsigned o = (unsigned)(qlo[g] + u) * (4 * ESZ);
            if constexpr (PM != 0) {
              if constexpr (STORE == SB_STORE_F32) {
                nx[g] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rs[g], o + 16u, 0, 0));
              } else {
                const unsigned n2 = __builtin_amdgcn_raw_buffer_load_b32(rs[g], o + 8u, 0, 0);
                if constexpr (STORE == SB_STORE_F16) {
                  typedef _Float16 h2v __attribute__((ext_vector_type(2)));
                  nx[g] = (float)__builtin_bit_cast(h2v, n2).x * sc[g];
                } else {
                  nx[g] = __uint_as_float(n2 << 16) * sc[g];
                }
              }
            }
            if constexpr (STORE == SB_STORE_F32) {
              v[g] = __builtin_amdgcn_raw_buffer_load_b128(rs[g], o, 0, 0);
            } else {
              const u2 d = __builtin_amdgcn_raw_buffer_load_b64(rs[g], o, 0, 0);
              // element by element: a bit_cast of a vector element reads element 0
              const unsigned d0 = d.x, d1 = d.y;
              if constexpr (STORE == SB_STORE_F16) {
                typedef _Float16 h2v __attribute__((ext_vector_type(2)));
                const h2v lo2 = __builtin_bit_cast(h2v, d0);
                const h2v hi2 = __builtin_bit_cast(h2v, d1);
                v[g] = f4{(float)lo2.x * sc[g], (float)lo2.y * sc[g], (float)hi2.x * sc[g], (float)hi2.y * sc[g]};
              } else {
                v[g] = f4{__uint_as_float(d0 << 16) * sc[g], __uint_as_float(d0 & 0xffff0000u) * sc[g],
                          __uint_as_float(d1 << 16) * sc[g], __uint_as_float(d1 & 0xffff0000u) * sc[g]};
              }
            }
          }
#pragma unroll
          for (int g = 0; g < G; g++) {
            if (u >= nq[g]) continue;
            typedef __attribute__((address_space(3))) f4 LF4;
            if (pg[g]) {
              const f4 q = v[g];
              const float d0 = __fsub_rn(q.y, q.x), d1 = __fsub_rn(q.z, q.y), d2 = __fsub_rn(q.w, q.z),
                          d3 = __fsub_rn(nx[g], q.w);
              *(LF4*)(size_t)(dq[g] + 32u * u) = f4{q.x, d0, q.y, d1};
              *(LF4*)(size_t)(dq[g] + 32u * u + 16u) = f4{q.z, d2, q.w, d3};
            } else {
              *(LF4*)(size_t)(dq[g] + 16u * u) = v[g];
            }
          }
        }
      }
      __syncthreads();
    }
    // ---- every wave takes groups of CC parameters in turn (the fully
    // staged case in a loop of its own: it holds only the LDS trace loops)
    if (K == S) {
      for (int p = wi.p0 + wave * CC; p < wi.p1; p += RP) {
        begin(p, wi.p1);
        if (!wlive) continue;
        accumulate(Kp, S);
        finish(p);
      }
    } else {
      for (int p = wi.p0 + wave * CC; p < wi.p1; p += RP) {
        begin(p, wi.p1);
        if (!wlive) continue;
        accumulate(Kp, K);
        finish(p);
      }
    }
  }

  __syncthreads();  // every wave is done with the tables and slabs
  unsigned long long* s_best = (unsigned long long*)(dyn + 64);
#pragma unroll
  for (int r = 0; r < R; r++) s_best[wave * 64 * R + r * 64 + lane] = best[r];
  __syncthreads();
  if (wave == 0) {
#pragma unroll
    for (int r = 0; r < R; r++) {
      unsigned long long m = s_best[r * 64 + lane];
#pragma unroll
      for (int q = 1; q < NW; q++) {
        const unsigned long long v = s_best[q * 64 * R + r * 64 + lane];
        m = v > m ? v : m;
      }
      const int t0 = t0min + r * 64 + lane;
      if (t0 < a.ns && m) atomicMax(&a.keys[(int64_t)(cdp - a.cdp_base) * a.ns + t0], m);
    }
  }
}

// ------------------------------------------------------------- launchers
template <int W, int R, int CC, int NW, int PM>
static hipError_t lds_store(const SearchArgs& a, int store, int nblocks, hipStream_t st) {
  const size_t sh = (size_t)a.lds_bytes;
  if (store == SB_STORE_F32)
    hipLaunchKernelGGL((k_search_lds<W, R, CC, SB_STORE_F32, NW, PM>), dim3(nblocks), dim3(64 * NW), sh, st, a);
  else if (store == SB_STORE_F16)
    hipLaunchKernelGGL((k_search_lds<W, R, CC, SB_STORE_F16, NW, PM>), dim3(nblocks), dim3(64 * NW), sh, st, a);
  else
    hipLaunchKernelGGL((k_search_lds<W, R, CC, SB_STORE_BF16, NW, PM>), dim3(nblocks), dim3(64 * NW), sh, st, a);
  return hipGetLastError();
}

template <int R, int CC, int NW, int PM = 0>
static hipError_t lds_w(const SearchArgs& a, int store, int nblocks, hipStream_t st) {
  switch (a.w) {
    case 1: return lds_store<1, R, CC, NW, PM>(a, store, nblocks, st);
    case 3: return lds_store<3, R, CC, NW, PM>(a, store, nblocks, st);
    case 5: return lds_store<5, R, CC, NW, PM>(a, store, nblocks, st);
    case 7: return lds_store<7, R, CC, NW, PM>(a, store, nblocks, st);
    default: return hipErrorInvalidValue;
  }
}

int lds_slab_offset(const FastVariant& v, int S) {
  switch (v.waves) {
    case 4: return LdsLayout<4>::slab(S);
    case 8: return LdsLayout<8>::slab(S);
    case 16: return LdsLayout<16>::slab(S);
    default: return -1;
  }
}

hipError_t launch_search_lds(const SearchArgs& a, int store, const FastVariant& v, int nblocks, hipStream_t st) {
  if (nblocks <= 0) return hipSuccess;
  if (a.method != SB_METHOD_CMP || a.lds_bytes <= 0 || a.lds_bytes > kLdsMaxBytes) return hipErrorInvalidValue;
  if (a.lds_bytes < 64 + v.waves * 64 * v.R * 8) return hipErrorInvalidValue;  // the keys at the end
  if (v.R == 1 && v.CC == 4 && v.waves == 8) {
    if (v.pairs == 1) return lds_w<1, 4, 8, 1>(a, store, nblocks, st);
    if (v.pairs == 2) return lds_w<1, 4, 8, 2>(a, store, nblocks, st);
    return lds_w<1, 4, 8>(a, store, nblocks, st);
  }
  if (v.R == 1 && v.CC == 4 && v.waves == 16) return lds_w<1, 4, 16>(a, store, nblocks, st);
  if (v.R == 2 && v.CC == 2 && v.waves == 8) return lds_w<2, 2, 8>(a, store, nblocks, st);
  if (v.R == 2 && v.CC == 2 && v.waves == 16) return lds_w<2, 2, 16>(a, store, nblocks, st);
  if (v.R == 2 && v.CC == 2 && v.waves == 4) return lds_w<2, 2, 4>(a, store, nblocks, st);
  return hipErrorInvalidValue;
}

}  // namespace sb
