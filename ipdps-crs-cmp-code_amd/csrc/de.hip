// CRS-DE: differential evolution over the CRS (A, B, C) box per (CDP, t0).
//
// The reference only describes this method (README.md:82-131: binary
// crs-ocl-de, flags -ngen/-azimuth); the algorithm is DE/rand/1/bin
// (Storn & Price 1997) with the CRS semblance of CRS/OpenMP/src/main.cpp:218-265
// as fitness (invalid -> 0), a synchronous generation update, clipping to the
// box and ">=" selection.  Random draws are a counter-based hash of
// (seed, CDP index, t0, generation, member, draw), so the CPU twin draws the
// same numbers.
//
// Mapping: one lane per population member; a segment of P2 = pow2 >= pop
// lanes holds one (CDP, t0), so a wave carries 64/P2 consecutive t0.  Each
// lane walks the CDP's supergather serially for its own candidate; the
// mutation reads r1/r2/r3 from sibling lanes with ds_bpermute.
#include <type_traits>

#include "device_common.hpp"
#include "launch.hpp"
#include "staged.hpp"

namespace sb {

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__device__ __forceinline__ uint32_t de_hash(uint64_t seed, uint32_t cdp, uint32_t t0, uint32_t gen, uint32_t member,
                                            uint32_t draw) {
  uint64_t k = mix64(seed ^ (0x9E3779B97F4A7C15ull * ((uint64_t)cdp + 1)));
  k = mix64(k ^ (((uint64_t)t0 << 32) | gen));
  k = mix64(k ^ (((uint64_t)member << 32) | draw));
  return (uint32_t)(k >> 32);
}

__device__ __forceinline__ float de_unit(uint32_t h) { return (float)(h >> 8) * (1.0f / 16777216.0f); }

// Index among n - nex candidates, mapped past the excluded members.
__device__ __forceinline__ int de_pick(uint32_t h, int n, int e0, int e1, int e2, int nex) {
  int v = (int)(h % (uint32_t)(n - nex));
  // sort the exclusions (nex <= 3)
  int x0 = e0, x1 = e1, x2 = e2;
  if (nex >= 2 && x1 < x0) {
    int t = x0;
    x0 = x1;
    x1 = t;
  }
  if (nex == 3) {
    if (x2 < x1) {
      int t = x1;
      x1 = x2;
      x2 = t;
    }
    if (x1 < x0) {
      int t = x0;
      x0 = x1;
      x1 = t;
    }
  }
  if (v >= x0) v++;
  if (nex >= 2 && v >= x1) v++;
  if (nex == 3 && v >= x2) v++;
  return v;
}

// Fitness in the oracle's FP32 operation order (no contraction).
template <class LD>
__device__ void de_eval_exact(const SearchArgs& a, const LD& ld, int k0, int k1, int t0, float pa, float pb, float pc,
                              float& sem, float& stk) {
#pragma clang fp contract(off)
  const int w = a.w;
  float num[SB_MAX_W];
#pragma unroll
  for (int j = 0; j < SB_MAX_W; j++) num[j] = 0.f;
  float den = 0.f, lin = 0.f, mm = 0.f;
  int err = 0;
  const float t0s = a.dt_f * (float)t0;
  const float lo = (float)a.tau, hi = (float)(a.ns - a.tau - 1);
  for (int k = k0; k < k1; k++) {
    const SlotF s = a.slots_p[k];
    float t = t0s + pa * s.m;
    float tsq = t * t;
    float bm = pb * s.m2;
    float ch = pc * s.h;
    t = tsq + bm;
    t = t + ch;
    t = t < 0.0f ? -1.0f : sqrtf(t);
    const float ti = t * a.idt_f;
    if (ti >= lo && ti < hi) {
      const int it = (int)ti;
      const float x = ti - (float)it;
      const int e = it - a.tau;
      float s1 = ld(s.row, e), s0;
#pragma unroll
      for (int j = 0; j < SB_MAX_W; j++) {
        if (j < w) {
          s0 = s1;
          s1 = ld(s.row, e + j + 1);
          float d = s1 - s0;
          float v = d * x + s0;
          num[j] += v;
          float vv = v * v;
          den += vv;
          lin += v;
        }
      }
      mm += 1.0f;
    } else {
      err++;
    }
  }
  float acsq = 0.f;
#pragma unroll
  for (int j = 0; j < SB_MAX_W; j++)
    if (j < w) {
      float q = num[j] * num[j];
      acsq += q;
    }
  if ((double)den > 1e-13 && (double)mm > 1e-13 && err < 2) {
    sem = acsq / (den * mm);
    stk = lin / ((float)w * mm);
  } else {
    sem = 0.f;
    stk = 0.f;
  }
}

// Fast fitness: slot geometry scaled to samples, contracted FMAs, hardware
// sqrt, window loads from a raw buffer (out-of-range lanes read zeros).
template <int W, int STORE>
__device__ void de_eval_fast(const SearchArgs& a, int k0, int k1, float T0, float pa, float pb, float pc, float& sem,
                             float& stk) {
  constexpr int ESZ = STORE == SB_STORE_F32 ? 4 : 2;
  const unsigned nwin = (unsigned)max(a.ns - a.w, 0);
  const int nrec = (int)(a.stride * ESZ);
  float num[W], den = 0.f;
  int cnt = 0;
#pragma unroll
  for (int j = 0; j < W; j++) num[j] = 0.f;
  // Software-pipelined by one trace: the window of trace k+1 is loaded
  // while trace k accumulates (one load per candidate and trace, so without
  // this every wave waits out each load's latency).  Same order of sums.
  // 16-bit windows are restored to FP32 (stored value times the trace's
  // power-of-two scale, exact) and evaluated like FP32 ones, so a 16-bit
  // store gives the FP32 store's results on the restored samples -- which is
  // what the search on a restored FP32 copy (SB_OPT_RESTORE_FP32) computes.
  const float T0sq = __fmul_rn(T0, T0);
  const float K1 = __fmul_rn(__fmul_rn(2.f, T0), pa), K2 = fmaf(pa, pa, pb);
  typedef Win<W + 1> WinT;
  auto issue = [&](const SlotF& s, WinT& sv, float& x, bool& ok) {
    // the slot carries its row's address: no per-trace address arithmetic
    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc((void*)s.addr, (short)0, nrec, kRsrcWord3);
    // (T0 + a m)^2 + b m^2 + c h expanded: three FMAs per trace
    const float tt = fmaf(K2, s.m2, fmaf(K1, s.m, fmaf(pc, s.h, T0sq)));
    // t^2 < 0 is invalid (the reference's t = -1): clamp to keep the
    // fraction finite; the window test rejects it unless tau = 0.
    const float ts = __builtin_amdgcn_sqrtf(fmaxf(tt, 0.f));
    const int e = (int)ts - a.tau;
    ok = (unsigned)e < nwin;
    if constexpr (W == 1) ok = ok && tt >= 0.f;
    x = __builtin_amdgcn_fractf(ts);
    const unsigned off = ok ? (STORE == SB_STORE_F32 ? (unsigned)e * 4u : ((unsigned)e & ~1u) * 2u) : kOob;
    sv = load_win<W + 1, STORE>(rs, off, ((unsigned)e & 1u) << 4);
  };
  auto consume = [&](const SlotF& s, const WinT& sv, float x, bool ok) {
    cnt += ok ? 1 : 0;
    float r[W + 1];
#pragma unroll
    for (int j = 0; j <= W; j++) r[j] = STORE == SB_STORE_F32 ? sv.v[j] : sv.v[j] * s.scale;
#pragma unroll
    for (int j = 0; j < W; j++) {
      const float v = fmaf(r[j + 1] - r[j], x, r[j]);
      num[j] += v;
      den = fmaf(v, v, den);
    }
  };
  // Slots are read two traces ahead of their use, so the scalar loads are in
  // flight while a trace accumulates.  The main loop runs while two more
  // slots exist, so it reads them by a walking pointer with no clamping; the
  // last one to three traces follow.  Traces are summed in order.
  if (k0 < k1) {
    WinT wa, wb;
    float xa, xb;
    bool oa, ob;
    const SlotF* __restrict__ sp = a.slots + k0;
    SlotF sa = sp[0];
    issue(sa, wa, xa, oa);
    if (k1 - k0 >= 2) {
      SlotF sb = sp[1];
      int k = k0;
      for (; k + 3 < k1; k += 2, sp += 2) {  // traces k, k+1 summed; k+2, k+3 exist
        const SlotF sc = sp[2];
        issue(sb, wb, xb, ob);
        consume(sa, wa, xa, oa);
        const SlotF sd = sp[3];
        issue(sc, wa, xa, oa);
        consume(sb, wb, xb, ob);
        sa = sc;
        sb = sd;
      }
      // traces k (issued), k + 1 and, when k + 2 < k1, k + 2
      issue(sb, wb, xb, ob);
      consume(sa, wa, xa, oa);
      if (k + 2 < k1) {
        const SlotF sc = sp[2];
        issue(sc, wa, xa, oa);
        consume(sb, wb, xb, ob);
        consume(sc, wa, xa, oa);
      } else {
        consume(sb, wb, xb, ob);
      }
    } else {
      consume(sa, wa, xa, oa);
    }
  }
  float acsq = 0.f, lin = 0.f;
#pragma unroll
  for (int j = 0; j < W; j++) {
    acsq = fmaf(num[j], num[j], acsq);
    lin += num[j];
  }
  const int S = k1 - k0;
  const float m = (float)cnt;
  if (den > 1e-13f && cnt > 0 && S - cnt < 2) {
    sem = acsq * __builtin_amdgcn_rcpf(den * m);
    stk = lin * __builtin_amdgcn_rcpf((float)W * m);
  } else {
    sem = 0.f;
    stk = 0.f;
  }
}

template <int W, bool EXACT, int STORE, class LD>
__global__ __launch_bounds__(256) void k_crs_de(SearchArgs a, DeArgs d, LD ld) {
#pragma clang fp contract(off)
  const WorkItem wi = a.work[blockIdx.x];
  if (wi.cdp < 0) return;
  const int cdp = wi.cdp;
  const int P2 = d.pop <= 4 ? 4 : d.pop <= 8 ? 8 : d.pop <= 16 ? 16 : d.pop <= 32 ? 32 : 64;
  const int per_block = 256 / P2;
  const int lane = threadIdx.x & 63;
  const int seg = (int)threadIdx.x / P2;      // t0 slot within the block
  const int member = (int)threadIdx.x % P2;
  const int seg_base = (lane / P2) * P2;       // first lane of my segment in the wave
  const int t0 = wi.tile * per_block + seg;
  const bool live = member < d.pop && t0 < a.ns;
  const int t0c = t0 < a.ns ? t0 : a.ns - 1;
  const int k0 = a.slot_off[cdp], k1 = a.slot_off[cdp + 1];

  auto eval = [&](float pa, float pb, float pc, float& sem, float& stk) {
    if constexpr (EXACT)
      de_eval_exact(a, ld, k0, k1, t0c, pa, pb, pc, sem, stk);
    else
      de_eval_fast<W, STORE>(a, k0, k1, (float)t0c, pa, pb, pc, sem, stk);
  };

  float x[3], fit = 0.f, stk = 0.f;
  const int mi = member < d.pop ? member : 0;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const float u = de_unit(de_hash(d.seed, cdp, t0c, 0, mi, k));
    const float span = d.hi[k] - d.lo[k];
    const float us = u * span;
    x[k] = d.lo[k] + us;
  }
  eval(x[0], x[1], x[2], fit, stk);

  for (int gen = 1; gen <= d.ngen; gen++) {
    const int r1 = de_pick(de_hash(d.seed, cdp, t0c, gen, mi, 0), d.pop, mi, 0, 0, 1);
    const int r2 = de_pick(de_hash(d.seed, cdp, t0c, gen, mi, 1), d.pop, mi, r1, 0, 2);
    const int r3 = de_pick(de_hash(d.seed, cdp, t0c, gen, mi, 2), d.pop, mi, r1, r2, 3);
    const int jr = (int)(de_hash(d.seed, cdp, t0c, gen, mi, 3) % 3u);
    float tr[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
      // every lane of the segment takes part in the shuffles
      const float x1 = __shfl(x[k], seg_base + r1, 64);
      const float x2 = __shfl(x[k], seg_base + r2, 64);
      const float x3 = __shfl(x[k], seg_base + r3, 64);
      const uint32_t u = de_hash(d.seed, cdp, t0c, gen, mi, 4 + k) >> 8;
      if (u < d.cr_thr || k == jr) {
        float diff = x2 - x3;
        float fd = d.F * diff;
        float v = x1 + fd;
        tr[k] = fminf(fmaxf(v, d.lo[k]), d.hi[k]);
      } else {
        tr[k] = x[k];
      }
    }
    float tf, ts;
    eval(tr[0], tr[1], tr[2], tf, ts);
    if (tf >= fit) {
      x[0] = tr[0];
      x[1] = tr[1];
      x[2] = tr[2];
      fit = tf;
      stk = ts;
    }
  }

  // Best member of the segment: largest fitness, lowest index on ties.
  float bf = live ? fit : -1.f;
  int bi = member;
  for (int o = 1; o < P2; o <<= 1) {
    const float of = __shfl_xor(bf, o, 64);
    const int oi = __shfl_xor(bi, o, 64);
    if (of > bf || (of == bf && oi < bi)) {
      bf = of;
      bi = oi;
    }
  }
  const float ba = __shfl(x[0], seg_base + bi, 64);
  const float bb = __shfl(x[1], seg_base + bi, 64);
  const float bc = __shfl(x[2], seg_base + bi, 64);
  const float bs = __shfl(stk, seg_base + bi, 64);
  if (member == 0 && t0 < a.ns) {
    const int64_t o = (int64_t)(cdp - a.cdp_base) * a.ns + t0;
    a.out_index[o] = -1;
    a.out_a[o] = ba;
    a.out_b[o] = bb;
    a.out_c[o] = (float)((double)bc / 1e6);
    a.out_coher[o] = bf;
    a.out_stack[o] = bs;
  }
}

// ----------------------------------------------------- LDS-staged CRS-DE
// The fast evaluator with windows from LDS.  k_crs_de gives every
// candidate of every trace its own scattered 16-byte window load, and the
// vector-memory data path (TD ~0.99 busy) co-bounds it with its VALU work.
// Here the workgroup -- NT = 64 NW / P2 consecutive t0 of one CDP, P2 lanes
// (members) each -- walks the supergather in batches of consecutive traces
// that all its candidates evaluate together:
//
// * Every trial stays in the (A, B, C) box (initial draws inside it, trials
//   clipped to it), so a trace's window starts over the whole run are bounded
//   by the box and the tile's t0 span: (T0 + a m)^2 + b m^2 + c h between
//   its smallest and largest values over that box, square-rooted, widened by
//   2 samples (the FP32 traveltime's rounding moves floor(t) by far less)
//   and clipped to the valid starts [0, ns - w).  The bound holds for every
//   generation, so the batches are planned once per workgroup: up to 64
//   consecutive traces whose ranges, as 16-byte quads aligned as in the row,
//   fit the slab; an LDS table keeps each trace's quads and window base.
// * A trace whose unclipped bound lies inside the valid starts has no gated
//   window for any candidate (and t^2 >= (tau + 2)^2 > 0), and a batch of
//   such traces runs without the window test, the clamp before the square
//   root and the per-trace count: 17 instead of 22 VALU a window.
// * Each batch is staged by NW threads per trace (global loads through
//   per-lane row pointers) as (sample, forward difference) pairs -- the FP32
//   subtraction the trace loop would make; one FMA a window sample, W
//   8-byte reads -- or as samples (PAIRS = false: 4 bytes, W + 1 reads, the
//   difference formed per window); every lane then runs the batch's traces
//   in order, the window at byte base + 8 floor(t) (or 4 floor(t)); gated
//   windows read the zeros at address 0.
//
// The arithmetic, the trace order of every sum and the gating are those of
// de_eval_fast (a paired slab holds the same difference, a gated window is
// zero either way), so both kernels produce the same bits: same fitness,
// same selections, same outputs.  FP32 rows only (16-bit stores run on their
// restored FP32 copy, or on k_crs_de).
template <int W, int NW, bool PAIRS>
__global__ __launch_bounds__(64 * NW) void k_crs_de_lds(SearchArgs a, DeArgs d) {
  // as k_crs_de: the initial draws and the mutation are separate roundings
  // (plain operators under this pragma: HIP's __fmul_rn / __fadd_rn are
  // header functions whose operations contraction would still fuse); the
  // trace loop spells out its FMAs
#pragma clang fp contract(off)
  constexpr int NT = 64 * NW;
  const WorkItem wi = a.work[blockIdx.x];
  if (wi.cdp < 0) return;  // the whole workgroup
  const int cdp = wi.cdp;
  const int P2 = d.pop <= 4 ? 4 : d.pop <= 8 ? 8 : d.pop <= 16 ? 16 : d.pop <= 32 ? 32 : 64;
  const int per_block = NT / P2;
  const int tid = threadIdx.x, lane = tid & 63;
  const int seg = tid / P2;
  const int member = tid % P2;
  const int seg_base = (lane / P2) * P2;
  const int t0first = wi.tile * per_block;
  const int t0 = t0first + seg;
  const bool live = member < d.pop && t0 < a.ns;
  const int t0c = t0 < a.ns ? t0 : a.ns - 1;
  const int k0 = a.slot_off[cdp];
  const int S = a.slot_off[cdp + 1] - k0;
  const float T0lo = (float)t0first, T0hi = (float)min(t0first + per_block - 1, a.ns - 1);
  const unsigned nwin = (unsigned)max(a.ns - a.w, 0);
  const SlotF* __restrict__ sl = a.slots + k0;

  // ---- plan (once): the batches and window bases
  const staged::Box box{{d.lo[0], d.lo[1], d.lo[2]}, {d.hi[0], d.hi[1], d.hi[2]}};
  const int nbatch = staged::plan<W, NT, PAIRS>(a, sl, S, T0lo, T0hi, box);
  auto stage = [&](int kb, int ke) {
    staged::stage<NW, PAIRS>(a, S, kb, ke, [&](int k) { return (const float*)(size_t)sl[k].addr; });
  };

  const float T0 = (float)t0c;
  const float T0sq = T0 * T0;
  float num[W], den;
  int cnt;
  // One batch's traces [kb, ke) for this lane's candidate, in order, with
  // de_eval_fast's arithmetic; lane j holds trace kb + j's window base.
  // The window at LDS byte addr into the sums: W (sample, difference) pairs
  // (8-byte aligned) or W + 1 samples (4-byte aligned: ds_read2_b32 pairs),
  // interpolated as de_eval_fast interpolates.
  auto accumulate = [&](unsigned addr, float x) {
    if constexpr (PAIRS) {
      typedef float f2 __attribute__((ext_vector_type(2)));
      f2 pr[W];
#pragma unroll
      for (int q = 0; q < W; q++) pr[q] = staged::at<f2>(addr + 8u * q);
#pragma unroll
      for (int q = 0; q < W; q++) {
        const float v = fmaf(pr[q].y, x, pr[q].x);
        num[q] += v;
        den = fmaf(v, v, den);
      }
    } else {
      float r[W + 1];
#pragma unroll
      for (int q = 0; q <= W; q++) r[q] = staged::at<float>(addr + 4u * q);
#pragma unroll
      for (int q = 0; q < W; q++) {
        const float v = fmaf(r[q + 1] - r[q], x, r[q]);
        num[q] += v;
        den = fmaf(v, v, den);
      }
    }
  };
  auto run_batch = [&](int kb, int ke, bool ungated, float K1, float K2, float pc) {
    typedef const __attribute__((address_space(4))) SlotF CSlot;
    const int K = ke - kb;
    const unsigned mybase = lane < K ? staged::at<unsigned>(staged::kTable + 8u * (unsigned)(kb + lane)) : 0u;
    CSlot* sk = (CSlot*)(size_t)(sl + kb);
    if (ungated) {
      // every window valid (the plan's bound): no test, no count per trace
      for (int j = 0; j < K; j++) {
        const float m = sk[j].m, m2 = sk[j].m2, h = sk[j].h;
        const unsigned base = __builtin_amdgcn_readlane(mybase, j);
        const float tt = fmaf(K2, m2, fmaf(K1, m, fmaf(pc, h, T0sq)));
        const float ts = __builtin_amdgcn_sqrtf(tt);
        const float x = __builtin_amdgcn_fractf(ts);
        accumulate(base + ((unsigned)(int)ts << (PAIRS ? 3 : 2)), x);
      }
      cnt += K;
      return;
    }
    for (int j = 0; j < K; j++) {
      const float m = sk[j].m, m2 = sk[j].m2, h = sk[j].h;
      const unsigned base = __builtin_amdgcn_readlane(mybase, j);
      const float tt = fmaf(K2, m2, fmaf(K1, m, fmaf(pc, h, T0sq)));
      const float ts = __builtin_amdgcn_sqrtf(fmaxf(tt, 0.f));
      const int it = (int)ts;
      bool ok = (unsigned)(it - a.tau) < nwin;
      if constexpr (W == 1) ok = ok && tt >= 0.f;
      const float x = __builtin_amdgcn_fractf(ts);
      const unsigned addr = ok ? base + ((unsigned)it << (PAIRS ? 3 : 2)) : 0u;
      cnt += ok ? 1 : 0;
      accumulate(addr, x);
    }
  };
  // Fitness of (pa, pb, pc) over the whole supergather.  A supergather that
  // is one batch is staged once for every generation.
  const int b0 = staged::batch_at(S, 0);
  if (nbatch == 1) {
    stage(0, S);
    __syncthreads();
  }
  auto fitness = [&](float pa, float pb, float pc, float& sem, float& stk) {
    const float K1 = (2.f * T0) * pa, K2 = fmaf(pa, pa, pb);
#pragma unroll
    for (int q = 0; q < W; q++) num[q] = 0.f;
    den = 0.f;
    cnt = 0;
    if (nbatch == 1) {
      run_batch(0, S, (b0 & 0x8000) != 0, K1, K2, pc);
    } else {
      for (int b = 0; b < nbatch; b++) {
        const int vb = staged::batch_at(S, b), kb = vb & 0x7fff, ke = staged::batch_at(S, b + 1) & 0x7fff;
        __syncthreads();  // every lane is done with the previous batch's slab
        stage(kb, ke);
        __syncthreads();
        run_batch(kb, ke, (vb & 0x8000) != 0, K1, K2, pc);
      }
    }
    float acsq = 0.f, lin = 0.f;
#pragma unroll
    for (int q = 0; q < W; q++) {
      acsq = fmaf(num[q], num[q], acsq);
      lin += num[q];
    }
    const float mm = (float)cnt;
    if (den > 1e-13f && cnt > 0 && S - cnt < 2) {
      sem = acsq * __builtin_amdgcn_rcpf(den * mm);
      stk = lin * __builtin_amdgcn_rcpf((float)W * mm);
    } else {
      sem = 0.f;
      stk = 0.f;
    }
  };

  float x[3], fit = 0.f, stk = 0.f;
  const int mi = member < d.pop ? member : 0;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const float u = de_unit(de_hash(d.seed, cdp, t0c, 0, mi, k));
    const float span = d.hi[k] - d.lo[k];
    const float us = u * span;
    x[k] = d.lo[k] + us;
  }
  fitness(x[0], x[1], x[2], fit, stk);

  for (int gen = 1; gen <= d.ngen; gen++) {
    const int r1 = de_pick(de_hash(d.seed, cdp, t0c, gen, mi, 0), d.pop, mi, 0, 0, 1);
    const int r2 = de_pick(de_hash(d.seed, cdp, t0c, gen, mi, 1), d.pop, mi, r1, 0, 2);
    const int r3 = de_pick(de_hash(d.seed, cdp, t0c, gen, mi, 2), d.pop, mi, r1, r2, 3);
    const int jr = (int)(de_hash(d.seed, cdp, t0c, gen, mi, 3) % 3u);
    float tr[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
      const float x1 = __shfl(x[k], seg_base + r1, 64);
      const float x2 = __shfl(x[k], seg_base + r2, 64);
      const float x3 = __shfl(x[k], seg_base + r3, 64);
      const uint32_t u = de_hash(d.seed, cdp, t0c, gen, mi, 4 + k) >> 8;
      if (u < d.cr_thr || k == jr) {
        const float diff = x2 - x3;
        const float fd = d.F * diff;
        const float v = x1 + fd;
        tr[k] = fminf(fmaxf(v, d.lo[k]), d.hi[k]);
      } else {
        tr[k] = x[k];
      }
    }
    float tf, ts;
    fitness(tr[0], tr[1], tr[2], tf, ts);
    if (tf >= fit) {
      x[0] = tr[0];
      x[1] = tr[1];
      x[2] = tr[2];
      fit = tf;
      stk = ts;
    }
  }

  float bf = live ? fit : -1.f;
  int bi = member;
  for (int o = 1; o < P2; o <<= 1) {
    const float of = __shfl_xor(bf, o, 64);
    const int oi = __shfl_xor(bi, o, 64);
    if (of > bf || (of == bf && oi < bi)) {
      bf = of;
      bi = oi;
    }
  }
  const float ba = __shfl(x[0], seg_base + bi, 64);
  const float bb = __shfl(x[1], seg_base + bi, 64);
  const float bc = __shfl(x[2], seg_base + bi, 64);
  const float bs = __shfl(stk, seg_base + bi, 64);
  if (member == 0 && t0 < a.ns) {
    const int64_t o = (int64_t)(cdp - a.cdp_base) * a.ns + t0;
    a.out_index[o] = -1;
    a.out_a[o] = ba;
    a.out_b[o] = bb;
    a.out_c[o] = (float)((double)bc / 1e6);
    a.out_coher[o] = bf;
    a.out_stack[o] = bs;
  }
}

int de_lds_min_bytes(int ns, int smax, int pairs) {
  return staged::min_bytes(ns, smax, pairs != 0);
}

bool de_lds_supported(int w, int pop, int waves, int pairs, int lds_bytes, int ns, int smax) {
  return de_supported(w, pop) && (w == 1 || w == 3 || w == 5 || w == 7) &&
         (waves == 4 || waves == 8 || waves == 16) && 64 * waves >= pop && smax < 32768 &&
         lds_bytes >= de_lds_min_bytes(ns, smax, pairs) && lds_bytes <= 64 * 1024;
}

template <int NW, bool PAIRS>
static hipError_t de_lds_w(const SearchArgs& a, const DeArgs& d, int nblocks, hipStream_t st) {
  const size_t sh = (size_t)a.lds_bytes;
  switch (a.w) {
    case 1: hipLaunchKernelGGL((k_crs_de_lds<1, NW, PAIRS>), dim3(nblocks), dim3(64 * NW), sh, st, a, d); break;
    case 3: hipLaunchKernelGGL((k_crs_de_lds<3, NW, PAIRS>), dim3(nblocks), dim3(64 * NW), sh, st, a, d); break;
    case 5: hipLaunchKernelGGL((k_crs_de_lds<5, NW, PAIRS>), dim3(nblocks), dim3(64 * NW), sh, st, a, d); break;
    case 7: hipLaunchKernelGGL((k_crs_de_lds<7, NW, PAIRS>), dim3(nblocks), dim3(64 * NW), sh, st, a, d); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_de_lds(const SearchArgs& a, const DeArgs& d, int waves, int pairs, int smax, int nblocks,
                         hipStream_t st) {
  if (nblocks <= 0) return hipSuccess;
  if (!de_lds_supported(a.w, d.pop, waves, pairs, a.lds_bytes, a.ns, smax)) return hipErrorInvalidValue;
  if (pairs) {
    if (waves == 4) return de_lds_w<4, true>(a, d, nblocks, st);
    if (waves == 8) return de_lds_w<8, true>(a, d, nblocks, st);
    return de_lds_w<16, true>(a, d, nblocks, st);
  }
  if (waves == 4) return de_lds_w<4, false>(a, d, nblocks, st);
  if (waves == 8) return de_lds_w<8, false>(a, d, nblocks, st);
  return de_lds_w<16, false>(a, d, nblocks, st);
}

bool de_supported(int w, int pop) { return pop >= 4 && pop <= 64 && w >= 1 && w <= SB_MAX_W; }

template <bool EXACT, int STORE, class LD>
static hipError_t de_w(const SearchArgs& a, const DeArgs& d, const LD& ld, int nblocks, hipStream_t st) {
  if constexpr (EXACT) {
    hipLaunchKernelGGL((k_crs_de<1, true, STORE, LD>), dim3(nblocks), dim3(256), 0, st, a, d, ld);
    return hipGetLastError();
  } else {
    switch (a.w) {
      case 1: hipLaunchKernelGGL((k_crs_de<1, false, STORE, LD>), dim3(nblocks), dim3(256), 0, st, a, d, ld); break;
      case 3: hipLaunchKernelGGL((k_crs_de<3, false, STORE, LD>), dim3(nblocks), dim3(256), 0, st, a, d, ld); break;
      case 5: hipLaunchKernelGGL((k_crs_de<5, false, STORE, LD>), dim3(nblocks), dim3(256), 0, st, a, d, ld); break;
      case 7: hipLaunchKernelGGL((k_crs_de<7, false, STORE, LD>), dim3(nblocks), dim3(256), 0, st, a, d, ld); break;
      default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
  }
}

hipError_t launch_de(const SearchArgs& a, const DeArgs& d, int store, int exact, int nblocks, hipStream_t st) {
  if (nblocks <= 0) return hipSuccess;
  if (store == SB_STORE_F32) {
    LoadF32 ld{(const float*)a.samples, a.stride};
    return exact ? de_w<true, SB_STORE_F32>(a, d, ld, nblocks, st) : de_w<false, SB_STORE_F32>(a, d, ld, nblocks, st);
  }
  if (store == SB_STORE_F16) {
    LoadF16 ld{(const __half*)a.samples, a.scale, a.stride};
    return exact ? de_w<true, SB_STORE_F16>(a, d, ld, nblocks, st) : de_w<false, SB_STORE_F16>(a, d, ld, nblocks, st);
  }
  LoadBF16 ld{(const uint16_t*)a.samples, a.scale, a.stride};
  return exact ? de_w<true, SB_STORE_BF16>(a, d, ld, nblocks, st) : de_w<false, SB_STORE_BF16>(a, d, ld, nblocks, st);
}

}  // namespace sb
