// Plain structs shared by the kernels (kernels.hip) and the host runtime
// (runtime.cpp).  No HIP types, so the host side compiles with g++.
#pragma once

#include <cstddef>
#include <cstdint>

namespace sb {

// One entry of a supergather (CRS) or of a CDP's own trace list (CMP), with
// geometry scaled to samples: m * idt, m^2 * idt^2, h * idt^2 where m is the
// midpoint distance to the CDP's reference point (m) and h = half-offset^2/1e6.
// Computed in FP64 on the device, stored FP32, with the row's amplitude
// scale (1 for FP32 storage) and, in CRS fast slots, the byte address of the
// row's first sample.  32 bytes: one scalar load.
struct SlotF {
  int32_t row;
  float m, m2, h;
  float scale, scale2;
  uint64_t addr;
};

// Reference-arithmetic slot for SB_MODE_EXACT CRS (CRS/OpenMP/src/main.cpp:189-198):
// m2 = dx*dx + dy*dy, m = sqrt(m2), h = h0 -- all double, unscaled.
struct SlotD {
  double m, m2, h;
  int32_t row;
  int32_t pad;
};

// Per-trace geometry as uploaded.
struct TraceGeom {
  int32_t cdp, scalco, sx, sy, gx, gy;
};

// One block's work: a CDP, a t0 tile and a parameter chunk.  cdp < 0: no-op.
struct WorkItem {
  int32_t cdp;
  int32_t tile;    // t0 tile index (tile * 64 * R is the first t0)
  int32_t p0, p1;  // parameter range [p0, p1)
};

// Bytes of device memory before row 0 of the samples (at least SB_MAX_W / 2
// FP32 samples, a multiple of 256 so rows stay 256-byte aligned).
constexpr int kSamplesLead = 256;

// Everything a search kernel reads.  Device pointers.
struct SearchArgs {
  const void* samples;     // [ntr][stride] in the store type
  const float* scale;      // [ntr] per-trace scale (F16/BF16) or null
  int64_t stride;          // elements per row
  const int32_t* slot_off; // [ncdp+1]
  const SlotF* slots;      // fast slots
  const SlotD* slots_d;    // exact CRS slots (or null)
  const SlotF* slots_p;    // unscaled FP32 CRS slots (CRS-DE exact mode)
  const float* h_ref32;    // exact CMP h per row (or null)
  const int32_t* cdp_off;  // [ncdp+1] rows of each CDP
  int ncdp, ns, tau, w;
  int cdp_base, nout;      // outputs cover CDPs [cdp_base, cdp_base + nout)
  int method;              // SB_METHOD_*
  // grid: CMP uses c only.  Float values for the fast search.
  float fa0, finc_a, fb0, finc_b, fc0, finc_c;
  int na, nb, nc, npar;
  // reference-precision grid and time constants
  double da0, dinc_a, db0, dinc_b, dc0, dinc_c;  // CRS: double
  float cf0, cfinc;                              // CMP: float c0*1e6, inc
  float dt_f, idt_f;                             // CMP float dt, idt
  double dt_d, idt_d;                            // CRS double dt, idt
  // CMP fast search: a window is valid iff t^2 = T0^2 + c h < hi2 (and its
  // start lies in [0, ns - w)), hi2 = (ns - tau - 1)^2 lowered by 2^-20 so
  // that the hardware square root of any t^2 below it starts a window
  // inside the trace.  A bound on t^2 is monotone in h, which the prefix
  // path of the search relies on.
  float hi2;
  // outputs
  unsigned long long* keys;  // [nout][ns] packed best (sem bits << 32 | ~p)
  int32_t* out_index;
  float *out_a, *out_b, *out_c, *out_coher, *out_stack;
  // work
  const WorkItem* work;
  int nwork;
  unsigned int* queue;       // persistent-kernel counters (8, one per XCD) or null
  int lds_bytes;             // dynamic LDS per block of the LDS-staged search
  int pair_share;            // mixed slabs: sixteenths of the slab bytes the paired slabs may take
};

// DE settings.
struct DeArgs {
  int pop, ngen;
  float F;
  uint32_t cr_thr;  // crossover when (hash >> 8) < cr_thr, i.e. CR * 2^24
  uint64_t seed;
  float lo[3], hi[3];  // box; c scaled by 1e6
};

}  // namespace sb
