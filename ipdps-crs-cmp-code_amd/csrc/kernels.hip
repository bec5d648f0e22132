// gfx950 kernels of the CMP / CRS semblance search.
//
// Layout in HBM: samples [ntr][stride] (trace-major, CDP-binned), one slot
// table per method (CMP: each CDP's own traces; CRS: each CDP's supergather of
// neighbour-CDP traces, CRS/OpenMP/src/main.cpp:179-204) holding the trace row
// and its scaled geometry, and per-(CDP, t0) result arrays.
//
// Hot kernel (k_search_fast): lanes of a wave are 64 consecutive t0 of one
// CDP; every lane walks the CDP's traces in order for a small chunk of
// parameters, so a wave's 64 sample-window loads for one trace fall on a
// few consecutive cache lines (moveout is smooth in t0) and the trace
// geometry and parameter products are wave-uniform.  Windowed sums stay in
// registers; the argmax over parameters is fused (64-bit packed key,
// atomicMax across parameter chunks), so the reference's num/stt arrays
// (CMP/OpenMP/src/main.cpp:93-94) are never materialised.  The winner of
// every (CDP, t0) is re-evaluated by k_finalize in the reference's precision
// and operation order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <type_traits>
#include <hip/hip_fp16.h>
#include <hip/hip_bf16.h>

#include "device_common.hpp"
#include "device_types.hpp"
#include "launch.hpp"
#include "semblance.h"

namespace sb {

// ------------------------------------------------------ per-trace geometry
// CMP reference arithmetic (float, CMP/OpenMP/src/main.cpp:126-136) and the
// FP64 geometry (the CRS reference arithmetic, CRS/OpenMP/src/main.cpp:164-177).
// msign: the sign of the reference's coordinate factor (-1 for a negative
// scalco, whose 1/scalco factor negates the midpoints; harmless for its
// distances, undone where a signed projection needs true coordinates).
__global__ void k_prep_traces(const TraceGeom* __restrict__ g, int ntr, float* __restrict__ h_ref32,
                              double* __restrict__ h0, double* __restrict__ m0x, double* __restrict__ m0y,
                              float* __restrict__ msign) {
#pragma clang fp contract(off)
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ntr) return;
  const TraceGeom t = g[i];
  {
    float s = (float)t.scalco;
    if (-1e-13 < (double)s && (double)s < 1e-13)
      s = 1.0f;
    else if (s < 0)
      s = 1.0f / s;
    float hx = ((float)t.gx - (float)t.sx) * s;
    float hy = ((float)t.gy - (float)t.sy) * s;
    float hxx = hx * hx;
    float hyy = hy * hy;
    float hh = hxx + hyy;
    h_ref32[i] = (float)(0.25 * (double)hh / 1e6);
  }
  {
    double s = (double)t.scalco;
    if (-1e-13 < s && s < 1e-13)
      s = 1.0;
    else if (s < 0)
      s = 1.0 / s;
    m0x[i] = ((double)t.gx + (double)t.sx) * s * 0.5;
    m0y[i] = ((double)t.gy + (double)t.sy) * s * 0.5;
    msign[i] = s < 0 ? -1.f : 1.f;
    double hx = ((double)t.gx - (double)t.sx) * s;
    double hy = ((double)t.gy - (double)t.sy) * s;
    double hxx = hx * hx;
    double hyy = hy * hy;
    h0[i] = 0.25 * (hxx + hyy) / 1e6;
  }
}

// CMP slot table of CDPs [cdp0, cdp0 + gridDim.x), one block per CDP: the
// CDP's traces ordered by increasing H = h * idt^2 (FP32, ties by row), so
// that the valid traces of any (t0, c >= 0) form a prefix of the list (the
// fast search's prefix path).  The fast search sums in this order; the
// reference-order evaluations read rows directly.
__global__ __launch_bounds__(256) void k_cmp_slots(int cdp0, const int32_t* __restrict__ cdp_off,
                                                   const double* __restrict__ h0, const float* __restrict__ scale,
                                                   double idt, SlotF* __restrict__ slots) {
  __shared__ float s_h[1024];
  const int i = cdp0 + blockIdx.x;
  const int r0 = cdp_off[i], r1 = cdp_off[i + 1];
  const int n = r1 - r0;
  for (int rb = 0; rb < n; rb += blockDim.x) {
    const int r = r0 + rb + (int)threadIdx.x;
    const bool mine = r < r1;
    const float h = mine ? (float)(h0[r] * idt * idt) : 0.f;
    int rank = 0;
    for (int qb = 0; qb < n; qb += 1024) {
      const int m = min(1024, n - qb);
      __syncthreads();
      for (int q = threadIdx.x; q < m; q += blockDim.x) s_h[q] = (float)(h0[r0 + qb + q] * idt * idt);
      __syncthreads();
      if (mine)
        for (int q = 0; q < m; q++) {
          const float hq = s_h[q];
          rank += (hq < h || (hq == h && qb + q < r - r0)) ? 1 : 0;
        }
    }
    if (mine) {
      SlotF s;
      s.row = r;
      s.m = 0.f;
      s.m2 = 0.f;
      s.h = h;
      s.scale = scale ? scale[r] : 1.f;
      s.scale2 = s.scale * s.scale;
      s.addr = 0;
      slots[r0 + rank] = s;
    }
  }
}

// CRS supergather slot tables, one block per CDP (CDPs cdp0 + blockIdx.x):
// neighbour CDPs in list order, traces in gather order; distances to the
// CDP's first trace (CRS/OpenMP/src/main.cpp:179-198, 287-311).  With az set
// (an extension the reference only names: crs-ocl-de's -azimuth,
// README.md:94) m is instead the signed projection of the displacement onto
// the azimuth direction, m = dx az_s + dy az_c, and m^2 = m m.
__global__ void k_crs_slots(int cdp0, const int32_t* __restrict__ cdp_off, const int32_t* __restrict__ nbr_off,
                            const int32_t* __restrict__ nbr, const int32_t* __restrict__ slot_off,
                            const double* __restrict__ m0x, const double* __restrict__ m0y,
                            const double* __restrict__ h0, const float* __restrict__ scale, double idt,
                            const char* samples, uint64_t row_bytes, SlotF* __restrict__ sf,
                            SlotD* __restrict__ sd, SlotF* __restrict__ sp, const float* __restrict__ msign, int az,
                            double az_s, double az_c) {
#pragma clang fp contract(off)
  const int i = cdp0 + blockIdx.x;
  const int ref = cdp_off[i];
  const double rx = m0x[ref], ry = m0y[ref];
  int base = slot_off[i];
  for (int q = nbr_off[i]; q < nbr_off[i + 1]; q++) {
    const int j = nbr[q];
    const int r0 = cdp_off[j], r1 = cdp_off[j + 1];
    for (int r = r0 + (int)threadIdx.x; r < r1; r += blockDim.x) {
      double dx = m0x[r] - rx;
      double dy = m0y[r] - ry;
      double m2, m;
      if (az) {
        // true coordinates: the reference's factor sign undone
        const double sr = (double)msign[ref], st = (double)msign[r];
        double tx = m0x[r] * st;
        double ty = m0y[r] * st;
        double rtx = rx * sr;
        double rty = ry * sr;
        dx = tx - rtx;
        dy = ty - rty;
        double px = dx * az_s;
        double py = dy * az_c;
        m = px + py;
        m2 = m * m;
      } else {
        double dxx = dx * dx;
        double dyy = dy * dy;
        m2 = dxx + dyy;
        m = sqrt(m2);
      }
      double h = h0[r];
      SlotD d;
      d.m = m;
      d.m2 = m2;
      d.h = h;
      d.row = r;
      d.pad = 0;
      sd[base + (r - r0)] = d;
      SlotF f;
      f.row = r;
      f.m = (float)(m * idt);
      f.m2 = (float)(m2 * idt * idt);
      f.h = (float)(h * idt * idt);
      f.scale = scale ? scale[r] : 1.f;
      f.scale2 = f.scale * f.scale;
      f.addr = (uint64_t)(samples + (uint64_t)r * row_bytes);
      sf[base + (r - r0)] = f;
      f.m = (float)m;
      f.m2 = (float)m2;
      f.h = (float)h;
      sp[base + (r - r0)] = f;
    }
    base += r1 - r0;
  }
}

// ------------------------------------------------------------ quantisation
// F16/BF16 store: per-trace power-of-two scale >= max|s|, so s/scale is exact
// before the final rounding and restoring is an exact FP32 multiply.
template <int STORE>
__global__ __launch_bounds__(256) void k_quantize(const float* __restrict__ src, int ns, int64_t stride,
                                                  void* __restrict__ dst, float* __restrict__ scale) {
  const int64_t row = blockIdx.x;
  const float* s = src + row * (int64_t)ns;
  float mx = 0.f;
  for (int i = threadIdx.x; i < ns; i += blockDim.x) mx = fmaxf(mx, fabsf(s[i]));
  __shared__ float red[256];
  red[threadIdx.x] = mx;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + o]);
    __syncthreads();
  }
  mx = red[0];
  float sc = 1.f;
  if (mx > 0.f) {
    int e;
    frexpf(mx, &e);  // mx = f * 2^e, f in [0.5, 1)
    sc = ldexpf(1.f, e);
  }
  const float inv = 1.f / sc;
  if (threadIdx.x == 0) scale[row] = sc;
  for (int i = threadIdx.x; i < (int)stride; i += blockDim.x) {
    float v = i < ns ? s[i] * inv : 0.f;
    if (STORE == SB_STORE_F16) {
      ((__half*)dst)[row * stride + i] = __float2half_rn(v);
    } else {
      ((__hip_bfloat16*)dst)[row * stride + i] = __float2bfloat16(v);
    }
  }
}

// Rows of a 16-bit store restored to FP32 -- the stored value times the
// row's power-of-two scale, the product every 16-bit search path forms --
// padding included, into a transient FP32 copy with the same row stride.
// Eight samples per thread (one 16-byte load, two 16-byte stores); rows of
// stride elements, stride a multiple of 64.
template <int STORE>
__global__ __launch_bounds__(256) void k_restore_rows(const uint4* __restrict__ src, const float* __restrict__ scale,
                                                      int64_t stride, int64_t nvec, float4* __restrict__ dst) {
  const int64_t v = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (v >= nvec) return;
  const float sc = scale[(v * 8) / stride];
  const uint4 q = src[v];
  float f[8];
  const unsigned w[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
  for (int k = 0; k < 4; k++) {
    if constexpr (STORE == SB_STORE_F16) {
      typedef _Float16 h2v __attribute__((ext_vector_type(2)));
      const h2v h = __builtin_bit_cast(h2v, w[k]);
      f[2 * k] = (float)h.x * sc;
      f[2 * k + 1] = (float)h.y * sc;
    } else {
      f[2 * k] = __uint_as_float(w[k] << 16) * sc;
      f[2 * k + 1] = __uint_as_float(w[k] & 0xffff0000u) * sc;
    }
  }
  dst[2 * v] = make_float4(f[0], f[1], f[2], f[3]);
  dst[2 * v + 1] = make_float4(f[4], f[5], f[6], f[7]);
}

// A CRS slot table whose row addresses point into a restored FP32 copy:
// addr = base + row * row_bytes (base: the copy's row-0 address).
__global__ __launch_bounds__(256) void k_rebase_slots(const SlotF* __restrict__ src, int64_t n, uint64_t base,
                                                      uint64_t row_bytes, SlotF* __restrict__ dst) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  SlotF s = src[i];
  s.addr = base + (uint64_t)s.row * row_bytes;
  dst[i] = s;
}

// Plain FP32 copy into the padded row layout.
__global__ __launch_bounds__(256) void k_pad_rows(const float* __restrict__ src, int ns, int64_t stride,
                                                  float* __restrict__ dst) {
  const int64_t row = blockIdx.x;
  for (int i = threadIdx.x; i < (int)stride; i += blockDim.x)
    dst[row * stride + i] = i < ns ? src[row * (int64_t)ns + i] : 0.f;
}

__global__ void k_fill_u64(unsigned long long* p, int64_t n, unsigned long long v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void k_fill_u32(unsigned int* p, int n, unsigned int v) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// Work table built on the device (the allocation-free entry point has no host
// staging): item slot b serves XCD b % 8 and takes item (b % 8) * per + b / 8
// of the (CDP, tile, chunk)-ordered list, so each XCD runs a contiguous run of
// CDPs (their blocks share that XCD's L2); slots past the list are no-ops.
__global__ void k_make_work(WorkItem* __restrict__ w, int nblocks, int cdp0, int ncdp, int ntiles, int nchunks,
                            int pchunk, int npar) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nblocks) return;
  const int64_t n = (int64_t)ncdp * ntiles * nchunks;
  const int64_t per = (n + 7) / 8;
  const int64_t i = (int64_t)(b % 8) * per + b / 8;
  WorkItem it{-1, 0, 0, 0};
  if (i < n) {
    const int64_t ct = i / nchunks;
    const int k = (int)(i - ct * nchunks);
    it.cdp = cdp0 + (int)(ct / ntiles);
    it.tile = (int)(ct - (ct / ntiles) * ntiles);
    it.p0 = k * pchunk;
    it.p1 = min(npar, (k + 1) * pchunk);
  }
  w[b] = it;
}

// The allocation-free entry point's scratch: one copy per device, alive until
// the process ends.  Host code never uses the name itself (on the host it is
// a stand-in, not a device address): scratch_address looks up the device
// address once per device and every launch takes that pointer.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

hipError_t scratch_address(int device, void** out) {
  *out = nullptr;
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress((void**)&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    if (!ptr) return hipErrorInvalidSymbol;
    g_scratch_ptr[device] = ptr;
  }
  *out = g_scratch_ptr[device];
  return hipSuccess;
}

// ---------------------------------------------- reference-order evaluation
// One (CDP, t0, c) of CMP exactly as CMP/OpenMP/src/main.cpp:150-203: float,
// no contraction, correctly rounded sqrt/div, traces in gather order.
// MAXW: the largest window the caller runs (w <= MAXW).
// KAHAN: the window, energy and stack sums carry a compensation term
// (Kahan summation) -- used to report 16-bit-store results, whose inputs
// already differ from the reference's, so that FP32 accumulation adds no
// error of its own; FP32 stores keep the reference's plain sums (and bits).
// EXACT_W: the window is exactly MAXW samples (the per-sample guards vanish).
template <int MAXW = SB_MAX_W, bool KAHAN = false, bool EXACT_W = false, class LD>
__device__ void eval_cmp_exact(const SearchArgs& a, const LD& ld, int cdp, int t0, int cid, float& sem,
                               float& stk) {
#pragma clang fp contract(off)
  const float c = a.cf0 + a.cfinc * (float)cid;
  float tt0 = a.dt_f * (float)t0;
  tt0 = tt0 * tt0;
  const float lo = (float)a.tau, hi = (float)(a.ns - a.tau - 1);
  const int w = EXACT_W ? MAXW : a.w;
  float num[MAXW];
#pragma unroll
  for (int j = 0; j < MAXW; j++) num[j] = 0.f;
  float den = 0.f, lin = 0.f, m = 0.f;
  float cnum[MAXW], cden = 0.f, clin = 0.f;  // Kahan compensations
#pragma unroll
  for (int j = 0; j < MAXW; j++) cnum[j] = 0.f;
  auto add = [](float& sum, float& comp, float v) {
    if constexpr (KAHAN) {
      const float y = v - comp;
      const float t = sum + y;
      comp = (t - sum) - y;
      sum = t;
    } else {
      sum += v;
    }
  };
  int err = 0;
  const int r0 = a.cdp_off[cdp], r1 = a.cdp_off[cdp + 1];
  // The next row's h is loaded while this row computes, and a window's
  // samples are all loaded before its sums (independent loads in flight;
  // the arithmetic and its order are unchanged).
  float h_next = a.h_ref32[r0];
  for (int row = r0; row < r1; row++) {
    const float h = h_next;
    if (row + 1 < r1) h_next = a.h_ref32[row + 1];
    float ch = c * h;
    float t = sqrtf(tt0 + ch) * a.idt_f;
    if (t >= lo && t < hi) {
      const int it = (int)t;
      const float x = t - (float)it;
      const int e = it - a.tau;
      float sv[MAXW + 1];
#pragma unroll
      for (int j = 0; j <= MAXW; j++)
        if (j <= w) sv[j] = ld(row, e + j);
      float s1 = sv[0], s0;
#pragma unroll
      for (int j = 0; j < MAXW; j++) {
        if (j < w) {
          s0 = s1;
          s1 = sv[j + 1];
          float d = s1 - s0;
          float v = d * x + s0;
          add(num[j], cnum[j], v);
          float vv = v * v;
          add(den, cden, vv);
          add(lin, clin, v);
        }
      }
      m += 1.0f;
    } else {
      err++;
    }
  }
  float acsq = 0.f;
#pragma unroll
  for (int j = 0; j < MAXW; j++)
    if (j < w) {
      float q = num[j] * num[j];
      acsq += q;
    }
  const float wf = (float)w;
  if ((double)den > 1e-13 && (double)m > 1e-13 && (double)wf > 1e-13 && err < 2) {
    sem = acsq / (den * m);
    stk = lin / (wf * m);
  } else {
    sem = -1.0f;
    stk = -1.0f;
  }
}

// One (CDP, t0, par) of CRS exactly as CRS/OpenMP/src/main.cpp:210-265: double.
template <int MAXW = SB_MAX_W, bool EXACT_W = false, class LD>
__device__ void eval_crs_exact(const SearchArgs& a, const LD& ld, int cdp, int t0, int p, double& sem,
                               double& stk) {
#pragma clang fp contract(off)
  const int nbnc = a.nc * a.nb;
  const double pa = a.da0 + a.dinc_a * (double)(p / nbnc);
  const double pb = a.db0 + a.dinc_b * (double)((p / a.nc) % a.nb);
  const double pc = a.dc0 + a.dinc_c * (double)(p % a.nc);
  const double tt0 = a.dt_d * (double)t0;
  const double lo = (double)a.tau, hi = (double)(a.ns - a.tau - 1);
  const int w = EXACT_W ? MAXW : a.w;
  double num[MAXW];
#pragma unroll
  for (int j = 0; j < MAXW; j++) num[j] = 0.0;
  double den = 0.0, lin = 0.0, mm = 0.0;
  int err = 0;
  const int k0 = a.slot_off[cdp], k1 = a.slot_off[cdp + 1];
  // Next slot loaded while this one computes; window samples loaded before
  // the sums (same arithmetic, same order).
  SlotD s_next = a.slots_d[k0 < k1 ? k0 : 0];
  for (int k = k0; k < k1; k++) {
    const SlotD s = s_next;
    if (k + 1 < k1) s_next = a.slots_d[k + 1];
    double t = tt0 + pa * s.m;
    double tsq = t * t;
    double bm = pb * s.m2;
    double ch = pc * s.h;
    t = tsq + bm;
    t = t + ch;
    t = t < 0.0 ? -1.0 : sqrt(t);
    const double ti = t * a.idt_d;
    if (ti >= lo && ti < hi) {
      const int it = (int)ti;
      const double x = ti - (double)it;
      const int e = it - a.tau;
      float sv[MAXW + 1];
#pragma unroll
      for (int j = 0; j <= MAXW; j++)
        if (j <= w) sv[j] = ld(s.row, e + j);
      double s1 = (double)sv[0], s0;
#pragma unroll
      for (int j = 0; j < MAXW; j++) {
        if (j < w) {
          s0 = s1;
          s1 = (double)sv[j + 1];
          double d = s1 - s0;
          double v = d * x + s0;
          num[j] += v;
          double vv = v * v;
          den += vv;
          lin += v;
        }
      }
      mm += 1.0;
    } else {
      err++;
    }
  }
  double acsq = 0.0;
#pragma unroll
  for (int j = 0; j < MAXW; j++)
    if (j < w) {
      double q = num[j] * num[j];
      acsq += q;
    }
  if (den > 1e-13 && mm > 1e-13 && (double)w > 1e-13 && err < 2) {
    sem = acsq / (den * mm);
    stk = lin / ((double)w * mm);
  } else {
    sem = 0.0;
    stk = 0.0;
  }
}

// Output encodings (CMP/OpenMP/src/main.cpp:219-221,253; CRS/OpenMP/src/main.cpp:282-284,348-350).
__device__ void write_cmp(const SearchArgs& a, int64_t o, int idx, float sem, float stk) {
#pragma clang fp contract(off)
  a.out_index[o] = idx;
  a.out_coher[o] = sem;
  a.out_stack[o] = stk;
  a.out_a[o] = 0.f;
  a.out_b[o] = 0.f;
  a.out_c[o] = idx < 0 ? 0.0f : (float)((double)(a.cf0 + a.cfinc * (float)idx) / 1e6);
}

__device__ void write_crs(const SearchArgs& a, int64_t o, int idx, double sem, double stk) {
#pragma clang fp contract(off)
  const int nbnc = a.nc * a.nb;
  a.out_index[o] = idx;
  a.out_coher[o] = (float)sem;
  a.out_stack[o] = (float)stk;
  a.out_a[o] = (float)(a.da0 + a.dinc_a * (double)(idx / nbnc));
  a.out_b[o] = (float)(a.db0 + a.dinc_b * (double)((idx / a.nc) % a.nb));
  a.out_c[o] = (float)((a.dc0 + a.dinc_c * (double)(idx % a.nc)) / 1e6);
}

// ------------------------------------------------------ SB_MODE_EXACT search
// One thread per (CDP, t0) scanning every parameter in order with the
// reference's strict '>' arg-max (CMP/OpenMP/src/main.cpp:207-222,
// CRS/OpenMP/src/main.cpp:270-285).  Work items carry (cdp, 64-t0 tile).
// MAXW: the largest window this instantiation runs (7 or SB_MAX_W), or with
// EXACT_W the window itself.
template <int MAXW, bool EXACT_W, class LD>
__global__ __launch_bounds__(64) void k_search_exact(SearchArgs a, LD ld) {
  const WorkItem wi = a.work[blockIdx.x];
  if (wi.cdp < 0) return;
  const int t0 = wi.tile * 64 + (int)threadIdx.x;
  if (t0 >= a.ns) return;
  const int64_t o = (int64_t)(wi.cdp - a.cdp_base) * a.ns + t0;
  if (a.method == SB_METHOD_CMP) {
    float max_sem = 0.f, best_stk = 0.f;
    int max_c = -1;
    for (int c = 0; c < a.nc; c++) {
      float sem, stk;
      eval_cmp_exact<MAXW, false, EXACT_W>(a, ld, wi.cdp, t0, c, sem, stk);
      if (sem > max_sem) {
        max_sem = sem;
        max_c = c;
        best_stk = stk;
      }
    }
    write_cmp(a, o, max_c, max_sem, max_c > -1 ? best_stk : 0.f);
  } else {
    double max_sem = 0.0, best_stk = 0.0;
    int max_p = -1;
    for (int p = 0; p < a.npar; p++) {
      double sem, stk;
      eval_crs_exact<MAXW, EXACT_W>(a, ld, wi.cdp, t0, p, sem, stk);
      if (sem > max_sem) {
        max_sem = sem;
        max_p = p;
        best_stk = stk;
      }
    }
    if (max_p < 0) {
      // No parameter beat 0: the reference keeps the absolute index 0, i.e.
      // (t0 = 0, par 0) -- stack = stt[0] of this CDP (quirk Q7).
      double sem0;
      eval_crs_exact<MAXW, EXACT_W>(a, ld, wi.cdp, 0, 0, sem0, best_stk);
      max_p = 0;
    }
    write_crs(a, o, max_p, max_sem, best_stk);
  }
}

// ------------------------------------------------------------- finalize
// Winners of windows up to 7 samples (the fast search's) are re-evaluated by
// the kFinalizeMaxW instantiation; wider windows (tiled reference-arithmetic
// search only) by the SB_MAX_W one.
constexpr int kFinalizeMaxW = 7;

// Decodes the packed best key of every (CDP, t0) and re-evaluates the winner
// in the reference's precision and order; writes the output arrays.
template <int MAXW, bool EXACT_W, class LD>
__global__ __launch_bounds__(256) void k_finalize(SearchArgs a, LD ld) {
  const int64_t n = (int64_t)a.nout * a.ns;
  const int64_t o = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= n) return;
  const int cdp = a.cdp_base + (int)(o / a.ns), t0 = (int)(o % a.ns);
  const unsigned long long key = a.keys[o];
  const int p = key ? (int)(~(uint32_t)(key & 0xffffffffu)) : -1;
  if (a.method == SB_METHOD_CMP) {
    float sem = 0.f, stk = 0.f;
    if (p >= 0) eval_cmp_exact<MAXW, LD::kQuantized, EXACT_W>(a, ld, cdp, t0, p, sem, stk);
    if (p >= 0 && sem > 0.f)
      write_cmp(a, o, p, sem, stk);
    else
      write_cmp(a, o, -1, 0.f, 0.f);
  } else {
    double sem = 0.0, stk = 0.0;
    if (p >= 0) eval_crs_exact<MAXW, EXACT_W>(a, ld, cdp, t0, p, sem, stk);
    if (p >= 0 && sem > 0.0) {
      write_crs(a, o, p, sem, stk);
    } else {
      double s0;
      eval_crs_exact<MAXW, EXACT_W>(a, ld, cdp, 0, 0, s0, stk);
      write_crs(a, o, 0, 0.0, stk);
    }
  }
}

// ---------------------------------------------------- SB_MODE_FAST search
// (x << S) + y in one v_lshl_add_u32 (the compiler would reassociate it into
// an add and a shift).  A VALU instruction only.
template <int S>
__device__ __forceinline__ unsigned lshl_add(int x, unsigned y) {
  unsigned d;
  asm("v_lshl_add_u32 %0, %1, %2, %3" : "=v"(d) : "v"(x), "i"(S), "s"(y));
  return d;
}

// One work item (CDP, t0 tile, parameter chunk) by one block of 4 waves.
template <int W, int R, int CC, bool CRS, int STORE>
__device__ __forceinline__ void search_item(const SearchArgs& a, const WorkItem wi,
                                            unsigned long long (&s_best)[4][64 * R]) {
  constexpr int ESZ = STORE == SB_STORE_F32 ? 4 : 2;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (wi.cdp < 0) return;
  const int cdp = wi.cdp;
  const int k0 = a.slot_off[cdp];
  const int S = a.slot_off[cdp + 1] - k0;
  if (S <= 0) return;  // no traces: nothing beats 0, keys stay empty
  const SlotF* __restrict__ sl = a.slots + k0;

  float T0[R], T0sq[R];
#pragma unroll
  for (int r = 0; r < R; r++) {
    T0[r] = (float)(wi.tile * 64 * R + r * 64 + lane);
    T0sq[r] = __fmul_rn(T0[r], T0[r]);
  }
  // Contiguous parameter sub-range of this wave, a multiple of CC long.
  const int np = wi.p1 - wi.p0;
  const int per = ((np + 4 * CC - 1) / (4 * CC)) * CC;
  const int pw0 = wi.p0 + wave * per;
  const int pw1 = min(pw0 + per, wi.p1);

  unsigned long long best[R];
#pragma unroll
  for (int r = 0; r < R; r++) best[r] = 0ull;

  const unsigned nwin = (unsigned)max(a.ns - a.w, 0);  // number of valid window starts
  const int nbnc = a.nb * a.nc;
  const int nrec = (int)(a.stride * ESZ);
  const char* __restrict__ base = (const char*)a.samples;

  // CMP tiles whose every t0 exceeds tau take the prefix path.  The CMP slot
  // table lists each CDP's traces by increasing h and t^2 = T0^2 + c h
  // (c >= 0) is monotone in h, so the valid traces of a (t0, c) are a prefix
  // of the list, and a result that tolerates one invalid trace can only lose
  // the last one.  The traces before it therefore run with no validity work
  // at all (a lane whose prefix is shorter is rejected afterwards by the
  // second-to-last trace's test, whatever its windows read), and only the
  // last trace is gated.  Sums, order and validity equal the general path's.
  const bool prefix = !CRS && wi.tile * 64 * R > a.tau;
  const int tau_even = (a.tau + 1) & ~1, even_adj = tau_even - a.tau;
  const unsigned adj2 = 2u * (unsigned)even_adj, adj16 = 16u * (unsigned)even_adj;

  for (int p = pw0; p < pw1; p += CC) {
    float pa[CC], pb[CC], pc[CC];
    bool act[CC];
#pragma unroll
    for (int ci = 0; ci < CC; ci++) {
      const int q = p + ci;
      act[ci] = q < pw1;
      const int qq = act[ci] ? q : p;
      if constexpr (CRS) {
        // a0 + inc i as one FMA each (what the compiler made of the
        // separate operations; k_search_crs_lds writes the same)
        const int ia = qq / nbnc, ib = (qq / a.nc) % a.nb, ic = qq % a.nc;
        pa[ci] = fmaf(a.finc_a, (float)ia, a.fa0);
        pb[ci] = fmaf(a.finc_b, (float)ib, a.fb0);
        pc[ci] = fmaf(a.finc_c, (float)ic, a.fc0);
      } else {
        pa[ci] = 0.f;
        pb[ci] = 0.f;
        pc[ci] = __fadd_rn(a.fc0, __fmul_rn(a.finc_c, (float)qq));
      }
    }
    // CRS: t^2 = T0^2 + (2 T0 a) m + (a^2 + b) m^2 + c h, the per-(row,
    // parameter) coefficients formed once per pass over the traces.
    float K1[R][CC], K2[CC];
#pragma unroll
    for (int ci = 0; ci < CC; ci++) {
      K2[ci] = CRS ? fmaf(pa[ci], pa[ci], pb[ci]) : 0.f;
#pragma unroll
      for (int r = 0; r < R; r++) K1[r][ci] = CRS ? __fmul_rn(__fmul_rn(2.f, T0[r]), pa[ci]) : 0.f;
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

    // One trace of the CDP for every (row, parameter) of the lane.
    // GATE 0: full validity test and count (general path); 1: none (prefix
    // path, all but the last trace); 2: the t^2 bound only (prefix path,
    // last trace).
    // SHAB: every parameter of the pass has the same (a, b) (c is the
    // fastest grid index, so most passes of CC consecutive parameters do),
    // and Q = (a, b, T0) part of the traveltime is formed once per trace.
    auto trace = [&](const SlotF& s, auto gate_tag, auto shab_tag) {
      constexpr int GATE = decltype(gate_tag)::value;
      constexpr bool SHAB = decltype(shab_tag)::value;
      float Qs[R];
      if constexpr (CRS && SHAB) {
#pragma unroll
        for (int r = 0; r < R; r++) Qs[r] = fmaf(K2[0], s.m2, fmaf(K1[r][0], s.m, T0sq[r]));
      }
      // Rows are addressed from `lead` bytes before their start (the
      // allocation has kSamplesLead bytes in front of row 0): FP32 from tau
      // samples, so a window's byte offset is 4 * floor(t); 16-bit stores
      // from the even sample count tau_even = tau rounded up to even, so the
      // window start floor(t) - tau + tau_even keeps the parity of
      // floor(t) - tau and its dword-aligned half needs no subtraction.
      const int lead = STORE == SB_STORE_F32 ? 4 * a.tau : 2 * tau_even;
      const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
          (void*)(base + (int64_t)s.row * a.stride * ESZ - lead), (short)0, nrec + lead, kRsrcWord3);
      // Every rounding step below is explicit (fmaf / __fmul_rn), so an
      // evaluation's bits do not depend on the launch shape or batch.
      unsigned off[R][CC], shf[R][CC];
      float xx[R][CC];
#pragma unroll
      for (int r = 0; r < R; r++) {
#pragma unroll
        for (int ci = 0; ci < CC; ci++) {
          float tt;
          if constexpr (CRS) {
            // (T0 + a m)^2 + b m^2 + c h expanded as c h + Q(a, b), Q =
            // (a^2 + b) m^2 + 2 T0 a m + T0^2: one FMA per evaluation when
            // the pass shares (a, b), three otherwise -- the same roundings
            // either way, so an evaluation's bits do not depend on the pass.
            if constexpr (SHAB)
              tt = fmaf(pc[ci], s.h, Qs[r]);
            else
              tt = fmaf(pc[ci], s.h, fmaf(K2[ci], s.m2, fmaf(K1[r][ci], s.m, T0sq[r])));
          } else {
            tt = fmaf(pc[ci], s.h, T0sq[r]);
          }
          float ts;
          if constexpr (CRS) {
            // t^2 < 0 is invalid (the reference sets t = -1): clamp so the
            // fraction stays finite, then the window test rejects it unless
            // tau = 0, where the sign test does.
            ts = __builtin_amdgcn_sqrtf(fmaxf(tt, 0.f));
          } else {
            ts = __builtin_amdgcn_sqrtf(tt);  // tt >= 0: the host checks c0, c1 >= 0
          }
          // Window start e = it - tau must lie in [0, ns - w): the
          // reference's it - tau >= 0 && it + tau + 1 < ns.
          const int it = (int)ts;
          const int e = it - a.tau;
          bool ok = true;
          if constexpr (GATE == 0) {
            ok = (unsigned)e < nwin;
            if constexpr (CRS && W == 1) ok = ok && tt >= 0.f;
            if constexpr (!CRS) ok = ok && tt < a.hi2;
          } else if constexpr (GATE == 2) {
            ok = tt < a.hi2;
          }
          if constexpr (GATE != 1) cnt[r][ci] += ok ? 1 : 0;
          xx[r][ci] = __builtin_amdgcn_fractf(ts);
          if constexpr (STORE == SB_STORE_F32) {
            off[r][ci] = ok ? (unsigned)it * 4u : kOob;
          } else {
            // e2 = window start + tau_even - tau = it + even_adj:
            // byte offset (2 e2) & ~3 and funnel shift 16 e2 (low 5 bits:
            // 16 (e2 & 1)), each shift-and-add one v_lshl_add_u32
            off[r][ci] = ok ? lshl_add<1>(it, adj2) & ~3u : kOob;
            shf[r][ci] = lshl_add<4>(it, adj16);
          }
        }
      }
      Win<W + 1> sv[R][CC];
#pragma unroll
      for (int r = 0; r < R; r++)
#pragma unroll
        for (int ci = 0; ci < CC; ci++) sv[r][ci] = load_win<W + 1, STORE>(rs, off[r][ci], shf[r][ci]);
      // 16-bit stores: samples restored exactly (stored value times the
      // trace's power-of-two scale), then the FP32 evaluation for every store
      // -- the LDS-staged search restores them once while staging.
      if constexpr (STORE != SB_STORE_F32) {
#pragma unroll
        for (int r = 0; r < R; r++)
#pragma unroll
          for (int ci = 0; ci < CC; ci++)
#pragma unroll
            for (int j = 0; j <= W; j++) sv[r][ci].v[j] *= s.scale;
      }
#pragma unroll
      for (int r = 0; r < R; r++) {
#pragma unroll
        for (int ci = 0; ci < CC; ci++) {
          const float x = xx[r][ci];
#pragma unroll
          for (int j = 0; j < W; j++) {
            const float v = fmaf(__fsub_rn(sv[r][ci].v[j + 1], sv[r][ci].v[j]), x, sv[r][ci].v[j]);
            num[r][ci][j] = __fadd_rn(num[r][ci][j], v);
            den[r][ci] = fmaf(v, v, den[r][ci]);
          }
        }
      }
    };

    // Whether each lane's traces before the last are all valid (prefix
    // path): the second-to-last trace's t^2 bound decides it.
    bool head_ok[R][CC];
    if (prefix) {
      const float h2 = S >= 2 ? sl[S - 2].h : 0.f;
      bool live = false;
#pragma unroll
      for (int r = 0; r < R; r++)
#pragma unroll
        for (int ci = 0; ci < CC; ci++) {
          head_ok[r][ci] = S < 2 || fmaf(pc[ci], h2, T0sq[r]) < a.hi2;
          live = live || (head_ok[r][ci] && act[ci]);
        }
      // A lane whose second-to-last trace is invalid has at least two
      // invalid traces and no result; when no lane of the wave has one for
      // any of its parameters, the trace loop is skipped (same keys).
      if (__builtin_amdgcn_ballot_w64(live) == 0ull) continue;
    }
    if (!prefix) {
      auto general = [&](auto shab_tag) {
        SlotF nx = sl[0];
        for (int k = 0; k < S; k++) {
          const SlotF s = nx;
          nx = sl[k + 1 < S ? k + 1 : k];  // next slot's scalar load in flight
          trace(s, std::integral_constant<int, 0>{}, shab_tag);
        }
      };
      // the pass's parameters p .. last share (a, b) iff their (ia, ib) index
      // p / nc does (wave-uniform)
      const int last = min(p + CC, pw1) - 1;
      if (CRS && p / a.nc == last / a.nc)
        general(std::true_type{});
      else
        general(std::false_type{});
#pragma unroll
      for (int r = 0; r < R; r++)
#pragma unroll
        for (int ci = 0; ci < CC; ci++) head_ok[r][ci] = (S - cnt[r][ci]) < 2;
    } else {
      SlotF nx = sl[0];
      for (int k = 0; k + 1 < S; k++) {
        const SlotF s = nx;
        nx = sl[k + 1];
        trace(s, std::integral_constant<int, 1>{}, std::false_type{});
      }
      trace(nx, std::integral_constant<int, 2>{}, std::false_type{});
#pragma unroll
      for (int r = 0; r < R; r++)
#pragma unroll
        for (int ci = 0; ci < CC; ci++) cnt[r][ci] += S - 1;
    }
#pragma unroll
    for (int r = 0; r < R; r++) {
#pragma unroll
      for (int ci = 0; ci < CC; ci++) {
        float acsq = 0.f;
#pragma unroll
        for (int j = 0; j < W; j++) acsq = fmaf(num[r][ci][j], num[r][ci][j], acsq);
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
  }

#pragma unroll
  for (int r = 0; r < R; r++) s_best[wave][r * 64 + lane] = best[r];
  __syncthreads();
  if (wave == 0) {
#pragma unroll
    for (int r = 0; r < R; r++) {
      unsigned long long m = s_best[0][r * 64 + lane];
#pragma unroll
      for (int q = 1; q < 4; q++) {
        const unsigned long long v = s_best[q][r * 64 + lane];
        m = v > m ? v : m;
      }
      const int t0 = wi.tile * 64 * R + r * 64 + lane;
      if (t0 < a.ns && m) atomicMax(&a.keys[(int64_t)(cdp - a.cdp_base) * a.ns + t0], m);
    }
  }
}

// One block per work item.
template <int W, int R, int CC, bool CRS, int STORE>
__global__ __launch_bounds__(256) void k_search_fast(SearchArgs a) {
  __shared__ unsigned long long s_best[4][64 * R];
  search_item<W, R, CC, CRS, STORE>(a, a.work[blockIdx.x], s_best);
}

// ------------------------------------------- SB_MODE_EXACT CMP, tiled
// The reference's CMP arithmetic (CMP/OpenMP/src/main.cpp:150-203: float, no
// contraction, correctly rounded sqrt and division, traces in gather order,
// the validity test on the float traveltime) in the tiled layout of
// k_search_fast: one workgroup per (CDP, 64-t0 tile, parameter chunk), each
// wave a contiguous run of the chunk, CC parameters per pass over the traces,
// one 16-byte window load per evaluation, and the first-max arg-max as the
// packed key (semblance bits, ~index).  k_finalize then evaluates each winner
// with eval_cmp_exact, the same arithmetic, so the outputs equal
// k_search_exact's.  FP32 store, windows of at most 7 samples.
template <int W, int CC>
__global__ __launch_bounds__(256) void k_search_exact_tiled(SearchArgs a) {
#pragma clang fp contract(off)
  __shared__ unsigned long long s_best[4][64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const WorkItem wi = a.work[blockIdx.x];
  if (wi.cdp < 0) return;
  const int cdp = wi.cdp;
  const int r0 = a.cdp_off[cdp], r1 = a.cdp_off[cdp + 1];
  const int S = r1 - r0;
  const int t0 = wi.tile * 64 + lane;
  float tt0 = a.dt_f * (float)t0;
  tt0 = tt0 * tt0;
  const float lo = (float)a.tau, hi = (float)(a.ns - a.tau - 1);
  const int np = wi.p1 - wi.p0;
  const int per = ((np + 4 * CC - 1) / (4 * CC)) * CC;
  const int pw0 = wi.p0 + wave * per;
  const int pw1 = min(pw0 + per, wi.p1);
  const int nrec = (int)(a.stride * 4);
  const char* __restrict__ base = (const char*)a.samples;
  unsigned long long best = 0ull;

  for (int p = pw0; p < pw1; p += CC) {
    float c[CC];
    bool act[CC];
#pragma unroll
    for (int ci = 0; ci < CC; ci++) {
      act[ci] = p + ci < pw1;
      c[ci] = a.cf0 + a.cfinc * (float)(act[ci] ? p + ci : p);
    }
    float num[CC][W], den[CC];
    int cnt[CC];
#pragma unroll
    for (int ci = 0; ci < CC; ci++) {
      den[ci] = 0.f;
      cnt[ci] = 0;
#pragma unroll
      for (int j = 0; j < W; j++) num[ci][j] = 0.f;
    }
    float h_next = S > 0 ? a.h_ref32[r0] : 0.f;
    for (int row = r0; row < r1; row++) {
      const float h = h_next;
      if (row + 1 < r1) h_next = a.h_ref32[row + 1];
      const __amdgpu_buffer_rsrc_t rs =
          __builtin_amdgcn_make_buffer_rsrc((void*)(base + (int64_t)row * a.stride * 4), (short)0, nrec, kRsrcWord3);
      unsigned off[CC];
      float xx[CC];
#pragma unroll
      for (int ci = 0; ci < CC; ci++) {
        const float ch = c[ci] * h;
        const float t = sqrtf(tt0 + ch) * a.idt_f;
        const bool ok = t >= lo && t < hi;
        const int it = ok ? (int)t : 0;
        const float x = t - (float)it;
        cnt[ci] += ok ? 1 : 0;
        xx[ci] = ok ? x : 0.f;  // invalid traces contribute exactly nothing
        off[ci] = ok ? (unsigned)(it - a.tau) * 4u : kOob;
      }
      Win<W + 1> sv[CC];
#pragma unroll
      for (int ci = 0; ci < CC; ci++) sv[ci] = load_win<W + 1, SB_STORE_F32>(rs, off[ci], 0u);
#pragma unroll
      for (int ci = 0; ci < CC; ci++) {
#pragma unroll
        for (int j = 0; j < W; j++) {
          const float d = sv[ci].v[j + 1] - sv[ci].v[j];
          const float v = d * xx[ci] + sv[ci].v[j];
          num[ci][j] += v;
          const float vv = v * v;
          den[ci] += vv;
        }
      }
    }
#pragma unroll
    for (int ci = 0; ci < CC; ci++) {
      float acsq = 0.f;
#pragma unroll
      for (int j = 0; j < W; j++) {
        const float q = num[ci][j] * num[ci][j];
        acsq += q;
      }
      const float m = (float)cnt[ci];
      const float wf = (float)W;
      if (act[ci] && (double)den[ci] > 1e-13 && (double)m > 1e-13 && (double)wf > 1e-13 && S - cnt[ci] < 2) {
        const float sem = acsq / (den[ci] * m);
        if (sem > 0.f) {
          const unsigned long long key =
              ((unsigned long long)__float_as_uint(sem) << 32) | (unsigned long long)(~(uint32_t)(p + ci));
          best = key > best ? key : best;
        }
      }
    }
  }
  s_best[wave][lane] = best;
  __syncthreads();
  if (wave == 0) {
    unsigned long long m = s_best[0][lane];
#pragma unroll
    for (int q = 1; q < 4; q++) m = s_best[q][lane] > m ? s_best[q][lane] : m;
    if (t0 < a.ns && m) atomicMax(&a.keys[(int64_t)(cdp - a.cdp_base) * a.ns + t0], m);
  }
}

// ------------------------------------------- SB_MODE_EXACT CRS, tiled
// The reference's CRS arithmetic (CRS/OpenMP/src/main.cpp:210-285: double, the
// supergather in list order, the validity test on the double traveltime) in
// the tiled layout: one workgroup per (CDP, 64-t0 tile) covering every
// parameter, each wave a contiguous quarter of the grid, CC parameters per
// pass over the supergather.  A double semblance and its index do not fit one
// 64-bit key, so each lane keeps its first maximum (strict '>' in parameter
// order), the waves' maxima are merged in LDS in wave order with the same
// rule, and the winner's index is written as a key for k_finalize, which
// re-evaluates it with eval_crs_exact (no winner: key 0, the reference's
// absolute-index-0 case).  FP32 store, windows of at most 7 samples.
template <int W, int CC>
__global__ __launch_bounds__(256) void k_search_exact_crs_tiled(SearchArgs a) {
#pragma clang fp contract(off)
  __shared__ double s_sem[4][64];
  __shared__ int s_idx[4][64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const WorkItem wi = a.work[blockIdx.x];
  if (wi.cdp < 0) return;
  const int cdp = wi.cdp;
  const int k0 = a.slot_off[cdp], k1 = a.slot_off[cdp + 1];
  const int S = k1 - k0;
  const int t0 = wi.tile * 64 + lane;
  const double tt0 = a.dt_d * (double)t0;
  const double lo = (double)a.tau, hi = (double)(a.ns - a.tau - 1);
  const int np = wi.p1 - wi.p0;
  const int per = ((np + 4 * CC - 1) / (4 * CC)) * CC;
  const int pw0 = wi.p0 + wave * per;
  const int pw1 = min(pw0 + per, wi.p1);
  const int nbnc = a.nc * a.nb;
  const int nrec = (int)(a.stride * 4);
  const char* __restrict__ base = (const char*)a.samples;
  double best = 0.0;  // the reference starts its maximum at 0
  int best_p = -1;

  for (int p = pw0; p < pw1; p += CC) {
    double pa[CC], pb[CC], pc[CC];
    bool act[CC];
#pragma unroll
    for (int ci = 0; ci < CC; ci++) {
      act[ci] = p + ci < pw1;
      const int q = act[ci] ? p + ci : p;
      pa[ci] = a.da0 + a.dinc_a * (double)(q / nbnc);
      pb[ci] = a.db0 + a.dinc_b * (double)((q / a.nc) % a.nb);
      pc[ci] = a.dc0 + a.dinc_c * (double)(q % a.nc);
    }
    double num[CC][W], den[CC];
    int cnt[CC];
#pragma unroll
    for (int ci = 0; ci < CC; ci++) {
      den[ci] = 0.0;
      cnt[ci] = 0;
#pragma unroll
      for (int j = 0; j < W; j++) num[ci][j] = 0.0;
    }
    SlotD s_next = a.slots_d[S > 0 ? k0 : 0];
    for (int k = k0; k < k1; k++) {
      const SlotD sd = s_next;
      if (k + 1 < k1) s_next = a.slots_d[k + 1];
      const __amdgpu_buffer_rsrc_t rs =
          __builtin_amdgcn_make_buffer_rsrc((void*)(base + (int64_t)sd.row * a.stride * 4), (short)0, nrec, kRsrcWord3);
      unsigned off[CC];
      double xx[CC];
#pragma unroll
      for (int ci = 0; ci < CC; ci++) {
        double t = tt0 + pa[ci] * sd.m;
        const double tsq = t * t;
        const double bm = pb[ci] * sd.m2;
        const double ch = pc[ci] * sd.h;
        t = tsq + bm;
        t = t + ch;
        t = t < 0.0 ? -1.0 : sqrt(t);
        const double ti = t * a.idt_d;
        const bool ok = ti >= lo && ti < hi;
        const int it = ok ? (int)ti : 0;
        cnt[ci] += ok ? 1 : 0;
        xx[ci] = ok ? ti - (double)it : 0.0;  // invalid traces contribute exactly nothing
        off[ci] = ok ? (unsigned)(it - a.tau) * 4u : kOob;
      }
      Win<W + 1> sv[CC];
#pragma unroll
      for (int ci = 0; ci < CC; ci++) sv[ci] = load_win<W + 1, SB_STORE_F32>(rs, off[ci], 0u);
#pragma unroll
      for (int ci = 0; ci < CC; ci++) {
        double s1 = (double)sv[ci].v[0], s0;
#pragma unroll
        for (int j = 0; j < W; j++) {
          s0 = s1;
          s1 = (double)sv[ci].v[j + 1];
          const double d = s1 - s0;
          const double v = d * xx[ci] + s0;
          num[ci][j] += v;
          const double vv = v * v;
          den[ci] += vv;
        }
      }
    }
#pragma unroll
    for (int ci = 0; ci < CC; ci++) {
      double acsq = 0.0;
#pragma unroll
      for (int j = 0; j < W; j++) {
        const double q = num[ci][j] * num[ci][j];
        acsq += q;
      }
      const double mm = (double)cnt[ci];
      double sem = 0.0;
      if (den[ci] > 1e-13 && mm > 1e-13 && (double)W > 1e-13 && S - cnt[ci] < 2) sem = acsq / (den[ci] * mm);
      if (act[ci] && sem > best) {
        best = sem;
        best_p = p + ci;
      }
    }
  }
  s_sem[wave][lane] = best;
  s_idx[wave][lane] = best_p;
  __syncthreads();
  if (wave == 0 && t0 < a.ns) {
    double m = s_sem[0][lane];
    int mp = s_idx[0][lane];
#pragma unroll
    for (int q = 1; q < 4; q++)
      if (s_sem[q][lane] > m) {  // later waves hold later parameters: strict '>' keeps the first maximum
        m = s_sem[q][lane];
        mp = s_idx[q][lane];
      }
    a.keys[(int64_t)(cdp - a.cdp_base) * a.ns + t0] =
        mp >= 0 ? ((1ull << 32) | (unsigned long long)(~(uint32_t)mp)) : 0ull;
  }
}

// Persistent grid (a separate kernel, so the per-item kernel's register
// allocation is not shaped by this loop): the blocks of XCD
// x = blockIdx % 8 (workgroups are dealt to the 8 XCDs round-robin) take the
// work table's items x, x+8, x+16, ... -- the CDPs dealt to that XCD -- in
// order from the XCD's own counter until the table is exhausted.
template <int W, int R, int CC, bool CRS, int STORE>
__global__ __launch_bounds__(256) void k_search_fast_persistent(SearchArgs a) {
  __shared__ unsigned long long s_best[4][64 * R];
  __shared__ unsigned int s_next;
  const unsigned x = blockIdx.x & 7u;
  for (;;) {
    if (threadIdx.x == 0) s_next = atomicAdd(&a.queue[x], 1u);
    __syncthreads();
    const int64_t idx = (int64_t)s_next * 8 + x;
    __syncthreads();  // s_next read by every thread before it is rewritten
    if (idx >= a.nwork) break;
    search_item<W, R, CC, CRS, STORE>(a, a.work[idx], s_best);
    __syncthreads();  // wave 0 has read s_best before the next item writes it
  }
}

// ------------------------------------------------------------- launchers
#define SB_CHECK_LAUNCH() (hipGetLastError())

hipError_t launch_prep_traces(const TraceGeom* g, int ntr, float* h_ref32, double* h0, double* m0x, double* m0y,
                              float* msign, hipStream_t st) {
  if (ntr <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_prep_traces, dim3((ntr + 255) / 256), dim3(256), 0, st, g, ntr, h_ref32, h0, m0x, m0y, msign);
  return SB_CHECK_LAUNCH();
}

hipError_t launch_cmp_slots(int cdp0, int ncdp, const int32_t* cdp_off, const double* h0, const float* scale,
                            double idt, SlotF* slots, hipStream_t st) {
  if (ncdp <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_cmp_slots, dim3(ncdp), dim3(256), 0, st, cdp0, cdp_off, h0, scale, idt, slots);
  return SB_CHECK_LAUNCH();
}

hipError_t launch_crs_slots(int cdp0, int ncdp, const int32_t* cdp_off, const int32_t* nbr_off, const int32_t* nbr,
                            const int32_t* slot_off, const double* m0x, const double* m0y, const double* h0,
                            const float* scale, double idt, const void* samples, uint64_t row_bytes, SlotF* sf,
                            SlotD* sd, SlotF* sp, const float* msign, int az, double az_s, double az_c,
                            hipStream_t st) {
  if (ncdp <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_crs_slots, dim3(ncdp), dim3(256), 0, st, cdp0, cdp_off, nbr_off, nbr, slot_off, m0x, m0y, h0,
                     scale, idt, (const char*)samples, row_bytes, sf, sd, sp, msign, az, az_s, az_c);
  return SB_CHECK_LAUNCH();
}

hipError_t launch_store(const float* src, int ntr, int ns, int64_t stride, int store, void* dst, float* scale,
                        hipStream_t st) {
  if (ntr <= 0) return hipSuccess;
  if (store == SB_STORE_F32)
    hipLaunchKernelGGL(k_pad_rows, dim3(ntr), dim3(256), 0, st, src, ns, stride, (float*)dst);
  else if (store == SB_STORE_F16)
    hipLaunchKernelGGL(k_quantize<SB_STORE_F16>, dim3(ntr), dim3(256), 0, st, src, ns, stride, dst, scale);
  else
    hipLaunchKernelGGL(k_quantize<SB_STORE_BF16>, dim3(ntr), dim3(256), 0, st, src, ns, stride, dst, scale);
  return SB_CHECK_LAUNCH();
}

hipError_t launch_restore_rows(const void* src, const float* scale, int64_t nrows, int64_t stride, int store,
                               float* dst, hipStream_t st) {
  if (nrows <= 0) return hipSuccess;
  if (store == SB_STORE_F32 || stride % 64 != 0) return hipErrorInvalidValue;
  const int64_t nvec = nrows * stride / 8;
  const dim3 grid((unsigned)((nvec + 255) / 256));
  if (store == SB_STORE_F16)
    hipLaunchKernelGGL(k_restore_rows<SB_STORE_F16>, grid, dim3(256), 0, st, (const uint4*)src, scale, stride, nvec,
                       (float4*)dst);
  else
    hipLaunchKernelGGL(k_restore_rows<SB_STORE_BF16>, grid, dim3(256), 0, st, (const uint4*)src, scale, stride, nvec,
                       (float4*)dst);
  return SB_CHECK_LAUNCH();
}

hipError_t launch_rebase_slots(const SlotF* src, int64_t n, uint64_t base, uint64_t row_bytes, SlotF* dst,
                               hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_rebase_slots, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, src, n, base, row_bytes, dst);
  return SB_CHECK_LAUNCH();
}

hipError_t launch_make_work(WorkItem* w, int nblocks, int cdp0, int ncdp, int ntiles, int nchunks, int pchunk,
                            int npar, hipStream_t st) {
  if (nblocks <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_make_work, dim3((nblocks + 255) / 256), dim3(256), 0, st, w, nblocks, cdp0, ncdp, ntiles,
                     nchunks, pchunk, npar);
  return SB_CHECK_LAUNCH();
}

hipError_t launch_fill_u64(unsigned long long* p, int64_t n, unsigned long long v, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill_u64, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, n, v);
  return SB_CHECK_LAUNCH();
}

template <class LD>
static hipError_t finalize_with(const SearchArgs& a, const LD& ld, hipStream_t st) {
  const int64_t n = (int64_t)a.nout * a.ns;
  const dim3 grid((unsigned)((n + 255) / 256));
  switch (a.w) {  // the common windows compiled exactly
    case 1: hipLaunchKernelGGL((k_finalize<1, true, LD>), grid, dim3(256), 0, st, a, ld); break;
    case 3: hipLaunchKernelGGL((k_finalize<3, true, LD>), grid, dim3(256), 0, st, a, ld); break;
    case 5: hipLaunchKernelGGL((k_finalize<5, true, LD>), grid, dim3(256), 0, st, a, ld); break;
    case 7: hipLaunchKernelGGL((k_finalize<7, true, LD>), grid, dim3(256), 0, st, a, ld); break;
    default:
      if (a.w <= kFinalizeMaxW)
        hipLaunchKernelGGL((k_finalize<kFinalizeMaxW, false, LD>), grid, dim3(256), 0, st, a, ld);
      else
        hipLaunchKernelGGL((k_finalize<SB_MAX_W, false, LD>), grid, dim3(256), 0, st, a, ld);
  }
  return SB_CHECK_LAUNCH();
}

template <class LD>
static hipError_t exact_with(const SearchArgs& a, const LD& ld, int nblocks, hipStream_t st) {
  switch (a.w) {  // the common windows compiled exactly
    case 1: hipLaunchKernelGGL((k_search_exact<1, true, LD>), dim3(nblocks), dim3(64), 0, st, a, ld); break;
    case 3: hipLaunchKernelGGL((k_search_exact<3, true, LD>), dim3(nblocks), dim3(64), 0, st, a, ld); break;
    case 5: hipLaunchKernelGGL((k_search_exact<5, true, LD>), dim3(nblocks), dim3(64), 0, st, a, ld); break;
    case 7: hipLaunchKernelGGL((k_search_exact<7, true, LD>), dim3(nblocks), dim3(64), 0, st, a, ld); break;
    default:
      if (a.w <= 7)
        hipLaunchKernelGGL((k_search_exact<7, false, LD>), dim3(nblocks), dim3(64), 0, st, a, ld);
      else
        hipLaunchKernelGGL((k_search_exact<SB_MAX_W, false, LD>), dim3(nblocks), dim3(64), 0, st, a, ld);
  }
  return SB_CHECK_LAUNCH();
}

hipError_t launch_finalize(const SearchArgs& a, int store, hipStream_t st) {
  if ((int64_t)a.nout * a.ns <= 0) return hipSuccess;
  if (store == SB_STORE_F32) return finalize_with(a, LoadF32{(const float*)a.samples, a.stride}, st);
  if (store == SB_STORE_F16) return finalize_with(a, LoadF16{(const __half*)a.samples, a.scale, a.stride}, st);
  return finalize_with(a, LoadBF16{(const uint16_t*)a.samples, a.scale, a.stride}, st);
}

hipError_t launch_search_exact(const SearchArgs& a, int store, int nblocks, hipStream_t st) {
  if (nblocks <= 0) return hipSuccess;
  if (store == SB_STORE_F32) return exact_with(a, LoadF32{(const float*)a.samples, a.stride}, nblocks, st);
  if (store == SB_STORE_F16)
    return exact_with(a, LoadF16{(const __half*)a.samples, a.scale, a.stride}, nblocks, st);
  return exact_with(a, LoadBF16{(const uint16_t*)a.samples, a.scale, a.stride}, nblocks, st);
}

bool exact_tiled_supported(int method, int store, int w) {
  return store == SB_STORE_F32 && w >= 1 && w <= SB_MAX_W && (w & 1);
}

hipError_t launch_search_exact_tiled(const SearchArgs& a, int nblocks, hipStream_t st) {
  if (nblocks <= 0) return hipSuccess;
  if (a.method == SB_METHOD_CRS) {
    switch (a.w) {
      case 1: hipLaunchKernelGGL((k_search_exact_crs_tiled<1, 4>), dim3(nblocks), dim3(256), 0, st, a); break;
      case 3: hipLaunchKernelGGL((k_search_exact_crs_tiled<3, 4>), dim3(nblocks), dim3(256), 0, st, a); break;
      case 5: hipLaunchKernelGGL((k_search_exact_crs_tiled<5, 2>), dim3(nblocks), dim3(256), 0, st, a); break;
      case 7: hipLaunchKernelGGL((k_search_exact_crs_tiled<7, 2>), dim3(nblocks), dim3(256), 0, st, a); break;
      case 9: hipLaunchKernelGGL((k_search_exact_crs_tiled<9, 1>), dim3(nblocks), dim3(256), 0, st, a); break;
      case 11: hipLaunchKernelGGL((k_search_exact_crs_tiled<11, 1>), dim3(nblocks), dim3(256), 0, st, a); break;
      case 13: hipLaunchKernelGGL((k_search_exact_crs_tiled<13, 1>), dim3(nblocks), dim3(256), 0, st, a); break;
      case 15: hipLaunchKernelGGL((k_search_exact_crs_tiled<15, 1>), dim3(nblocks), dim3(256), 0, st, a); break;
      default: return hipErrorInvalidValue;
    }
    return SB_CHECK_LAUNCH();
  }
  switch (a.w) {
    case 1: hipLaunchKernelGGL((k_search_exact_tiled<1, 4>), dim3(nblocks), dim3(256), 0, st, a); break;
    case 3: hipLaunchKernelGGL((k_search_exact_tiled<3, 4>), dim3(nblocks), dim3(256), 0, st, a); break;
    case 5: hipLaunchKernelGGL((k_search_exact_tiled<5, 4>), dim3(nblocks), dim3(256), 0, st, a); break;
    case 7: hipLaunchKernelGGL((k_search_exact_tiled<7, 4>), dim3(nblocks), dim3(256), 0, st, a); break;
    case 9: hipLaunchKernelGGL((k_search_exact_tiled<9, 2>), dim3(nblocks), dim3(256), 0, st, a); break;
    case 11: hipLaunchKernelGGL((k_search_exact_tiled<11, 2>), dim3(nblocks), dim3(256), 0, st, a); break;
    case 13: hipLaunchKernelGGL((k_search_exact_tiled<13, 2>), dim3(nblocks), dim3(256), 0, st, a); break;
    case 15: hipLaunchKernelGGL((k_search_exact_tiled<15, 2>), dim3(nblocks), dim3(256), 0, st, a); break;
    default: return hipErrorInvalidValue;
  }
  return SB_CHECK_LAUNCH();
}

template <int W, int R, int CC, bool CRS, int STORE>
static hipError_t fast_launch(const SearchArgs& a, int nblocks, int num_cus, hipStream_t st) {
  if (!a.queue) {
    hipLaunchKernelGGL((k_search_fast<W, R, CC, CRS, STORE>), dim3(nblocks), dim3(256), 0, st, a);
    return SB_CHECK_LAUNCH();
  }
  auto kern = k_search_fast_persistent<W, R, CC, CRS, STORE>;
  int per = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kern, 256, 0) != hipSuccess || per < 1) per = 1;
  const int full = std::max(8, num_cus * per / 8 * 8);
  const int grid = std::min((nblocks + 7) / 8 * 8, full);
  hipLaunchKernelGGL(kern, dim3(grid), dim3(256), 0, st, a);
  return SB_CHECK_LAUNCH();
}

template <int W, int R, int CC, bool CRS>
static hipError_t fast_store(const SearchArgs& a, int store, int nblocks, int num_cus, hipStream_t st) {
  if (store == SB_STORE_F32) return fast_launch<W, R, CC, CRS, SB_STORE_F32>(a, nblocks, num_cus, st);
  if (store == SB_STORE_F16) return fast_launch<W, R, CC, CRS, SB_STORE_F16>(a, nblocks, num_cus, st);
  return fast_launch<W, R, CC, CRS, SB_STORE_BF16>(a, nblocks, num_cus, st);
}

template <int R, int CC, bool CRS>
static hipError_t fast_w(const SearchArgs& a, int store, int nblocks, int num_cus, hipStream_t st) {
  switch (a.w) {
    case 1: return fast_store<1, R, CC, CRS>(a, store, nblocks, num_cus, st);
    case 3: return fast_store<3, R, CC, CRS>(a, store, nblocks, num_cus, st);
    case 5: return fast_store<5, R, CC, CRS>(a, store, nblocks, num_cus, st);
    case 7: return fast_store<7, R, CC, CRS>(a, store, nblocks, num_cus, st);
    default: return hipErrorInvalidValue;
  }
}

bool fast_supported(int w) { return w == 1 || w == 3 || w == 5 || w == 7; }

const FastVariant kFastVariants[] = {{2, 4, 0, 4, 0},  {1, 4, 0, 4, 0},  {4, 2, 0, 4, 0},  {2, 2, 0, 4, 0},
                                     {1, 4, 48, 8, 0}, {1, 4, 52, 8, 0}, {1, 4, 64, 8, 0}, {1, 4, 40, 8, 0},
                                     {1, 4, 64, 8, 1}, {1, 4, 52, 8, 1}, {1, 4, 40, 8, 1}, {1, 4, 52, 8, 2, 16},
                                     {1, 4, 52, 8, 2, 8}};
const int kNumFastVariants = sizeof(kFastVariants) / sizeof(kFastVariants[0]);

hipError_t launch_search_fast(const SearchArgs& a, int store, int variant, int nblocks, int num_cus, hipStream_t st) {
  if (nblocks <= 0) return hipSuccess;
  const bool crs = a.method == SB_METHOD_CRS;
  const int n = nblocks, c = num_cus;
  switch (variant) {
    case 0: return crs ? fast_w<2, 4, true>(a, store, n, c, st) : fast_w<2, 4, false>(a, store, n, c, st);
    case 1: return crs ? fast_w<1, 4, true>(a, store, n, c, st) : fast_w<1, 4, false>(a, store, n, c, st);
    case 2: return crs ? fast_w<4, 2, true>(a, store, n, c, st) : fast_w<4, 2, false>(a, store, n, c, st);
    case 3: return crs ? fast_w<2, 2, true>(a, store, n, c, st) : fast_w<2, 2, false>(a, store, n, c, st);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace sb
