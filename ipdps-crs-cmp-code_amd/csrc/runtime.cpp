// Device context, resident gathers, results and the search entry points of
// libsemblance (C ABI in include/semblance.h).
//
// Replaces the reference's per-CDP host loops (CMP/CUDA/src/main.cu:276-296,
// CRS/CUDA/src/main.cu:420-450): the whole binned gather is resident in HBM
// and one launch covers every (CDP, t0 tile, parameter chunk).
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include "device_types.hpp"
#include "launch.hpp"
#include "runtime.hpp"
#include "sb_internal.hpp"

using namespace sb;

namespace sb {

int hip_fail(hipError_t e, const char* what) {
  return fail(SB_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

#define HIPR(call)                                   \
  do {                                               \
    hipError_t _e = (call);                          \
    if (_e != hipSuccess) return hip_fail(_e, #call); \
  } while (0)

// Caching device allocator: freed blocks are kept per ctx and reused for
// requests of up to twice their size, so steady-state searches and serving
// do no hipMalloc/hipFree (hipFree synchronises the device).  Cached bytes
// count against the budget and are released first when it runs short.
static size_t round_block(size_t bytes) {
  if (bytes <= (1u << 20)) {
    size_t r = 256;
    while (r < bytes) r <<= 1;
    return r;
  }
  return (bytes + (2u << 20) - 1) / (2u << 20) * (2u << 20);
}

// Blocks of transient allocations (restored FP32 copies, sized per request)
// above 32 MiB come in power-of-two classes, so that a cached block serves
// every later copy of its class: sizes that vary per request would otherwise
// miss the cache, allocate anew and push older blocks out (hipFree).
size_t transient_block_size(size_t bytes) {
  size_t want = round_block(bytes);
  if (want > (size_t(32) << 20)) {
    size_t c = size_t(32) << 20;
    while (c < want) c <<= 1;
    want = c;
  }
  return want;
}

// The free-event of a cached block goes back to the pool once the block is
// reused or released (hipFree waits for the device anyway).
static void drop_event(sb_ctx* ctx, hipEvent_t ev) {
  if (ev) ctx->event_pool.push_back(ev);
}

static void release_cache(sb_ctx* ctx) {
  for (auto& kv : ctx->cache) {
    (void)hipFree(kv.second.p);
    drop_event(ctx, kv.second.freed);
    ctx->used -= kv.first;
  }
  ctx->cache.clear();
}

// transient: the block is used only by work queued on the ctx stream, which
// is ordered after the free of a cached block (no host wait), and a
// shortage returns SB_ERR_BUDGET without an error message (the caller has
// another way).
static int alloc_block(sb_ctx* ctx, size_t bytes, void** p, bool transient) {
  *p = nullptr;
  if (bytes == 0) return SB_OK;
  const size_t want = transient ? transient_block_size(bytes) : round_block(bytes);
  std::lock_guard<std::mutex> lk(ctx->mem_mu);
  auto it = ctx->cache.lower_bound(want);
  if (it != ctx->cache.end() && it->first <= 2 * want) {
    // The new owner may write the block from any stream (or the host): wait
    // until the work queued on the ctx stream before the free has finished.
    if (it->second.freed) {
      if (!transient) {
        const hipError_t e = hipEventSynchronize(it->second.freed);
        if (e != hipSuccess) return hip_fail(e, "cached block");
      }
      drop_event(ctx, it->second.freed);
    }
    *p = it->second.p;
    ctx->sizes[*p] = it->first;
    ctx->cache.erase(it);
    return SB_OK;
  }
  if (ctx->used + want > ctx->budget) release_cache(ctx);
  if (ctx->used + want > ctx->budget) {
    if (transient) return SB_ERR_BUDGET;
    return fail(SB_ERR_BUDGET, "HBM budget exceeded: need " + std::to_string(want) + " B, " +
                                   std::to_string(ctx->budget - ctx->used) + " B left");
  }
  hipError_t e = hipMalloc(p, want);
  if (e != hipSuccess && !ctx->cache.empty()) {
    (void)hipGetLastError();
    release_cache(ctx);
    e = hipMalloc(p, want);
  }
  if (e != hipSuccess) {
    *p = nullptr;
    if (transient) {
      (void)hipGetLastError();  // leave no error for the next launch check
      return SB_ERR_BUDGET;
    }
    return hip_fail(e, "hipMalloc");
  }
  ctx->used += want;
  ctx->sizes[*p] = want;
  return SB_OK;
}

int ctx_alloc(sb_ctx* ctx, size_t bytes, void** p) { return alloc_block(ctx, bytes, p, false); }

int ctx_alloc_transient(sb_ctx* ctx, size_t bytes, void** p) { return alloc_block(ctx, bytes, p, true); }

// Returns a block to the cache, stream-ordered: an event recorded on the ctx
// stream marks the point after which no queued work of this ctx reads it.
// (No block is freed while the stream is being captured into a graph; if
// one were, or the event cannot be recorded, the stream is drained instead.)
void ctx_free(sb_ctx* ctx, void* p) {
  if (!p) return;
  std::lock_guard<std::mutex> lk(ctx->mem_mu);
  auto it = ctx->sizes.find(p);
  if (it == ctx->sizes.end()) return;
  hipEvent_t ev = nullptr;
  if (!ctx->capturing) {
    if (!ctx->event_pool.empty()) {
      ev = ctx->event_pool.back();
      ctx->event_pool.pop_back();
    } else if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) {
      ev = nullptr;
    }
    if (ev && hipEventRecord(ev, ctx->stream) != hipSuccess) {
      ctx->event_pool.push_back(ev);
      ev = nullptr;
    }
    if (!ev) (void)hipStreamSynchronize(ctx->stream);
  }
  ctx->cache.insert({it->second, sb_ctx::Cached{p, ev}});
  ctx->sizes.erase(it);
}

void ctx_trim(sb_ctx* ctx, size_t keep) {
  std::lock_guard<std::mutex> lk(ctx->mem_mu);
  size_t cached = 0;
  for (auto& kv : ctx->cache) cached += kv.first;
  while (cached > keep && !ctx->cache.empty()) {
    auto it = std::prev(ctx->cache.end());  // largest first
    (void)hipFree(it->second.p);
    drop_event(ctx, it->second.freed);
    ctx->used -= it->first;
    cached -= it->first;
    ctx->cache.erase(it);
  }
}

void ctx_release_all(sb_ctx* ctx) {
  std::lock_guard<std::mutex> lk(ctx->mem_mu);
  release_cache(ctx);
  for (auto& kv : ctx->sizes) (void)hipFree(kv.first);
  ctx->sizes.clear();
  ctx->used = 0;
  for (hipEvent_t ev : ctx->event_pool) (void)hipEventDestroy(ev);
  ctx->event_pool.clear();
}

}  // namespace sb

// ------------------------------------------------------------------ status
extern "C" const char* sb_strerror(int code) {
  switch (code) {
    case SB_OK: return "ok";
    case SB_ERR_INVALID: return "invalid argument";
    case SB_ERR_IO: return "i/o error";
    case SB_ERR_FORMAT: return "bad input format";
    case SB_ERR_HIP: return "HIP runtime error";
    case SB_ERR_BUDGET: return "HBM budget exceeded";
    case SB_ERR_NOMEM: return "out of host memory";
    case SB_ERR_UNSUPPORTED: return "unsupported configuration";
    case SB_ERR_STATE: return "invalid state";
    case SB_ERR_DEADLINE: return "admission deadline passed";
    default: return "unknown error";
  }
}

extern "C" const char* sb_version(void) { return "libsemblance 0.1 (gfx950)"; }

// ----------------------------------------------------------------- context
extern "C" int sb_ctx_create(int device, size_t hbm_budget, sb_ctx** out) {
  if (!out) return fail(SB_ERR_INVALID, "sb_ctx_create: null out");
  *out = nullptr;
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n <= 0) return fail(SB_ERR_HIP, "no HIP device available");
  if (device < 0 || device >= n) return fail(SB_ERR_INVALID, "sb_ctx_create: bad device index");
  HIPR(hipSetDevice(device));
  std::unique_ptr<sb_ctx> c(new sb_ctx);
  c->device = device;
  size_t fr = 0, tot = 0;
  HIPR(hipMemGetInfo(&fr, &tot));
  c->budget = hbm_budget ? hbm_budget : (size_t)(0.9 * (double)fr);
  HIPR(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  for (auto& ev : c->ev) HIPR(hipEventCreate(&ev));
  hipDeviceProp_t prop;
  HIPR(hipGetDeviceProperties(&prop, device));
  c->num_cus = prop.multiProcessorCount;
  // The allocation-free entry point's static scratch: its device address on
  // this device (looked up once per device, with the device current).
  HIPR(scratch_address(device, &c->scratch));
  *out = c.release();
  return SB_OK;
}

extern "C" void sb_ctx_destroy(sb_ctx* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  ctx_release_all(ctx);
  if (ctx->work_host) (void)hipHostFree(ctx->work_host);
  for (void* p : ctx->work_host_retired) (void)hipHostFree(p);
  if (ctx->work_host_ev) (void)hipEventDestroy(ctx->work_host_ev);
  for (auto& ev : ctx->ev) (void)hipEventDestroy(ev);
  (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

extern "C" int sb_ctx_memory(const sb_ctx* ctx, size_t* used, size_t* budget) {
  if (!ctx) return fail(SB_ERR_INVALID, "null ctx");
  if (used) *used = ctx->used;
  if (budget) *budget = ctx->budget;
  return SB_OK;
}

extern "C" int sb_ctx_sync(sb_ctx* ctx) {
  if (!ctx) return fail(SB_ERR_INVALID, "null ctx");
  HIPR(hipSetDevice(ctx->device));
  HIPR(hipStreamSynchronize(ctx->stream));
  return SB_OK;
}

// ----------------------------------------------------------------- gathers
extern "C" void sb_gather_destroy(sb_gather* g) {
  if (!g) return;
  sb_ctx* c = g->ctx;
  (void)hipSetDevice(c->device);
  // Searches from other contexts may still read the gather; its blocks go
  // back to the cache for reuse, so wait for the whole device.
  (void)hipDeviceSynchronize();
  for (void* p : {(void*)g->d_samples_mem, (void*)g->d_scale, (void*)g->d_cdp_off, (void*)g->d_h_ref32, (void*)g->d_h0,
                  (void*)g->d_m0x, (void*)g->d_m0y, (void*)g->d_msign, (void*)g->d_cmp_slots, (void*)g->d_crs_slot_off,
                  (void*)g->d_crs_slots, (void*)g->d_crs_slots_d, (void*)g->d_crs_slots_p})
    ctx_free(c, p);
  delete g;
}

namespace {

int upload(sb_ctx* c, void* dst, const void* src, size_t bytes) {
  if (!bytes) return SB_OK;
  HIPR(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
  return SB_OK;
}

}  // namespace

namespace sb {

// Incremental construction of a gather: begin() validates, allocates and
// prepares the geometry (and the CRS supergather offsets); batch() uploads
// the samples of a CDP range on stream `cs` -- staged, padded or quantised --
// with their CMP slots, then the CRS slot tables of every CDP whose
// neighbours have all arrived, and records an event; finish() waits and
// hands the gather over.  sb_gather_create runs the batches back to back;
// sb_search_streamed overlaps them with range searches.
struct GatherBuild {
  sb_ctx* ctx = nullptr;
  std::unique_ptr<sb_gather, void (*)(sb_gather*)> G{nullptr, sb_gather_destroy};
  const float* samples = nullptr;
  hipStream_t cs = nullptr;
  void* stage[2] = {nullptr, nullptr};
  int64_t stage_rows = 0;
  int which = 0;
  void *dnoff = nullptr, *dnbr = nullptr;  // CRS neighbour lists on the device
  std::vector<int32_t> nbr_reach;          // max neighbour index over CDPs [0, i]
  int crs_ready = 0;                       // CDPs [0, crs_ready) have slot tables
  double idt = 0;

  ~GatherBuild() {
    if (cs) (void)hipStreamSynchronize(cs);  // staged copies done before the buffers are reused
    for (void* p : stage) ctx_free(ctx, p);
    ctx_free(ctx, dnoff);
    ctx_free(ctx, dnbr);
  }

  int begin(sb_ctx* c, int ncdp, const int32_t* cdp_off, int ns, int dt_us, const sb_trace_geom* geom,
            const float* smp, int store, const int32_t* nbr_off, const int32_t* nbr, hipStream_t copy_stream,
            int64_t batch_rows);
  int batch(int cb, int ce, hipEvent_t ev, int* ready);
  int finish(sb_gather** out);
};

int GatherBuild::begin(sb_ctx* c, int ncdp, const int32_t* cdp_off, int ns, int dt_us, const sb_trace_geom* geom,
                       const float* smp, int store, const int32_t* nbr_off, const int32_t* nbr,
                       hipStream_t copy_stream, int64_t batch_rows) {
  ctx = c;
  cs = copy_stream;
  samples = smp;
  if (!ctx || !cdp_off || ncdp <= 0 || ns <= 0 || dt_us <= 0 || !geom || !samples)
    return fail(SB_ERR_INVALID, "gather: bad arguments");
  if (store != SB_STORE_F32 && store != SB_STORE_F16 && store != SB_STORE_BF16)
    return fail(SB_ERR_INVALID, "gather: bad store type");
  if (cdp_off[0] != 0) return fail(SB_ERR_INVALID, "cdp_off[0] must be 0");
  for (int i = 0; i < ncdp; i++)
    if (cdp_off[i + 1] <= cdp_off[i]) return fail(SB_ERR_INVALID, "every CDP needs at least one trace");
  const int ntr = cdp_off[ncdp];
  int64_t nnbr = 0;
  if (nbr_off) {
    if (!nbr || nbr_off[0] != 0) return fail(SB_ERR_INVALID, "bad neighbour lists");
    for (int i = 0; i < ncdp; i++)
      if (nbr_off[i + 1] <= nbr_off[i]) return fail(SB_ERR_INVALID, "every CDP needs a non-empty neighbour list");
    nnbr = nbr_off[ncdp];
    for (int64_t q = 0; q < nnbr; q++)
      if (nbr[q] < 0 || nbr[q] >= ncdp) return fail(SB_ERR_INVALID, "neighbour index out of range");
  }
  HIPR(hipSetDevice(ctx->device));
  G.reset(new sb_gather);
  G->ctx = ctx;
  G->serial = ++ctx->gather_serial;
  G->ncdp = ncdp;
  G->ntr = ntr;
  G->ns = ns;
  G->dt_us = dt_us;
  G->store = store;
  G->dm_azimuth_mdeg = ctx->dm_azimuth_mdeg;
  // >= 8 zero elements of padding after every row: 16-bit window loads may
  // read one dword past the last sample.
  G->stride = ((int64_t)ns + 8 + 63) / 64 * 64;
  G->cdp_off.assign(cdp_off, cdp_off + ncdp + 1);
  const size_t esz = store == SB_STORE_F32 ? 4 : 2;
  int rc;
  // kSamplesLead bytes before row 0: the FP32 search addresses a row from
  // tau samples before its start.
  if ((rc = ctx_alloc(ctx, (size_t)ntr * G->stride * esz + kSamplesLead, &G->d_samples_mem))) return rc;
  G->d_samples = (char*)G->d_samples_mem + kSamplesLead;
  if (store != SB_STORE_F32 && (rc = ctx_alloc(ctx, (size_t)ntr * 4, (void**)&G->d_scale))) return rc;
  if ((rc = ctx_alloc(ctx, (size_t)(ncdp + 1) * 4, (void**)&G->d_cdp_off))) return rc;
  if ((rc = ctx_alloc(ctx, (size_t)ntr * 4, (void**)&G->d_h_ref32))) return rc;
  if ((rc = ctx_alloc(ctx, (size_t)ntr * 8, (void**)&G->d_h0))) return rc;
  if ((rc = ctx_alloc(ctx, (size_t)ntr * 8, (void**)&G->d_m0x))) return rc;
  if ((rc = ctx_alloc(ctx, (size_t)ntr * 8, (void**)&G->d_m0y))) return rc;
  if ((rc = ctx_alloc(ctx, (size_t)ntr * 4, (void**)&G->d_msign))) return rc;
  if ((rc = ctx_alloc(ctx, (size_t)ntr * sizeof(SlotF), (void**)&G->d_cmp_slots))) return rc;
  stage_rows = std::max<int64_t>(1, std::min<int64_t>(batch_rows, ntr));
  for (auto& st : stage)
    if ((rc = ctx_alloc(ctx, (size_t)stage_rows * ns * 4, &st))) return rc;
  idt = 1.0 / ((double)dt_us / 1e6);

  // Geometry (FP64, and the CMP float reference arithmetic) on the device,
  // on the copy stream ahead of every batch.
  HIPR(hipMemcpyAsync(G->d_cdp_off, cdp_off, (size_t)(ncdp + 1) * 4, hipMemcpyHostToDevice, cs));
  {
    std::vector<TraceGeom> tg(ntr);
    for (int i = 0; i < ntr; i++)
      tg[i] = TraceGeom{geom[i].cdp, geom[i].scalco, geom[i].sx, geom[i].sy, geom[i].gx, geom[i].gy};
    void* dg = nullptr;
    if ((rc = ctx_alloc(ctx, (size_t)ntr * sizeof(TraceGeom), &dg))) return rc;
    hipError_t e = hipMemcpyAsync(dg, tg.data(), (size_t)ntr * sizeof(TraceGeom), hipMemcpyHostToDevice, cs);
    if (e == hipSuccess)
      e = launch_prep_traces((const TraceGeom*)dg, ntr, G->d_h_ref32, G->d_h0, G->d_m0x, G->d_m0y, G->d_msign, cs);
    // tg is read by the copy: wait before it goes out of scope.
    if (e == hipSuccess) e = hipStreamSynchronize(cs);
    ctx_free(ctx, dg);
    if (e != hipSuccess) return hip_fail(e, "geometry");
  }

  // CRS supergathers: offsets on the host; the tables follow the samples.
  if (nbr_off) {
    G->has_crs = true;
    G->crs_slot_off.resize(ncdp + 1);
    G->crs_slot_off[0] = 0;
    nbr_reach.resize(ncdp);
    G->crs_row_lo.resize(ncdp);
    G->crs_row_hi.resize(ncdp);
    int32_t reach = 0;
    for (int i = 0; i < ncdp; i++) {
      int64_t sz = 0;
      int32_t lo = ncdp, hi = -1;
      for (int q = nbr_off[i]; q < nbr_off[i + 1]; q++) {
        sz += cdp_off[nbr[q] + 1] - cdp_off[nbr[q]];
        reach = std::max(reach, nbr[q]);
        lo = std::min(lo, nbr[q]);
        hi = std::max(hi, nbr[q]);
      }
      nbr_reach[i] = reach;
      G->crs_row_lo[i] = cdp_off[lo];
      G->crs_row_hi[i] = cdp_off[hi + 1];
      if (G->crs_slot_off[i] + sz > INT32_MAX) return fail(SB_ERR_INVALID, "supergathers too large");
      G->crs_slot_off[i + 1] = G->crs_slot_off[i] + (int32_t)sz;
    }
    G->ncrs_slots = G->crs_slot_off[ncdp];
    if ((rc = ctx_alloc(ctx, (size_t)(ncdp + 1) * 4, (void**)&G->d_crs_slot_off))) return rc;
    if ((rc = ctx_alloc(ctx, (size_t)G->ncrs_slots * sizeof(SlotF), (void**)&G->d_crs_slots))) return rc;
    if ((rc = ctx_alloc(ctx, (size_t)G->ncrs_slots * sizeof(SlotD), (void**)&G->d_crs_slots_d))) return rc;
    if ((rc = ctx_alloc(ctx, (size_t)G->ncrs_slots * sizeof(SlotF), (void**)&G->d_crs_slots_p))) return rc;
    if ((rc = ctx_alloc(ctx, (size_t)(ncdp + 1) * 4, &dnoff))) return rc;
    if ((rc = ctx_alloc(ctx, (size_t)std::max<int64_t>(nnbr, 1) * 4, &dnbr))) return rc;
    HIPR(hipMemcpyAsync(G->d_crs_slot_off, G->crs_slot_off.data(), (size_t)(ncdp + 1) * 4, hipMemcpyHostToDevice,
                        cs));
    HIPR(hipMemcpyAsync(dnoff, nbr_off, (size_t)(ncdp + 1) * 4, hipMemcpyHostToDevice, cs));
    if (nnbr) HIPR(hipMemcpyAsync(dnbr, nbr, (size_t)nnbr * 4, hipMemcpyHostToDevice, cs));
    // crs_slot_off (a member) outlives the copies; the caller's lists are
    // read before the first batch's work.
    HIPR(hipStreamSynchronize(cs));
  }
  return SB_OK;
}

// Samples and slots of CDPs [cb, ce) (rows of at most stage_rows per copy).
// *ready = the number of leading CDPs searchable once `ev` has fired.
int GatherBuild::batch(int cb, int ce, hipEvent_t ev, int* ready) {
  const int ns = G->ns, store = G->store;
  const size_t esz = store == SB_STORE_F32 ? 4 : 2;
  const int64_t rb = G->cdp_off[cb], re = G->cdp_off[ce];
  for (int64_t r0 = rb; r0 < re; r0 += stage_rows) {
    const int64_t nr = std::min<int64_t>(stage_rows, re - r0);
    void* st = stage[which];
    which ^= 1;  // the next copy fills the other buffer while this one is converted
    HIPR(hipMemcpyAsync(st, samples + r0 * ns, (size_t)nr * ns * 4, hipMemcpyHostToDevice, cs));
    HIPR(launch_store((const float*)st, (int)nr, ns, G->stride, store, (char*)G->d_samples + (size_t)r0 * G->stride * esz,
                      G->d_scale ? G->d_scale + r0 : nullptr, cs));
  }
  HIPR(launch_cmp_slots(cb, ce - cb, G->d_cdp_off, G->d_h0, G->d_scale, idt, G->d_cmp_slots, cs));
  int done = ce;
  if (G->has_crs) {
    // CDPs whose every neighbour lies below ce (the reach is a prefix maximum).
    int r = crs_ready;
    while (r < G->ncdp && nbr_reach[r] < ce) r++;
    const bool az = G->dm_azimuth_mdeg >= 0;
    const double th = az ? (double)G->dm_azimuth_mdeg / 1000.0 * (M_PI / 180.0) : 0.0;
    HIPR(launch_crs_slots(crs_ready, r - crs_ready, G->d_cdp_off, (const int32_t*)dnoff, (const int32_t*)dnbr,
                          G->d_crs_slot_off, G->d_m0x, G->d_m0y, G->d_h0, G->d_scale, idt, G->d_samples,
                          (uint64_t)G->stride * esz, G->d_crs_slots, G->d_crs_slots_d, G->d_crs_slots_p, G->d_msign,
                          az ? 1 : 0, std::sin(th), std::cos(th), cs));
    crs_ready = r;
    done = r;
  }
  if (ev) HIPR(hipEventRecord(ev, cs));
  *ready = done;
  return SB_OK;
}

int GatherBuild::finish(sb_gather** out) {
  HIPR(hipStreamSynchronize(cs));
  sb_gather* g = G.get();
  g->bytes = 0;
  {
    std::lock_guard<std::mutex> lk(ctx->mem_mu);
    for (void* p : {(void*)g->d_samples_mem, (void*)g->d_scale, (void*)g->d_cdp_off, (void*)g->d_h_ref32, (void*)g->d_h0,
                    (void*)g->d_m0x, (void*)g->d_m0y, (void*)g->d_msign, (void*)g->d_cmp_slots, (void*)g->d_crs_slot_off,
                    (void*)g->d_crs_slots, (void*)g->d_crs_slots_d, (void*)g->d_crs_slots_p}) {
      auto it = ctx->sizes.find(p);
      if (it != ctx->sizes.end()) g->bytes += it->second;
    }
  }
  *out = G.release();
  return SB_OK;
}

}  // namespace sb

extern "C" int sb_gather_create(sb_ctx* ctx, int ncdp, const int32_t* cdp_off, int ns, int dt_us,
                                const sb_trace_geom* geom, const float* samples, int store, const int32_t* nbr_off,
                                const int32_t* nbr, sb_gather** out) {
  if (!ctx || !out) return fail(SB_ERR_INVALID, "sb_gather_create: bad arguments");
  *out = nullptr;
  GatherBuild b;
  const int64_t rows = std::max<int64_t>(1, (int64_t)(64 << 20) / ((int64_t)std::max(ns, 1) * 4));
  int rc = b.begin(ctx, ncdp, cdp_off, ns, dt_us, geom, samples, store, nbr_off, nbr, ctx->stream, rows);
  if (rc) return rc;
  int ready = 0;
  if ((rc = b.batch(0, ncdp, nullptr, &ready))) return rc;
  return b.finish(out);
}

extern "C" int sb_gather_info(const sb_gather* g, int* ncdp, int* ntr, int* ns, int* store, size_t* bytes) {
  if (!g) return fail(SB_ERR_INVALID, "null gather");
  if (ncdp) *ncdp = g->ncdp;
  if (ntr) *ntr = g->ntr;
  if (ns) *ns = g->ns;
  if (store) *store = g->store;
  if (bytes) *bytes = g->bytes;
  return SB_OK;
}

// ----------------------------------------------------------------- results
extern "C" int sb_result_create(sb_ctx* ctx, int ncdp, int ns, sb_result** out) {
  if (!ctx || !out || ncdp <= 0 || ns <= 0) return fail(SB_ERR_INVALID, "sb_result_create: bad arguments");
  *out = nullptr;
  HIPR(hipSetDevice(ctx->device));
  std::unique_ptr<sb_result> r(new sb_result);
  r->ctx = ctx;
  r->ncdp = ncdp;
  r->ns = ns;
  const size_t n = (size_t)ncdp * ns;
  int rc;
  void* p = nullptr;
  if ((rc = ctx_alloc(ctx, n * (6 * 4 + 8), &p))) return rc;
  r->base = p;
  r->index = (int32_t*)p;
  r->a = (float*)((char*)p + n * 4);
  r->b = (float*)((char*)p + n * 8);
  r->c = (float*)((char*)p + n * 12);
  r->coher = (float*)((char*)p + n * 16);
  r->stack = (float*)((char*)p + n * 20);
  r->keys = (unsigned long long*)((char*)p + n * 24);
  *out = r.release();
  return SB_OK;
}

extern "C" void sb_result_destroy(sb_result* r) {
  if (!r) return;
  (void)hipSetDevice(r->ctx->device);
  (void)hipStreamSynchronize(r->ctx->stream);
  ctx_free(r->ctx, r->base);
  delete r;
}

extern "C" int sb_result_download(sb_ctx* ctx, const sb_result* r, sb_host_result* out) {
  if (!ctx || !r || !out) return fail(SB_ERR_INVALID, "sb_result_download: bad arguments");
  HIPR(hipSetDevice(ctx->device));
  const size_t n = (size_t)r->ncdp * r->ns;
  struct {
    void* h;
    const void* d;
  } m[] = {{out->index, r->index}, {out->a, r->a},         {out->b, r->b},
           {out->c, r->c},         {out->coher, r->coher}, {out->stack, r->stack}};
  // The device arrays are consecutive; host arrays laid out the same way
  // (as the serving loop's staging is) take one copy instead of six.
  bool one = true;
  for (int i = 0; i < 6; i++)
    one = one && m[i].h && (char*)m[i].h == (char*)m[0].h + n * 4 * i &&
          (const char*)m[i].d == (const char*)m[0].d + n * 4 * i;
  if (one) {
    HIPR(hipMemcpyAsync(m[0].h, m[0].d, n * 24, hipMemcpyDeviceToHost, ctx->stream));
  } else {
    for (auto& x : m)
      if (x.h) HIPR(hipMemcpyAsync(x.h, x.d, n * 4, hipMemcpyDeviceToHost, ctx->stream));
  }
  HIPR(hipStreamSynchronize(ctx->stream));
  return SB_OK;
}

// --------------------------------------------------------------- searches
namespace sb {

// Work table of CDPs [cb, ce): every CDP contributes ntiles x nchunks items
// (tile-major, parameter chunks inner).  CDPs are dealt to the 8 XCDs by
// decreasing cost onto the least-loaded one, and each XCD's items go to the
// block indices of that XCD (b % 8: consecutive workgroups land on
// consecutive XCDs), so every block of a CDP shares one L2 and heavy CDPs
// start first on every XCD.  Written straight into the pinned staging
// buffer and copied to the device on the ctx stream.
int build_work(sb_ctx* ctx, const std::vector<int64_t>& cost, int cb, int ce, int ntiles, int npar, int pchunk,
               int* nblocks, const std::string& key) {
  if (!key.empty() && key == ctx->work_key) {
    *nblocks = ctx->work_n;
    return SB_OK;
  }
  ctx->work_key.clear();
  constexpr int kXcd = 8;
  const int nchunks = std::max(1, (npar + pchunk - 1) / pchunk);
  const int64_t per = (int64_t)ntiles * nchunks;  // items per CDP
  const int ncdp = ce - cb;
  std::vector<int> order(ncdp);
  std::iota(order.begin(), order.end(), cb);
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return cost[x] > cost[y]; });
  std::vector<std::vector<int>> lists(kXcd);
  std::vector<int64_t> load(kXcd, 0);
  for (int i : order) {
    const int x = (int)(std::min_element(load.begin(), load.end()) - load.begin());
    load[x] += cost[i];
    lists[x].push_back(i);
  }
  // Item j of XCD x sits at index j * 8 + x; the table ends after the last
  // real item (the shorter lists' tails in between are no-ops).
  int64_t n = 0;
  for (int x = 0; x < kXcd; x++)
    if (!lists[x].empty()) n = std::max<int64_t>(n, ((int64_t)lists[x].size() * per - 1) * kXcd + x + 1);
  const size_t bytes = (size_t)std::max<int64_t>(n, 1) * sizeof(WorkItem);
  if (bytes > ctx->work_cap) {
    ctx_free(ctx, ctx->work_buf);
    ctx->work_buf = nullptr;
    ctx->work_cap = 0;
    int rc = ctx_alloc(ctx, bytes * 2, &ctx->work_buf);
    if (rc) return rc;
    ctx->work_cap = bytes * 2;
  }
  if (n > 0) {
    // The staging buffer is rewritten only once the previous table's copy
    // has completed (work_host_ev).
    const size_t nb = (size_t)n * sizeof(WorkItem);
    if (ctx->work_host_ev) HIPR(hipEventSynchronize(ctx->work_host_ev));
    if (nb > ctx->work_host_cap) {
      // at least doubles, from 1 MiB (43690 work items)
      const size_t cap = std::max({nb * 2, ctx->work_host_cap * 2, size_t(1) << 20});
      void* p = nullptr;
      HIPR(hipHostMalloc(&p, cap, hipHostMallocDefault));
      if (ctx->work_host) ctx->work_host_retired.push_back(ctx->work_host);
      ctx->work_host = p;
      ctx->work_host_cap = cap;
    }
    WorkItem* items = (WorkItem*)ctx->work_host;
    for (int x = 0; x < kXcd; x++) {
      int64_t j = 0;
      for (int i : lists[x])
        for (int t = 0; t < ntiles; t++)
          for (int k = 0; k < nchunks; k++, j++)
            items[j * kXcd + x] = WorkItem{i, t, k * pchunk, std::min(npar, (k + 1) * pchunk)};
      for (; j * kXcd + x < n; j++) items[j * kXcd + x] = WorkItem{-1, 0, 0, 0};
    }
    HIPR(hipMemcpyAsync(ctx->work_buf, ctx->work_host, nb, hipMemcpyHostToDevice, ctx->stream));
    if (!ctx->work_host_ev) HIPR(hipEventCreateWithFlags(&ctx->work_host_ev, hipEventDisableTiming));
    HIPR(hipEventRecord(ctx->work_host_ev, ctx->stream));
  }
  *nblocks = (int)n;
  ctx->work_key = key;
  ctx->work_n = *nblocks;
  return SB_OK;
}

}  // namespace sb

namespace {

// Identifies a work table: gather object + its generation, method, kernel
// family and tiling, so repeated calls on one gather reuse the uploaded table.
std::string work_key(const sb_gather* g, int method, int family, int ntiles, int npar, int pchunk, int cb, int ce) {
  char buf[192];
  std::snprintf(buf, sizeof buf, "%p/%llu/%d/%d/%d/%d/%d/%d/%d", (const void*)g, (unsigned long long)g->serial,
                method, family, ntiles, npar, pchunk, cb, ce);
  return buf;
}

void base_args(const sb_gather* g, sb_result* r, int cb, SearchArgs& a) {
  std::memset(&a, 0, sizeof a);
  a.cdp_base = cb;
  a.nout = r->ncdp;
  a.samples = g->d_samples;
  a.scale = g->d_scale;
  a.stride = g->stride;
  a.cdp_off = g->d_cdp_off;
  a.h_ref32 = g->d_h_ref32;
  a.ncdp = g->ncdp;
  a.ns = g->ns;
  a.keys = r->keys;
  a.out_index = r->index;
  a.out_a = r->a;
  a.out_b = r->b;
  a.out_c = r->c;
  a.out_coher = r->coher;
  a.out_stack = r->stack;
}

// A result covers CDPs [cb, ce) of the gather; gather and result may belong
// to different contexts of the same device (a resident gather is shared).
int check_range(const sb_ctx* ctx, const sb_gather* g, const sb_result* r, int cb, int ce) {
  if (cb < 0 || ce > g->ncdp || cb >= ce) return fail(SB_ERR_INVALID, "bad CDP range");
  if (r->ncdp != ce - cb || r->ns != g->ns) return fail(SB_ERR_INVALID, "result size does not match the CDP range");
  if (r->ctx != ctx || g->ctx->device != ctx->device)
    return fail(SB_ERR_INVALID, "result must belong to the ctx, gather to the same device");
  return SB_OK;
}

// Timing events are not recorded while a graph is being captured.
hipError_t record(sb_ctx* ctx, int i) { return ctx->capturing ? hipSuccess : hipEventRecord(ctx->ev[i], ctx->stream); }

// SB_OPT_RESTORE_FP32: the rows a fast search of a 16-bit store reads,
// restored to FP32 in a transient copy (same row stride, kSamplesLead bytes
// in front as in an FP32 store), so that the FP32 kernels run on the values
// the 16-bit kernels restore after each load; CRS-DE also gets a copy of its
// slot table with row addresses in the copy.  The blocks go back to the ctx
// cache (stream-ordered) when this goes out of scope -- or, in the eager run
// of sb_graph_create, to the graph, whose capture then restores into them.
struct Restored {
  sb_ctx* ctx = nullptr;
  bool owned = false;  // the blocks are this search's (else the graph's)
  void* mem = nullptr;
  void* slots = nullptr;
  int64_t r0 = 0, r1 = 0;  // rows of the copy
  int so0 = 0, so1 = 0;    // rebased slots
  ~Restored() {
    if (!owned) return;
    if (ctx->restore_keep) {  // handed to the graph being created
      ctx->kept_mem = mem;
      ctx->kept_slots = slots;
      ctx->kept_mem_bytes = mem_bytes();
      ctx->kept_slots_bytes = slots_bytes();
      return;
    }
    ctx_free(ctx, mem);
    ctx_free(ctx, slots);
  }
  size_t mem_bytes() const { return (size_t)kSamplesLead + (size_t)(r1 - r0) * stride * 4; }
  size_t slots_bytes() const { return (size_t)(so1 - so0) * sizeof(SlotF); }
  int64_t stride = 0;
  // Sets up the copy for CDPs [cb, ce) (CMP: their rows; CRS: the rows of
  // their supergathers); false, with nothing held, when it does not apply
  // (FP32 store, option off, a graph capture without blocks kept for it) or
  // does not fit the budget.
  bool prepare(sb_ctx* c, const sb_gather* g, int cb, int ce, bool crs, bool rebase) {
    if (g->store == SB_STORE_F32 || !c->restore_fp32) return false;
    stride = g->stride;
    if (crs) {
      r0 = g->crs_row_lo[cb];
      r1 = g->crs_row_hi[cb];
      for (int i = cb + 1; i < ce; i++) {
        r0 = std::min<int64_t>(r0, g->crs_row_lo[i]);
        r1 = std::max<int64_t>(r1, g->crs_row_hi[i]);
      }
    } else {
      r0 = g->cdp_off[cb];
      r1 = g->cdp_off[ce];
    }
    if (rebase) {
      so0 = g->crs_slot_off[cb];
      so1 = g->crs_slot_off[ce];
    }
    if (c->capturing) {
      // no allocation under capture: the blocks the eager run kept
      if (!c->graph_rest_mem || c->graph_rest_mem_bytes < mem_bytes()) return false;
      if (rebase && (!c->graph_rest_slots || c->graph_rest_slots_bytes < slots_bytes())) return false;
      ctx = c;
      mem = c->graph_rest_mem;
      slots = rebase ? c->graph_rest_slots : nullptr;
      return true;
    }
    if (ctx_alloc_transient(c, mem_bytes(), &mem) != SB_OK) return false;
    if (rebase && ctx_alloc_transient(c, slots_bytes(), &slots) != SB_OK) {
      ctx_free(c, mem);
      mem = nullptr;
      return false;
    }
    ctx = c;
    owned = true;
    return true;
  }
  // Row 0 of the copy as the kernels address it (rows below r0 are never read).
  uintptr_t row0(const sb_gather* g) const {
    return (uintptr_t)mem + (uintptr_t)kSamplesLead - (uintptr_t)(r0 * g->stride * 4);
  }
  // The search's arguments on the copy (FP32 store from here on).
  void apply(const sb_gather* g, SearchArgs& a) const {
    a.samples = (const void*)row0(g);
    if (slots) a.slots = (const SlotF*)((uintptr_t)slots - (uintptr_t)so0 * sizeof(SlotF));
  }
  hipError_t launch(const sb_gather* g) const {
    hipError_t e = launch_restore_rows((const char*)g->d_samples + r0 * g->stride * 2, g->d_scale + r0, r1 - r0,
                                       g->stride, g->store, (float*)((char*)mem + kSamplesLead), ctx->stream);
    if (e == hipSuccess && slots)
      e = launch_rebase_slots(g->d_crs_slots + so0, so1 - so0, (uint64_t)row0(g), (uint64_t)g->stride * 4,
                              (SlotF*)slots, ctx->stream);
    return e;
  }
};

// The work table a launch reads: a graph being captured gets its own copy.
const WorkItem* work_table(const sb_ctx* ctx) {
  return (const WorkItem*)(ctx->graph_work ? ctx->graph_work : ctx->work_buf);
}

int choose_pchunk(const sb_ctx* ctx, int ncdp, int ntiles, int npar, int unit) {
  // Enough blocks to fill every CU several times, and at most ~24 passes
  // over the traces per wave (long blocks leave the CUs unevenly loaded at
  // the end of a launch: the 11^3 CRS grid over 500 CDPs ran 438 ms in one
  // chunk, 404 ms in four); chunks a multiple of `unit` (waves per block x
  // parameters per wave pass).
  const int64_t want = (int64_t)ctx->num_cus * 8;
  const int64_t have = (int64_t)ncdp * ntiles;
  const int64_t by_fill = (want + have - 1) / have;
  const int64_t by_length = (npar + 24 * unit - 1) / (24 * unit);
  int nchunks = (int)std::min<int64_t>(std::max(by_fill, by_length), (npar + unit - 1) / unit);
  nchunks = std::max(1, nchunks);
  int pchunk = (npar + nchunks - 1) / nchunks;
  pchunk = (pchunk + unit - 1) / unit * unit;
  return pchunk;
}

// Parameters per workgroup of the LDS-staged search when untuned: four, two
// or one groups of CC per wave -- the most that still leave eight rounds of
// three workgroups per CU (one staging then serves more evaluations; small
// searches keep enough workgroups to spread over the CUs).
int lds_default_pchunk(const sb_ctx* ctx, int ncdp, int ntiles, int npar, int unit) {
  const int64_t want = (int64_t)ctx->num_cus * 3 * 8;
  for (int groups = 4; groups > 1; groups /= 2) {
    const int pc = groups * unit;
    if ((int64_t)ncdp * ntiles * ((npar + pc - 1) / pc) >= want) return pc;
  }
  return unit;
}

// Whether k_search_lds can run every CDP of [cb, ce): the per-trace tables
// of the largest CDP must fit its LDS with room for slabs.  (A workgroup
// whose window ranges do not fit the rest reads its windows from global
// memory, with the same results.)
bool lds_fits(const sb_gather* g, const SearchArgs& a, const FastVariant& v, int cb, int ce) {
  (void)a;
  int smax = 0;
  for (int i = cb; i < ce; i++) smax = std::max(smax, g->cdp_off[i + 1] - g->cdp_off[i]);
  if (smax > kLdsMaxTraces) return false;
  const int smp = lds_slab_offset(v, smax);
  return smp >= 0 && (int64_t)smp + 4096 <= (int64_t)v.lds_kb * 1024;
}

// Where a search's scratch lives: the ctx's cached work table and the
// result's key array (allocated through the ctx), or, for the
// allocation-free entry point, the static device array: the work table built
// on the device in its first bytes, the keys after it.
struct Scratch {
  void* base = nullptr;  // null: the ctx's buffers
};

int run_search(sb_ctx* ctx, const sb_gather* g, SearchArgs& a, int mode, double params_per_tuple,
               const std::vector<int64_t>& S, int cb, int ce, Scratch sc = Scratch{}) {
  HIPR(hipSetDevice(ctx->device));
  std::lock_guard<std::mutex> lk(ctx->run_mu);
  int rc;
  int nblocks = 0;
  const int store = g->store;
  const std::vector<int64_t>& cost = S;
  double tuples = 0;
  for (int i = cb; i < ce; i++) tuples += (double)S[i] * g->ns * params_per_tuple;
  int variant = -1;
  bool lds = false;
  bool crs_lds = false;  // the LDS-staged CRS grid search
  int smax = 0;
  FastVariant v{};
  // Windows the fast kernel is not compiled for (w > 7) run in reference
  // arithmetic: slower, same parity bar.
  if (mode == SB_MODE_FAST && !fast_supported(a.w)) mode = SB_MODE_EXACT;
  // Reference arithmetic on the tiled layout (keys + winner re-evaluation)
  // where it is compiled; otherwise one thread per (CDP, t0).
  const bool tiled_exact = mode == SB_MODE_EXACT && exact_tiled_supported(a.method, store, a.w);
  // A 16-bit store's fast search on a restored FP32 copy (SB_OPT_RESTORE_FP32;
  // not for the allocation-free entry point): kstore is the store the search
  // kernel reads.  The winners are re-evaluated from the store itself.
  Restored rest;
  const bool restored =
      mode == SB_MODE_FAST && !sc.base && rest.prepare(ctx, g, cb, ce, a.method == SB_METHOD_CRS, false);
  const int kstore = restored ? SB_STORE_F32 : store;
  // The work table's tiling: family (for the cache key), t0 tiles and
  // parameters per workgroup.
  int family, ntiles, npar_w, pchunk;
  if (mode == SB_MODE_FAST) {
    variant = ctx->fast_variant >= 0 ? ctx->fast_variant
              : kstore == SB_STORE_F32 ? kDefaultFastVariantF32
                                       : kDefaultFastVariant;
    v = kFastVariants[variant];
    if (v.lds_kb > 0 && a.method == SB_METHOD_CMP && lds_fits(g, a, v, cb, ce)) {
      lds = true;
      a.lds_bytes = v.lds_kb * 1024;
      a.pair_share = v.share;
    }
    // The global-memory kernel runs kGlobalFallbackVariant when the LDS
    // slabs do not apply (CRS, or CDPs too large to stage).  Both kernels
    // read the same work table.
    if (v.lds_kb > 0 && !lds) {
      variant = kGlobalFallbackVariant;
      v = kFastVariants[variant];
    }
    family = 0;
    ntiles = (g->ns + 64 * v.R - 1) / (64 * v.R);
    npar_w = a.npar;
    pchunk = ctx->pchunk_override > 0 ? ctx->pchunk_override
             : lds ? lds_default_pchunk(ctx, ce - cb, ntiles, a.npar, v.waves * v.CC)
                   : choose_pchunk(ctx, ce - cb, ntiles, a.npar, v.waves * v.CC);
    // CRS on FP32 rows: the LDS-staged search (SB_OPT_CRS_LDS), its work items
    // (CDP, 64-t0 tile, whole (a, b) pairs of the grid)
    if (a.method == SB_METHOD_CRS && ctx->crs_lds != 0 && kstore == SB_STORE_F32 && !sc.base) {
      const int lb = ctx->crs_lds < 0 ? kCrsLdsDefaultBytes : ctx->crs_lds;
      for (int i = cb; i < ce; i++) smax = std::max<int>(smax, (int)S[i]);
      if (crs_lds_supported(a.w, a.nc, ctx->crs_cpw, kCrsLdsWaves, ctx->crs_pairs, lb, g->ns, smax)) {
        crs_lds = true;
        a.lds_bytes = lb;
        family = 3;
        ntiles = (g->ns + 63) / 64;
        pchunk = crs_lds_pairs_per_block(a.nc, ctx->crs_cpw, kCrsLdsWaves) * a.nc;
      }
    }
  } else if (tiled_exact) {
    family = 2;
    ntiles = (g->ns + 63) / 64;
    npar_w = a.npar;
    pchunk = a.method == SB_METHOD_CRS ? a.npar : choose_pchunk(ctx, ce - cb, ntiles, a.npar, 16);
  } else {
    family = 1;
    ntiles = (g->ns + 63) / 64;
    npar_w = 1;
    pchunk = 1;
  }
  const int64_t nout = (int64_t)(ce - cb) * g->ns;
  if (!sc.base) {
    const std::string key = work_key(g, a.method, family, ntiles, npar_w, pchunk, cb, ce);
    if ((rc = build_work(ctx, cost, cb, ce, ntiles, npar_w, pchunk, &nblocks, key))) return rc;
    a.work = work_table(ctx);
  } else {
    const int nchunks = std::max(1, (npar_w + pchunk - 1) / pchunk);
    const int64_t items = (int64_t)(ce - cb) * ntiles * nchunks;
    const int64_t nb = (items + 7) / 8 * 8;
    if (nb > kScratchWorkItems || nout > kScratchKeys)
      return fail(SB_ERR_UNSUPPORTED, "search too large for the static scratch: " + std::to_string(nb) +
                                          " work items (max " + std::to_string(kScratchWorkItems) + "), " +
                                          std::to_string(nout) + " (CDP, t0) keys (max " +
                                          std::to_string(kScratchKeys) + ")");
    nblocks = (int)nb;
    WorkItem* w = (WorkItem*)sc.base;
    a.keys = (unsigned long long*)((char*)sc.base + kScratchKeysOffset);
    const hipError_t e = launch_make_work(w, nblocks, cb, ce - cb, ntiles, nchunks, pchunk, npar_w, ctx->stream);
    if (e != hipSuccess) return hip_fail(e, "work table");
    a.work = w;
  }
  a.nwork = nblocks;
  HIPR(record(ctx, 0));
  hipError_t e = hipSuccess;
  if (mode == SB_MODE_FAST || tiled_exact) e = launch_fill_u64(a.keys, nout, 0ull, ctx->stream);
  if (e != hipSuccess) return hip_fail(e, "fill");
  const bool persistent = mode == SB_MODE_FAST && !lds && !crs_lds && ctx->persistent && !sc.base;
  if (persistent) {
    // Persistent grid: 8 per-XCD item counters, cleared by a kernel (not a
    // memset, so that the sequence can be captured into a graph).
    if (!ctx->queue_buf && (rc = ctx_alloc(ctx, 64, &ctx->queue_buf))) return rc;
    a.queue = (unsigned int*)ctx->queue_buf;
    e = launch_fill_u64((unsigned long long*)ctx->queue_buf, 4, 0ull, ctx->stream);
    if (e != hipSuccess) return hip_fail(e, "queue clear");
  }
  HIPR(record(ctx, 1));
  SearchArgs as = a;  // the search kernel's view of the samples
  if (restored) {
    rest.apply(g, as);
    e = rest.launch(g);
    if (e != hipSuccess) return hip_fail(e, "FP32 restore");
  }
  if (mode == SB_MODE_FAST && crs_lds)
    e = launch_search_crs_lds(as, ctx->crs_cpw, kCrsLdsWaves, ctx->crs_pairs, smax, nblocks, ctx->stream);
  else if (mode == SB_MODE_FAST && lds)
    e = launch_search_lds(as, kstore, v, nblocks, ctx->stream);
  else if (mode == SB_MODE_FAST)
    e = launch_search_fast(as, kstore, variant, nblocks, ctx->num_cus, ctx->stream);
  else if (tiled_exact)
    e = launch_search_exact_tiled(a, nblocks, ctx->stream);
  else
    e = launch_search_exact(a, store, nblocks, ctx->stream);
  if (e != hipSuccess) return hip_fail(e, "search launch");
  HIPR(record(ctx, 2));
  if (mode == SB_MODE_FAST || tiled_exact) {
    e = launch_finalize(a, store, ctx->stream);
    if (e != hipSuccess) return hip_fail(e, "finalize launch");
  }
  HIPR(record(ctx, 3));
  ctx->last = sb_timing{};
  ctx->last.tuples = tuples;
  ctx->last.grid_blocks = nblocks;
  ctx->last.block_threads =
      mode == SB_MODE_FAST || tiled_exact ? (crs_lds ? 64 * kCrsLdsWaves : lds ? 64 * v.waves : 256) : 64;
  ctx->last.variant = crs_lds ? -2 : variant;
  ctx->last.lds_staged = lds || crs_lds ? 1 : 0;
  ctx->last.restored = restored ? 1 : 0;
  ctx->timing_pending = true;
  ctx->timing_graph = false;
  return SB_OK;
}

// CMP parameter set-up exactly as CMP/OpenMP/src/main.cpp:68-110 (float):
// grid, window and the fast search's t^2 bound; *mode = the arithmetic that
// runs (a negative C range runs in reference arithmetic).
int cmp_setup(const sb_gather* g, const sb_cmp_params* p, SearchArgs& a, int* mode) {
  if (p->nc <= 0) return fail(SB_ERR_INVALID, "nc must be positive");
  if (p->mode != SB_MODE_FAST && p->mode != SB_MODE_EXACT) return fail(SB_ERR_INVALID, "bad mode");
  a.method = SB_METHOD_CMP;
  a.slot_off = g->d_cdp_off;
  a.slots = g->d_cmp_slots;
  // c0 = stof(c0) * 1e6, ...
  const float c0 = (float)((double)(float)p->c0 * 1e6);
  const float c1 = (float)((double)(float)p->c1 * 1e6);
  const float itau = (float)p->tau;
  const float inc = (c1 - c0) * (1.0f / (float)p->nc);
  float dt = (float)g->dt_us;
  dt = dt / 1000000.0f;
  const float idt = 1.0f / dt;
  const float tt = itau * idt;
  const int tau = (int)tt > 0 ? (int)tt : 0;
  a.tau = tau;
  a.w = 2 * tau + 1;
  if (a.w > SB_MAX_W) return fail(SB_ERR_UNSUPPORTED, "window 2*tau+1 exceeds SB_MAX_W");
  a.cf0 = c0;
  a.cfinc = inc;
  a.fc0 = c0;
  a.finc_c = inc;
  a.nc = p->nc;
  a.na = a.nb = 1;
  a.npar = p->nc;
  a.dt_f = dt;
  a.idt_f = idt;
  {
    const int hi = g->ns - tau - 1;
    a.hi2 = hi > 0 ? (float)((double)hi * (double)hi * (1.0 - 0x1p-20)) : 0.f;
  }
  // The fast kernel assumes t0^2 + C h^2 >= 0; a negative C range (NaN
  // traveltimes, all invalid in the reference) runs in reference arithmetic.
  *mode = (c0 < 0.f || c1 < 0.f) ? SB_MODE_EXACT : p->mode;
  return SB_OK;
}

std::vector<int64_t> cmp_cost(const sb_gather* g) {
  std::vector<int64_t> S(g->ncdp);
  for (int i = 0; i < g->ncdp; i++) S[i] = g->cdp_off[i + 1] - g->cdp_off[i];
  return S;
}

}  // namespace

extern "C" int sb_cmp_run(sb_ctx* ctx, const sb_gather* g, const sb_cmp_params* p, sb_result* r) {
  if (!g) return fail(SB_ERR_INVALID, "sb_cmp_run: null gather");
  return sb_cmp_run_range(ctx, g, p, 0, g->ncdp, r);
}

extern "C" int sb_cmp_run_range(sb_ctx* ctx, const sb_gather* g, const sb_cmp_params* p, int cb, int ce,
                                sb_result* r) {
  if (!ctx || !g || !p || !r) return fail(SB_ERR_INVALID, "sb_cmp_run: null argument");
  int rc;
  if ((rc = check_range(ctx, g, r, cb, ce))) return rc;
  SearchArgs a;
  base_args(g, r, cb, a);
  int mode;
  if ((rc = cmp_setup(g, p, a, &mode))) return rc;
  return run_search(ctx, g, a, mode, (double)p->nc, cmp_cost(g), cb, ce);
}

extern "C" int sb_cmp_run_noalloc(sb_ctx* ctx, const sb_gather* g, const sb_cmp_params* p, int cb, int ce,
                                  const sb_device_result* out) {
  if (!ctx || !g || !p || !out) return fail(SB_ERR_INVALID, "sb_cmp_run_noalloc: null argument");
  if (!out->index || !out->a || !out->b || !out->c || !out->coher || !out->stack)
    return fail(SB_ERR_INVALID, "sb_cmp_run_noalloc: every output array is required");
  if (cb < 0 || ce > g->ncdp || cb >= ce) return fail(SB_ERR_INVALID, "bad CDP range");
  if (g->ctx->device != ctx->device) return fail(SB_ERR_INVALID, "gather on another device");
  if (!ctx->scratch) return fail(SB_ERR_STATE, "static scratch not initialised on this device");
  int rc;
  SearchArgs a;
  std::memset(&a, 0, sizeof a);
  a.cdp_base = cb;
  a.nout = ce - cb;
  a.samples = g->d_samples;
  a.scale = g->d_scale;
  a.stride = g->stride;
  a.cdp_off = g->d_cdp_off;
  a.h_ref32 = g->d_h_ref32;
  a.ncdp = g->ncdp;
  a.ns = g->ns;
  a.out_index = out->index;
  a.out_a = out->a;
  a.out_b = out->b;
  a.out_c = out->c;
  a.out_coher = out->coher;
  a.out_stack = out->stack;
  int mode;
  if ((rc = cmp_setup(g, p, a, &mode))) return rc;
  Scratch sc;
  sc.base = ctx->scratch;
  return run_search(ctx, g, a, mode, (double)p->nc, cmp_cost(g), cb, ce, sc);
}

extern "C" int sb_crs_run(sb_ctx* ctx, const sb_gather* g, const sb_crs_params* p, sb_result* r) {
  if (!g) return fail(SB_ERR_INVALID, "sb_crs_run: null gather");
  return sb_crs_run_range(ctx, g, p, 0, g->ncdp, r);
}

extern "C" int sb_crs_run_range(sb_ctx* ctx, const sb_gather* g, const sb_crs_params* p, int cb, int ce,
                                sb_result* r) {
  if (!ctx || !g || !p || !r) return fail(SB_ERR_INVALID, "sb_crs_run: null argument");
  int rc;
  if ((rc = check_range(ctx, g, r, cb, ce))) return rc;
  if (!g->has_crs) return fail(SB_ERR_STATE, "gather was created without neighbour lists");
  if (p->na <= 0 || p->nb <= 0 || p->nc <= 0) return fail(SB_ERR_INVALID, "na, nb, nc must be positive");
  if ((int64_t)p->na * p->nb * p->nc > INT32_MAX / 2) return fail(SB_ERR_INVALID, "parameter grid too large");
  if (p->mode != SB_MODE_FAST && p->mode != SB_MODE_EXACT) return fail(SB_ERR_INVALID, "bad mode");
  SearchArgs a;
  base_args(g, r, cb, a);
  a.method = SB_METHOD_CRS;
  a.slot_off = g->d_crs_slot_off;
  a.slots = g->d_crs_slots;
  a.slots_d = g->d_crs_slots_d;
  // CRS/OpenMP/src/main.cpp:85-138, all in double.
  const double c0 = p->c0 * 1e6, c1 = p->c1 * 1e6;
  const double inc_a = (p->a1 - p->a0) * (1.0 / (double)p->na);
  const double inc_b = (p->b1 - p->b0) * (1.0 / (double)p->nb);
  const double inc_c = (c1 - c0) * (1.0 / (double)p->nc);
  double dt = (double)g->dt_us;
  dt = dt / 1000000.0;
  const double idt = 1.0 / dt;
  const double tt = p->tau * idt;
  const int tau = (int)tt > 0 ? (int)tt : 0;
  a.tau = tau;
  a.w = 2 * tau + 1;
  if (a.w > SB_MAX_W) return fail(SB_ERR_UNSUPPORTED, "window 2*tau+1 exceeds SB_MAX_W");
  a.na = p->na;
  a.nb = p->nb;
  a.nc = p->nc;
  a.npar = p->na * p->nb * p->nc;
  a.da0 = p->a0;
  a.dinc_a = inc_a;
  a.db0 = p->b0;
  a.dinc_b = inc_b;
  a.dc0 = c0;
  a.dinc_c = inc_c;
  a.fa0 = (float)p->a0;
  a.finc_a = (float)inc_a;
  a.fb0 = (float)p->b0;
  a.finc_b = (float)inc_b;
  a.fc0 = (float)c0;
  a.finc_c = (float)inc_c;
  a.dt_d = dt;
  a.idt_d = idt;
  std::vector<int64_t> S(g->ncdp);
  for (int i = 0; i < g->ncdp; i++) S[i] = g->crs_slot_off[i + 1] - g->crs_slot_off[i];
  return run_search(ctx, g, a, p->mode, (double)a.npar, S, cb, ce);
}

extern "C" int sb_crs_de_run(sb_ctx* ctx, const sb_gather* g, const sb_de_params* p, sb_result* r) {
  if (!g) return fail(SB_ERR_INVALID, "sb_crs_de_run: null gather");
  return sb_crs_de_run_range(ctx, g, p, 0, g->ncdp, r);
}

extern "C" int sb_crs_de_run_range(sb_ctx* ctx, const sb_gather* g, const sb_de_params* p, int cb, int ce,
                                   sb_result* r) {
  if (!ctx || !g || !p || !r) return fail(SB_ERR_INVALID, "sb_crs_de_run: null argument");
  int rc;
  if ((rc = check_range(ctx, g, r, cb, ce))) return rc;
  if (!g->has_crs) return fail(SB_ERR_STATE, "gather was created without neighbour lists");
  if (p->pop < 4 || p->pop > 64) return fail(SB_ERR_INVALID, "population must be in [4, 64]");
  if (p->ngen < 0) return fail(SB_ERR_INVALID, "ngen must be >= 0");
  if (!(p->CR >= 0 && p->CR <= 1) || !(p->F >= 0 && p->F <= 2)) return fail(SB_ERR_INVALID, "bad F or CR");
  if (p->mode != SB_MODE_FAST && p->mode != SB_MODE_EXACT) return fail(SB_ERR_INVALID, "bad mode");
  SearchArgs a;
  base_args(g, r, cb, a);
  a.method = SB_METHOD_CRS;
  a.slot_off = g->d_crs_slot_off;
  a.slots = g->d_crs_slots;
  a.slots_d = g->d_crs_slots_d;
  a.slots_p = g->d_crs_slots_p;
  // Window as the CRS reference computes it (double); fitness time base in
  // float (the precision of the reference's FP32 CRS variants).
  double dtd = (double)g->dt_us;
  dtd = dtd / 1000000.0;
  const double idtd = 1.0 / dtd;
  const double tt = p->tau * idtd;
  a.tau = (int)tt > 0 ? (int)tt : 0;
  a.w = 2 * a.tau + 1;
  if (a.w > SB_MAX_W) return fail(SB_ERR_UNSUPPORTED, "window 2*tau+1 exceeds SB_MAX_W");
  // Windows the fast evaluator is not compiled for (w > 7) use the
  // reference-order evaluator.
  const int de_mode = p->mode == SB_MODE_FAST && !fast_supported(a.w) ? SB_MODE_EXACT : p->mode;
  float dt = (float)g->dt_us;
  dt = dt / 1000000.0f;
  a.dt_f = dt;
  a.idt_f = 1.0f / dt;
  DeArgs d;
  std::memset(&d, 0, sizeof d);
  d.pop = p->pop;
  d.ngen = p->ngen;
  d.F = (float)p->F;
  d.cr_thr = (uint32_t)(p->CR * 16777216.0);
  d.seed = p->seed;
  d.lo[0] = (float)p->a0;
  d.hi[0] = (float)p->a1;
  d.lo[1] = (float)p->b0;
  d.hi[1] = (float)p->b1;
  d.lo[2] = (float)(p->c0 * 1e6);
  d.hi[2] = (float)(p->c1 * 1e6);
  const int P2 = d.pop <= 4 ? 4 : d.pop <= 8 ? 8 : d.pop <= 16 ? 16 : d.pop <= 32 ? 32 : 64;

  HIPR(hipSetDevice(ctx->device));
  std::lock_guard<std::mutex> lk(ctx->run_mu);
  // A 16-bit store's fast search on a restored FP32 copy of its supergather
  // rows, with a slot table addressing the copy (SB_OPT_RESTORE_FP32).
  Restored rest;
  const bool restored = de_mode == SB_MODE_FAST && rest.prepare(ctx, g, cb, ce, true, true);
  // The LDS-staged evaluator (SB_OPT_DE_LDS) for fast searches of FP32 rows.
  const int lds_bytes = ctx->de_lds < 0 ? kDeLdsDefaultBytes : ctx->de_lds;
  int smax = 0;
  for (int i = cb; i < ce; i++) smax = std::max(smax, g->crs_slot_off[i + 1] - g->crs_slot_off[i]);
  const bool use_lds = de_mode == SB_MODE_FAST && ctx->de_lds != 0 && (restored || g->store == SB_STORE_F32) &&
                       de_lds_supported(a.w, d.pop, ctx->de_waves, ctx->de_pairs, lds_bytes, g->ns, smax);
  const int threads = use_lds ? 64 * ctx->de_waves : 256;
  const int per_block = threads / P2;
  const int ntiles = (g->ns + per_block - 1) / per_block;
  std::vector<int64_t> cost(g->ncdp);
  double tuples = 0;
  for (int i = 0; i < g->ncdp; i++) {
    cost[i] = g->crs_slot_off[i + 1] - g->crs_slot_off[i];
    if (i >= cb && i < ce) tuples += (double)cost[i] * g->ns * (double)p->pop * (p->ngen + 1);
  }
  int nblocks = 0;
  if ((rc = build_work(ctx, cost, cb, ce, ntiles, 1, 1, &nblocks, work_key(g, SB_METHOD_CRS, 2, ntiles, 1, 1, cb, ce))))
    return rc;
  a.work = work_table(ctx);
  a.nwork = nblocks;
  HIPR(record(ctx, 0));
  HIPR(record(ctx, 1));
  SearchArgs as = a;
  hipError_t e = hipSuccess;
  if (restored) {
    rest.apply(g, as);
    e = rest.launch(g);
    if (e != hipSuccess) return hip_fail(e, "FP32 restore");
  }
  if (use_lds) {
    as.lds_bytes = lds_bytes;
    e = launch_de_lds(as, d, ctx->de_waves, ctx->de_pairs, smax, nblocks, ctx->stream);
  } else {
    e = launch_de(as, d, restored ? SB_STORE_F32 : g->store, de_mode == SB_MODE_EXACT, nblocks, ctx->stream);
  }
  if (e != hipSuccess) return hip_fail(e, "CRS-DE launch");
  HIPR(record(ctx, 2));
  HIPR(record(ctx, 3));
  ctx->last = sb_timing{};
  ctx->last.tuples = tuples;
  ctx->last.grid_blocks = nblocks;
  ctx->last.block_threads = threads;
  ctx->last.lds_staged = use_lds ? 1 : 0;
  ctx->last.variant = -1;
  ctx->last.restored = restored ? 1 : 0;
  ctx->timing_pending = true;
  ctx->timing_graph = false;
  return SB_OK;
}

extern "C" int sb_last_timing(const sb_ctx* cctx, sb_timing* t) {
  sb_ctx* ctx = const_cast<sb_ctx*>(cctx);
  if (!ctx || !t) return fail(SB_ERR_INVALID, "sb_last_timing: bad arguments");
  if (ctx->timing_pending) {
    HIPR(hipSetDevice(ctx->device));
    HIPR(hipEventSynchronize(ctx->ev[3]));
    float a = 0, b = 0, c = 0;
    HIPR(hipEventElapsedTime(&c, ctx->ev[0], ctx->ev[3]));
    if (ctx->timing_graph) {  // a graph replay: only its whole span is timed
      a = c;
    } else {
      HIPR(hipEventElapsedTime(&a, ctx->ev[1], ctx->ev[2]));
      HIPR(hipEventElapsedTime(&b, ctx->ev[2], ctx->ev[3]));
    }
    ctx->last.search_ms = a;
    ctx->last.finalize_ms = b;
    ctx->last.total_ms = c;
    ctx->timing_pending = false;
  }
  *t = ctx->last;
  return SB_OK;
}

namespace {
template <class Run, class P>
int run_and_download(sb_ctx* ctx, const sb_gather* g, const P* p, sb_host_result* out, Run run) {
  if (!ctx || !g || !p || !out) return fail(SB_ERR_INVALID, "null argument");
  sb_result* r = nullptr;
  int rc = sb_result_create(ctx, g->ncdp, g->ns, &r);
  if (rc) return rc;
  rc = run(ctx, g, p, r);
  if (!rc) rc = sb_result_download(ctx, r, out);
  sb_result_destroy(r);
  return rc;
}
}  // namespace

extern "C" int sb_cmp(sb_ctx* ctx, const sb_gather* g, const sb_cmp_params* p, sb_host_result* out) {
  return run_and_download(ctx, g, p, out, sb_cmp_run);
}
extern "C" int sb_crs(sb_ctx* ctx, const sb_gather* g, const sb_crs_params* p, sb_host_result* out) {
  return run_and_download(ctx, g, p, out, sb_crs_run);
}
extern "C" int sb_crs_de(sb_ctx* ctx, const sb_gather* g, const sb_de_params* p, sb_host_result* out) {
  return run_and_download(ctx, g, p, out, sb_crs_de_run);
}


// ------------------------------------------------------------- options
extern "C" int sb_ctx_set_option(sb_ctx* ctx, int option, int value) {
  if (!ctx) return fail(SB_ERR_INVALID, "null ctx");
  if (option == SB_OPT_PERSISTENT) {
    ctx->persistent = value ? 1 : 0;
    return SB_OK;
  }
  if (option == SB_OPT_STREAM_ROWS) {
    if (value < 0) return fail(SB_ERR_INVALID, "SB_OPT_STREAM_ROWS must be >= 0");
    ctx->stream_rows = value;
    return SB_OK;
  }
  if (option == SB_OPT_DM_AZIMUTH) {
    if (value < -1 || value >= 360000) return fail(SB_ERR_INVALID, "SB_OPT_DM_AZIMUTH must be -1 or in [0, 360000)");
    ctx->dm_azimuth_mdeg = value;
    return SB_OK;
  }
  if (option == SB_OPT_RESTORE_FP32) {
    ctx->restore_fp32 = value ? 1 : 0;
    return SB_OK;
  }
  if (option == SB_OPT_DE_LDS) {
    if (value != -1 && value != 0 && (value < 4096 || value > kLdsMaxBytes || value % 16))
      return fail(SB_ERR_INVALID, "SB_OPT_DE_LDS must be -1, 0 or a multiple of 16 in [4096, 65536]");
    ctx->de_lds = value;
    return SB_OK;
  }
  if (option == SB_OPT_DE_WAVES) {
    if (value != 4 && value != 8 && value != 16) return fail(SB_ERR_INVALID, "SB_OPT_DE_WAVES must be 4, 8 or 16");
    ctx->de_waves = value;
    return SB_OK;
  }
  if (option == SB_OPT_DE_PAIRS) {
    ctx->de_pairs = value ? 1 : 0;
    return SB_OK;
  }
  if (option == SB_OPT_CRS_LDS) {
    if (value != -1 && value != 0 && (value < 4096 || value > kLdsMaxBytes || value % 16))
      return fail(SB_ERR_INVALID, "SB_OPT_CRS_LDS must be -1, 0 or a multiple of 16 in [4096, 65536]");
    ctx->crs_lds = value;
    return SB_OK;
  }
  if (option == SB_OPT_CRS_PAIRS) {
    ctx->crs_pairs = value ? 1 : 0;
    return SB_OK;
  }
  if (option == SB_OPT_CRS_CPW) {
    if (value != 0 && value != 6 && value != 12) return fail(SB_ERR_INVALID, "SB_OPT_CRS_CPW must be 0, 6 or 12");
    ctx->crs_cpw = value;
    return SB_OK;
  }
  return fail(SB_ERR_INVALID, "unknown option");
}

extern "C" int sb_ctx_get_option(const sb_ctx* ctx, int option, int* value) {
  if (!ctx || !value) return fail(SB_ERR_INVALID, "null argument");
  if (option == SB_OPT_PERSISTENT) {
    *value = ctx->persistent;
    return SB_OK;
  }
  if (option == SB_OPT_STREAM_ROWS) {
    *value = ctx->stream_rows;
    return SB_OK;
  }
  if (option == SB_OPT_DM_AZIMUTH) {
    *value = ctx->dm_azimuth_mdeg;
    return SB_OK;
  }
  if (option == SB_OPT_RESTORE_FP32) {
    *value = ctx->restore_fp32;
    return SB_OK;
  }
  if (option == SB_OPT_DE_LDS) {
    *value = ctx->de_lds;
    return SB_OK;
  }
  if (option == SB_OPT_DE_WAVES) {
    *value = ctx->de_waves;
    return SB_OK;
  }
  if (option == SB_OPT_DE_PAIRS) {
    *value = ctx->de_pairs;
    return SB_OK;
  }
  if (option == SB_OPT_CRS_LDS) {
    *value = ctx->crs_lds;
    return SB_OK;
  }
  if (option == SB_OPT_CRS_PAIRS) {
    *value = ctx->crs_pairs;
    return SB_OK;
  }
  if (option == SB_OPT_CRS_CPW) {
    *value = ctx->crs_cpw;
    return SB_OK;
  }
  return fail(SB_ERR_INVALID, "unknown option");
}

// -------------------------------------------------------------- graphs
struct sb_graph {
  sb_ctx* ctx = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  void* work = nullptr;  // the graph's copy of the work table
  // a 16-bit store's restored FP32 copy (and rebased CRS-DE slots), which
  // every replay rewrites (SB_OPT_RESTORE_FP32)
  void* rest_mem = nullptr;
  void* rest_slots = nullptr;
  size_t rest_mem_bytes = 0, rest_slots_bytes = 0;
};

extern "C" void sb_graph_destroy(sb_graph* gr) {
  if (!gr) return;
  (void)hipSetDevice(gr->ctx->device);
  (void)hipStreamSynchronize(gr->ctx->stream);
  if (gr->exec) (void)hipGraphExecDestroy(gr->exec);
  if (gr->graph) (void)hipGraphDestroy(gr->graph);
  ctx_free(gr->ctx, gr->work);
  ctx_free(gr->ctx, gr->rest_mem);
  ctx_free(gr->ctx, gr->rest_slots);
  delete gr;
}

extern "C" int sb_graph_create(sb_ctx* ctx, const sb_gather* g, int kind, const void* params, int cb, int ce,
                               sb_result* r, sb_graph** out) {
  if (!ctx || !g || !params || !r || !out) return fail(SB_ERR_INVALID, "sb_graph_create: null argument");
  *out = nullptr;
  auto run = [&]() -> int {
    if (kind == SB_METHOD_CMP) return sb_cmp_run_range(ctx, g, (const sb_cmp_params*)params, cb, ce, r);
    if (kind == SB_METHOD_CRS) return sb_crs_run_range(ctx, g, (const sb_crs_params*)params, cb, ce, r);
    if (kind == SB_GRAPH_CRS_DE) return sb_crs_de_run_range(ctx, g, (const sb_de_params*)params, cb, ce, r);
    return fail(SB_ERR_INVALID, "sb_graph_create: bad kind");
  };
  // 1. eager run: validates the arguments and builds the work table (and
  // keeps a 16-bit store's restored copy for the graph).
  ctx->restore_keep = true;
  ctx->kept_mem = ctx->kept_slots = nullptr;
  int rc = run();
  ctx->restore_keep = false;
  std::unique_ptr<sb_graph, void (*)(sb_graph*)> G(new sb_graph, sb_graph_destroy);
  G->ctx = ctx;
  G->rest_mem = ctx->kept_mem;
  G->rest_slots = ctx->kept_slots;
  G->rest_mem_bytes = ctx->kept_mem_bytes;
  G->rest_slots_bytes = ctx->kept_slots_bytes;
  ctx->kept_mem = ctx->kept_slots = nullptr;
  if (rc) return rc;
  HIPR(hipSetDevice(ctx->device));
  HIPR(hipStreamSynchronize(ctx->stream));
  const size_t bytes = std::max<size_t>((size_t)ctx->work_n, 1) * sizeof(WorkItem);
  if ((rc = ctx_alloc(ctx, bytes, &G->work))) return rc;
  HIPR(hipMemcpy(G->work, ctx->work_buf, bytes, hipMemcpyDeviceToDevice));
  // 2. capture the same calls: the table is cached (same key), so nothing
  // but kernel launches reaches the stream.
  ctx->graph_work = G->work;
  ctx->graph_rest_mem = G->rest_mem;
  ctx->graph_rest_slots = G->rest_slots;
  ctx->graph_rest_mem_bytes = G->rest_mem_bytes;
  ctx->graph_rest_slots_bytes = G->rest_slots_bytes;
  ctx->capturing = true;
  hipError_t e = hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal);
  if (e == hipSuccess) {
    rc = run();
    hipGraph_t graph = nullptr;
    const hipError_t e2 = hipStreamEndCapture(ctx->stream, &graph);
    G->graph = graph;
    if (e2 != hipSuccess && !rc) rc = hip_fail(e2, "graph capture");
  } else {
    rc = hip_fail(e, "graph capture");
  }
  ctx->capturing = false;
  ctx->graph_work = nullptr;
  ctx->graph_rest_mem = ctx->graph_rest_slots = nullptr;
  ctx->graph_rest_mem_bytes = ctx->graph_rest_slots_bytes = 0;
  ctx->timing_pending = false;
  if (rc) return rc;
  HIPR(hipGraphInstantiate(&G->exec, G->graph, nullptr, nullptr, 0));
  *out = G.release();
  return SB_OK;
}

extern "C" int sb_graph_launch(sb_graph* gr) {
  if (!gr) return fail(SB_ERR_INVALID, "null graph");
  sb_ctx* ctx = gr->ctx;
  HIPR(hipSetDevice(ctx->device));
  std::lock_guard<std::mutex> lk(ctx->run_mu);
  HIPR(hipEventRecord(ctx->ev[0], ctx->stream));
  HIPR(hipGraphLaunch(gr->exec, ctx->stream));
  HIPR(hipEventRecord(ctx->ev[3], ctx->stream));
  ctx->timing_pending = true;
  ctx->timing_graph = true;
  return SB_OK;
}

// ------------------------------------------------------ streamed search
extern "C" int sb_search_streamed(sb_ctx* ctx, int kind, const void* params, int ncdp, const int32_t* cdp_off, int ns,
                                  int dt_us, const sb_trace_geom* geom, const float* samples, int store,
                                  const int32_t* nbr_off, const int32_t* nbr, sb_host_result* out, sb_gather** keep) {
  if (!ctx || !params || !out || !cdp_off || ncdp <= 0 || ns <= 0)
    return fail(SB_ERR_INVALID, "sb_search_streamed: bad arguments");
  if (keep) *keep = nullptr;
  if (kind != SB_METHOD_CMP && kind != SB_METHOD_CRS && kind != SB_GRAPH_CRS_DE)
    return fail(SB_ERR_INVALID, "sb_search_streamed: bad kind");
  if (kind != SB_METHOD_CMP && (!nbr_off || !nbr))
    return fail(SB_ERR_INVALID, "sb_search_streamed: CRS needs neighbour lists");
  HIPR(hipSetDevice(ctx->device));
  const auto t_call = std::chrono::steady_clock::now();
  hipStream_t cs = nullptr;
  HIPR(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
  std::vector<hipEvent_t> evs;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> spans;  // device span of each range search
  double tuples = 0;
  sb_result* r = nullptr;
  int rc = SB_OK;
  {
    GatherBuild b;
    // Batches of about 32 MiB of FP32 samples: small enough that the first
    // search starts early, large enough to keep launches few.
    const int64_t rows = ctx->stream_rows > 0 ? ctx->stream_rows : std::max<int64_t>(1, (int64_t)(32 << 20) / ((int64_t)ns * 4));
    rc = b.begin(ctx, ncdp, cdp_off, ns, dt_us, geom, samples, store, kind == SB_METHOD_CMP ? nullptr : nbr_off,
                 kind == SB_METHOD_CMP ? nullptr : nbr, cs, rows);
    if (!rc) rc = sb_result_create(ctx, ncdp, ns, &r);
    int cb = 0, searched = 0;
    while (!rc && cb < ncdp) {
      int ce = cb + 1;
      while (ce < ncdp && (int64_t)cdp_off[ce + 1] - cdp_off[cb] <= rows) ce++;
      hipEvent_t ev = nullptr;
      if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) {
        rc = fail(SB_ERR_HIP, "event");
        break;
      }
      evs.push_back(ev);
      int ready = 0;
      if ((rc = b.batch(cb, ce, ev, &ready))) break;
      if (ready > searched) {
        // The search of the CDPs that just became complete waits for their
        // batch only; the next batch's copy proceeds meanwhile.
        if (hipStreamWaitEvent(ctx->stream, ev, 0) != hipSuccess) {
          rc = fail(SB_ERR_HIP, "stream wait");
          break;
        }
        sb_result view = *r;  // rows [searched, ready)
        const size_t shift = (size_t)searched * ns;
        view.ncdp = ready - searched;
        view.index += shift;
        view.a += shift;
        view.b += shift;
        view.c += shift;
        view.coher += shift;
        view.stack += shift;
        view.keys += shift;
        const sb_gather* g = b.G.get();
        hipEvent_t e0 = nullptr, e1 = nullptr;
        if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
          if (e0) (void)hipEventDestroy(e0);
          rc = fail(SB_ERR_HIP, "event");
          break;
        }
        spans.push_back({e0, e1});
        (void)hipEventRecord(e0, ctx->stream);  // fires once the batch's copy has landed
        if (kind == SB_METHOD_CMP)
          rc = sb_cmp_run_range(ctx, g, (const sb_cmp_params*)params, searched, ready, &view);
        else if (kind == SB_METHOD_CRS)
          rc = sb_crs_run_range(ctx, g, (const sb_crs_params*)params, searched, ready, &view);
        else
          rc = sb_crs_de_run_range(ctx, g, (const sb_de_params*)params, searched, ready, &view);
        (void)hipEventRecord(e1, ctx->stream);
        if (!rc) tuples += ctx->last.tuples;
        searched = ready;
      }
      cb = ce;
    }
    if (!rc && searched != ncdp) rc = fail(SB_ERR_STATE, "streamed search left CDPs unsearched");
    if (!rc) rc = sb_result_download(ctx, r, out);
    sb_gather* g = nullptr;
    if (!rc) rc = b.finish(&g);
    if (!rc && keep)
      *keep = g;
    else if (g)
      sb_gather_destroy(g);
  }
  if (r) sb_result_destroy(r);
  (void)hipStreamSynchronize(ctx->stream);
  // sb_last_timing after a streamed search: the summed device time of its
  // range searches (key clear, search, winner re-evaluation; the uploads
  // overlap them) and the whole call.
  double dev_ms = 0;
  for (auto& sp : spans) {
    float ms = 0;
    if (!rc && hipEventElapsedTime(&ms, sp.first, sp.second) == hipSuccess) dev_ms += ms;
    (void)hipEventDestroy(sp.first);
    (void)hipEventDestroy(sp.second);
  }
  if (!rc) {
    ctx->last = sb_timing{};
    ctx->last.search_ms = dev_ms;
    ctx->last.total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_call).count();
    ctx->last.tuples = tuples;
    ctx->last.variant = -1;
    ctx->timing_pending = false;
  }
  for (hipEvent_t ev : evs) (void)hipEventDestroy(ev);
  (void)hipStreamDestroy(cs);
  return rc;
}
