// Resident request loop for interactive velocity analysis.
//
// Continuous batching on one GPU:
//   - sb_server_submit queues a request and returns a ticket at once;
//   - kSlots worker threads, each owning an sb_ctx (its own HIP stream), take
//     requests in arrival order as soon as a slot is idle and the request's
//     device footprint fits the remaining HBM budget, so requests that arrive
//     while others run are admitted as memory frees up and their kernels
//     overlap on the device;
//   - a slot merges the waiting requests that share its request's shape
//     (method, ns, dt, parameters, store) into one gather and one launch, up
//     to max_batch CDPs, and splits the results back per request;
//   - sb_server_wait blocks for one ticket and copies its results out;
//   - with an admission deadline (sb_server_set_deadline) a reaper thread
//     fails requests that waited longer than it without being admitted
//     (load shedding: under overload the admitted requests keep a bounded
//     latency instead of every request queueing behind the backlog).
// Lines registered with sb_server_add_gather stay on the device; requests
// that name one and a CDP range skip the upload, and the requests merged
// into one batch run as range searches on the slot's stream into one result
// buffer, downloaded once.  Device blocks come from each slot ctx's caching
// allocator, so steady-state requests do no hipMalloc/hipFree.
// A request's results do not depend on what it was batched with: every
// evaluation's arithmetic is fixed (no shape-dependent contraction) and the
// arg-max key is exact.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "runtime.hpp"
#include "sb_internal.hpp"

using namespace sb;

namespace {

constexpr int kSlots = 4;  // matches the 4 hardware queues a process gets by default
// Slot 0 only takes small requests (at most kSmallEvals evaluations, about
// a quarter of a millisecond of GPU time, e.g. one 1-CDP panel), so
// interactive panels never queue behind a longer search.
constexpr double kSmallEvals = 5e8;
// Slot 1 takes the oldest small request instead once one has waited this
// long (slot 0 is behind), otherwise it serves like slots 2 and 3.
constexpr double kSmallHelpMs = 4.0;
// Device bytes each slot keeps cached between requests (outside the budget):
// enough for a few restored copies of the largest classes, so that 16-bit
// requests reuse them instead of freeing (hipFree waits for the device).
constexpr size_t kSlotCache = size_t(2) << 30;
// Pinned download staging per slot at start: 64 CDPs x 5461 samples of
// results (24 bytes each).
constexpr size_t kPinStart = size_t(8) << 20;

double now_ms(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

// A line kept on the device for range requests.
struct Resident {
  sb_gather* g = nullptr;
  size_t bytes = 0;
  int users = 0;  // requests queued or running on it
  bool removing = false;
};

struct Req {
  uint64_t ticket = 0;
  sb_request r{};
  Resident* res = nullptr;  // resident gather, or null when the request carries traces
  int ntr = 0;
  size_t bytes = 0;  // device footprint estimate
  double evals = 0;  // CDP x t0 x parameter x trace evaluations
  double t_submit = 0, t_admit = 0, t_done = 0;
  int status = -1;  // -1 pending/running, else final status
  bool running = false;
  std::vector<int32_t> index;
  std::vector<float> a, b, c, coher, stack;
};

// Same launch shape: requests with equal keys can share one gather and launch.
bool same_shape(const sb_request& x, const sb_request& y) {
  if (x.method != y.method || x.ns != y.ns || x.dt_us != y.dt_us || x.store != y.store) return false;
  if (x.gather_id != y.gather_id) return false;
  if (x.method == SB_METHOD_CMP)
    return x.cmp.c0 == y.cmp.c0 && x.cmp.c1 == y.cmp.c1 && x.cmp.nc == y.cmp.nc && x.cmp.tau == y.cmp.tau &&
           x.cmp.mode == y.cmp.mode;
  const sb_crs_params &p = x.crs, &q = y.crs;
  return p.a0 == q.a0 && p.a1 == q.a1 && p.b0 == q.b0 && p.b1 == q.b1 && p.c0 == q.c0 && p.c1 == q.c1 &&
         p.na == q.na && p.nb == q.nb && p.nc == q.nc && p.tau == q.tau && p.mode == q.mode;
}

// Device bytes a request needs: what sb_gather_create and sb_result_create allocate.
size_t footprint(const sb_request& r, int ntr) {
  const int64_t stride = ((int64_t)r.ns + 8 + 63) / 64 * 64;
  const size_t esz = r.store == SB_STORE_F32 ? 4 : 2;
  size_t b = (size_t)ntr * stride * esz + (size_t)ntr * (4 + 4 + 24 + 32) + (size_t)(r.ncdp + 1) * 4;
  // transient upload staging (chunks of <= 64 MiB of samples) and geometry
  b += std::min<size_t>((size_t)ntr * r.ns * 4, (size_t)(64 << 20) + (size_t)r.ns * 4) + (size_t)ntr * 24;
  if (r.method == SB_METHOD_CRS && r.nbr_off) {
    int64_t slots = 0;
    for (int i = 0; i < r.ncdp; i++)
      for (int q = r.nbr_off[i]; q < r.nbr_off[i + 1]; q++) slots += r.cdp_off[r.nbr[q] + 1] - r.cdp_off[r.nbr[q]];
    b += (size_t)slots * (32 + 32 + 32) + (size_t)(r.ncdp + 1) * 8 + (size_t)r.nbr_off[r.ncdp] * 4;
  }
  b += (size_t)r.ncdp * r.ns * 32 + (1 << 20);  // results + work table
  return b;
}

// A 16-bit store's search restores the rows it reads to FP32 in a transient
// copy (SB_OPT_RESTORE_FP32, on for the slot contexts): nrows rows, in the
// allocator's block class.
size_t restore_bytes(int store, int ns, int64_t nrows) {
  if (store == SB_STORE_F32) return 0;
  const int64_t stride = ((int64_t)ns + 8 + 63) / 64 * 64;
  return transient_block_size((size_t)nrows * stride * 4 + 256);
}

// Range request on a resident gather: the results, the work table and a
// 16-bit store's restored rows (CMP: the range's; CRS: its supergathers').
size_t footprint_range(const sb_request& r, const sb_gather* g) {
  const int cb = r.cdp_begin, ce = r.cdp_end;
  int64_t rows = 0;
  if (r.method == SB_METHOD_CMP) {
    rows = g->cdp_off[ce] - g->cdp_off[cb];
  } else {
    int64_t lo = g->crs_row_lo[cb], hi = g->crs_row_hi[cb];
    for (int i = cb + 1; i < ce; i++) {
      lo = std::min<int64_t>(lo, g->crs_row_lo[i]);
      hi = std::max<int64_t>(hi, g->crs_row_hi[i]);
    }
    rows = hi - lo;
  }
  return (size_t)(ce - cb) * r.ns * 32 + (1 << 20) + restore_bytes(g->store, g->ns, rows);
}

double evaluations_range(const sb_request& r, const sb_gather* g) {
  const int cb = r.cdp_begin, ce = r.cdp_end;
  if (r.method == SB_METHOD_CMP) return (double)(g->cdp_off[ce] - g->cdp_off[cb]) * r.ns * r.cmp.nc;
  return (double)(g->crs_slot_off[ce] - g->crs_slot_off[cb]) * r.ns * (double)r.crs.na * r.crs.nb * r.crs.nc;
}

double evaluations(const sb_request& r) {
  double traces = 0;
  if (r.method == SB_METHOD_CMP) {
    traces = r.cdp_off[r.ncdp];
    return traces * r.ns * r.cmp.nc;
  }
  for (int i = 0; i < r.ncdp; i++)
    for (int q = r.nbr_off[i]; q < r.nbr_off[i + 1]; q++) traces += r.cdp_off[r.nbr[q] + 1] - r.cdp_off[r.nbr[q]];
  return traces * r.ns * (double)r.crs.na * r.crs.nb * r.crs.nc;
}

}  // namespace

struct sb_server {
  int device = 0;
  size_t budget = 0, reserved = 0;
  int max_batch = 64;
  std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<std::shared_ptr<Req>> pending;
  std::map<uint64_t, std::shared_ptr<Req>> all;
  uint64_t next_ticket = 1, done = 0, launches = 0;
  bool stop = false;
  double deadline_ms = 0;  // 0: no admission deadline
  std::condition_variable cv_reap;
  std::vector<std::thread> workers;
  std::vector<sb_ctx*> ctx;
  sb_ctx* res_ctx = nullptr;  // owns the resident gathers
  std::vector<void*> pin;     // per slot: pinned download staging
  std::vector<size_t> pin_cap;
  // Outgrown staging buffers, freed with the server: hipHostFree waits for
  // the device, which would stall the slot behind the other slots' searches
  // for milliseconds.
  std::vector<void*> pin_retired;
  std::mutex pin_mu;  // guards pin_retired

  // Grows slot `slot`'s staging to at least `bytes` (at least doubling).
  int grow_pin(int slot, size_t bytes) {
    size_t cap = std::max(bytes, 2 * pin_cap[slot]);
    cap = (cap + (size_t(1) << 20) - 1) & ~((size_t(1) << 20) - 1);
    void* p = nullptr;
    if (hipHostMalloc(&p, cap, hipHostMallocDefault) != hipSuccess)
      return fail(SB_ERR_HIP, "pinned staging allocation failed");
    if (pin[slot]) {
      std::lock_guard<std::mutex> lk(pin_mu);
      pin_retired.push_back(pin[slot]);
    }
    pin[slot] = p;
    pin_cap[slot] = cap;
    return SB_OK;
  }

  // Downloads n = rows * ns results of `res` through the slot's pinned
  // buffer (a true DMA; pageable copies serialise through the runtime's
  // staging) and hands out the six arrays, valid until the slot's next
  // download.
  int download(int slot, const sb_result* res, size_t n, sb_host_result* h) {
    const size_t bytes = n * 24;
    if (bytes > pin_cap[slot]) {
      const int rc = grow_pin(slot, bytes);
      if (rc) return rc;
    }
    char* p = (char*)pin[slot];
    *h = sb_host_result{(int32_t*)p, (float*)(p + n * 4), (float*)(p + n * 8), (float*)(p + n * 12),
                        (float*)(p + n * 16), (float*)(p + n * 20)};
    return sb_result_download(ctx[slot], res, h);
  }
  std::map<int, Resident> resident;
  int next_gid = 1;

  void worker(int slot);
  void reaper();
  int process(int slot, std::vector<std::shared_ptr<Req>>& batch);
  int process_range(int slot, std::vector<std::shared_ptr<Req>>& batch);
};

// Batch of range requests on one resident gather: one result buffer, one
// range search per request into its rows, one download.
int sb_server::process_range(int slot, std::vector<std::shared_ptr<Req>>& batch) {
  sb_ctx* c = ctx[slot];
  const sb_request& r0 = batch[0]->r;
  const sb_gather* g = batch[0]->res->g;
  int rows = 0;
  for (auto& q : batch) rows += q->r.cdp_end - q->r.cdp_begin;
  sb_result* res = nullptr;
  int rc = sb_result_create(c, rows, r0.ns, &res);
  if (rc) return rc;
  int o = 0;
  for (auto& q : batch) {
    const int cb = q->r.cdp_begin, ce = q->r.cdp_end;
    sb_result view = *res;  // rows [o, o + ce - cb) of res
    const size_t shift = (size_t)o * r0.ns;
    view.ncdp = ce - cb;
    view.index += shift;
    view.a += shift;
    view.b += shift;
    view.c += shift;
    view.coher += shift;
    view.stack += shift;
    view.keys += shift;
    rc = r0.method == SB_METHOD_CMP ? sb_cmp_run_range(c, g, &q->r.cmp, cb, ce, &view)
                                    : sb_crs_run_range(c, g, &q->r.crs, cb, ce, &view);
    if (rc) break;
    o += ce - cb;
  }
  if (!rc) {
    sb_host_result h;
    rc = download(slot, res, (size_t)rows * r0.ns, &h);
    size_t k = 0;
    for (auto& q : batch) {
      const size_t m = (size_t)(q->r.cdp_end - q->r.cdp_begin) * r0.ns;
      q->index.assign(h.index + k, h.index + k + m);
      q->a.assign(h.a + k, h.a + k + m);
      q->b.assign(h.b + k, h.b + k + m);
      q->c.assign(h.c + k, h.c + k + m);
      q->coher.assign(h.coher + k, h.coher + k + m);
      q->stack.assign(h.stack + k, h.stack + k + m);
      k += m;
    }
  }
  sb_result_destroy(res);
  return rc;
}

int sb_server::process(int slot, std::vector<std::shared_ptr<Req>>& batch) {
  if (batch[0]->res) return process_range(slot, batch);
  sb_ctx* c = ctx[slot];
  const sb_request& r0 = batch[0]->r;
  // Concatenate the batch into one binned gather (CRS neighbour indices shifted).
  int ncdp = 0, ntr = 0;
  for (auto& q : batch) {
    ncdp += q->r.ncdp;
    ntr += q->ntr;
  }
  const bool merged = batch.size() > 1;
  std::vector<int32_t> off, nbr_off, nbr;
  std::vector<sb_trace_geom> geom;
  std::vector<float> samples;
  const int32_t* p_off = r0.cdp_off;
  const sb_trace_geom* p_geom = r0.geom;
  const float* p_samples = r0.samples;
  const int32_t* p_nbr_off = r0.nbr_off;
  const int32_t* p_nbr = r0.nbr;
  if (merged) {
    off.reserve(ncdp + 1);
    off.push_back(0);
    geom.reserve(ntr);
    samples.reserve((size_t)ntr * r0.ns);
    if (r0.method == SB_METHOD_CRS) nbr_off.push_back(0);
    int cdp_base = 0;
    for (auto& q : batch) {
      const sb_request& r = q->r;
      for (int i = 0; i < r.ncdp; i++) off.push_back(off.back() + (r.cdp_off[i + 1] - r.cdp_off[i]));
      geom.insert(geom.end(), r.geom, r.geom + q->ntr);
      samples.insert(samples.end(), r.samples, r.samples + (size_t)q->ntr * r.ns);
      if (r.method == SB_METHOD_CRS) {
        for (int i = 0; i < r.ncdp; i++) {
          for (int k = r.nbr_off[i]; k < r.nbr_off[i + 1]; k++) nbr.push_back(r.nbr[k] + cdp_base);
          nbr_off.push_back((int32_t)nbr.size());
        }
      }
      cdp_base += r.ncdp;
    }
    p_off = off.data();
    p_geom = geom.data();
    p_samples = samples.data();
    p_nbr_off = r0.method == SB_METHOD_CRS ? nbr_off.data() : nullptr;
    p_nbr = r0.method == SB_METHOD_CRS ? nbr.data() : nullptr;
  }
  sb_gather* g = nullptr;
  int rc = sb_gather_create(c, ncdp, p_off, r0.ns, r0.dt_us, p_geom, p_samples, r0.store,
                            r0.method == SB_METHOD_CRS ? p_nbr_off : nullptr,
                            r0.method == SB_METHOD_CRS ? p_nbr : nullptr, &g);
  if (rc) return rc;
  sb_result* res = nullptr;
  rc = sb_result_create(c, ncdp, r0.ns, &res);
  if (!rc) rc = r0.method == SB_METHOD_CMP ? sb_cmp_run(c, g, &r0.cmp, res) : sb_crs_run(c, g, &r0.crs, res);
  if (!rc) {
    sb_host_result h;
    rc = download(slot, res, (size_t)ncdp * r0.ns, &h);
    size_t o = 0;
    for (auto& q : batch) {
      const size_t m = (size_t)q->r.ncdp * r0.ns;
      q->index.assign(h.index + o, h.index + o + m);
      q->a.assign(h.a + o, h.a + o + m);
      q->b.assign(h.b + o, h.b + o + m);
      q->c.assign(h.c + o, h.c + o + m);
      q->coher.assign(h.coher + o, h.coher + o + m);
      q->stack.assign(h.stack + o, h.stack + o + m);
      o += m;
    }
  }
  if (res) sb_result_destroy(res);
  sb_gather_destroy(g);
  return rc;
}

// Fails the pending requests that waited past the admission deadline
// (under the server lock).
void sb_server::reaper() {
  std::unique_lock<std::mutex> lk(mu);
  while (!stop) {
    if (deadline_ms <= 0) {
      cv_reap.wait(lk);
      continue;
    }
    const double now = now_ms(t0);
    bool any = false;
    for (auto it = pending.begin(); it != pending.end();) {
      auto& q = *it;
      if (now - q->t_submit > deadline_ms) {
        q->status = SB_ERR_DEADLINE;
        q->t_admit = q->t_done = now;
        if (q->res) q->res->users--;
        done++;
        any = true;
        it = pending.erase(it);
      } else {
        ++it;
      }
    }
    if (any) cv_done.notify_all();
    cv_reap.wait_for(lk, std::chrono::duration<double, std::milli>(std::max(0.25, deadline_ms / 8)));
  }
}

void sb_server::worker(int slot) {
  (void)hipSetDevice(device);
  std::unique_lock<std::mutex> lk(mu);
  while (true) {
    std::shared_ptr<Req> head;
    cv_work.wait(lk, [&] {
      if (stop) return true;
      // Slots 1..: the oldest request that fits the free budget.  Slot 0:
      // the smallest small request that fits (shortest job first, so 1-CDP
      // panels do not queue behind larger small ones), unless one waited a
      // second.  Slot 1 first: the oldest small request that fits once it
      // has waited kSmallHelpMs.  A request alone larger than the whole
      // budget is rejected.
      const double now = now_ms(t0);
      if (slot == 1) {
        for (auto& q : pending) {
          const bool fits = q->bytes > budget || reserved + q->bytes <= budget;
          if (!fits && now - q->t_submit > 1000.0) break;  // nothing overtakes it (as below)
          if (q->evals > kSmallEvals || !fits) continue;
          if (now - q->t_submit <= kSmallHelpMs) break;  // the rest are newer
          head = q;
          return true;
        }
      }
      std::shared_ptr<Req> best;
      for (auto& q : pending) {
        if (slot == 0 && q->evals > kSmallEvals) continue;
        const bool old = now - q->t_submit > 1000.0;
        if (q->bytes > budget || reserved + q->bytes <= budget) {
          if (slot != 0 || old) {
            head = q;
            return true;
          }
          if (!best || q->evals < best->evals) best = q;
        } else if (old) {
          break;  // do not let newer requests overtake one that waited a second
        }
      }
      head = best;
      return head != nullptr;
    });
    if (stop && !head) return;
    if (!head) continue;
    pending.erase(std::find(pending.begin(), pending.end(), head));
    if (head->bytes > budget) {
      head->status = SB_ERR_BUDGET;
      head->t_done = now_ms(t0);
      if (head->res) head->res->users--;
      done++;
      cv_done.notify_all();
      continue;
    }
    std::vector<std::shared_ptr<Req>> batch{head};
    size_t bytes = head->bytes;
    int cdps = head->r.ncdp;
    for (auto it = pending.begin(); it != pending.end() && cdps < max_batch;) {
      auto& q = *it;
      if (same_shape(q->r, head->r) && cdps + q->r.ncdp <= max_batch && reserved + bytes + q->bytes <= budget &&
          (slot != 0 || q->evals <= kSmallEvals)) {
        batch.push_back(q);
        bytes += q->bytes;
        cdps += q->r.ncdp;
        it = pending.erase(it);
      } else {
        ++it;
      }
    }
    reserved += bytes;
    const double ta = now_ms(t0);
    for (auto& q : batch) {
      q->t_admit = ta;
      q->running = true;
    }
    launches++;
    lk.unlock();
    const int rc = process(slot, batch);
    ctx_trim(ctx[slot], kSlotCache);
    lk.lock();
    reserved -= bytes;
    const double td = now_ms(t0);
    for (auto& q : batch) {
      q->status = rc;
      q->t_done = td;
      q->running = false;
      if (q->res) q->res->users--;
      done++;
    }
    cv_done.notify_all();
    cv_work.notify_all();
  }
}

extern "C" int sb_server_create(int device, size_t hbm_budget, int max_batch, sb_server** out) {
  if (!out || max_batch < 1) return fail(SB_ERR_INVALID, "sb_server_create: bad arguments");
  *out = nullptr;
  std::unique_ptr<sb_server> s(new sb_server);
  s->device = device;
  s->max_batch = max_batch;
  for (int i = 0; i < kSlots; i++) {
    sb_ctx* c = nullptr;
    int rc = sb_ctx_create(device, hbm_budget, &c);
    if (rc) {
      for (auto* x : s->ctx) sb_ctx_destroy(x);
      return rc;
    }
    s->ctx.push_back(c);
  }
  {
    int rc = sb_ctx_create(device, hbm_budget, &s->res_ctx);
    if (rc) {
      for (auto* x : s->ctx) sb_ctx_destroy(x);
      return rc;
    }
  }
  s->pin.assign(kSlots, nullptr);
  s->pin_cap.assign(kSlots, 0);
  for (int i = 0; i < kSlots; i++) {
    const int rc = s->grow_pin(i, kPinStart);
    if (rc) {
      for (void* p : s->pin)
        if (p) (void)hipHostFree(p);
      for (auto* x : s->ctx) sb_ctx_destroy(x);
      sb_ctx_destroy(s->res_ctx);
      return rc;
    }
  }
  // Every slot ctx may use the whole budget; the server keeps the sum within it.
  size_t b = 0;
  sb_ctx_memory(s->ctx[0], nullptr, &b);
  s->budget = b;
  for (int i = 0; i < kSlots; i++) s->workers.emplace_back([p = s.get(), i] { p->worker(i); });
  s->workers.emplace_back([p = s.get()] { p->reaper(); });
  *out = s.release();
  return SB_OK;
}

extern "C" int sb_server_submit(sb_server* s, const sb_request* req, uint64_t* ticket) {
  if (!s || !req || !ticket) return fail(SB_ERR_INVALID, "sb_server_submit: bad arguments");
  const sb_request& r = *req;
  if (r.method != SB_METHOD_CMP && r.method != SB_METHOD_CRS)
    return fail(SB_ERR_INVALID, "sb_server_submit: bad method");
  if (r.gather_id != 0) {
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->resident.find(r.gather_id);
    if (it == s->resident.end() || it->second.removing)
      return fail(SB_ERR_INVALID, "sb_server_submit: unknown resident gather");
    Resident& R = it->second;
    if (r.cdp_begin < 0 || r.cdp_end > R.g->ncdp || r.cdp_begin >= r.cdp_end)
      return fail(SB_ERR_INVALID, "sb_server_submit: bad CDP range");
    if (r.method == SB_METHOD_CRS && !R.g->has_crs)
      return fail(SB_ERR_STATE, "sb_server_submit: gather has no neighbour lists");
    if (s->stop) return fail(SB_ERR_STATE, "server is shutting down");
    auto q = std::make_shared<Req>();
    q->r = r;
    q->r.ncdp = r.cdp_end - r.cdp_begin;
    q->r.ns = R.g->ns;
    q->r.dt_us = R.g->dt_us;
    q->r.store = R.g->store;
    q->r.cdp_off = nullptr;
    q->r.geom = nullptr;
    q->r.samples = nullptr;
    q->r.nbr_off = q->r.nbr = nullptr;
    q->res = &R;
    q->bytes = footprint_range(q->r, R.g);
    q->evals = evaluations_range(q->r, R.g);
    R.users++;
    q->ticket = s->next_ticket++;
    q->t_submit = now_ms(s->t0);
    s->pending.push_back(q);
    s->all[q->ticket] = q;
    *ticket = q->ticket;
    s->cv_work.notify_all();
    return SB_OK;
  }
  if (r.ncdp <= 0 || r.ns <= 0 || r.dt_us <= 0 || !r.cdp_off || !r.geom || !r.samples ||
      (r.method == SB_METHOD_CRS && (!r.nbr_off || !r.nbr)))
    return fail(SB_ERR_INVALID, "sb_server_submit: bad request");
  if (r.cdp_off[0] != 0) return fail(SB_ERR_INVALID, "cdp_off[0] must be 0");
  for (int i = 0; i < r.ncdp; i++)
    if (r.cdp_off[i + 1] <= r.cdp_off[i]) return fail(SB_ERR_INVALID, "every CDP needs at least one trace");
  auto q = std::make_shared<Req>();
  q->r = r;
  q->ntr = r.cdp_off[r.ncdp];
  q->bytes = footprint(r, q->ntr) + restore_bytes(r.store, r.ns, q->ntr);
  q->evals = evaluations(r);
  std::lock_guard<std::mutex> lk(s->mu);
  if (s->stop) return fail(SB_ERR_STATE, "server is shutting down");
  q->ticket = s->next_ticket++;
  q->t_submit = now_ms(s->t0);
  s->pending.push_back(q);
  s->all[q->ticket] = q;
  *ticket = q->ticket;
  s->cv_work.notify_all();
  return SB_OK;
}

extern "C" int sb_server_wait(sb_server* s, uint64_t ticket, double timeout_ms, sb_host_result* out,
                              sb_request_stats* st) {
  if (!s) return fail(SB_ERR_INVALID, "null server");
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->all.find(ticket);
  if (it == s->all.end()) return fail(SB_ERR_INVALID, "unknown ticket");
  std::shared_ptr<Req> q = it->second;
  auto ready = [&] { return q->status >= 0; };
  if (timeout_ms < 0)
    s->cv_done.wait(lk, ready);
  else if (!s->cv_done.wait_for(lk, std::chrono::duration<double, std::milli>(timeout_ms), ready))
    return fail(SB_ERR_STATE, "timeout");
  s->all.erase(it);
  lk.unlock();
  if (st) *st = sb_request_stats{q->t_submit, q->t_admit, q->t_done, q->status};
  if (q->status != SB_OK) return fail(q->status, "request failed");
  if (out) {
    const size_t n = q->index.size();
    if (out->index) std::memcpy(out->index, q->index.data(), n * 4);
    if (out->a) std::memcpy(out->a, q->a.data(), n * 4);
    if (out->b) std::memcpy(out->b, q->b.data(), n * 4);
    if (out->c) std::memcpy(out->c, q->c.data(), n * 4);
    if (out->coher) std::memcpy(out->coher, q->coher.data(), n * 4);
    if (out->stack) std::memcpy(out->stack, q->stack.data(), n * 4);
  }
  return SB_OK;
}

extern "C" int sb_server_add_gather(sb_server* s, int ncdp, const int32_t* cdp_off, int ns, int dt_us,
                                    const sb_trace_geom* geom, const float* samples, int store,
                                    const int32_t* nbr_off, const int32_t* nbr, int* gather_id) {
  if (!s || !gather_id || ncdp <= 0 || !cdp_off) return fail(SB_ERR_INVALID, "sb_server_add_gather: bad arguments");
  *gather_id = 0;
  sb_request est{};
  est.method = nbr_off ? SB_METHOD_CRS : SB_METHOD_CMP;
  est.ncdp = ncdp;
  est.ns = ns;
  est.store = store;
  est.cdp_off = cdp_off;
  est.nbr_off = nbr_off;
  est.nbr = nbr;
  if (cdp_off[0] != 0) return fail(SB_ERR_INVALID, "cdp_off[0] must be 0");
  for (int i = 0; i < ncdp; i++)
    if (cdp_off[i + 1] <= cdp_off[i]) return fail(SB_ERR_INVALID, "every CDP needs at least one trace");
  if (nbr_off)
    for (int i = 0; i < ncdp; i++)
      for (int k = nbr_off[i]; k < nbr_off[i + 1]; k++)
        if (nbr[k] < 0 || nbr[k] >= ncdp) return fail(SB_ERR_INVALID, "neighbour index out of range");
  const size_t want = footprint(est, cdp_off[ncdp]);
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stop) return fail(SB_ERR_STATE, "server is shutting down");
    if (s->reserved + want > s->budget) return fail(SB_ERR_BUDGET, "resident gather does not fit the budget");
    s->reserved += want;  // held while uploading
  }
  sb_gather* g = nullptr;
  const int rc = sb_gather_create(s->res_ctx, ncdp, cdp_off, ns, dt_us, geom, samples, store, nbr_off, nbr, &g);
  ctx_trim(s->res_ctx, 0);  // upload staging
  std::lock_guard<std::mutex> lk(s->mu);
  s->reserved -= want;
  if (rc) return rc;
  const int id = s->next_gid++;
  Resident& R = s->resident[id];
  R.g = g;
  R.bytes = g->bytes;
  s->reserved += R.bytes;
  *gather_id = id;
  return SB_OK;
}

extern "C" int sb_server_remove_gather(sb_server* s, int gather_id) {
  if (!s) return fail(SB_ERR_INVALID, "null server");
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->resident.find(gather_id);
  if (it == s->resident.end() || it->second.removing) return fail(SB_ERR_INVALID, "unknown resident gather");
  Resident& R = it->second;
  R.removing = true;
  s->cv_done.wait(lk, [&] { return R.users == 0; });
  sb_gather* g = R.g;
  const size_t bytes = R.bytes;
  s->resident.erase(it);
  lk.unlock();
  sb_gather_destroy(g);  // waits for the device
  lk.lock();
  s->reserved -= bytes;
  s->cv_work.notify_all();
  return SB_OK;
}

extern "C" int sb_server_set_deadline(sb_server* s, double max_queue_ms) {
  if (!s || !(max_queue_ms >= 0)) return fail(SB_ERR_INVALID, "sb_server_set_deadline: bad arguments");
  std::lock_guard<std::mutex> lk(s->mu);
  s->deadline_ms = max_queue_ms;
  s->cv_reap.notify_all();
  return SB_OK;
}

extern "C" int sb_server_stats(const sb_server* cs, uint64_t* done, uint64_t* launches, size_t* hbm_used) {
  sb_server* s = const_cast<sb_server*>(cs);
  if (!s) return fail(SB_ERR_INVALID, "null server");
  std::lock_guard<std::mutex> lk(s->mu);
  if (done) *done = s->done;
  if (launches) *launches = s->launches;
  if (hbm_used) *hbm_used = s->reserved;
  return SB_OK;
}

extern "C" void sb_server_destroy(sb_server* s) {
  if (!s) return;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
    // queued requests that never started fail
    for (auto& q : s->pending) {
      q->status = SB_ERR_STATE;
      if (q->res) q->res->users--;
      s->done++;
    }
    s->pending.clear();
  }
  s->cv_work.notify_all();
  s->cv_done.notify_all();
  s->cv_reap.notify_all();
  for (auto& t : s->workers) t.join();
  for (auto& kv : s->resident) sb_gather_destroy(kv.second.g);
  s->resident.clear();
  for (auto* c : s->ctx) sb_ctx_destroy(c);
  sb_ctx_destroy(s->res_ctx);
  for (void* p : s->pin)
    if (p) (void)hipHostFree(p);
  for (void* p : s->pin_retired) (void)hipHostFree(p);
  delete s;
}
