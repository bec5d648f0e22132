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
//   - sb_server_wait blocks for one ticket and copies its results out.
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

double now_ms(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

struct Req {
  uint64_t ticket = 0;
  sb_request r{};
  int ntr = 0;
  size_t bytes = 0;  // device footprint estimate
  double t_submit = 0, t_admit = 0, t_done = 0;
  int status = -1;  // -1 pending/running, else final status
  bool running = false;
  std::vector<int32_t> index;
  std::vector<float> a, b, c, coher, stack;
};

// Same launch shape: requests with equal keys can share one gather and launch.
bool same_shape(const sb_request& x, const sb_request& y) {
  if (x.method != y.method || x.ns != y.ns || x.dt_us != y.dt_us || x.store != y.store) return false;
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
  std::vector<std::thread> workers;
  std::vector<sb_ctx*> ctx;

  void worker(int slot);
  int process(int slot, std::vector<std::shared_ptr<Req>>& batch);
};

int sb_server::process(int slot, std::vector<std::shared_ptr<Req>>& batch) {
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
    const size_t n = (size_t)ncdp * r0.ns;
    std::vector<int32_t> idx(n);
    std::vector<float> a(n), b(n), cc(n), coh(n), stk(n);
    sb_host_result h{idx.data(), a.data(), b.data(), cc.data(), coh.data(), stk.data()};
    rc = sb_result_download(c, res, &h);
    size_t o = 0;
    for (auto& q : batch) {
      const size_t m = (size_t)q->r.ncdp * r0.ns;
      q->index.assign(idx.begin() + o, idx.begin() + o + m);
      q->a.assign(a.begin() + o, a.begin() + o + m);
      q->b.assign(b.begin() + o, b.begin() + o + m);
      q->c.assign(cc.begin() + o, cc.begin() + o + m);
      q->coher.assign(coh.begin() + o, coh.begin() + o + m);
      q->stack.assign(stk.begin() + o, stk.begin() + o + m);
      o += m;
    }
  }
  if (res) sb_result_destroy(res);
  sb_gather_destroy(g);
  return rc;
}

void sb_server::worker(int slot) {
  (void)hipSetDevice(device);
  std::unique_lock<std::mutex> lk(mu);
  while (true) {
    std::shared_ptr<Req> head;
    cv_work.wait(lk, [&] {
      if (stop) return true;
      // Oldest request that fits the free budget; a request alone larger
      // than the whole budget is rejected.
      for (auto& q : pending) {
        if (q->bytes > budget || reserved + q->bytes <= budget) {
          head = q;
          return true;
        }
        // do not let small requests overtake one that waited a second
        if (now_ms(t0) - q->t_submit > 1000.0) return false;
      }
      return false;
    });
    if (stop && !head) return;
    if (!head) continue;
    pending.erase(std::find(pending.begin(), pending.end(), head));
    if (head->bytes > budget) {
      head->status = SB_ERR_BUDGET;
      head->t_done = now_ms(t0);
      done++;
      cv_done.notify_all();
      continue;
    }
    std::vector<std::shared_ptr<Req>> batch{head};
    size_t bytes = head->bytes;
    int cdps = head->r.ncdp;
    for (auto it = pending.begin(); it != pending.end() && cdps < max_batch;) {
      auto& q = *it;
      if (same_shape(q->r, head->r) && cdps + q->r.ncdp <= max_batch && reserved + bytes + q->bytes <= budget) {
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
    lk.lock();
    reserved -= bytes;
    const double td = now_ms(t0);
    for (auto& q : batch) {
      q->status = rc;
      q->t_done = td;
      q->running = false;
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
  // Every slot ctx may use the whole budget; the server keeps the sum within it.
  size_t b = 0;
  sb_ctx_memory(s->ctx[0], nullptr, &b);
  s->budget = b;
  for (int i = 0; i < kSlots; i++) s->workers.emplace_back([p = s.get(), i] { p->worker(i); });
  *out = s.release();
  return SB_OK;
}

extern "C" int sb_server_submit(sb_server* s, const sb_request* req, uint64_t* ticket) {
  if (!s || !req || !ticket) return fail(SB_ERR_INVALID, "sb_server_submit: bad arguments");
  const sb_request& r = *req;
  if ((r.method != SB_METHOD_CMP && r.method != SB_METHOD_CRS) || r.ncdp <= 0 || r.ns <= 0 || r.dt_us <= 0 ||
      !r.cdp_off || !r.geom || !r.samples || (r.method == SB_METHOD_CRS && (!r.nbr_off || !r.nbr)))
    return fail(SB_ERR_INVALID, "sb_server_submit: bad request");
  if (r.cdp_off[0] != 0) return fail(SB_ERR_INVALID, "cdp_off[0] must be 0");
  for (int i = 0; i < r.ncdp; i++)
    if (r.cdp_off[i + 1] <= r.cdp_off[i]) return fail(SB_ERR_INVALID, "every CDP needs at least one trace");
  auto q = std::make_shared<Req>();
  q->r = r;
  q->ntr = r.cdp_off[r.ncdp];
  q->bytes = footprint(r, q->ntr);
  std::lock_guard<std::mutex> lk(s->mu);
  if (s->stop) return fail(SB_ERR_STATE, "server is shutting down");
  q->ticket = s->next_ticket++;
  q->t_submit = now_ms(s->t0);
  s->pending.push_back(q);
  s->all[q->ticket] = q;
  *ticket = q->ticket;
  s->cv_work.notify_one();
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
      s->done++;
    }
    s->pending.clear();
  }
  s->cv_work.notify_all();
  s->cv_done.notify_all();
  for (auto& t : s->workers) t.join();
  for (auto* c : s->ctx) sb_ctx_destroy(c);
  delete s;
}
