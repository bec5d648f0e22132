// Autotuner of the SB_MODE_FAST launch shape.
//
// The tile of one workgroup is (64 * R t0) x (pchunk parameters) x (all traces
// of the CDP or supergather).  R and the per-pass parameter count CC are
// compile-time shapes (kFastVariants); pchunk is a runtime split of the
// parameter axis across workgroups.  sb_tune times every (variant, pchunk)
// candidate on the caller's own gather with HIP events, then times the three
// fastest again in interleaved rounds (so that a drift of the clock over the
// first pass does not decide between candidates a few percent apart) and
// keeps the one with the lowest median on the ctx; results are cached per
// (method, store, w, ncdp, ns, npar, ntr, SB_OPT_RESTORE_FP32).
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cfloat>
#include <string>
#include <vector>

#include "launch.hpp"
#include "runtime.hpp"
#include "sb_internal.hpp"

using namespace sb;

extern "C" int sb_tune_set(sb_ctx* ctx, int variant, int pchunk) {
  if (!ctx || variant < -1 || variant >= kNumFastVariants || pchunk < 0) return fail(SB_ERR_INVALID, "bad tuning");
  ctx->fast_variant = variant;
  ctx->pchunk_override = pchunk;
  return SB_OK;
}

extern "C" int sb_tune_get(const sb_ctx* ctx, int* variant, int* pchunk, int* nvariants) {
  if (!ctx) return fail(SB_ERR_INVALID, "null ctx");
  if (variant) *variant = ctx->fast_variant;
  if (pchunk) *pchunk = ctx->pchunk_override;
  if (nvariants) *nvariants = kNumFastVariants;
  return SB_OK;
}

extern "C" int sb_tune(sb_ctx* ctx, const sb_gather* g, int method, const void* params, int reps) {
  if (!ctx || !g || !params || reps < 1) return fail(SB_ERR_INVALID, "sb_tune: bad arguments");
  if (method != SB_METHOD_CMP && method != SB_METHOD_CRS) return fail(SB_ERR_INVALID, "sb_tune: bad method");
  int npar, w_key, tau;
  bool fast;  // whether the search runs a fast-mode kernel at all
  if (method == SB_METHOD_CMP) {
    const sb_cmp_params* p = (const sb_cmp_params*)params;
    npar = p->nc;
    w_key = (int)(p->tau * 1e6);
    fast = p->mode == SB_MODE_FAST && p->c0 >= 0 && p->c1 >= 0;
    // the window as sb_cmp_run computes it (float)
    float dt = (float)g->dt_us;
    dt = dt / 1000000.0f;
    const float tt = (float)p->tau * (1.0f / dt);
    tau = (int)tt > 0 ? (int)tt : 0;
  } else {
    const sb_crs_params* p = (const sb_crs_params*)params;
    npar = p->na * p->nb * p->nc;
    w_key = (int)(p->tau * 1e6);
    fast = p->mode == SB_MODE_FAST;
    // the window as sb_crs_run computes it (double)
    double dt = (double)g->dt_us;
    dt = dt / 1000000.0;
    const double tt = p->tau * (1.0 / dt);
    tau = (int)tt > 0 ? (int)tt : 0;
  }
  if (npar <= 0) return fail(SB_ERR_INVALID, "sb_tune: empty parameter grid");
  // Searches that run in reference arithmetic (SB_MODE_EXACT, windows above
  // 7 samples, CMP grids with c < 0) do not use the tuned shape: nothing to
  // time, and nothing cached under a key that fast searches would reuse.
  if (!fast || !fast_supported(2 * tau + 1)) return SB_OK;
  // CRS searches the LDS-staged kernel serves (SB_OPT_CRS_LDS) have no
  // launch shape to choose: their work items are fixed by the grid.
  if (method == SB_METHOD_CRS && ctx->crs_lds != 0 && (g->store == SB_STORE_F32 || ctx->restore_fp32)) {
    const sb_crs_params* p = (const sb_crs_params*)params;
    int smax = 0;
    for (int i = 0; i < g->ncdp; i++) smax = std::max(smax, g->crs_slot_off[i + 1] - g->crs_slot_off[i]);
    const int lb = ctx->crs_lds < 0 ? kCrsLdsDefaultBytes : ctx->crs_lds;
    if (crs_lds_supported(2 * tau + 1, p->nc, ctx->crs_cpw, kCrsLdsWaves, ctx->crs_pairs, lb, g->ns, smax))
      return SB_OK;
  }
  const std::string key = std::to_string(method) + "/" + std::to_string(g->store) + "/" + std::to_string(w_key) + "/" +
                          std::to_string(g->ncdp) + "/" + std::to_string(g->ns) + "/" + std::to_string(npar) + "/" +
                          std::to_string(g->ntr) + "/fast" +
                          (g->store != SB_STORE_F32 && ctx->restore_fp32 ? "/restored" : "");
  auto hit = ctx->tune_cache.find(key);
  if (hit != ctx->tune_cache.end()) {
    ctx->fast_variant = hit->second.first;
    ctx->pchunk_override = hit->second.second;
    return SB_OK;
  }
  sb_result* r = nullptr;
  int rc = sb_result_create(ctx, g->ncdp, g->ns, &r);
  if (rc) return rc;
  const int save_v = ctx->fast_variant, save_p = ctx->pchunk_override;
  // One timed search of the current shape.
  auto time_once = [&](double* ms) {
    int e = method == SB_METHOD_CMP ? sb_cmp_run(ctx, g, (const sb_cmp_params*)params, r)
                                    : sb_crs_run(ctx, g, (const sb_crs_params*)params, r);
    sb_timing t;
    if (!e) e = sb_last_timing(ctx, &t);
    if (!e) *ms = t.search_ms;
    return e;
  };
  struct Cand {
    double first;  // summed time of the first pass
    int v, pc;
  };
  std::vector<Cand> pool;
  for (int v = 0; v < kNumFastVariants && !rc; v++) {
    const FastVariant& fv = kFastVariants[v];
    const bool lds = fv.lds_kb > 0;
    if (lds && method != SB_METHOD_CMP) continue;  // CMP only
    // Global-memory shapes: automatic split or fractions of the grid.
    // LDS-staged shapes: one, two or four groups of CC parameters per wave.
    const int cands[] = {0, npar, npar / 2, npar / 4, npar / 8, npar / 16};
    const int lds_cands[] = {0, 2 * fv.waves * fv.CC, 4 * fv.waves * fv.CC};
    for (int pc : lds ? std::vector<int>(lds_cands, lds_cands + 3) : std::vector<int>(cands, cands + 6)) {
      if (pc != 0 && pc < 4 * fv.CC) continue;
      ctx->fast_variant = v;
      ctx->pchunk_override = pc;
      double tsum = 0, ms = 0;
      for (int i = 0; i < reps + 1 && !rc; i++) {
        rc = time_once(&ms);
        if (!rc && i > 0) tsum += ms;  // first run is a warm-up
      }
      if (rc) break;
      pool.push_back(Cand{tsum, v, pc});
    }
  }
  // The three fastest of the first pass, re-timed round by round (one run
  // of each per round), ranked by their medians.
  int best_v = save_v, best_p = save_p;
  if (!rc && !pool.empty()) {
    std::stable_sort(pool.begin(), pool.end(), [](const Cand& x, const Cand& y) { return x.first < y.first; });
    const size_t nf = std::min<size_t>(3, pool.size());
    const int rounds = 2 * reps + 1;
    std::vector<std::vector<double>> ts(nf);
    for (int i = 0; i < rounds && !rc; i++)
      for (size_t f = 0; f < nf && !rc; f++) {
        ctx->fast_variant = pool[f].v;
        ctx->pchunk_override = pool[f].pc;
        double ms = 0;
        rc = time_once(&ms);
        ts[f].push_back(ms);
      }
    double best = DBL_MAX;
    for (size_t f = 0; f < nf && !rc; f++) {
      std::vector<double>& t = ts[f];
      std::nth_element(t.begin(), t.begin() + t.size() / 2, t.end());
      const double med = t[t.size() / 2];
      if (med < best) {
        best = med;
        best_v = pool[f].v;
        best_p = pool[f].pc;
      }
    }
  }
  sb_result_destroy(r);
  if (rc) {
    ctx->fast_variant = save_v;
    ctx->pchunk_override = save_p;
    return rc;
  }
  ctx->fast_variant = best_v;
  ctx->pchunk_override = best_p;
  ctx->tune_cache[key] = {best_v, best_p};
  return SB_OK;
}
