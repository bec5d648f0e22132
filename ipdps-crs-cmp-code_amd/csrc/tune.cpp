// Autotuner of the SB_MODE_FAST launch shape.
//
// The tile of one workgroup is (64 * R t0) x (pchunk parameters) x (all traces
// of the CDP or supergather).  R and the per-pass parameter count CC are
// compile-time shapes (kFastVariants); pchunk is a runtime split of the
// parameter axis across workgroups.  sb_tune times every (variant, pchunk)
// candidate on the caller's own gather with HIP events and keeps the fastest
// on the ctx; results are cached per (method, store, w, ncdp, ns, npar).
#include <hip/hip_runtime_api.h>

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
  const std::string key = std::to_string(method) + "/" + std::to_string(g->store) + "/" + std::to_string(w_key) + "/" +
                          std::to_string(g->ncdp) + "/" + std::to_string(g->ns) + "/" + std::to_string(npar) + "/" +
                          std::to_string(g->ntr) + "/fast";
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
  double best = DBL_MAX;
  int best_v = save_v, best_p = save_p;
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
      double tsum = 0;
      for (int i = 0; i < reps + 1 && !rc; i++) {
        rc = method == SB_METHOD_CMP ? sb_cmp_run(ctx, g, (const sb_cmp_params*)params, r)
                                     : sb_crs_run(ctx, g, (const sb_crs_params*)params, r);
        sb_timing t;
        if (!rc) rc = sb_last_timing(ctx, &t);
        if (!rc && i > 0) tsum += t.search_ms;  // first run is a warm-up
      }
      if (rc) break;
      if (tsum < best) {
        best = tsum;
        best_v = v;
        best_p = pc;
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
