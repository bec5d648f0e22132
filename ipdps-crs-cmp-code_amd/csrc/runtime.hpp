// Internal object layouts of the device runtime.
#pragma once

#include <hip/hip_runtime_api.h>

#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "device_types.hpp"
#include "launch.hpp"
#include "semblance.h"

struct sb_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  size_t budget = 0, used = 0;
  std::mutex mem_mu, run_mu;
  std::map<void*, size_t> sizes;           // live blocks
  // Freed blocks kept for reuse, each with an event recorded on `stream` when
  // it was freed: work queued before the free may still read the block, so it
  // is handed out again only once that event has completed.
  struct Cached {
    void* p;
    hipEvent_t freed;  // null: nothing queued could still use it
  };
  std::multimap<size_t, Cached> cache;
  std::vector<hipEvent_t> event_pool;      // completed free-events for reuse
  hipEvent_t ev[4] = {};
  sb_timing last{};
  bool timing_pending = false;
  bool timing_graph = false;  // the pending timing is a graph replay
  int num_cus = 256;
  // launch shape of the fast search (autotuned or set by sb_tune_set)
  int fast_variant = -1;  // -1: by store (kDefaultFastVariantF32 / kDefaultFastVariant)
  int pchunk_override = 0;
  std::map<std::string, std::pair<int, int>> tune_cache;
  // reusable work table / queue counters
  void* work_buf = nullptr;
  size_t work_cap = 0;
  std::string work_key;  // identifies the table in work_buf
  void* work_host = nullptr;  // pinned staging of the table upload
  size_t work_host_cap = 0;
  // Outgrown staging buffers, freed with the ctx (hipHostFree waits for the
  // device: other contexts' queued work would stall this one's search).
  std::vector<void*> work_host_retired;
  hipEvent_t work_host_ev = nullptr;  // the last upload from work_host has completed
  int work_n = 0;
  unsigned long long gather_serial = 0;
  void* queue_buf = nullptr;   // persistent-grid item counters
  int persistent = 0;          // SB_OPT_PERSISTENT
  int stream_rows = 0;         // SB_OPT_STREAM_ROWS (0: automatic)
  int dm_azimuth_mdeg = -1;    // SB_OPT_DM_AZIMUTH (-1: the reference's unsigned distance)
  int restore_fp32 = 1;        // SB_OPT_RESTORE_FP32
  int de_lds = -1;             // SB_OPT_DE_LDS (-1: automatic slab size, 0: global-memory evaluator)
  int de_waves = 8;            // SB_OPT_DE_WAVES
  int de_pairs = 1;            // SB_OPT_DE_PAIRS
  int crs_lds = -1;            // SB_OPT_CRS_LDS (-1: automatic slab size, 0: global-memory search)
  int crs_pairs = 1;           // SB_OPT_CRS_PAIRS
  int crs_cpw = 0;             // SB_OPT_CRS_CPW (0: automatic)
  void* scratch = nullptr;     // device address of the static scratch on this device
  bool capturing = false;      // a graph is being captured on `stream`
  void* graph_work = nullptr;  // work table copy owned by the graph being captured
  // SB_OPT_RESTORE_FP32 under sb_graph_create: the eager run keeps its
  // restored copy (restore_keep -> kept_*) for the graph, whose capture
  // replays the restore into those blocks (graph_rest_*).
  bool restore_keep = false;
  void* kept_mem = nullptr;
  void* kept_slots = nullptr;
  size_t kept_mem_bytes = 0, kept_slots_bytes = 0;
  void* graph_rest_mem = nullptr;
  void* graph_rest_slots = nullptr;
  size_t graph_rest_mem_bytes = 0, graph_rest_slots_bytes = 0;
};

struct sb_gather {
  sb_ctx* ctx = nullptr;
  unsigned long long serial = 0;  // unique per ctx, keys cached work tables
  int ncdp = 0, ntr = 0, ns = 0, dt_us = 0, store = 0;
  int64_t stride = 0;
  bool has_crs = false;
  void* d_samples = nullptr;      // row 0; kSamplesLead bytes into d_samples_mem
  void* d_samples_mem = nullptr;
  float* d_scale = nullptr;
  int32_t* d_cdp_off = nullptr;
  float* d_h_ref32 = nullptr;
  double *d_h0 = nullptr, *d_m0x = nullptr, *d_m0y = nullptr;
  float* d_msign = nullptr;  // sign of each trace's reference coordinate factor
  sb::SlotF* d_cmp_slots = nullptr;
  int32_t* d_crs_slot_off = nullptr;
  sb::SlotF* d_crs_slots = nullptr;
  sb::SlotD* d_crs_slots_d = nullptr;
  sb::SlotF* d_crs_slots_p = nullptr;  // unscaled FP32 geometry (CRS-DE exact mode)
  int64_t ncrs_slots = 0;
  int dm_azimuth_mdeg = -1;  // midpoint displacement of the slot tables (SB_OPT_DM_AZIMUTH at creation)
  std::vector<int32_t> cdp_off, crs_slot_off;  // host copies
  std::vector<int32_t> crs_row_lo, crs_row_hi;  // rows [lo, hi) spanning each CDP's supergather
  size_t bytes = 0;
};

struct sb_result {
  sb_ctx* ctx = nullptr;
  int ncdp = 0, ns = 0;
  void* base = nullptr;
  int32_t* index = nullptr;
  float *a = nullptr, *b = nullptr, *c = nullptr, *coher = nullptr, *stack = nullptr;
  unsigned long long* keys = nullptr;
};

namespace sb {
int ctx_alloc(sb_ctx* ctx, size_t bytes, void** p);
// A block for work queued on the ctx stream only (a cached block is reused
// with no host wait: the stream orders the new work after the old); no
// error message when it does not fit (SB_ERR_BUDGET: the caller falls back).
int ctx_alloc_transient(sb_ctx* ctx, size_t bytes, void** p);
void ctx_free(sb_ctx* ctx, void* p);
void ctx_release_all(sb_ctx* ctx);
// Frees cached blocks, largest first, until at most keep bytes stay cached.
void ctx_trim(sb_ctx* ctx, size_t keep);
// Device bytes a transient allocation of `bytes` takes (its block class).
size_t transient_block_size(size_t bytes);
int hip_fail(hipError_t e, const char* what);
int build_work(sb_ctx* ctx, const std::vector<int64_t>& cost, int cb, int ce, int ntiles, int npar, int pchunk,
               int* nblocks, const std::string& key);
}  // namespace sb
