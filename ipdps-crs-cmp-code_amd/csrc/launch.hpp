// Host-callable launchers of the kernels in kernels.hip / de.hip.
#pragma once

#include <hip/hip_runtime_api.h>

#include "device_types.hpp"
#include "semblance.h"

namespace sb {

struct FastVariant {
  int R;       // t0 rows per lane (block covers 64*R t0)
  int CC;      // parameters evaluated together per trace pass
  int lds_kb;  // 0: windows from global memory (k_search_fast); else KiB of
               // LDS slab per block (k_search_lds, CMP only)
  int waves;   // waves per block (a block's waves split its parameters)
  int pairs;   // k_search_lds slabs: 0 samples (4 bytes a sample); 1 (sample, forward difference)
               // pairs (8 bytes a sample); 2 pairs for the nearest traces, samples after them
  int share;   // pairs == 2: sixteenths of the slab bytes the pairs may take when the samples of
               // every trace do not fit
};

// LDS-staged CMP search: at most this many dynamic LDS bytes per block and
// traces per CDP (the per-trace tables live in the same LDS).
constexpr int kLdsMaxBytes = 64 * 1024;
constexpr int kLdsMaxTraces = 2048;
// Byte offset of the slabs in an LDS-staged block for a CDP of S traces
// (-1: no such shape); the slabs hold FP32 samples (or FP32 pairs).
int lds_slab_offset(const FastVariant& v, int S);
extern const FastVariant kFastVariants[];
extern const int kNumFastVariants;
// Shapes of an untuned ctx (CMP; searches they cannot stage run
// kGlobalFallbackVariant): LDS-staged, 8 waves, 52 KiB, with (sample,
// difference) slabs for FP32 (one FMA per window sample) and sample slabs for
// 16-bit stores (whose staging would convert twice as many values into half
// as many staged traces).
constexpr int kDefaultFastVariantF32 = 9;
constexpr int kDefaultFastVariant = 5;
// Global-memory shape an LDS-staged variant falls back to (R = 1, CC = 4:
// the fastest global shape on the CRS grid and on the CMP line).
constexpr int kGlobalFallbackVariant = 1;

bool fast_supported(int w);

// Static device scratch of the allocation-free entry point
// (sb_cmp_run_noalloc), one copy per device: the work table (16-byte items)
// in its first kScratchKeysOffset bytes, the arg-max keys (8 bytes per
// (CDP, t0)) after them.
constexpr size_t kScratchBytes = SB_NOALLOC_SCRATCH_BYTES;
constexpr size_t kScratchKeysOffset = 1u << 20;
constexpr int64_t kScratchWorkItems = (int64_t)(kScratchKeysOffset / sizeof(WorkItem));
constexpr int64_t kScratchKeys = (int64_t)((kScratchBytes - kScratchKeysOffset) / 8);
// Device address of the scratch on `device` (which must be current), looked
// up with hipGetSymbolAddress on the first call for that device and kept.
hipError_t scratch_address(int device, void** out);
// Work table of CDPs [cdp0, cdp0 + ncdp) built on the device: nblocks (a
// multiple of 8) items, every (CDP, t0 tile, parameter chunk) once, in that
// order, the XCD of block b (b % 8) taking its own contiguous eighth of the
// list; the rest no-ops.
hipError_t launch_make_work(WorkItem* w, int nblocks, int cdp0, int ncdp, int ntiles, int nchunks, int pchunk,
                            int npar, hipStream_t st);

hipError_t launch_prep_traces(const TraceGeom* g, int ntr, float* h_ref32, double* h0, double* m0x, double* m0y,
                              float* msign, hipStream_t st);
// CMP slot tables of CDPs [cdp0, cdp0 + ncdp): each CDP's traces by increasing h.
hipError_t launch_cmp_slots(int cdp0, int ncdp, const int32_t* cdp_off, const double* h0, const float* scale,
                            double idt, SlotF* slots, hipStream_t st);
// CRS slot tables; the fast slots carry the byte address of their row
// (samples + row * row_bytes).  az != 0: m is the signed projection of the
// midpoint displacement onto the direction (az_s, az_c) (sine and cosine of
// the azimuth clockwise from +y) in true coordinates (msign: the sign of each
// trace's reference coordinate factor), m^2 = m m; else the distance |dm|.
hipError_t launch_crs_slots(int cdp0, int ncdp, const int32_t* cdp_off, const int32_t* nbr_off, const int32_t* nbr,
                            const int32_t* slot_off, const double* m0x, const double* m0y, const double* h0,
                            const float* scale, double idt, const void* samples, uint64_t row_bytes, SlotF* sf,
                            SlotD* sd, SlotF* sp, const float* msign, int az, double az_s, double az_c,
                            hipStream_t st);
hipError_t launch_store(const float* src, int ntr, int ns, int64_t stride, int store, void* dst, float* scale,
                        hipStream_t st);
hipError_t launch_fill_u64(unsigned long long* p, int64_t n, unsigned long long v, hipStream_t st);
// nrows rows of a 16-bit store (src: the first row, scale: its scale; rows
// of stride elements, a multiple of 64) restored to FP32 into dst, padding
// included: the values every 16-bit search path restores.
hipError_t launch_restore_rows(const void* src, const float* scale, int64_t nrows, int64_t stride, int store,
                               float* dst, hipStream_t st);
// n CRS fast slots copied with addr = base + row * row_bytes.
hipError_t launch_rebase_slots(const SlotF* src, int64_t n, uint64_t base, uint64_t row_bytes, SlotF* dst,
                               hipStream_t st);

// a.queue set: persistent grid (at most what the CUs hold at once, a
// multiple of 8) pulling the nblocks work items; else one block per item.
hipError_t launch_search_fast(const SearchArgs& a, int store, int variant, int nblocks, int num_cus, hipStream_t st);
hipError_t launch_search_lds(const SearchArgs& a, int store, const FastVariant& v, int nblocks, hipStream_t st);
hipError_t launch_search_exact(const SearchArgs& a, int store, int nblocks, hipStream_t st);
// Reference-arithmetic search in the tiled layout (work items as the fast
// search's with 64-t0 tiles; keys, then launch_finalize).  FP32 store, w in
// {1, 3, 5, 7}.  CMP: any parameter chunking (first-max keys, atomicMax);
// CRS: one work item per (CDP, tile) covering every parameter (the double
// semblances are compared in the workgroup, not through keys).
bool exact_tiled_supported(int method, int store, int w);
hipError_t launch_search_exact_tiled(const SearchArgs& a, int nblocks, hipStream_t st);
hipError_t launch_finalize(const SearchArgs& a, int store, hipStream_t st);

// CRS-DE (de.hip).  One work item = (cdp, tile of t0).
bool de_supported(int w, int pop);
hipError_t launch_de(const SearchArgs& a, const DeArgs& d, int store, int exact, int nblocks, hipStream_t st);
// LDS-staged CRS grid search (FP32 rows; same keys as launch_search_fast's
// CRS kernel): workgroups of `waves` waves over (CDP, 64-t0 tile) and
// crs_lds_pairs_per_block (a, b) pairs of the grid (work items' parameter
// ranges cover whole pairs), a.lds_bytes of LDS for the plan and the slab;
// smax = the longest supergather searched.
constexpr int kCrsLdsDefaultBytes = 40 * 1024;
constexpr int kCrsLdsWaves = 8;
// cpw: c values a wave holds (6 or 12; 0: automatic).
int crs_lds_min_bytes(int ns, int smax, int pairs, int waves);
bool crs_lds_supported(int w, int nc, int cpw, int waves, int pairs, int lds_bytes, int ns, int smax);
int crs_lds_pairs_per_block(int nc, int cpw, int waves);
hipError_t launch_search_crs_lds(const SearchArgs& a, int cpw, int waves, int pairs, int smax, int nblocks,
                                 hipStream_t st);
// LDS-staged fast CRS-DE (FP32 rows; same results as launch_de's fast
// kernel): workgroups of `waves` waves over 64 * waves / P2 consecutive t0
// (P2 = the population rounded up to a power of two), a.lds_bytes of LDS
// holding a 10-byte-a-trace plan of the supergather and the slab; smax =
// the longest supergather searched.
constexpr int kDeLdsDefaultBytes = 40 * 1024;
// pairs: (sample, difference) slabs (one FMA a window sample) or samples.
int de_lds_min_bytes(int ns, int smax, int pairs);
bool de_lds_supported(int w, int pop, int waves, int pairs, int lds_bytes, int ns, int smax);
hipError_t launch_de_lds(const SearchArgs& a, const DeArgs& d, int waves, int pairs, int smax, int nblocks,
                         hipStream_t st);

}  // namespace sb
