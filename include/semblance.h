/*
 * semblance.h -- C ABI of libsemblance, the MI355X (gfx950) CMP / CRS / CRS-DE
 * semblance search.
 *
 * The reference ships no library: its seams are the OpenMP loops and the CUDA
 * kernels inside each CLI program.  Every entry point below names the
 * reference interface it replaces (paths relative to the reference tree).
 *
 * Conventions
 *   - Plain C types only.  Every call returns an int status (SB_OK = 0);
 *     sb_last_error() gives a thread-local detail string.  Nothing exits the
 *     process (the reference exits on every error: CMP/OpenMP/src/log.cpp:44-46).
 *   - Host arrays are caller-owned.  Device memory is owned by an sb_ctx and
 *     the objects created from it.
 *   - One sb_ctx per host thread.  Calls on one ctx are serialised on the
 *     ctx's HIP stream.
 *   - Result arrays are [ncdp][ns], CDP-major, in the gather's CDP order.
 */
#ifndef SEMBLANCE_H
#define SEMBLANCE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ status */
enum {
  SB_OK = 0,
  SB_ERR_INVALID = 1,     /* bad argument or inconsistent sizes            */
  SB_ERR_IO = 2,          /* file missing, short read ("Bad input data")   */
  SB_ERR_FORMAT = 3,      /* mixed ns in one file, empty gather, ...       */
  SB_ERR_HIP = 4,         /* HIP runtime error / no device                 */
  SB_ERR_BUDGET = 5,      /* request exceeds the ctx HBM budget            */
  SB_ERR_NOMEM = 6,       /* host allocation failed                        */
  SB_ERR_UNSUPPORTED = 7, /* e.g. window w = 2*tau+1 above SB_MAX_W        */
  SB_ERR_STATE = 8,       /* object used after shutdown / wrong method     */
  SB_ERR_DEADLINE = 9     /* serving: not admitted within the deadline     */
};

#define SB_MAX_W 15       /* largest coherence window (2*tau+1) supported   */

const char* sb_strerror(int code);
const char* sb_last_error(void);
const char* sb_version(void);

/* ------------------------------------------------------------ enumerations */
enum { SB_METHOD_CMP = 0, SB_METHOD_CRS = 1 };

/* Device storage of trace amplitudes.  F16/BF16 keep a per-trace power-of-two
 * scale; amplitudes are restored in FP32 before any accumulation. */
enum { SB_STORE_F32 = 0, SB_STORE_F16 = 1, SB_STORE_BF16 = 2 };

/* Arithmetic of the search.
 *  SB_MODE_FAST : FP32 search (contracted FMAs, hardware sqrt, window sums
 *                 reassociated); the winning parameter of every (CDP, t0) is
 *                 then re-evaluated in the reference's precision and
 *                 operation order (CMP float, CRS double).  Windows of more
 *                 than 7 samples (and CMP grids with c < 0) run as
 *                 SB_MODE_EXACT.
 *  SB_MODE_EXACT: the whole search in the reference's precision and
 *                 operation order (CMP float, CRS double), no contraction,
 *                 correctly rounded sqrt/div.  Bit-identical to the
 *                 reference OpenMP programs on the same F32 input. */
enum { SB_MODE_FAST = 0, SB_MODE_EXACT = 1 };

/* --------------------------------------------------- SU files (host side) */
/* One trace's header fields used by the path, at their SU byte offsets. */
typedef struct {
  int32_t cdp;    /* @20  */
  int32_t scalco; /* @70 (short in the file) */
  int32_t sx;     /* @72  */
  int32_t sy;     /* @76  */
  int32_t gx;     /* @80  */
  int32_t gy;     /* @84  */
} sb_trace_geom;

typedef struct sb_su_gather sb_su_gather; /* a binned SU file in host memory */

/* Reads a native-endian SU file, drops traces whose half-offset exceeds aph,
 * bins by the cdp header and orders CDPs like the reference:
 *   SB_METHOD_CMP: CDPs sorted by fold, descending (std::sort)
 *                  -- replaces su_gather::su_gather, CMP/OpenMP/src/su_gather.cpp:33-69
 *   SB_METHOD_CRS: CDPs by ascending key, plus for each CDP the list of CDPs
 *                  whose first-trace midpoint lies within apm
 *                  -- replaces CRS/OpenMP/src/su_gather.cpp:33-90
 * and linearizes samples to [ntr][ns] (su_gather::linearize,
 * CMP/OpenMP/src/su_gather.cpp:73-106). */
int sb_su_load(const char* path, int method, int aph, int apm, sb_su_gather** out);
/* The same from a SEG-Y file (3200-byte textual + 400-byte binary header,
 * optional rev 1 extended textual headers, big-endian trace headers and
 * samples in format 1 = IBM float, 2 = int32, 3 = int16, 5 = IEEE float or
 * 8 = int8; a trace's ns/dt of 0 take the binary header's).  Headers are
 * converted to the SU (native-endian) layout, so sections written with
 * sb_su_write_section are SU files as the reference writes them.  An
 * extension: the reference reads SU only. */
int sb_segy_load(const char* path, int method, int aph, int apm, sb_su_gather** out);
void sb_su_free(sb_su_gather* g);

/* Sizes of a loaded gather.  nnbr = total neighbour-list length (CRS). */
int sb_su_info(const sb_su_gather* g, int* ncdp, int* ntr, int* ns, int* dt_us, int* nnbr);
/* Borrowed views, valid until sb_su_free. */
const int32_t* sb_su_cdp_offsets(const sb_su_gather* g);   /* [ncdp+1] */
const float* sb_su_samples(const sb_su_gather* g);         /* [ntr][ns] */
const sb_trace_geom* sb_su_geom(const sb_su_gather* g);    /* [ntr] */
const int32_t* sb_su_nbr_offsets(const sb_su_gather* g);   /* [ncdp+1] or NULL */
const int32_t* sb_su_nbr(const sb_su_gather* g);           /* [nnbr] or NULL */

/* Writes one output section: one trace per CDP carrying the header of the
 * CDP's first trace and data = values[cdp][0..ns).  With rewrite_header = 1
 * (parameter sections: CMP c, CRS a/b/c) the header gets offset = 0,
 * sx = gx = (sx+gx)>>1, sy = gy = (sy+gy)>>1; coherence and stack sections
 * (rewrite_header = 0) keep it as read.  Replaces the output loops
 * CMP/OpenMP/src/main.cpp:244-260 and CRS/OpenMP/src/main.cpp:331-359. */
int sb_su_write_section(const sb_su_gather* g, const char* path, const float* values, int rewrite_header);

/* Writes traces with a zeroed 240-byte header carrying the given fields plus
 * ns, dt (us), tracl/tracr = 1-based trace number and offset = the signed
 * source-receiver distance in header units (su_trace::fputtr,
 * CMP/OpenMP/src/su_trace.cpp:55-66). */
int sb_su_write_traces(const char* path, int ntr, int ns, int dt_us,
                       const sb_trace_geom* geom, const float* samples);

/* Reference neighbour lists (CRS/OpenMP/src/su_gather.cpp:68-87) computed
 * from the first trace of each CDP.  Call with nbr = NULL to get *nnbr. */
int sb_crs_neighbors(int ncdp, const sb_trace_geom* first_geom, int apm,
                     int32_t* nbr_off, int32_t* nbr, int* nnbr);

/* Keep-test of the aph aperture exactly as su_gather applies it: the
 * half-offset from su_trace::halfoffset_x/y (float), squared and summed in
 * float for CMP (CMP/OpenMP/src/su_gather.cpp:44-47) and in double for CRS
 * (CRS/OpenMP/src/su_gather.cpp:44-47), against aph*aph.  Returns 1 = keep. */
int sb_keep_trace(const sb_trace_geom* g, int method, int aph);

/* ------------------------------------------------------- device context */
typedef struct sb_ctx sb_ctx;

/* hbm_budget = bytes this ctx may hold on the device (0 = 90% of free HBM). */
int sb_ctx_create(int device, size_t hbm_budget, sb_ctx** out);
void sb_ctx_destroy(sb_ctx* ctx);
/* Bytes currently held / budget. */
int sb_ctx_memory(const sb_ctx* ctx, size_t* used, size_t* budget);
/* Waits for all work queued on the ctx stream. */
int sb_ctx_sync(sb_ctx* ctx);

/* ---------------------------------------------- device-resident gathers */
typedef struct sb_gather sb_gather;

/* Uploads a binned gather (rows [cdp_off[i], cdp_off[i+1]) belong to CDP i),
 * computes per-trace geometry on the device (FP64, then stored FP32, or the
 * reference's own float/double arithmetic for SB_MODE_EXACT searches) and,
 * for CRS (nbr_off != NULL), the supergather slot table of every CDP.
 * samples: [ntr][ns] FP32, quantised on the device when store != F32.
 * Replaces the setup of CMP/CUDA/src/main.cu:232-274 (init_half) and
 * CRS/CUDA/src/main.cu:407-434 (init_mid, compute_points_for_gather). */
int sb_gather_create(sb_ctx* ctx, int ncdp, const int32_t* cdp_off, int ns, int dt_us,
                     const sb_trace_geom* geom, const float* samples, int store,
                     const int32_t* nbr_off, const int32_t* nbr, sb_gather** out);
void sb_gather_destroy(sb_gather* g);
int sb_gather_info(const sb_gather* g, int* ncdp, int* ntr, int* ns, int* store, size_t* bytes);

/* ------------------------------------------------------------- searches */
typedef struct {
  double c0, c1;  /* search range of C in s^2/m^2, as parsed (-c0/-c1)       */
  int nc;         /* number of C values; grid excludes c1 (Q1)              */
  double tau;     /* half window in seconds (-tau)                          */
  int mode;       /* SB_MODE_*                                              */
} sb_cmp_params;

typedef struct {
  double a0, a1, b0, b1, c0, c1; /* -a0 -a1 -b0 -b1 -c0 -c1                */
  int na, nb, nc;                /* -na -nb -nc                            */
  double tau;                    /* -tau                                   */
  int mode;                      /* SB_MODE_*                              */
} sb_crs_params;

typedef struct {
  double a0, a1, b0, b1, c0, c1; /* search box, as for CRS                  */
  double tau;
  int pop;        /* population per (CDP, t0), 4..64                         */
  int ngen;       /* generations (-ngen)                                     */
  double F, CR;   /* differential weight and crossover rate                 */
  uint64_t seed;
  int mode;       /* SB_MODE_FAST: FP32 contracted; SB_MODE_EXACT: FP32 in
                     the oracle's operation order (bit-identical CPU twin)  */
} sb_de_params;

/* Per-(CDP, t0) results.  index: CMP c index (-1: none), CRS flat
 * (ia*nb+ib)*nc+ic, DE -1.  a/b/c: parameter values written to the
 * reference's output files (CMP: c only, in s^2/m^2). */
typedef struct {
  int32_t* index;
  float* a;
  float* b;
  float* c;
  float* coher;
  float* stack;
} sb_host_result;

typedef struct sb_result sb_result; /* device-resident result arrays */
int sb_result_create(sb_ctx* ctx, int ncdp, int ns, sb_result** out);
void sb_result_destroy(sb_result* r);
/* Copies a result to host arrays (any pointer may be NULL to skip it).  Host
 * arrays that are consecutive in the order index, a, b, c, coher, stack
 * (ncdp * ns elements each) are filled by one copy; pinned ones copy by DMA. */
int sb_result_download(sb_ctx* ctx, const sb_result* r, sb_host_result* out);

/* CMP search: replaces compute_semblances + redux_semblances,
 * CMP/CUDA/src/main.cu:87-173 (OpenMP: CMP/OpenMP/src/main.cpp:119-222). */
int sb_cmp_run(sb_ctx* ctx, const sb_gather* g, const sb_cmp_params* p, sb_result* r);
/* CRS grid search: replaces compute_points_for_gather + compute_semblances +
 * redux_semblances, CRS/CUDA/src/main.cu:114-264 (OpenMP: CRS/OpenMP/src/main.cpp:155-285). */
int sb_crs_run(sb_ctx* ctx, const sb_gather* g, const sb_crs_params* p, sb_result* r);
/* CRS-DE search (described only in README.md:82-131 of the reference). */
int sb_crs_de_run(sb_ctx* ctx, const sb_gather* g, const sb_de_params* p, sb_result* r);

/* The same searches over CDPs [cdp_begin, cdp_end) of a resident gather
 * (velocity panels); r holds (cdp_end - cdp_begin) x ns results.  The gather
 * may belong to another ctx of the same device.  No reference counterpart
 * (the reference always searches whole files). */
int sb_cmp_run_range(sb_ctx* ctx, const sb_gather* g, const sb_cmp_params* p, int cdp_begin, int cdp_end,
                     sb_result* r);
int sb_crs_run_range(sb_ctx* ctx, const sb_gather* g, const sb_crs_params* p, int cdp_begin, int cdp_end,
                     sb_result* r);
int sb_crs_de_run_range(sb_ctx* ctx, const sb_gather* g, const sb_de_params* p, int cdp_begin, int cdp_end,
                        sb_result* r);

/* Caller-owned device arrays of (cdp_end - cdp_begin) x ns results, laid out
 * as sb_host_result's; all six are written. */
typedef struct {
  int32_t* index;
  float* a;
  float* b;
  float* c;
  float* coher;
  float* stack;
} sb_device_result;

/* Bytes of the static device scratch behind sb_cmp_run_noalloc (per device,
 * held from the first sb_ctx_create on that device until the process ends):
 * 1 MiB of work table (65 536 work items) and 3 MiB of arg-max keys (393 216
 * (CDP, t0) results). */
#define SB_NOALLOC_SCRATCH_BYTES (4u << 20)

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): sb_cmp_run_range with the results written straight to
 * caller-owned device arrays and every scratch buffer (work table, arg-max
 * keys) in a static device array of SB_NOALLOC_SCRATCH_BYTES, so the call
 * allocates nothing -- no hipMalloc, hipFree or host staging; the work table
 * is built on the device.  Same kernels and results as sb_cmp_run_range.
 * Searches whose work table or keys exceed the scratch fail with
 * SB_ERR_UNSUPPORTED (e.g. more than 393 216 CDP x ns results).  Every call
 * on one GPU shares that GPU's scratch: calls that overlap on one device
 * (two ctxs, two host threads) race, so serialise them.  Runs on the ctx
 * stream; the SB_OPT_PERSISTENT option does not apply. */
int sb_cmp_run_noalloc(sb_ctx* ctx, const sb_gather* g, const sb_cmp_params* p, int cdp_begin, int cdp_end,
                       const sb_device_result* out);

/* Convenience: search and download in one call. */
int sb_cmp(sb_ctx* ctx, const sb_gather* g, const sb_cmp_params* p, sb_host_result* out);
int sb_crs(sb_ctx* ctx, const sb_gather* g, const sb_crs_params* p, sb_host_result* out);
int sb_crs_de(sb_ctx* ctx, const sb_gather* g, const sb_de_params* p, sb_host_result* out);

/* Timing of the last *_run on this ctx, from HIP events recorded on the ctx
 * stream around each kernel.  tuples = sum over CDPs of traces x ns x params
 * (the reference's "Semblances-Trace" count).  After sb_search_streamed:
 * search_ms = the summed device time of its searches (the counterpart of the
 * CUDA programs' summed kernel events, CMP/CUDA/src/main.cu:306-326),
 * finalize_ms = 0, total_ms = the whole call. */
typedef struct {
  double search_ms;    /* main search kernel                     */
  double finalize_ms;  /* winner re-evaluation / output kernel   */
  double total_ms;     /* first to last event of the call        */
  double tuples;
  int grid_blocks;     /* main kernel launch size                */
  int block_threads;
  int variant;         /* launch shape that ran (sb_tune_get numbering; -1: exact mode) */
  int lds_staged;      /* 1 if the LDS-staged kernel ran (windows from LDS slabs; traces
                          whose ranges did not fit read from global memory) */
  int restored;        /* 1 if the search read a transient FP32 copy of a 16-bit store's
                          rows (SB_OPT_RESTORE_FP32; search_ms includes the copy) */
} sb_timing;
int sb_last_timing(const sb_ctx* ctx, sb_timing* t);

/* Launch shape of the SB_MODE_FAST search on this ctx: variant indexes the
 * compiled (t0 rows per lane, parameters per pass) shapes, -1 (a new ctx's
 * setting) picks by store: the LDS-staged shape with (sample, difference)
 * slabs for FP32, the one with sample slabs for 16-bit stores; pchunk =
 * parameters per workgroup (0 = automatic).  sb_tune measures every shape on
 * the given gather (reps timed runs each), re-times the three fastest in
 * 2 reps + 1 interleaved rounds and keeps the one with the lowest median
 * (the project's autotuner); sb_last_timing reports the shape that ran. */
int sb_tune_set(sb_ctx* ctx, int variant, int pchunk);
int sb_tune_get(const sb_ctx* ctx, int* variant, int* pchunk, int* nvariants);
int sb_tune(sb_ctx* ctx, const sb_gather* g, int method, const void* params, int reps);

/* Context options.  SB_OPT_PERSISTENT = 1: the global-memory fast search
 * runs as a persistent grid (as many workgroups as the CUs hold at once,
 * each XCD's workgroups pulling that XCD's work items from their own
 * counter) instead of one workgroup per work item.  Results are the same.
 * No reference counterpart (the reference launches one grid per CDP,
 * CMP/CUDA/src/main.cu:271-304). */
/* SB_OPT_DM_AZIMUTH: the midpoint displacement m of the CRS supergathers of
 * gathers created afterwards on this ctx (sb_gather_create,
 * sb_search_streamed).  -1 (default): the reference's unsigned distance
 * m = |dm| to the CDP's first trace (CRS/OpenMP/src/main.cpp:189-198);
 * 0..359999: the signed projection of dm onto the azimuth, in millidegrees
 * clockwise from +y (north): m = dx sin(az) + dy cos(az), m^2 = m m, so the
 * A term tells dips on either side apart.  An extension behind crs-ocl-de's
 * -azimuth flag, which the reference documents but does not ship
 * (README.md:94,115,131). */
/* SB_OPT_RESTORE_FP32 = 1 (default): an SB_MODE_FAST search of a 16-bit
 * store (CMP, CRS grid, CRS-DE) first restores the rows it reads to FP32 --
 * the stored value times the trace's scale, the values the 16-bit kernels
 * restore after every load -- in a transient copy on the ctx (counted against
 * its budget, released when the search's work is done), then runs the FP32
 * kernels on it.  0, or a copy that does not fit the budget, or a search
 * being captured into a graph: the 16-bit kernels read the store directly,
 * restoring every window after its load.  Both evaluate the same FP32
 * values with the same FP32 arithmetic: the results are the same either
 * way.  No reference counterpart (the reference has no 16-bit store). */
/* SB_OPT_CRS_LDS: the CRS grid fast search's window source (FP32 rows and
 * 16-bit stores' restored copies).  -1 (default): the LDS-staged search with
 * 40 KiB per workgroup; 0: every window read from global memory (the launch
 * shape sb_tune / sb_tune_set choose); 4096..65536 (a multiple of 16): the
 * LDS-staged search with that many bytes of LDS per workgroup.  A workgroup
 * covers one (CDP, 64-t0 tile) and several (a, b) pairs of the grid with
 * their c values, and walks the supergather in batches of consecutive
 * traces whose window ranges over its parameter box it stages in LDS.
 * SB_OPT_CRS_PAIRS (1 = default): the slab holds (sample, difference)
 * pairs; 0: samples.  SB_OPT_CRS_CPW: c values each wave holds (6 or 12; 0 =
 * default: 6).  Results are the same bits for every setting; a search
 * the staging cannot serve (more than 48 c values, ns and the longest
 * supergather too large for the slab) reads global memory.  No reference
 * counterpart. */
/* SB_OPT_DE_LDS: the CRS-DE fast search's window source.  -1 (default): the
 * LDS-staged evaluator with 40 KiB per workgroup; 0: every window read
 * from global memory; 4096..65536 (a multiple of 16): the LDS-staged
 * evaluator with that many bytes of LDS per workgroup.  The workgroup walks
 * the supergather in batches of consecutive traces whose window ranges over
 * its t0 span and the (A, B, C) box it stages in LDS.  SB_OPT_DE_WAVES (4, 8
 * = default, 16): waves per workgroup of the LDS-staged evaluator.
 * SB_OPT_DE_PAIRS (1 = default): its slab holds every sample with its
 * forward difference (8 bytes, one FMA a window sample); 0: samples only (4
 * bytes, the difference formed per window).  Results
 * are the same bits for every setting; a search the staging cannot serve (a
 * 16-bit store read directly, ns too long for the slab) reads global memory.
 * No reference counterpart (the reference ships no CRS-DE source). */
enum {
  SB_OPT_PERSISTENT = 1,
  SB_OPT_STREAM_ROWS = 2, /* traces per sb_search_streamed batch, 0 = auto */
  SB_OPT_DM_AZIMUTH = 3,
  SB_OPT_RESTORE_FP32 = 4,
  SB_OPT_DE_LDS = 5,
  SB_OPT_DE_WAVES = 6,
  SB_OPT_DE_PAIRS = 7,
  SB_OPT_CRS_LDS = 8,
  SB_OPT_CRS_PAIRS = 9,
  SB_OPT_CRS_CPW = 10
};
int sb_ctx_set_option(sb_ctx* ctx, int option, int value);
int sb_ctx_get_option(const sb_ctx* ctx, int option, int* value);

/* ------------------------------------------------------ streamed search */
/* One-shot search of a line whose traces are still in host memory, with the
 * upload overlapped: the traces go up in CDP batches on a copy stream
 * (staged, padded or quantised, slot tables built) while each batch's CDPs
 * -- for CRS / CRS-DE, each CDP once all its neighbour CDPs have arrived --
 * are searched on the ctx stream.  kind: SB_METHOD_CMP (sb_cmp_params),
 * SB_METHOD_CRS (sb_crs_params) or SB_GRAPH_CRS_DE (sb_de_params); nbr_off /
 * nbr as for sb_gather_create (ignored for CMP).  out receives ncdp x ns
 * results, the same values as sb_gather_create followed by one search.
 * keep != NULL receives the resident gather for further searches.  Replaces
 * the reference's read-then-compute sequence (CMP/OpenMP/src/main.cpp:77-236,
 * CRS/OpenMP/src/main.cpp:98-317). */
int sb_search_streamed(sb_ctx* ctx, int kind, const void* params, int ncdp, const int32_t* cdp_off, int ns,
                       int dt_us, const sb_trace_geom* geom, const float* samples, int store,
                       const int32_t* nbr_off, const int32_t* nbr, sb_host_result* out, sb_gather** keep);

/* --------------------------------------------------------------- graphs */
/* One search captured into a HIP graph and replayed without host work:
 * sb_graph_create runs the search once (building the work table), then
 * captures every kernel of it -- key clear, search, winner re-evaluation --
 * on the ctx stream; sb_graph_launch replays it there, reading the gather
 * and writing r as an eager *_run_range call would (download with
 * sb_result_download).  kind: SB_METHOD_CMP (sb_cmp_params), SB_METHOD_CRS
 * (sb_crs_params) or SB_GRAPH_CRS_DE (sb_de_params).  The graph keeps its
 * own copy of the work table; ctx, gather and result must outlive it.  No
 * reference counterpart (the reference relaunches per CDP). */
enum { SB_GRAPH_CRS_DE = 2 };
typedef struct sb_graph sb_graph;
int sb_graph_create(sb_ctx* ctx, const sb_gather* g, int kind, const void* params, int cdp_begin, int cdp_end,
                    sb_result* r, sb_graph** out);
int sb_graph_launch(sb_graph* gr);
void sb_graph_destroy(sb_graph* gr);

/* --------------------------------------------------- device measurement */
/* The FP32 vector FMA rate the device sustains, in TFLOP/s (2 FLOP per FMA):
 * independent FMA chains in every lane of a grid that fills every SIMD, no
 * memory traffic.  The search kernels run no packed FP32, so this -- not the
 * published peak, which counts packed FP32 -- is the ceiling of their
 * arithmetic.  No reference counterpart (a measurement for the roofline). */
int sb_device_fp32_fma_rate(int device, double* tflops);
/* Interpolated window samples per second (the search's inner step: one
 * interpolation, its window sum and its square) for four instruction forms,
 * written out in every lane of a grid that fills every SIMD:
 *   [0] FP32 samples: sub, fma, add, fma (what every store runs)
 *   [1] FP16 sample and FP16 difference: v_fma_mix_f32, fma, fma
 *   [2] FP16 sample pair, FP16 weights: v_dot2_f32_f16, fma, fma
 *   [3] FP32 sample and FP32 difference: fma, fma, fma
 * Forms 1 and 2 are the mixed-precision steps a 16-bit store could take
 * instead of restoring its samples to FP32.  No reference counterpart. */
int sb_device_window_rates(int device, double samples_per_s[4]);

/* ------------------------------------------------------ synthetic data */
/* Seeded synthetic gathers for the benchmark configurations (host, OpenMP).
 * kind 0: 2D line, CMP events t^2 = t0^2 + C h^2 (h = half-offset);
 * kind 1: 3D-style CDPs, CRS events t^2 = (t0 + A dm)^2 + t0 B dm^2 + C h^2. */
typedef struct {
  int kind;
  int ncdp;
  int fold_min, fold_max;   /* uniform integer fold per CDP            */
  int ns, dt_us;
  double cdp_spacing;       /* m                                       */
  double h_max;             /* half-offsets in [0, h_max] m            */
  double h_jitter;          /* kind 0: +- jitter on the regular offsets */
  int regular_offsets;      /* kind 0: 1 = h_j = j * h_step            */
  double h_step;            /* kind 0 regular step                     */
  double mid_half_width;    /* kind 1: midpoints uniform +- this (x,y) */
  int nevents;
  double freq;              /* Ricker peak frequency, Hz               */
  double snr;               /* 0 = no noise                            */
  int aph;                  /* keep only traces passing the aph test (0 = keep all) */
  uint64_t seed;
} sb_synth_spec;

/* Two-call pattern: samples == NULL returns the kept trace count in *ntr and
 * the CDP count in *ncdp; then pass arrays of that size.  Traces come out
 * binned by CDP in ascending key order with cdp_off[ncdp+1]. */
int sb_synth(const sb_synth_spec* spec, int* ncdp, int* ntr, int32_t* cdp_off,
             sb_trace_geom* geom, float* samples);

/* ------------------------------------------------------------- serving */
/* A resident request loop for interactive velocity analysis (continuous
 * batching on one GPU).  Worker slots, each with its own HIP stream, take
 * requests in arrival order as soon as a slot is idle and the request's
 * device footprint fits the free HBM budget, so requests that arrive while
 * others run are admitted as memory frees up and their kernels overlap on
 * the device; waiting requests of the same shape (method, ns, dt, parameters,
 * store) are merged into one gather and one launch.  A request's results do
 * not depend on what it was batched with.  No reference counterpart: the
 * reference is a batch CLI. */
typedef struct sb_server sb_server;

/* A request either carries its traces (gather_id == 0: cdp_off, geom,
 * samples, nbr_off/nbr and store describe a binned panel, uploaded for this
 * request) or names a resident gather (gather_id from sb_server_add_gather)
 * and the CDPs [cdp_begin, cdp_end) of it to search; then only method and
 * the parameters are read besides those three fields, and the results have
 * cdp_end - cdp_begin rows. */
typedef struct {
  int method;              /* SB_METHOD_CMP or SB_METHOD_CRS               */
  int ncdp, ns, dt_us;
  const int32_t* cdp_off;  /* [ncdp+1]                                     */
  const sb_trace_geom* geom;
  const float* samples;    /* [ntr][ns]                                    */
  const int32_t* nbr_off;  /* CRS only                                     */
  const int32_t* nbr;
  sb_cmp_params cmp;
  sb_crs_params crs;
  int store;
  int gather_id;           /* 0: traces carried above; else resident gather */
  int cdp_begin, cdp_end;  /* CDP range of the resident gather              */
} sb_request;

typedef struct {
  double t_submit_ms, t_admit_ms, t_done_ms; /* server clock                */
  int status;
} sb_request_stats;

/* hbm_budget: bytes the server may hold (0 = 90% of free HBM); max_batch:
 * most CDPs merged into one launch. */
int sb_server_create(int device, size_t hbm_budget, int max_batch, sb_server** out);
/* Queues a request and returns its ticket without blocking.  The request's
 * arrays are read in place: keep them valid until sb_server_wait returns. */
int sb_server_submit(sb_server* s, const sb_request* req, uint64_t* ticket);
/* Blocks until the ticket is done (timeout_ms < 0: forever); copies results
 * into out (arrays [ncdp][ns]) and releases the request.  A request larger
 * than the whole budget fails with SB_ERR_BUDGET. */
int sb_server_wait(sb_server* s, uint64_t ticket, double timeout_ms,
                   sb_host_result* out, sb_request_stats* st);
/* Uploads a binned line once and keeps it on the device (arguments as for
 * sb_gather_create); requests then name it by *gather_id (> 0) with a CDP
 * range.  Its bytes count against the server budget until it is removed. */
int sb_server_add_gather(sb_server* s, int ncdp, const int32_t* cdp_off, int ns, int dt_us,
                         const sb_trace_geom* geom, const float* samples, int store,
                         const int32_t* nbr_off, const int32_t* nbr, int* gather_id);
/* Rejects new requests naming the gather, waits for those queued or running,
 * then frees it. */
int sb_server_remove_gather(sb_server* s, int gather_id);
/* Load shedding: a request that has waited max_queue_ms since its
 * submission without being admitted fails with SB_ERR_DEADLINE (its
 * sb_server_wait returns that status, within about max_queue_ms / 8 of the
 * deadline) instead of queueing; admitted requests always complete.
 * 0 (default): no deadline, every request waits its turn.  Requests already
 * queued are held to the new value. */
int sb_server_set_deadline(sb_server* s, double max_queue_ms);
/* hbm_used: bytes held by running requests and resident gathers. */
int sb_server_stats(const sb_server* s, uint64_t* done, uint64_t* launches, size_t* hbm_used);
/* Stops the workers; queued requests that have not started fail. */
void sb_server_destroy(sb_server* s);

#ifdef __cplusplus
}
#endif
#endif /* SEMBLANCE_H */
