"""ctypes binding of libsemblance (include/semblance.h).

The product is the C ABI; this module only marshals numpy arrays across it
for the tests, the benchmark and interactive use.  It never computes a
semblance itself: if the library is missing, every call raises.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libsemblance.so")

METHOD_CMP, METHOD_CRS = 0, 1
STORE_F32, STORE_F16, STORE_BF16 = 0, 1, 2
MODE_FAST, MODE_EXACT = 0, 1
OPT_PERSISTENT = 1
OPT_STREAM_ROWS = 2
OPT_DM_AZIMUTH = 3
OPT_RESTORE_FP32 = 4
OPT_DE_LDS = 5
OPT_DE_WAVES = 6
OPT_DE_PAIRS = 7
OPT_CRS_LDS = 8
OPT_CRS_PAIRS = 9
OPT_CRS_CPW = 10
GRAPH_CRS_DE = 2
ERR_DEADLINE = 9
STORES = {"f32": STORE_F32, "f16": STORE_F16, "bf16": STORE_BF16}

GEOM_DTYPE = np.dtype([("cdp", "<i4"), ("scalco", "<i4"), ("sx", "<i4"), ("sy", "<i4"), ("gx", "<i4"),
                       ("gy", "<i4")])


class SbError(RuntimeError):
    pass


class TraceGeom(C.Structure):
    _fields_ = [("cdp", C.c_int32), ("scalco", C.c_int32), ("sx", C.c_int32), ("sy", C.c_int32),
                ("gx", C.c_int32), ("gy", C.c_int32)]


class CmpParams(C.Structure):
    _fields_ = [("c0", C.c_double), ("c1", C.c_double), ("nc", C.c_int), ("tau", C.c_double), ("mode", C.c_int)]


class CrsParams(C.Structure):
    _fields_ = [("a0", C.c_double), ("a1", C.c_double), ("b0", C.c_double), ("b1", C.c_double),
                ("c0", C.c_double), ("c1", C.c_double), ("na", C.c_int), ("nb", C.c_int), ("nc", C.c_int),
                ("tau", C.c_double), ("mode", C.c_int)]


class DeParams(C.Structure):
    _fields_ = [("a0", C.c_double), ("a1", C.c_double), ("b0", C.c_double), ("b1", C.c_double),
                ("c0", C.c_double), ("c1", C.c_double), ("tau", C.c_double), ("pop", C.c_int),
                ("ngen", C.c_int), ("F", C.c_double), ("CR", C.c_double), ("seed", C.c_uint64),
                ("mode", C.c_int)]


class HostResult(C.Structure):
    _fields_ = [("index", C.POINTER(C.c_int32)), ("a", C.POINTER(C.c_float)), ("b", C.POINTER(C.c_float)),
                ("c", C.POINTER(C.c_float)), ("coher", C.POINTER(C.c_float)), ("stack", C.POINTER(C.c_float))]


class DeviceArrays(C.Structure):
    _fields_ = [("index", C.c_void_p), ("a", C.c_void_p), ("b", C.c_void_p), ("c", C.c_void_p),
                ("coher", C.c_void_p), ("stack", C.c_void_p)]


class Timing(C.Structure):
    _fields_ = [("search_ms", C.c_double), ("finalize_ms", C.c_double), ("total_ms", C.c_double),
                ("tuples", C.c_double), ("grid_blocks", C.c_int), ("block_threads", C.c_int),
                ("variant", C.c_int), ("lds_staged", C.c_int), ("restored", C.c_int)]


class SynthSpec(C.Structure):
    _fields_ = [("kind", C.c_int), ("ncdp", C.c_int), ("fold_min", C.c_int), ("fold_max", C.c_int),
                ("ns", C.c_int), ("dt_us", C.c_int), ("cdp_spacing", C.c_double), ("h_max", C.c_double),
                ("h_jitter", C.c_double), ("regular_offsets", C.c_int), ("h_step", C.c_double),
                ("mid_half_width", C.c_double), ("nevents", C.c_int), ("freq", C.c_double), ("snr", C.c_double),
                ("aph", C.c_int), ("seed", C.c_uint64)]


class Request(C.Structure):
    _fields_ = [("method", C.c_int), ("ncdp", C.c_int), ("ns", C.c_int), ("dt_us", C.c_int),
                ("cdp_off", C.POINTER(C.c_int32)), ("geom", C.c_void_p), ("samples", C.POINTER(C.c_float)),
                ("nbr_off", C.POINTER(C.c_int32)), ("nbr", C.POINTER(C.c_int32)), ("cmp", CmpParams),
                ("crs", CrsParams), ("store", C.c_int), ("gather_id", C.c_int), ("cdp_begin", C.c_int),
                ("cdp_end", C.c_int)]


class RequestStats(C.Structure):
    _fields_ = [("t_submit_ms", C.c_double), ("t_admit_ms", C.c_double), ("t_done_ms", C.c_double),
                ("status", C.c_int)]


_lib = None

# Every symbol include/semblance.h declares (checked by the tests).
EXPORTS = [
    "sb_strerror", "sb_last_error", "sb_version", "sb_su_load", "sb_segy_load", "sb_su_free", "sb_su_info", "sb_su_cdp_offsets",
    "sb_su_samples", "sb_su_geom", "sb_su_nbr_offsets", "sb_su_nbr", "sb_su_write_section", "sb_su_write_traces",
    "sb_crs_neighbors", "sb_keep_trace", "sb_ctx_create", "sb_ctx_destroy", "sb_ctx_memory", "sb_ctx_sync",
    "sb_gather_create", "sb_gather_destroy", "sb_gather_info", "sb_result_create", "sb_result_destroy",
    "sb_result_download", "sb_cmp_run", "sb_crs_run", "sb_crs_de_run", "sb_cmp_run_range", "sb_crs_run_range",
    "sb_crs_de_run_range", "sb_cmp", "sb_crs", "sb_crs_de",
    "sb_last_timing", "sb_tune_set", "sb_tune_get", "sb_tune", "sb_synth", "sb_server_create", "sb_server_submit",
    "sb_server_wait", "sb_server_add_gather", "sb_server_remove_gather", "sb_server_stats", "sb_server_destroy",
    "sb_ctx_set_option", "sb_ctx_get_option", "sb_graph_create", "sb_graph_launch", "sb_graph_destroy",
    "sb_search_streamed", "sb_cmp_run_noalloc", "sb_device_fp32_fma_rate", "sb_device_window_rates",
    "sb_server_set_deadline",
]


def lib():
    """Loads libsemblance.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise SbError("libsemblance.so is not built (run __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        i32p = C.POINTER(C.c_int32)
        L.sb_strerror.restype = C.c_char_p
        L.sb_last_error.restype = C.c_char_p
        L.sb_version.restype = C.c_char_p
        L.sb_su_load.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int, C.POINTER(P)]
        L.sb_segy_load.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int, C.POINTER(P)]
        L.sb_su_free.argtypes = [P]
        L.sb_su_free.restype = None
        L.sb_su_info.argtypes = [P] + [C.POINTER(C.c_int)] * 5
        for n in ("sb_su_cdp_offsets", "sb_su_nbr_offsets", "sb_su_nbr"):
            getattr(L, n).argtypes = [P]
            getattr(L, n).restype = i32p
        L.sb_su_samples.argtypes = [P]
        L.sb_su_samples.restype = C.POINTER(C.c_float)
        L.sb_su_geom.argtypes = [P]
        L.sb_su_geom.restype = P
        L.sb_su_write_section.argtypes = [P, C.c_char_p, C.POINTER(C.c_float), C.c_int]
        L.sb_su_write_traces.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int, P, C.POINTER(C.c_float)]
        L.sb_crs_neighbors.argtypes = [C.c_int, P, C.c_int, i32p, i32p, C.POINTER(C.c_int)]
        L.sb_keep_trace.argtypes = [P, C.c_int, C.c_int]
        L.sb_ctx_create.argtypes = [C.c_int, C.c_size_t, C.POINTER(P)]
        L.sb_ctx_destroy.argtypes = [P]
        L.sb_ctx_destroy.restype = None
        L.sb_ctx_memory.argtypes = [P, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]
        L.sb_ctx_sync.argtypes = [P]
        L.sb_gather_create.argtypes = [P, C.c_int, i32p, C.c_int, C.c_int, P, C.POINTER(C.c_float), C.c_int, i32p,
                                       i32p, C.POINTER(P)]
        L.sb_gather_destroy.argtypes = [P]
        L.sb_gather_destroy.restype = None
        L.sb_gather_info.argtypes = [P, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int),
                                     C.POINTER(C.c_int), C.POINTER(C.c_size_t)]
        L.sb_result_create.argtypes = [P, C.c_int, C.c_int, C.POINTER(P)]
        L.sb_result_destroy.argtypes = [P]
        L.sb_result_destroy.restype = None
        L.sb_result_download.argtypes = [P, P, C.POINTER(HostResult)]
        L.sb_cmp_run.argtypes = [P, P, C.POINTER(CmpParams), P]
        L.sb_crs_run.argtypes = [P, P, C.POINTER(CrsParams), P]
        L.sb_crs_de_run.argtypes = [P, P, C.POINTER(DeParams), P]
        L.sb_cmp_run_range.argtypes = [P, P, C.POINTER(CmpParams), C.c_int, C.c_int, P]
        L.sb_crs_run_range.argtypes = [P, P, C.POINTER(CrsParams), C.c_int, C.c_int, P]
        L.sb_crs_de_run_range.argtypes = [P, P, C.POINTER(DeParams), C.c_int, C.c_int, P]
        L.sb_cmp.argtypes = [P, P, C.POINTER(CmpParams), C.POINTER(HostResult)]
        L.sb_crs.argtypes = [P, P, C.POINTER(CrsParams), C.POINTER(HostResult)]
        L.sb_crs_de.argtypes = [P, P, C.POINTER(DeParams), C.POINTER(HostResult)]
        L.sb_last_timing.argtypes = [P, C.POINTER(Timing)]
        L.sb_tune_set.argtypes = [P, C.c_int, C.c_int]
        L.sb_tune_get.argtypes = [P, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.sb_tune.argtypes = [P, P, C.c_int, P, C.c_int]
        L.sb_synth.argtypes = [C.POINTER(SynthSpec), C.POINTER(C.c_int), C.POINTER(C.c_int), i32p, P,
                               C.POINTER(C.c_float)]
        L.sb_server_create.argtypes = [C.c_int, C.c_size_t, C.c_int, C.POINTER(P)]
        L.sb_server_submit.argtypes = [P, C.POINTER(Request), C.POINTER(C.c_uint64)]
        L.sb_server_wait.argtypes = [P, C.c_uint64, C.c_double, C.POINTER(HostResult), C.POINTER(RequestStats)]
        L.sb_server_add_gather.argtypes = [P, C.c_int, i32p, C.c_int, C.c_int, P, C.POINTER(C.c_float), C.c_int,
                                           i32p, i32p, C.POINTER(C.c_int)]
        L.sb_server_remove_gather.argtypes = [P, C.c_int]
        L.sb_server_stats.argtypes = [P, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64), C.POINTER(C.c_size_t)]
        L.sb_server_destroy.argtypes = [P]
        L.sb_server_destroy.restype = None
        L.sb_ctx_set_option.argtypes = [P, C.c_int, C.c_int]
        L.sb_ctx_get_option.argtypes = [P, C.c_int, C.POINTER(C.c_int)]
        L.sb_graph_create.argtypes = [P, P, C.c_int, P, C.c_int, C.c_int, P, C.POINTER(P)]
        L.sb_graph_launch.argtypes = [P]
        L.sb_search_streamed.argtypes = [P, C.c_int, P, C.c_int, i32p, C.c_int, C.c_int, P, C.POINTER(C.c_float),
                                         C.c_int, i32p, i32p, C.POINTER(HostResult), C.POINTER(P)]
        L.sb_graph_destroy.argtypes = [P]
        L.sb_graph_destroy.restype = None
        L.sb_cmp_run_noalloc.argtypes = [P, P, C.POINTER(CmpParams), C.c_int, C.c_int, C.POINTER(DeviceArrays)]
        L.sb_device_fp32_fma_rate.argtypes = [C.c_int, C.POINTER(C.c_double)]
        L.sb_device_window_rates.argtypes = [C.c_int, C.POINTER(C.c_double)]
        L.sb_server_set_deadline.argtypes = [P, C.c_double]
        _lib = L
    return _lib


def check(rc):
    if rc != 0:
        L = lib()
        raise SbError("%s: %s" % (L.sb_strerror(rc).decode(), L.sb_last_error().decode()))


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _ptr(a, t):
    return a.ctypes.data_as(C.POINTER(t))


def fp32_fma_rate(device=0):
    """sb_device_fp32_fma_rate: the device's sustained FP32 FMA rate, TFLOP/s."""
    t = C.c_double()
    check(lib().sb_device_fp32_fma_rate(device, C.byref(t)))
    return t.value


WINDOW_FORMS = ("fp32_sub_fma", "fp16_fma_mix", "fp16_dot2", "fp32_pair_fma")


def window_rates(device=0):
    """sb_device_window_rates: interpolated window samples per second for the
    four instruction forms of the search's inner step (WINDOW_FORMS)."""
    r = (C.c_double * 4)()
    check(lib().sb_device_window_rates(device, r))
    return dict(zip(WINDOW_FORMS, list(r)))


# ---------------------------------------------------------------- host side
def synth(kind, ncdp, fold, ns, dt_us, seed, h_max=1500.0, h_jitter=0.0, regular_offsets=False, h_step=25.0,
          cdp_spacing=12.5, mid_half_width=50.0, nevents=8, freq=25.0, snr=5.0, aph=0):
    """Seeded synthetic gather: returns (cdp_off, geom, samples)."""
    L = lib()
    fmin, fmax = (fold, fold) if isinstance(fold, int) else fold
    spec = SynthSpec(kind, ncdp, fmin, fmax, ns, dt_us, cdp_spacing, h_max, h_jitter, int(regular_offsets), h_step,
                     mid_half_width, nevents, freq, snr, aph, seed)
    nc, nt = C.c_int(), C.c_int()
    check(L.sb_synth(C.byref(spec), C.byref(nc), C.byref(nt), None, None, None))
    cdp_off = np.zeros(nc.value + 1, np.int32)
    geom = np.zeros(nt.value, GEOM_DTYPE)
    samples = np.zeros((nt.value, ns), np.float32)
    check(L.sb_synth(C.byref(spec), C.byref(nc), C.byref(nt), _ptr(cdp_off, C.c_int32), geom.ctypes.data,
                     _ptr(samples, C.c_float)))
    return cdp_off, geom, samples


def neighbors(first_geom, apm):
    L = lib()
    g = np.ascontiguousarray(first_geom, GEOM_DTYPE)
    n = len(g)
    nn = C.c_int()
    check(L.sb_crs_neighbors(n, g.ctypes.data, apm, None, None, C.byref(nn)))
    off = np.zeros(n + 1, np.int32)
    nbr = np.zeros(max(nn.value, 1), np.int32)
    check(L.sb_crs_neighbors(n, g.ctypes.data, apm, _ptr(off, C.c_int32), _ptr(nbr, C.c_int32), C.byref(nn)))
    return off, nbr[:nn.value]


def write_traces(path, geom, samples, dt_us):
    L = lib()
    g = np.ascontiguousarray(geom, GEOM_DTYPE)
    s = np.ascontiguousarray(samples, np.float32)
    check(L.sb_su_write_traces(path.encode(), len(g), s.shape[1], dt_us, g.ctypes.data, _ptr(s, C.c_float)))


class SuGather:
    """A binned SU file (sb_su_load), or a SEG-Y file with segy=True
    (sb_segy_load)."""

    def __init__(self, path, method, aph, apm=0, segy=False):
        L = lib()
        h = C.c_void_p()
        check((L.sb_segy_load if segy else L.sb_su_load)(path.encode(), method, aph, apm, C.byref(h)))
        self.h = h
        v = [C.c_int() for _ in range(5)]
        check(L.sb_su_info(h, *[C.byref(x) for x in v]))
        self.ncdp, self.ntr, self.ns, self.dt_us, nn = [x.value for x in v]
        self.cdp_off = np.ctypeslib.as_array(L.sb_su_cdp_offsets(h), (self.ncdp + 1,)).copy()
        self.samples = np.ctypeslib.as_array(L.sb_su_samples(h), (self.ntr, self.ns)).copy()
        gp = L.sb_su_geom(h)
        self.geom = np.frombuffer((C.c_char * (self.ntr * GEOM_DTYPE.itemsize)).from_address(gp),
                                  GEOM_DTYPE).copy()
        if method == METHOD_CRS:
            self.nbr_off = np.ctypeslib.as_array(L.sb_su_nbr_offsets(h), (self.ncdp + 1,)).copy()
            self.nbr = (np.ctypeslib.as_array(L.sb_su_nbr(h), (nn,)).copy() if nn else np.zeros(0, np.int32))
        else:
            self.nbr_off = self.nbr = None

    def write_section(self, path, values, rewrite_header):
        v = np.ascontiguousarray(values, np.float32).reshape(self.ncdp, self.ns)
        check(lib().sb_su_write_section(self.h, path.encode(), _ptr(v, C.c_float), int(rewrite_header)))

    def close(self):
        if self.h:
            lib().sb_su_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# -------------------------------------------------------------- device side
class Context:
    def __init__(self, device=0, hbm_budget=0):
        self.h = C.c_void_p()
        check(lib().sb_ctx_create(device, hbm_budget, C.byref(self.h)))

    def memory(self):
        u, b = C.c_size_t(), C.c_size_t()
        check(lib().sb_ctx_memory(self.h, C.byref(u), C.byref(b)))
        return u.value, b.value

    def sync(self):
        check(lib().sb_ctx_sync(self.h))

    def timing(self):
        t = Timing()
        check(lib().sb_last_timing(self.h, C.byref(t)))
        return {k: getattr(t, k) for k, _ in Timing._fields_}

    def set_option(self, option, value):
        check(lib().sb_ctx_set_option(self.h, option, int(value)))

    def get_option(self, option):
        v = C.c_int()
        check(lib().sb_ctx_get_option(self.h, option, C.byref(v)))
        return v.value

    def tune_set(self, variant, pchunk=0):
        check(lib().sb_tune_set(self.h, variant, pchunk))

    def tune_get(self):
        v, p, n = C.c_int(), C.c_int(), C.c_int()
        check(lib().sb_tune_get(self.h, C.byref(v), C.byref(p), C.byref(n)))
        return v.value, p.value, n.value

    def close(self):
        if self.h:
            lib().sb_ctx_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class Gather:
    def __init__(self, ctx, cdp_off, geom, samples, dt_us, store=STORE_F32, nbr_off=None, nbr=None):
        self.ctx = ctx
        self.cdp_off = _i32(cdp_off)
        self.geom = np.ascontiguousarray(geom, GEOM_DTYPE)
        s = np.ascontiguousarray(samples, np.float32)
        self.ncdp = len(self.cdp_off) - 1
        self.ns = s.shape[1]
        self.h = C.c_void_p()
        no = _i32(nbr_off) if nbr_off is not None else None
        nb = _i32(nbr if len(nbr) else np.zeros(1, np.int32)) if nbr_off is not None else None
        check(lib().sb_gather_create(ctx.h, self.ncdp, _ptr(self.cdp_off, C.c_int32), self.ns, dt_us,
                                     self.geom.ctypes.data, _ptr(s, C.c_float), store,
                                     _ptr(no, C.c_int32) if no is not None else None,
                                     _ptr(nb, C.c_int32) if nb is not None else None, C.byref(self.h)))

    @classmethod
    def adopt(cls, ctx, handle, cdp_off, geom, ncdp, ns):
        """Wraps a gather the library created (sb_search_streamed's keep)."""
        self = cls.__new__(cls)
        self.ctx, self.cdp_off, self.geom, self.ncdp, self.ns, self.h = ctx, cdp_off, geom, ncdp, ns, handle
        return self

    def info(self):
        v = [C.c_int() for _ in range(4)]
        b = C.c_size_t()
        check(lib().sb_gather_info(self.h, *[C.byref(x) for x in v], C.byref(b)))
        return dict(ncdp=v[0].value, ntr=v[1].value, ns=v[2].value, store=v[3].value, bytes=b.value)

    def close(self):
        if self.h:
            lib().sb_gather_destroy(self.h)
            self.h = None


class DeviceResult:
    def __init__(self, ctx, ncdp, ns):
        self.ctx, self.ncdp, self.ns = ctx, ncdp, ns
        self.h = C.c_void_p()
        check(lib().sb_result_create(ctx.h, ncdp, ns, C.byref(self.h)))

    def download(self, contiguous=False):
        """The six result arrays.  contiguous: views of one host buffer laid
        out as the device arrays are (the library then makes one copy)."""
        keys = ("index", "a", "b", "c", "coher", "stack")
        if contiguous:
            buf = np.zeros((6, self.ncdp, self.ns), np.float32)
            out = {k: buf[i].view(np.int32) if k == "index" else buf[i] for i, k in enumerate(keys)}
        else:
            out = {k: np.zeros((self.ncdp, self.ns), np.int32 if k == "index" else np.float32) for k in keys}
        hr = HostResult(*[_ptr(out[k], C.c_int32 if k == "index" else C.c_float) for k in keys])
        check(lib().sb_result_download(self.ctx.h, self.h, C.byref(hr)))
        return out

    def close(self):
        if self.h:
            lib().sb_result_destroy(self.h)
            self.h = None


def cmp_params(c0, c1, nc, tau, mode=MODE_FAST):
    return CmpParams(c0, c1, nc, tau, mode)


def crs_params(a0, a1, na, b0, b1, nb, c0, c1, nc, tau, mode=MODE_FAST):
    return CrsParams(a0, a1, b0, b1, c0, c1, na, nb, nc, tau, mode)


def de_params(a0, a1, b0, b1, c0, c1, tau, pop=32, ngen=30, F=0.5, CR=0.9, seed=1, mode=MODE_FAST):
    return DeParams(a0, a1, b0, b1, c0, c1, tau, pop, ngen, F, CR, seed, mode)


def run(ctx, gather, params, result, cdp_range=None):
    """Queues one search (CMP, CRS or CRS-DE by params type) into result,
    over all CDPs or over cdp_range = (begin, end)."""
    L = lib()
    kind = "cmp" if isinstance(params, CmpParams) else "crs" if isinstance(params, CrsParams) else "crs_de"
    if cdp_range is None:
        check(getattr(L, "sb_%s_run" % kind)(ctx.h, gather.h, C.byref(params), result.h))
    else:
        check(getattr(L, "sb_%s_run_range" % kind)(ctx.h, gather.h, C.byref(params), int(cdp_range[0]),
                                                   int(cdp_range[1]), result.h))


def search(ctx, gather, params, cdp_range=None):
    n = gather.ncdp if cdp_range is None else cdp_range[1] - cdp_range[0]
    r = DeviceResult(ctx, n, gather.ns)
    try:
        run(ctx, gather, params, r, cdp_range)
        return r.download()
    finally:
        r.close()


def search_streamed(ctx, params, cdp_off, geom, samples, dt_us, store=STORE_F32, nbr_off=None, nbr=None,
                    keep=False):
    """sb_search_streamed: uploads the line in CDP batches overlapped with the
    search; returns the results (and the resident Gather when keep=True)."""
    off = _i32(cdp_off)
    g = np.ascontiguousarray(geom, GEOM_DTYPE)
    s = np.ascontiguousarray(samples, np.float32)
    ncdp, ns = len(off) - 1, s.shape[1]
    kind = METHOD_CMP if isinstance(params, CmpParams) else METHOD_CRS if isinstance(params, CrsParams) \
        else GRAPH_CRS_DE
    no = nb = None
    if nbr_off is not None:
        no, nb = _i32(nbr_off), _i32(nbr if len(nbr) else [0])
    out = {k: np.zeros((ncdp, ns), np.int32 if k == "index" else np.float32)
           for k in ("index", "a", "b", "c", "coher", "stack")}
    hr = HostResult(*[_ptr(out[k], C.c_int32 if k == "index" else C.c_float)
                      for k in ("index", "a", "b", "c", "coher", "stack")])
    gh = C.c_void_p()
    check(lib().sb_search_streamed(ctx.h, kind, C.cast(C.pointer(params), C.c_void_p), ncdp, _ptr(off, C.c_int32), ns,
                                   dt_us, g.ctypes.data, _ptr(s, C.c_float), store,
                                   None if no is None else _ptr(no, C.c_int32),
                                   None if nb is None else _ptr(nb, C.c_int32), C.byref(hr),
                                   C.byref(gh) if keep else None))
    if not keep:
        return out
    return out, Gather.adopt(ctx, gh, off, g, ncdp, ns)


class Graph:
    """sb_graph: one search captured into a HIP graph; launch() replays it
    into `result` (download with result.download())."""

    def __init__(self, ctx, gather, params, result, cdp_range=None):
        kind = METHOD_CMP if isinstance(params, CmpParams) else METHOD_CRS if isinstance(params, CrsParams) \
            else GRAPH_CRS_DE
        cb, ce = (0, gather.ncdp) if cdp_range is None else cdp_range
        self.ctx, self._keep = ctx, (gather, params, result)
        self.h = C.c_void_p()
        check(lib().sb_graph_create(ctx.h, gather.h, kind, C.cast(C.pointer(params), C.c_void_p), int(cb), int(ce),
                                    result.h, C.byref(self.h)))

    def launch(self):
        check(lib().sb_graph_launch(self.h))

    def close(self):
        if self.h:
            lib().sb_graph_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def tune(ctx, gather, params, reps=2):
    method = METHOD_CMP if isinstance(params, CmpParams) else METHOD_CRS
    check(lib().sb_tune(ctx.h, gather.h, method, C.cast(C.pointer(params), C.c_void_p), reps))
    return ctx.tune_get()


class Server:
    """sb_server: resident request loop.  Arrays passed to submit() are kept
    referenced here until wait() returns (the C side reads them in place)."""

    def __init__(self, device=0, hbm_budget=0, max_batch=64):
        self.h = C.c_void_p()
        check(lib().sb_server_create(device, hbm_budget, max_batch, C.byref(self.h)))
        self._keep = {}
        self._gathers = {}

    def submit(self, cdp_off, geom, samples, dt_us, params, store=STORE_F32, nbr_off=None, nbr=None):
        off = _i32(cdp_off)
        g = np.ascontiguousarray(geom, GEOM_DTYPE)
        s = np.ascontiguousarray(samples, np.float32)
        req = Request()
        req.ncdp, req.ns, req.dt_us, req.store = len(off) - 1, s.shape[1], dt_us, store
        req.cdp_off = _ptr(off, C.c_int32)
        req.geom = g.ctypes.data
        req.samples = _ptr(s, C.c_float)
        keep = [off, g, s]
        if isinstance(params, CmpParams):
            req.method, req.cmp = METHOD_CMP, params
        else:
            req.method, req.crs = METHOD_CRS, params
            no, nb = _i32(nbr_off), _i32(nbr if len(nbr) else [0])
            req.nbr_off, req.nbr = _ptr(no, C.c_int32), _ptr(nb, C.c_int32)
            keep += [no, nb]
        t = C.c_uint64()
        check(lib().sb_server_submit(self.h, C.byref(req), C.byref(t)))
        self._keep[t.value] = (keep, req.ncdp, req.ns)
        return t.value

    def add_gather(self, cdp_off, geom, samples, dt_us, store=STORE_F32, nbr_off=None, nbr=None):
        """Uploads a binned line once; returns its id for submit_range()."""
        off = _i32(cdp_off)
        g = np.ascontiguousarray(geom, GEOM_DTYPE)
        s = np.ascontiguousarray(samples, np.float32)
        no = nb = None
        if nbr_off is not None:
            no, nb = _i32(nbr_off), _i32(nbr if len(nbr) else [0])
        gid = C.c_int()
        check(lib().sb_server_add_gather(self.h, len(off) - 1, _ptr(off, C.c_int32), s.shape[1], dt_us,
                                         g.ctypes.data, _ptr(s, C.c_float), store,
                                         None if no is None else _ptr(no, C.c_int32),
                                         None if nb is None else _ptr(nb, C.c_int32), C.byref(gid)))
        self._gathers[gid.value] = s.shape[1]
        return gid.value

    def remove_gather(self, gid):
        check(lib().sb_server_remove_gather(self.h, gid))
        self._gathers.pop(gid, None)

    def submit_range(self, gid, cdp_begin, cdp_end, params):
        """Queues a search of CDPs [cdp_begin, cdp_end) of a resident gather."""
        req = Request()
        req.gather_id, req.cdp_begin, req.cdp_end = gid, cdp_begin, cdp_end
        if isinstance(params, CmpParams):
            req.method, req.cmp = METHOD_CMP, params
        else:
            req.method, req.crs = METHOD_CRS, params
        t = C.c_uint64()
        check(lib().sb_server_submit(self.h, C.byref(req), C.byref(t)))
        self._keep[t.value] = ([], cdp_end - cdp_begin, self._gathers[gid])
        return t.value

    def wait(self, ticket, timeout_ms=-1.0):
        keep, ncdp, ns = self._keep[ticket]
        keys = ("index", "a", "b", "c", "coher", "stack")
        out = {k: np.zeros((ncdp, ns), np.int32 if k == "index" else np.float32) for k in keys}
        hr = HostResult(*[_ptr(out[k], C.c_int32 if k == "index" else C.c_float) for k in keys])
        st = RequestStats()
        rc = lib().sb_server_wait(self.h, ticket, timeout_ms, C.byref(hr), C.byref(st))
        self._keep.pop(ticket, None)
        stats = {k: getattr(st, k) for k, _ in RequestStats._fields_}
        if rc != 0:
            return None, dict(stats, error=lib().sb_strerror(rc).decode())
        return out, stats

    def set_deadline(self, max_queue_ms):
        """Requests not admitted within max_queue_ms fail (0: no deadline)."""
        check(lib().sb_server_set_deadline(self.h, float(max_queue_ms)))

    def stats(self):
        d, l, u = C.c_uint64(), C.c_uint64(), C.c_size_t()
        check(lib().sb_server_stats(self.h, C.byref(d), C.byref(l), C.byref(u)))
        return dict(done=d.value, launches=l.value, hbm_reserved=u.value)

    def close(self):
        if self.h:
            lib().sb_server_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
