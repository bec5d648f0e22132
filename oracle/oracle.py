"""ctypes binding of the C oracle (oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  The product never uses it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_DIR = os.path.join(HERE, "_ref")

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE, "oracle"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        d, i = C.c_double, C.c_int
        L.orc_cmp.argtypes = [i, P, i, i, P, P, d, d, i, d, P, P, P, P, P, P]
        L.orc_crs.argtypes = [i, P, i, i, P, P, P, P, d, d, i, d, d, i, d, d, i, d, P, P, P, P, P, P, P, P]
        L.orc_crs_range.argtypes = L.orc_crs.argtypes[:18] + [i, i] + L.orc_crs.argtypes[18:]
        L.orc_crs_range_az.argtypes = L.orc_crs.argtypes[:18] + [i, i, i] + L.orc_crs.argtypes[18:]
        L.orc_crs_de_range.argtypes = [i, P, i, i, P, P, P, P, d, d, d, d, d, d, d, i, i, d, d, C.c_uint64, i, i,
                                       P, P, P, P, P]
        L.orc_cmp_f64.argtypes = [i, P, i, i, P, P, d, d, i, d, P, P, P, P]
        L.orc_crs_eval_f64.argtypes = [i, P, i, i, P, P, P, P, d, d, d, d, i, P, P]
        L.orc_crs_de.argtypes = [i, P, i, i, P, P, P, P, d, d, d, d, d, d, d, i, i, d, d, C.c_uint64, P, P, P, P, P]
        L.orc_de_history.argtypes = [i, P, i, i, P, P, P, P, d, d, d, d, d, d, d, i, i, d, d, C.c_uint64, i, i,
                                     P, P, P, P, P, P]
        L.orc_de_hash.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32]
        L.orc_de_hash.restype = C.c_uint32
        _lib = L
    return _lib


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def _p(a):
    return a.ctypes.data if a is not None else None


def _geom(g):
    g = np.ascontiguousarray(g)
    assert g.dtype.itemsize == 24
    return g


def cmp(cdp_off, geom, samples, dt_us, c0, c1, nc, tau, panel=False):
    cdp_off, samples, geom = _c(cdp_off, np.int32), _c(samples, np.float32), _geom(geom)
    ncdp, ns = len(cdp_off) - 1, samples.shape[1]
    out = {k: np.zeros((ncdp, ns), np.int32 if k == "index" else np.float32) for k in ("index", "c", "coher", "stack")}
    ps = np.zeros((ncdp, ns, nc), np.float32) if panel else None
    pk = np.zeros((ncdp, ns, nc), np.float32) if panel else None
    rc = lib().orc_cmp(ncdp, _p(cdp_off), ns, dt_us, _p(geom), _p(samples), c0, c1, nc, tau, _p(out["index"]),
                       _p(out["c"]), _p(out["coher"]), _p(out["stack"]), _p(ps), _p(pk))
    assert rc == 0
    if panel:
        out["panel_sem"], out["panel_stk"] = ps, pk
    return out


def crs(cdp_off, geom, samples, dt_us, nbr_off, nbr, a0, a1, na, b0, b1, nb, c0, c1, nc, tau, panel=False,
        cdp_range=None, az_mdeg=-1):
    """orc_crs; cdp_range = (begin, end) searches only those CDPs (rows of the
    outputs), their supergathers reaching any CDP of the gather; az_mdeg >= 0
    uses the signed projection of the midpoint displacement onto that
    azimuth (orc_crs_range_az)."""
    cdp_off, samples, geom = _c(cdp_off, np.int32), _c(samples, np.float32), _geom(geom)
    nbr_off, nbr = _c(nbr_off, np.int32), _c(nbr if len(nbr) else [0], np.int32)
    ncdp, ns = len(cdp_off) - 1, samples.shape[1]
    cb, ce = (0, ncdp) if cdp_range is None else cdp_range
    npar = na * nb * nc
    out = {k: np.zeros((ce - cb, ns), np.int32 if k == "index" else np.float32)
           for k in ("index", "a", "b", "c", "coher", "stack")}
    ps = np.zeros((ce - cb, ns, npar), np.float32) if panel else None
    pk = np.zeros((ce - cb, ns, npar), np.float32) if panel else None
    rc = lib().orc_crs_range_az(ncdp, _p(cdp_off), ns, dt_us, _p(geom), _p(samples), _p(nbr_off), _p(nbr), a0, a1,
                                na, b0, b1, nb, c0, c1, nc, tau, az_mdeg, cb, ce, _p(out["index"]), _p(out["a"]),
                                _p(out["b"]), _p(out["c"]), _p(out["coher"]), _p(out["stack"]), _p(ps), _p(pk))
    assert rc == 0
    if panel:
        out["panel_sem"], out["panel_stk"] = ps, pk
    return out


def cmp_f64(cdp_off, geom, samples, dt_us, c0, c1, nc, tau, panel=False):
    cdp_off, samples, geom = _c(cdp_off, np.int32), _c(samples, np.float32), _geom(geom)
    ncdp, ns = len(cdp_off) - 1, samples.shape[1]
    idx = np.zeros((ncdp, ns), np.int32)
    coh = np.zeros((ncdp, ns), np.float64)
    stk = np.zeros((ncdp, ns), np.float64)
    ps = np.zeros((ncdp, ns, nc), np.float64) if panel else None
    rc = lib().orc_cmp_f64(ncdp, _p(cdp_off), ns, dt_us, _p(geom), _p(samples), c0, c1, nc, tau, _p(idx), _p(coh),
                           _p(stk), _p(ps))
    assert rc == 0
    out = {"index": idx, "coher": coh, "stack": stk}
    if panel:
        out["panel_sem"] = ps
    return out


def crs_eval_f64(cdp, cdp_off, geom, samples, dt_us, nbr_off, nbr, a, b, c, tau, t0):
    cdp_off, samples, geom = _c(cdp_off, np.int32), _c(samples, np.float32), _geom(geom)
    nbr_off, nbr = _c(nbr_off, np.int32), _c(nbr if len(nbr) else [0], np.int32)
    sem, stk = C.c_double(), C.c_double()
    lib().orc_crs_eval_f64(cdp, _p(cdp_off), samples.shape[1], dt_us, _p(geom), _p(samples), _p(nbr_off), _p(nbr),
                           a, b, c, tau, t0, C.byref(sem), C.byref(stk))
    return sem.value, stk.value


def crs_de(cdp_off, geom, samples, dt_us, nbr_off, nbr, a0, a1, b0, b1, c0, c1, tau, pop, ngen, F, CR, seed,
           cdp_range=None):
    cdp_off, samples, geom = _c(cdp_off, np.int32), _c(samples, np.float32), _geom(geom)
    nbr_off, nbr = _c(nbr_off, np.int32), _c(nbr if len(nbr) else [0], np.int32)
    ncdp, ns = len(cdp_off) - 1, samples.shape[1]
    cb, ce = (0, ncdp) if cdp_range is None else cdp_range
    out = {k: np.zeros((ce - cb, ns), np.float32) for k in ("a", "b", "c", "coher", "stack")}
    rc = lib().orc_crs_de_range(ncdp, _p(cdp_off), ns, dt_us, _p(geom), _p(samples), _p(nbr_off), _p(nbr), a0, a1, b0,
                                b1, c0, c1, tau, pop, ngen, F, CR, seed, cb, ce, _p(out["a"]), _p(out["b"]),
                                _p(out["c"]), _p(out["coher"]), _p(out["stack"]))
    assert rc == 0
    return out


def de_history(cdp_off, geom, samples, dt_us, nbr_off, nbr, a0, a1, b0, b1, c0, c1, tau, pop, ngen, F, CR, seed,
               cdp, t0):
    """Trajectory of the CRS-DE twin at one (cdp, t0): pop_x [ngen+1][pop][3]
    and pop_f [ngen+1][pop] (population after g generations, c scaled by 1e6),
    trial_x [ngen][pop][3], trial_f [ngen][pop], picks [ngen][pop][4] = (r1, r2,
    r3, jr) of the trials made from pop_x[g], and best = (a, b, c scaled,
    fitness, stack) as reported."""
    cdp_off, samples, geom = _c(cdp_off, np.int32), _c(samples, np.float32), _geom(geom)
    nbr_off, nbr = _c(nbr_off, np.int32), _c(nbr if len(nbr) else [0], np.int32)
    ncdp, ns = len(cdp_off) - 1, samples.shape[1]
    h = {"pop_x": np.zeros((ngen + 1, pop, 3), np.float32), "pop_f": np.zeros((ngen + 1, pop), np.float32),
         "trial_x": np.zeros((max(ngen, 1), pop, 3), np.float32), "trial_f": np.zeros((max(ngen, 1), pop), np.float32),
         "picks": np.zeros((max(ngen, 1), pop, 4), np.int32), "best": np.zeros(5, np.float32)}
    rc = lib().orc_de_history(ncdp, _p(cdp_off), ns, dt_us, _p(geom), _p(samples), _p(nbr_off), _p(nbr), a0, a1, b0,
                              b1, c0, c1, tau, pop, ngen, F, CR, seed, cdp, t0, _p(h["pop_x"]), _p(h["pop_f"]),
                              _p(h["trial_x"]), _p(h["trial_f"]), _p(h["picks"]), _p(h["best"]))
    assert rc == 0
    for k in ("trial_x", "trial_f", "picks"):
        h[k] = h[k][:ngen]
    return h


def de_hash(seed, cdp, t0, gen, member, draw):
    return lib().orc_de_hash(seed, cdp, t0, gen, member, draw)


# ------------------------------------------------------------ SU reading
SU_HDR = 240


def read_su(path):
    """Reads an SU file with numpy: (headers [n][240] uint8, data [n][ns] float32)."""
    raw = np.fromfile(path, dtype=np.uint8)
    if raw.size == 0:
        return np.zeros((0, SU_HDR), np.uint8), np.zeros((0, 0), np.float32)
    ns = int(raw[114:116].view(np.uint16)[0])
    rec = SU_HDR + 4 * ns
    assert raw.size % rec == 0, "ragged SU file"
    r = raw.reshape(-1, rec)
    return r[:, :SU_HDR].copy(), r[:, SU_HDR:].copy().view(np.float32)


def hdr_i32(h, off):
    return h[:, off:off + 4].copy().view(np.int32).ravel()


def hdr_i16(h, off):
    return h[:, off:off + 2].copy().view(np.int16).ravel()
