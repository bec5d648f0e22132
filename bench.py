"""Benchmark: CMP semblance search at production size (2000 CDPs, fold 30-120,
ns 1501, dt 2 ms, nc 256, tau 2 ms) on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

A step is one full CMP search over the resident gather: the fused search
kernel over every (CDP, t0 tile, parameter chunk) plus the winner
re-evaluation / output kernel.  value = semblance evaluations per second
(CDP x t0 x param x trace, the reference programs' own count) with FP32
trace storage; the FP16-store figure, the 1-CDP panel latency and the CPU
baseline ride along on the same JSON line.
"""
import argparse
import importlib.util
import json
import os
import re
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
SEED = 0x5EB1A9CE
FLOP_PER_TUPLE_CMP = 5 + 7 * 3        # traveltime 5 + 7 per window sample, w = 3
PEAK_FP32_VECTOR_TFLOPS = 157.3       # MI355X published FP32 vector peak
CFG = dict(ncdp=2000, fold=(30, 120), ns=1501, dt_us=2000, h_max=1500.0, h_jitter=5.0, nevents=8, snr=5.0,
           c0=1.98e-7, c1=1.77e-6, nc=256, tau=0.002, aph=2000)
FLOP_PER_TUPLE_CRS = 10 + 7 * 3
BOX = dict(a0=-0.7e-3, a1=0.7e-3, b0=-1e-7, b1=1e-7, c0=1.975e-7, c1=1.384e-6)
# CRS grid: 500 CDPs, 1000 traces each (midpoints +-50 m, half-offsets 0-2600 m), aph 600, apm 50, 11^3 grid.
CRS_CFG = dict(ncdp=500, fold=(1000, 1000), ns=1001, dt_us=4000, h_max=2600.0, mid=50.0, nevents=6, snr=3.0,
               aph=600, apm=50, n=11, tau=0.004)
# CRS-DE: 2000 CDPs of fold 200-1000, otherwise as the CRS grid; population 32, 30 generations.
DE_CFG = dict(ncdp=2000, fold=(200, 1000), ns=1001, dt_us=4000, h_max=2600.0, mid=50.0, nevents=6, snr=3.0,
              aph=600, apm=50, tau=0.004, pop=32, ngen=30, F=0.5, CR=0.9)


def _load(name, path):
    spec = importlib.util.spec_from_file_location(name, path)
    m = importlib.util.module_from_spec(spec)
    sys.modules[name] = m
    spec.loader.exec_module(m)
    return m


def timed_steps(sb, ctx, g, prm, steps, warmup):
    r = sb.DeviceResult(ctx, g.ncdp, g.ns)
    for _ in range(warmup):
        sb.run(ctx, g, prm, r)
    ctx.sync()
    kernel_ms = []
    t0 = time.perf_counter()
    for _ in range(steps):
        sb.run(ctx, g, prm, r)
        kernel_ms.append(ctx.timing())  # reads the HIP events of this step
    ctx.sync()
    wall = time.perf_counter() - t0
    r.close()
    return wall, kernel_ms


def panel_latency(sb, ctx, cdp_off, geom, samples, n=100):
    """Latency of single-CDP CMP panels through the full request path:
    upload, search, download (ms)."""
    prm = sb.cmp_params(CFG["c0"], CFG["c1"], CFG["nc"], CFG["tau"])
    lat = []
    folds = np.diff(cdp_off)
    order = np.argsort(-folds)
    for i in range(n):
        cdp = int(order[i % len(order)])
        r0, r1 = cdp_off[cdp], cdp_off[cdp + 1]
        t = time.perf_counter()
        g = sb.Gather(ctx, np.array([0, r1 - r0], np.int32), geom[r0:r1], samples[r0:r1], CFG["dt_us"])
        sb.search(ctx, g, prm)
        g.close()
        lat.append((time.perf_counter() - t) * 1e3)
    lat = np.array(lat[5:])
    return float(np.percentile(lat, 50)), float(np.percentile(lat, 99))


# Launch shapes of the search in the library's variant numbering: rows per
# lane R, parameters per pass CC, waves per workgroup; variants >= 4 are
# LDS-staged (k_search_lds: one group of CC parameters per wave and pass).
GLOBAL_SHAPES = {0: (2, 4), 1: (1, 4), 2: (4, 2), 3: (2, 2)}
LDS_SHAPES = {v: (1, 4, 8) for v in range(4, 13)}


def effective_pchunk(res):
    """Parameters per workgroup of a timed run: the tuned split, or for the
    automatic split (0) the one the launch used, from its workgroup count
    (CDPs x t0 tiles x parameter chunks, chunks of equal length)."""
    v, pc = res["tune"][0], res["tune"][1]
    if pc > 0:
        return pc
    if v in GLOBAL_SHAPES:
        R, CC = GLOBAL_SHAPES[v]
        unit = 4 * CC  # 4 waves per workgroup
    else:
        R, CC, NW = LDS_SHAPES[v]
        unit = NW * CC
    ntiles = -(-CFG["ns"] // (64 * R))
    nchunks = max(1, res["blocks"] // (CFG["ncdp"] * ntiles))
    return -(-(-(-CFG["nc"] // nchunks)) // unit) * unit


def skipped_fraction(cdp_off, geom, variant, pchunk, nblocks):
    """Fraction of the reference's evaluations the CMP fast search never runs:
    a wave skips a parameter group when each of its (t0, c) already has two
    invalid traces -- the second-largest half-offset's t^2 = t0^2 + c h
    reaches (ns - tau - 1)^2 -- in a tile past tau.  Those (t0, c) have no
    result in any arithmetic.  Estimated in float64 from the geometry and the
    launch shape (the kernel decides in FP32)."""
    if variant in GLOBAL_SHAPES:
        R, CC = GLOBAL_SHAPES[variant]
        NW = 0
    elif variant in LDS_SHAPES:
        R, CC, NW = LDS_SHAPES[variant]
    else:
        return 0.0
    ns, nc, dt = CFG["ns"], CFG["nc"], CFG["dt_us"] * 1e-6
    tau = int(np.float32(CFG["tau"]) * np.float32(1.0 / np.float32(dt)))
    hi = ns - tau - 1
    hi2 = float(np.float32(hi * hi * (1.0 - 2.0 ** -20)))
    c0 = np.float32(np.float64(np.float32(CFG["c0"])) * 1e6)
    c1 = np.float32(np.float64(np.float32(CFG["c1"])) * 1e6)
    inc = np.float32((c1 - c0) * np.float32(1.0 / nc))
    c = (c0 + inc * np.arange(nc, dtype=np.float32)).astype(np.float64)
    s = np.where(geom["scalco"] == 0, 1.0, np.where(geom["scalco"] > 0, geom["scalco"], 1.0 / geom["scalco"]))
    hx = (geom["gx"].astype(np.float64) - geom["sx"]) * s
    hy = (geom["gy"].astype(np.float64) - geom["sy"]) * s
    h = (0.25 * (hx * hx + hy * hy) / 1e6 / dt / dt).astype(np.float32).astype(np.float64)
    tile = 64 * R
    ntiles = (ns + tile - 1) // tile
    ncdp = len(cdp_off) - 1
    if pchunk <= 0:  # automatic split: as many chunks as the launch used
        pchunk = max(1, -(-nc // max(1, nblocks // (ncdp * ntiles))))
    # parameter groups of one block's chunk, per wave, CC at a time: a
    # contiguous run per wave (global-memory kernel) or CC-groups dealt to
    # the waves in turn (LDS-staged kernel)
    groups = []
    for p0 in range(0, nc, pchunk):
        p1 = min(nc, p0 + pchunk)
        if NW:
            for p in range(p0, p1, CC):
                groups.append((p, min(p + CC, p1)))
            continue
        per = -(-(p1 - p0) // (4 * CC)) * CC
        for w in range(4):
            pw0, pw1 = p0 + w * per, min(p0 + (w + 1) * per, p1)
            for p in range(pw0, pw1, CC):
                groups.append((p, min(p + CC, pw1)))
    starts = np.arange(ntiles) * tile
    rows = np.minimum(starts + tile, ns) - starts
    skipped = total = 0.0
    for i in range(ncdp):
        hs = np.sort(h[cdp_off[i]:cdp_off[i + 1]])
        S = len(hs)
        total += float(S) * ns * nc
        if S < 2:
            continue
        # first t0 from which (t0, c) has no result: t0^2 >= hi2 - c h2
        t0star = np.ceil(np.sqrt(np.maximum(hi2 - c * hs[-2], 0.0)))
        for p, q in groups:
            gmax = t0star[p:q].max()
            dead = (starts >= gmax) & (starts > tau)
            skipped += float(S) * (q - p) * rows[dead].sum()
    return skipped / total


def cpu_baseline(sb, cdp_off, geom, samples, every=4):
    """The reference OpenMP program (oracle/_ref/cmp-omp2) on every `every`-th
    CDP of the same synthetic line, all host cores; its own INFO line."""
    ref = os.path.join(ROOT, "oracle", "_ref", "cmp-omp2")
    cores = len(os.sched_getaffinity(0))
    sel = np.arange(0, len(cdp_off) - 1, every)
    rows = np.concatenate([np.arange(cdp_off[i], cdp_off[i + 1]) for i in sel])
    tuples = float(len(rows)) * CFG["ns"] * CFG["nc"]
    sample = "every %dth CDP (%d CDPs, %d traces, %.3g evals)" % (every, len(sel), len(rows), tuples)
    with tempfile.TemporaryDirectory() as td:
        path = os.path.join(td, "sample.su")
        sb.write_traces(path, geom[rows], samples[rows], CFG["dt_us"])
        if os.path.exists(ref):
            cmd = [ref, "-c0", repr(CFG["c0"]), "-c1", repr(CFG["c1"]), "-nc", str(CFG["nc"]), "-aph",
                   str(CFG["aph"]), "-tau", repr(CFG["tau"]), "-i", path, "-v", "2"]
            t = time.perf_counter()
            p = subprocess.run(cmd, cwd=td, capture_output=True, text=True,
                               env=dict(os.environ, OMP_NUM_THREADS=str(cores), OMP_PROC_BIND="close"))
            wall = time.perf_counter() - t
            m = re.search(r"Semblances-Trace/s: ([0-9.eE+-]+)", p.stdout)
            if p.returncode == 0 and m:
                return dict(value=float(m.group(1)) * 1e9, unit="semblance evals/s", cores=cores, kind="reference",
                            sample=sample + "; reference INFO line (timer excludes its last CDP)",
                            wall_s=round(wall, 3), wall_value=tuples / wall)
        # No reference binary on this host: time the C restatement instead.
        orc = _load("sb_oracle", os.path.join(ROOT, "oracle", "oracle.py"))
        off = np.concatenate([[0], np.cumsum(np.diff(cdp_off)[sel])]).astype(np.int32)
        os.environ["OMP_NUM_THREADS"] = str(cores)
        t = time.perf_counter()
        orc.cmp(off, geom[rows], samples[rows], CFG["dt_us"], CFG["c0"], CFG["c1"], CFG["nc"], CFG["tau"])
        wall = time.perf_counter() - t
        return dict(value=tuples / wall, unit="semblance evals/s", cores=cores, kind="port", sample=sample)


def cli_cmp_200cdp(sb):
    """The reference's own CPU-runnable case end to end through the drop-in
    CLI: a 200-CDP line, fold 60 at half-offsets 0-1475 m (25 m steps), ns
    1001, dt 2 ms, nc 101, tau 2 ms, written as an SU file.  Runs the `cmp`
    program in fast and exact mode and, when built, the reference's cmp-omp2
    on the same file and cores; reports each INFO line's rate and wall time,
    and whether the exact-mode output files equal the reference's byte for
    byte."""
    import filecmp
    cores = len(os.sched_getaffinity(0))
    lib = os.path.join(ROOT, "ipdps-crs-cmp-code_amd", "lib", "cmp")
    ref = os.path.join(ROOT, "oracle", "_ref", "cmp-omp2")
    off, geom, s = sb.synth(0, 200, 60, 1001, 2000, SEED, h_max=1500.0, regular_offsets=True, h_step=25.0,
                            nevents=8, snr=5.0)
    args = ["-c0", "1.98e-7", "-c1", "1.77e-6", "-nc", "101", "-aph", "1500", "-tau", "0.002", "-v", "2"]
    out = dict(workload="CMP 200 CDPs fold 60 ns 1001 nc 101 (%.4g evals), SU file -> SU sections" %
               (float(len(geom)) * 1001 * 101))
    names = ("cmp.c.su", "cmp.coher.su", "cmp.stack.su")
    with tempfile.TemporaryDirectory() as td:
        path = os.path.join(td, "line.su")
        sb.write_traces(path, geom, s, 2000)
        runs = [("fast", lib, {"SB_MODE": "fast"}), ("exact", lib, {"SB_MODE": "exact"})]
        if os.path.exists(ref):
            runs.append(("reference", ref, {"OMP_NUM_THREADS": str(cores), "OMP_PROC_BIND": "close"}))
        for name, prog, env in runs:
            d = os.path.join(td, name)
            os.mkdir(d)
            t = time.perf_counter()
            p = subprocess.run([prog] + args + ["-i", path], cwd=d, capture_output=True, text=True,
                               env=dict(os.environ, **env))
            wall = time.perf_counter() - t
            m = re.search(r"Semblances-Trace/s: ([0-9.eE+-]+)", p.stdout)
            out[name] = dict(rc=p.returncode, info_evals_per_s=float(m.group(1)) * 1e9 if m else None,
                             wall_s=round(wall, 3))
        if "reference" in out and out["reference"]["rc"] == 0 and out["exact"]["rc"] == 0 and out["fast"]["rc"] == 0:
            out["reference"]["cores"] = cores
            out["exact_files_identical"] = all(
                filecmp.cmp(os.path.join(td, "exact", n), os.path.join(td, "reference", n), shallow=False)
                for n in names)
            a = np.fromfile(os.path.join(td, "fast", "cmp.c.su"), np.uint8)
            b = np.fromfile(os.path.join(td, "reference", "cmp.c.su"), np.uint8)
            out["fast_c_section_identical"] = bool(a.shape == b.shape and np.array_equal(a, b))
    return out


def crs_workload(sb, cfg, seed):
    cdp_off, geom, samples = sb.synth(1, cfg["ncdp"], cfg["fold"], cfg["ns"], cfg["dt_us"], seed, h_max=cfg["h_max"],
                                      mid_half_width=cfg["mid"], nevents=cfg["nevents"], snr=cfg["snr"],
                                      aph=cfg["aph"])
    nbr_off, nbr = sb.neighbors(geom[cdp_off[:-1]], cfg["apm"])
    return cdp_off, geom, samples, nbr_off, nbr


def crs_cpu_baseline(sb, orc, cdp_off, geom, samples, nbr_off, nbr, every=50, block=6):
    """CPU rate of the CRS grid search on all host cores.  Headline: the C
    restatement of the reference's double-precision search (kind "port") on
    every `every`-th CDP of the line with its full supergather (a stratified
    sample: line edges and interior alike).  Beside it, the reference crs-omp2
    on the first `block` CDPs written as an SU file (their supergathers
    clipped at that range), its own INFO line."""
    cores = len(os.sched_getaffinity(0))
    os.environ["OMP_NUM_THREADS"] = str(cores)
    b = BOX
    sel = list(range(0, len(cdp_off) - 1, every))
    sizes = [sum(int(cdp_off[j + 1] - cdp_off[j]) for j in nbr[nbr_off[i]:nbr_off[i + 1]]) for i in sel]
    tuples = float(sum(sizes)) * CRS_CFG["ns"] * CRS_CFG["n"] ** 3
    t = time.perf_counter()
    for i in sel:
        orc.crs(cdp_off, geom, samples, CRS_CFG["dt_us"], nbr_off, nbr, b["a0"], b["a1"], 11, b["b0"], b["b1"], 11,
                b["c0"], b["c1"], 11, CRS_CFG["tau"], cdp_range=(i, i + 1))
    wall = time.perf_counter() - t
    out = dict(value=tuples / wall, unit="semblance evals/s", cores=cores, kind="port",
               sample="every %dth CDP (%d CDPs, supergathers of %d-%d traces, %.3g evals) at full ns and grid, the "
                      "oracle's double-precision restatement of crs-omp2" % (every, len(sel), min(sizes), max(sizes),
                                                                             tuples))
    out["reference_block"] = _crs_reference_block(sb, cdp_off, geom, samples, block)
    return out


def _crs_reference_block(sb, cdp_off, geom, samples, n):
    """The reference crs-omp2 on the first n CDPs of the line (their
    supergathers restricted to that range), all host cores, its INFO line."""
    ref = os.path.join(ROOT, "oracle", "_ref", "crs-omp2")
    cores = len(os.sched_getaffinity(0))
    rows = np.arange(cdp_off[0], cdp_off[n])
    sample = "first %d CDPs (%d traces), full ns and grid" % (n, len(rows))
    if not os.path.exists(ref):
        return None
    with tempfile.TemporaryDirectory() as td:
        path = os.path.join(td, "sample.su")
        sb.write_traces(path, geom[rows], samples[rows], CRS_CFG["dt_us"])
        b = BOX
        cmd = [ref, "-a0", repr(b["a0"]), "-a1", repr(b["a1"]), "-na", "11", "-b0", repr(b["b0"]), "-b1", repr(b["b1"]),
               "-nb", "11", "-c0", repr(b["c0"]), "-c1", repr(b["c1"]), "-nc", "11", "-aph", str(CRS_CFG["aph"]),
               "-apm", str(CRS_CFG["apm"]), "-tau", repr(CRS_CFG["tau"]), "-i", path, "-v", "2"]
        p = subprocess.run(cmd, cwd=td, capture_output=True, text=True,
                           env=dict(os.environ, OMP_NUM_THREADS=str(cores), OMP_PROC_BIND="close"))
        m = re.search(r"Semblances-Trace/s: ([0-9.eE+-]+)", p.stdout)
        if p.returncode != 0 or not m:
            return None
        return dict(value=float(m.group(1)) * 1e9, unit="semblance evals/s", cores=cores, kind="reference",
                    sample=sample + "; reference INFO line")


def de_cpu_baseline(orc, cdp_off, geom, samples, nbr_off, nbr, every=250):
    """The oracle's CPU twin of CRS-DE (OpenMP, FP32) on a stratified sample:
    every `every`-th CDP (from every/2), each at its full supergather and ns."""
    cores = len(os.sched_getaffinity(0))
    os.environ["OMP_NUM_THREADS"] = str(cores)
    c = DE_CFG
    ncdp = len(cdp_off) - 1
    cdps = list(range(every // 2, ncdp, every))
    folds = np.diff(cdp_off)
    sizes = [int(sum(folds[j] for j in nbr[nbr_off[i]:nbr_off[i + 1]])) for i in cdps]
    tuples = float(sum(sizes)) * c["ns"] * c["pop"] * (c["ngen"] + 1)
    t = time.perf_counter()
    for i in cdps:
        orc.crs_de(cdp_off, geom, samples, c["dt_us"], nbr_off, nbr, BOX["a0"], BOX["a1"], BOX["b0"], BOX["b1"],
                   BOX["c0"], BOX["c1"], c["tau"], c["pop"], c["ngen"], c["F"], c["CR"], 1, cdp_range=(i, i + 1))
    wall = time.perf_counter() - t
    return dict(value=tuples / wall, unit="semblance evals/s", cores=cores, kind="port",
                sample="every %dth CDP (%d CDPs, supergathers of %d-%d traces, %.3g evals) at full ns, oracle CPU "
                       "twin of CRS-DE, FP32" % (every, len(cdps), min(sizes), max(sizes), tuples))


def side_configs(sb, args):
    """CRS grid (500 CDPs) and CRS-DE (2000 CDPs) at their full sizes."""
    orc = _load("sb_oracle", os.path.join(ROOT, "oracle", "oracle.py"))
    out = crs_grid_side(sb, args, orc)
    out.update(crs_de_side(sb, args, orc))
    return out


def crs_grid_side(sb, args, orc, ncdp=500, lds=-1):
    """The CRS grid configuration (the first ncdp CDPs of its line), 11^3
    grid: the default search (lds = SB_OPT_CRS_LDS), and the global-memory
    search at its tuned launch shape beside it."""
    out = {}
    print("bench: generating CRS grid workload", flush=True)
    w = crs_workload(sb, dict(CRS_CFG, ncdp=ncdp), SEED + 2)
    cdp_off, geom, samples, nbr_off, nbr = w
    b = BOX
    prm = sb.crs_params(b["a0"], b["a1"], 11, b["b0"], b["b1"], 11, b["c0"], b["c1"], 11, CRS_CFG["tau"],
                        sb.MODE_FAST)
    with sb.Context(0) as ctx:
        ctx.set_option(sb.OPT_CRS_LDS, lds)
        g = sb.Gather(ctx, cdp_off, geom, samples, CRS_CFG["dt_us"], nbr_off=nbr_off, nbr=nbr)
        if not args.no_tune:
            sb.tune(ctx, g, prm, reps=1)
        wall, tk = timed_steps(sb, ctx, g, prm, 3, 1)
        ms = float(np.mean([t["search_ms"] for t in tk]))
        tuples = tk[0]["tuples"]
        out["crs_grid"] = dict(workload="CRS %d CDPs fold 1000 (aph 600 -> %d traces) ns 1001 11x11x11 apm 50 FP32" %
                               (ncdp, len(geom)), value=tuples * 3 / wall, ms_per_step=wall / 3 * 1e3, tuples=tuples,
                               search_ms=ms, roofline_frac=tuples * FLOP_PER_TUPLE_CRS / (ms * 1e-3) / 1e12 /
                               PEAK_FP32_VECTOR_TFLOPS, lds_staged=tk[0]["lds_staged"])
        print("bench: crs grid %.1f ms (LDS-staged %d)" % (ms, tk[0]["lds_staged"]), flush=True)
        if lds != 0 and not args.crs_only:
            # the global-memory search (one window load per evaluation) at its tuned shape
            ctx.set_option(sb.OPT_CRS_LDS, 0)
            try:
                if not args.no_tune:
                    sb.tune(ctx, g, prm, reps=1)
                wall2, tk2 = timed_steps(sb, ctx, g, prm, 3, 1)
            finally:
                ctx.set_option(sb.OPT_CRS_LDS, lds)
            ms2 = float(np.mean([t["search_ms"] for t in tk2]))
            out["crs_grid"]["global"] = dict(value=tuples * 3 / wall2, search_ms=ms2,
                                             roofline_frac=tuples * FLOP_PER_TUPLE_CRS / (ms2 * 1e-3) / 1e12 /
                                             PEAK_FP32_VECTOR_TFLOPS, variant=tk2[0]["variant"])
            print("bench: crs grid global-memory search %.1f ms" % ms2, flush=True)
        g.close()
    if not args.no_cpu:
        print("bench: CRS CPU baseline", flush=True)
        out["crs_grid"]["cpu_baseline"] = crs_cpu_baseline(sb, orc, cdp_off, geom, samples, nbr_off, nbr)
    return out


def crs_de_side(sb, args, orc, ncdp=2000, stores=("f16", "f32"), lds=-1):
    """The CRS-DE configuration (the first ncdp CDPs of its line): each store
    on the default evaluator (lds = SB_OPT_DE_LDS), FP32 also on the
    global-memory one, 16-bit stores also read directly."""
    out = {}
    b = BOX
    print("bench: generating CRS-DE workload", flush=True)
    cdp_off, geom, samples, nbr_off, nbr = crs_workload(sb, dict(DE_CFG, ncdp=ncdp), SEED + 3)
    c = DE_CFG
    out["crs_de"] = dict(workload="CRS-DE %d CDPs fold 200-1000 (aph 600 -> %d traces) ns 1001 pop 32 ngen 30" %
                         (ncdp, len(geom)))
    with sb.Context(0) as ctx:
        ctx.set_option(sb.OPT_DE_LDS, lds)
        for store in stores:
            g = sb.Gather(ctx, cdp_off, geom, samples, c["dt_us"], store=sb.STORES[store], nbr_off=nbr_off, nbr=nbr)
            prm = sb.de_params(b["a0"], b["a1"], b["b0"], b["b1"], b["c0"], b["c1"], c["tau"], c["pop"], c["ngen"],
                               c["F"], c["CR"], 1, sb.MODE_FAST)
            wall, tk = timed_steps(sb, ctx, g, prm, 2, 1)
            ms = float(np.mean([t["search_ms"] for t in tk]))
            tuples = tk[0]["tuples"]
            out["crs_de"][store] = dict(value=tuples * 2 / wall, ms_per_step=wall / 2 * 1e3, tuples=tuples,
                                        search_ms=ms, roofline_frac=tuples * FLOP_PER_TUPLE_CRS / (ms * 1e-3) /
                                        1e12 / PEAK_FP32_VECTOR_TFLOPS, restored=tk[0]["restored"],
                                        lds_staged=tk[0]["lds_staged"], block_threads=tk[0]["block_threads"])
            print("bench: crs-de %s %.4g evals/s (%.1f ms, LDS-staged %d)" %
                  (store, out["crs_de"][store]["value"], ms, tk[0]["lds_staged"]), flush=True)
            if store == "f32" and lds != 0 and not args.de_only:
                # the global-memory evaluator (every window its own load) on the same step
                ctx.set_option(sb.OPT_DE_LDS, 0)
                try:
                    wall3, tk3 = timed_steps(sb, ctx, g, prm, 2, 1)
                finally:
                    ctx.set_option(sb.OPT_DE_LDS, lds)
                ms3 = float(np.mean([t["search_ms"] for t in tk3]))
                out["crs_de"]["f32_global"] = dict(value=tuples * 2 / wall3, search_ms=ms3,
                                                   roofline_frac=tuples * FLOP_PER_TUPLE_CRS / (ms3 * 1e-3) / 1e12 /
                                                   PEAK_FP32_VECTOR_TFLOPS)
                print("bench: crs-de f32 global-memory evaluator %.1f ms" % ms3, flush=True)
            if store != "f32" and not args.de_only:
                ctx.set_option(sb.OPT_RESTORE_FP32, 0)
                try:
                    wall2, tk2 = timed_steps(sb, ctx, g, prm, 2, 1)
                finally:
                    ctx.set_option(sb.OPT_RESTORE_FP32, 1)
                ms2 = float(np.mean([t["search_ms"] for t in tk2]))
                out["crs_de"][store + "_direct"] = dict(value=tuples * 2 / wall2, ms_per_step=wall2 / 2 * 1e3,
                                                        search_ms=ms2)
                print("bench: crs-de %s without the FP32 copy %.4g evals/s (%.1f ms)" %
                      (store, out["crs_de"][store + "_direct"]["value"], ms2), flush=True)
            g.close()
    if not args.no_cpu:
        cpu = de_cpu_baseline(orc, cdp_off, geom, samples, nbr_off, nbr)
        out["crs_de"]["cpu_baseline"] = cpu
        if "f16" in out["crs_de"]:
            out["crs_de"]["f16_vs_cpu"] = out["crs_de"]["f16"]["value"] / cpu["value"]
    if "f16" in out["crs_de"] and "f32" in out["crs_de"]:
        out["crs_de"]["f16_vs_f32"] = out["crs_de"]["f16"]["value"] / out["crs_de"]["f32"]["value"]
    return out


def serving_lines(sb, seed=SEED + 4):
    """The resident lines of the serving workload: a CMP line (96 CDPs of fold
    24-1000, ns 3001, cut to ns 501/1001/1501/2001/3001) and a CRS line (32
    CDPs, ns 1001, cut to 501/1001)."""
    print("bench: generating serving lines", flush=True)
    cmp_off, cmp_geom, cmp_s = sb.synth(0, 96, (24, 1000), 3001, 2000, seed, h_max=1500.0, h_jitter=5.0, nevents=8)
    crs_off, crs_geom, crs_s = sb.synth(1, 32, (24, 1000), 1001, 4000, seed + 1, h_max=800.0, mid_half_width=50.0,
                                        nevents=6)
    crs_no, crs_nb = sb.neighbors(crs_geom[crs_off[:-1]], 50)
    lines = {}
    for n in (501, 1001, 1501, 2001, 3001):
        lines[("cmp", n)] = (cmp_off, cmp_geom, np.ascontiguousarray(cmp_s[:, :n]), 2000, {})
    for n in (501, 1001):
        lines[("crs", n)] = (crs_off, crs_geom, np.ascontiguousarray(crs_s[:, :n]), 4000,
                             dict(nbr_off=crs_no, nbr=crs_nb))
    return lines


def serving_bench(sb, lines, seconds=4.0, rate=200.0, probe_rate=100.0, seed=SEED + 4, store=0, deadline_ms=0.0):
    """Serving loop under load on the resident lines of serving_lines():
    requests name a line and a CDP range.  Poisson arrivals at `rate` of CMP
    (70 %) and CRS (30 %) panel requests plus a probe stream of 1-CDP CMP
    panels (ns 1501, nc 256) at `probe_rate`; latency = server done - submit
    time.  CMP requests: 1-64 CDPs, nc 21-128.  CRS requests: 1-8 CDPs, 5-11
    values per parameter (the 21-128 per parameter of the full
    specification offers ~10^3 times one GPU's capacity at 60 CRS
    requests/s)."""
    rng = np.random.default_rng(seed)
    b = BOX
    events = []
    t = 0.0
    while t < seconds:
        t += rng.exponential(1.0 / rate)
        if rng.random() < 0.7:
            n = int(rng.integers(1, 65))
            i0 = int(rng.integers(0, 96 - n + 1))
            prm = sb.cmp_params(CFG["c0"], CFG["c1"], int(rng.integers(21, 129)), 0.002)
            events.append((t, "cmp", ("cmp", int(rng.choice([501, 1001, 2001, 3001]))), i0, i0 + n, prm))
        else:
            n = int(rng.integers(1, 9))
            i0 = int(rng.integers(0, 32 - n + 1))
            n3 = [int(x) for x in rng.integers(5, 12, 3)]
            prm = sb.crs_params(b["a0"], b["a1"], n3[0], b["b0"], b["b1"], n3[1], b["c0"], b["c1"], n3[2], 0.004)
            events.append((t, "crs", ("crs", int(rng.choice([501, 1001]))), i0, i0 + n, prm))
    t = 0.0
    while t < seconds:
        t += rng.exponential(1.0 / probe_rate)
        i0 = int(rng.integers(0, 96))
        events.append((t, "probe", ("cmp", 1501), i0, i0 + 1, sb.cmp_params(CFG["c0"], CFG["c1"], 256, 0.002)))
    events.sort(key=lambda e: e[0])
    tickets = []
    with sb.Server(0, max_batch=64) as srv:
        gid = {key: srv.add_gather(l[0], l[1], l[2], l[3], store=store, **l[4]) for key, l in lines.items()}
        resident = srv.stats()["hbm_reserved"]
        # warm-up: one request of each kind
        for ev in events[:4]:
            srv.wait(srv.submit_range(gid[ev[2]], ev[3], ev[4], ev[5]))
        srv.set_deadline(deadline_ms)
        print("bench: serving %d requests over %.1f s (%.0f + %.0f req/s offered)" %
              (len(events), seconds, rate, probe_rate), flush=True)
        start = time.perf_counter()
        for ev in events:
            d = ev[0] - (time.perf_counter() - start)
            if d > 0:
                time.sleep(d)
            tickets.append((ev[1], srv.submit_range(gid[ev[2]], ev[3], ev[4], ev[5])))
        lat = {"cmp": [], "crs": [], "probe": []}
        wait = {"cmp": [], "crs": [], "probe": []}
        failed = 0
        shed = {"cmp": 0, "crs": 0, "probe": 0}
        for kind, tk in tickets:
            out, st = srv.wait(tk)
            if out is not None:
                lat[kind].append(st["t_done_ms"] - st["t_submit_ms"])
                wait[kind].append(st["t_admit_ms"] - st["t_submit_ms"])
            elif st["status"] == sb.ERR_DEADLINE:
                shed[kind] += 1
            else:
                failed += 1
        wall = time.perf_counter() - start
        stats = srv.stats()
    nshed = sum(shed.values())
    res = dict(store="f32" if store == 0 else "f16", offered_rps=rate + probe_rate, duration_s=round(wall, 3),
               requests=len(events), served=stats["done"] - 4 - nshed - failed, shed=nshed, failed=failed,
               served_rps=round((stats["done"] - 4 - nshed - failed) / wall, 1), launches=stats["launches"] - 4,
               resident_gb=round(resident / 1e9, 2), deadline_ms=deadline_ms)
    for k, v in lat.items():
        v = np.array(v)
        if len(v):
            w = np.array(wait[k])
            res[k] = dict(n=len(v), p50_ms=float(np.percentile(v, 50)), p99_ms=float(np.percentile(v, 99)),
                          max_ms=float(v.max()), queue_p50_ms=float(np.percentile(w, 50)),
                          queue_p99_ms=float(np.percentile(w, 99)))
    return res


def serving_sweep(sb, lines, rates=(200, 500, 1000, 2000), seconds=2.0, deadline_ms=25.0):
    """The serving workload over its arrival-rate range: per offered rate,
    served (and shed) vs offered, p50/p99 per request kind of the served
    requests, and the lowest offered rate at which the 1-CDP probe's p99
    exceeds 50 ms -- for both stores with every request queued, and for FP32
    with an admission deadline (load shedding)."""
    out = {}
    for name, store, dl in (("f32", 0, 0.0), ("f16", 1, 0.0), ("f32_deadline", 0, deadline_ms)):
        curve = []
        for r in rates:
            res = serving_bench(sb, lines, seconds=seconds, rate=float(r), store=store, deadline_ms=dl)
            curve.append(dict(offered_rps=res["offered_rps"], requests=res["requests"], served=res["served"],
                              shed=res["shed"], failed=res["failed"], served_rps=res["served_rps"],
                              duration_s=res["duration_s"],
                              **{k: {q: round(res[k][q], 2) for q in ("p50_ms", "p99_ms", "queue_p99_ms")}
                                 for k in ("probe", "cmp", "crs") if k in res}))
        crossing = next((c["offered_rps"] for c in curve if c.get("probe", {}).get("p99_ms", 0) > 50.0), None)
        out[name] = dict(curve=curve, deadline_ms=dl, probe_p99_over_50ms_from_rps=crossing)
    out["what"] = ("offered = mixed CMP/CRS panel requests at the named rate + 100 req/s of 1-CDP CMP probes, %.0f s "
                   "of arrivals per point; latency = done - submit of the served requests; f32_deadline: requests "
                   "not admitted within %.0f ms are shed (SB_ERR_DEADLINE)" % (seconds, deadline_ms))
    return out


def launch_modes(sb, ctx, g, prm, steps=3):
    """Search ms of the same step under the other launch modes (same
    results): the global-memory kernel per work item and as a persistent
    per-XCD grid (R=1, CC=4, 16 parameters per workgroup), the LDS-staged
    kernel (best of its shapes) and a replayed HIP graph of the tuned shape
    (whole step)."""
    v0, pc0, nv = ctx.tune_get()
    r = sb.DeviceResult(ctx, g.ncdp, g.ns)
    out = {}
    try:
        def time_it():
            sb.run(ctx, g, prm, r)
            ms = []
            for _ in range(steps):
                sb.run(ctx, g, prm, r)
                ms.append(ctx.timing()["search_ms"])
            return float(np.median(ms))
        ctx.tune_set(1, 16)
        out["global_search_ms"] = time_it()
        ctx.set_option(sb.OPT_PERSISTENT, 1)
        out["persistent_search_ms"] = time_it()
        ctx.set_option(sb.OPT_PERSISTENT, 0)
        best = None
        for v in range(4, nv):
            ctx.tune_set(v, 0)
            t = time_it()
            if ctx.timing()["lds_staged"] and (best is None or t < best[1]):
                best = (v, t)
        if best:
            out["lds_staged_search_ms"], out["lds_variant"] = best[1], best[0]
        ctx.tune_set(v0, pc0)
        with sb.Graph(ctx, g, prm, r) as gr:
            gr.launch()
            ctx.sync()
            t0 = time.perf_counter()
            for _ in range(steps):
                gr.launch()
            ctx.sync()
            out["graph_ms_per_step"] = (time.perf_counter() - t0) / steps * 1e3
    finally:
        ctx.set_option(sb.OPT_PERSISTENT, 0)
        ctx.tune_set(v0, pc0)
        r.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-tune", action="store_true")
    ap.add_argument("--stores", default="f32,f16", help="trace storage types to time")
    ap.add_argument("--variant", type=int, default=-1, help="fixed launch shape (skips the autotuner)")
    ap.add_argument("--pchunk", type=int, default=0, help="parameters per workgroup with --variant")
    ap.add_argument("--no-latency", action="store_true", help="skip the 1-CDP panel latency probe")
    ap.add_argument("--no-side", action="store_true", help="skip the CRS grid and CRS-DE side measurements")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving-loop load test")
    ap.add_argument("--no-modes", action="store_true", help="skip the persistent / LDS / graph launch-mode timings")
    ap.add_argument("--no-sweep", action="store_true", help="skip the serving sweep over arrival rates")
    ap.add_argument("--de-only", action="store_true",
                    help="time only the CRS-DE configuration (for profiling): --de-cdps, --de-lds, --stores")
    ap.add_argument("--de-cdps", type=int, default=2000, help="with --de-only: CDPs of the CRS-DE line")
    ap.add_argument("--de-lds", type=int, default=-1, help="with --de-only: SB_OPT_DE_LDS (0: global memory)")
    ap.add_argument("--crs-only", action="store_true",
                    help="time only the CRS grid configuration (for profiling): --crs-cdps, --crs-lds")
    ap.add_argument("--crs-cdps", type=int, default=500, help="with --crs-only: CDPs of the CRS line")
    ap.add_argument("--crs-lds", type=int, default=-1, help="with --crs-only: SB_OPT_CRS_LDS (0: global memory)")
    ap.add_argument("--serve-only", action="store_true",
                    help="run only the serving workload once (for profiling): --serve-rate, --serve-deadline, "
                         "--serve-store")
    ap.add_argument("--serve-rate", type=float, default=1000.0,
                    help="with --serve-only: mixed requests/s offered (plus 100/s 1-CDP probes)")
    ap.add_argument("--serve-deadline", type=float, default=0.0, help="with --serve-only: admission deadline, ms")
    ap.add_argument("--serve-store", default="f32", choices=("f32", "f16"), help="with --serve-only: line store")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU; --gpus %d is not supported (the search shards by CDP range "
              "with no exchange, see the sharding notes)" % args.gpus, file=sys.stderr)
        sys.exit(2)
    sb = _load("sbamd", os.path.join(ROOT, "ipdps-crs-cmp-code_amd", "__init__.py"))
    if args.crs_only:
        orc = _load("sb_oracle", os.path.join(ROOT, "oracle", "oracle.py"))
        if args.no_cpu:
            print(json.dumps(crs_grid_side(sb, args, orc, args.crs_cdps, args.crs_lds)), flush=True)
            return
        print("bench.py: --crs-only times the GPU only; pass --no-cpu", file=sys.stderr)
        sys.exit(2)
    if args.serve_only:
        r = serving_bench(sb, serving_lines(sb), seconds=2.0, rate=args.serve_rate,
                          deadline_ms=args.serve_deadline, store=0 if args.serve_store == "f32" else 1)
        print(json.dumps(r), flush=True)
        return
    if args.de_only:
        orc = _load("sb_oracle", os.path.join(ROOT, "oracle", "oracle.py"))
        if args.no_cpu:
            print(json.dumps(crs_de_side(sb, args, orc, args.de_cdps, tuple(args.stores.split(",")), args.de_lds)),
                  flush=True)
            return
        print("bench.py: --de-only times the GPU only; pass --no-cpu", file=sys.stderr)
        sys.exit(2)
    print("bench: generating synthetic line", flush=True)
    cdp_off, geom, samples = sb.synth(0, CFG["ncdp"], CFG["fold"], CFG["ns"], CFG["dt_us"], SEED + 1,
                                      h_max=CFG["h_max"], h_jitter=CFG["h_jitter"], nevents=CFG["nevents"],
                                      snr=CFG["snr"])
    prm = sb.cmp_params(CFG["c0"], CFG["c1"], CFG["nc"], CFG["tau"], sb.MODE_FAST)
    out = {}
    with sb.Context(0) as ctx:
        results = {}
        for store in args.stores.split(","):
            print("bench: uploading gather (%s)" % store, flush=True)
            g = sb.Gather(ctx, cdp_off, geom, samples, CFG["dt_us"], store=sb.STORES[store])
            if args.variant >= 0:
                ctx.tune_set(args.variant, args.pchunk)
            elif not args.no_tune:
                v, pc, _ = sb.tune(ctx, g, prm, reps=4)
                print("bench: tuned variant %d pchunk %d (%s)" % (v, pc, store), flush=True)
            wall, tk = timed_steps(sb, ctx, g, prm, args.steps, args.warmup)
            tuples = tk[0]["tuples"]
            search_ms = float(np.mean([t["search_ms"] for t in tk]))
            fin_ms = float(np.mean([t["finalize_ms"] for t in tk]))
            results[store] = dict(value=tuples * args.steps / wall, ms_per_step=wall / args.steps * 1e3,
                                  search_ms=search_ms, finalize_ms=fin_ms, tuples=tuples, tune=(tk[0]["variant"], ctx.tune_get()[1]),
                                  blocks=tk[0]["grid_blocks"], lds_staged=tk[0]["lds_staged"],
                                  restored=tk[0]["restored"])
            print("bench: %s %.4g evals/s, search %.3f ms, finalize %.3f ms" %
                  (store, results[store]["value"], search_ms, fin_ms), flush=True)
            if store != "f32":
                # the same step with the 16-bit kernels reading the store
                # (no transient FP32 copy), tuned for that
                retune = args.variant < 0 and not args.no_tune
                ctx.set_option(sb.OPT_RESTORE_FP32, 0)
                try:
                    if retune:
                        sb.tune(ctx, g, prm, reps=4)
                    wall2, tk2 = timed_steps(sb, ctx, g, prm, args.steps, args.warmup)
                    direct_launch = {"variant": tk2[0]["variant"], "pchunk": ctx.tune_get()[1]}
                finally:
                    ctx.set_option(sb.OPT_RESTORE_FP32, 1)
                    if retune:
                        sb.tune(ctx, g, prm, reps=4)  # the shape tuned with the copy (cached)
                results[store]["direct"] = dict(value=tuples * args.steps / wall2,
                                                ms_per_step=wall2 / args.steps * 1e3, launch=direct_launch)
                print("bench: %s without the FP32 copy %.4g evals/s" % (store, results[store]["direct"]["value"]),
                      flush=True)
            if not args.no_modes:
                results[store]["modes"] = launch_modes(sb, ctx, g, prm)
                print("bench: %s launch modes %s" % (store, results[store]["modes"]), flush=True)
            g.close()
        p50, p99 = (None, None) if args.no_latency else panel_latency(sb, ctx, cdp_off, geom, samples)
    if "f32" not in results:
        print(json.dumps({k: v for k, v in results.items()}), flush=True)
        return
    f32 = results["f32"]
    f16 = results.get("f16")
    achieved = f32["tuples"] * FLOP_PER_TUPLE_CMP / (f32["search_ms"] * 1e-3) / 1e12
    fma_rate = sb.fp32_fma_rate(0)  # the device's sustained FP32 FMA rate without packed FP32
    win_rates = sb.window_rates(0)  # the inner step's instruction forms, window samples/s
    skip = skipped_fraction(cdp_off, geom, f32["tune"][0], effective_pchunk(f32), f32["blocks"])
    lds = bool(f32["lds_staged"])
    kernel = "k_search_lds" if lds else "k_search_fast"
    # HBM bytes per launch of the timed shape, from rocprofv3 FETCH_SIZE /
    # WRITE_SIZE passes over the same workload and launch shape (null when
    # that shape was not profiled)
    traffic = None
    tf = os.path.join(ROOT, "profiles", "traffic_%s_cmp_2000cdp_v%d_p%d.json" % (kernel, f32["tune"][0],
                                                                                effective_pchunk(f32)))
    if os.path.exists(tf):
        with open(tf) as f:
            traffic = json.load(f).get("bytes_per_launch")
    line = {
        "metric": "semblance evals/s (CDP*t0*param*trace), FP16-store vs FP32 GPU; p99 panel ms",
        "value": f32["value"],
        "unit": "semblance evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": f32["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded Ricker events + noise, generated in-process)",
        "config": {"workload": "CMP 2000 CDPs fold 30-120 ns 1501 dt 2ms nc 256 tau 2ms (FP32 store)",
                   "ncdp": CFG["ncdp"], "ns": CFG["ns"], "nc": CFG["nc"], "tuples_per_step": f32["tuples"],
                   "launch": {"variant": f32["tune"][0], "pchunk": f32["tune"][1], "blocks": f32["blocks"],
                              "lds_staged": f32["lds_staged"]}},
        "fp16_store": None if f16 is None else {
            "value": f16["value"], "ms_per_step": f16["ms_per_step"], "speedup_vs_f32": f16["value"] / f32["value"],
            "launch": {"variant": f16["tune"][0], "pchunk": f16["tune"][1]}, "restored": f16["restored"],
            "direct": f16.get("direct"),
            "what": "restored: the search read a transient FP32 copy of the rows (SB_OPT_RESTORE_FP32, the "
                    "default; its time is inside the step); direct: the 16-bit kernels reading the store"},
        "launch_modes_ms": {k: results[k].get("modes") for k in results},
        "panel_latency_ms": {"p50": p50, "p99": p99, "what": "1-CDP CMP panel, upload+search+download, sequential"},
        "value_evaluated": f32["value"] * (1.0 - skip),
        "roofline": {"bound": "valu", "achieved": achieved * (1.0 - skip), "peak": PEAK_FP32_VECTOR_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved * (1.0 - skip) / PEAK_FP32_VECTOR_TFLOPS, "traffic": traffic,
                     "kernel": kernel + " (CMP, W=3)", "kernel_ms": f32["search_ms"],
                     "flop_per_eval": FLOP_PER_TUPLE_CMP, "skipped_frac": skip,
                     "achieved_reference_count": achieved,
                     "frac_reference_count": achieved / PEAK_FP32_VECTOR_TFLOPS,
                     "fp32_fma_rate_measured": fma_rate,
                     "frac_of_measured_fma_rate": achieved * (1.0 - skip) / fma_rate,
                     "window_sample_rates": win_rates,
                     "note": "gather-reduce on the FP32 vector ALUs (no MFMA); peak = MI355X FP32 157.3 TFLOP/s, "
                             "which counts packed FP32 -- the kernels run none, and fp32_fma_rate_measured is the "
                             "FMA rate this device sustains without it (independent FMA chains, full grid); "
                             "achieved counts the evaluations the kernel runs: the reference's count minus "
                             "skipped_frac of it, (t0, c) cells that have two invalid traces and no result in any "
                             "arithmetic (achieved_reference_count counts all of them, as the reference's rate "
                             "does); window_sample_rates: the device's rate of the search's inner step "
                             "(interpolate one window sample, add it and its square) in FP32 as the kernels run "
                             "it and in the mixed-precision forms a 16-bit store could use instead -- these are "
                             "no faster, so the 16-bit stores restore their samples to FP32; "
                             "value is the reference's count per second; " + (
                                 "windows come from LDS (a workgroup stages the ranges of as many of its traces "
                                 "as fit, once, as FP32 samples or, in the paired shapes, each sample with its "
                                 "forward difference so that a window sample is one FMA; the rest are read from "
                                 "global memory): bound by VALU issue and the LDS data path, the staging latency "
                                 "partly hidden"
                                 if lds else "the kernel is co-bound by the vector-memory address path (16-byte "
                                 "window loads) and VALU issue")},
    }
    if not args.no_cpu:
        print("bench: CPU baseline", flush=True)
        line["cpu_baseline"] = cpu_baseline(sb, cdp_off, geom, samples)
    if not args.no_side:
        del cdp_off, geom, samples
        line.update(side_configs(sb, args))
        print("bench: CMP 200-CDP line through the cmp program", flush=True)
        line["cli_cmp_200cdp"] = cli_cmp_200cdp(sb)
    if not args.no_serving:
        lines = serving_lines(sb)
        line["serving"] = serving_bench(sb, lines)
        line["serving_fp16_store"] = serving_bench(sb, lines, store=sb.STORE_F16)
        if "probe" in line["serving"]:
            line["panel_latency_ms"]["p99_under_load"] = line["serving"]["probe"]["p99_ms"]
        if not args.no_sweep:
            line["serving_sweep"] = serving_sweep(sb, lines)
        del lines
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
