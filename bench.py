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
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU; --gpus %d is not supported (the search shards by CDP range "
              "with no exchange, see the sharding notes)" % args.gpus, file=sys.stderr)
        sys.exit(2)
    sb = _load("sbamd", os.path.join(ROOT, "ipdps-crs-cmp-code_amd", "__init__.py"))
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
                v, pc, _ = sb.tune(ctx, g, prm)
                print("bench: tuned variant %d pchunk %d (%s)" % (v, pc, store), flush=True)
            wall, tk = timed_steps(sb, ctx, g, prm, args.steps, args.warmup)
            tuples = tk[0]["tuples"]
            search_ms = float(np.mean([t["search_ms"] for t in tk]))
            fin_ms = float(np.mean([t["finalize_ms"] for t in tk]))
            results[store] = dict(value=tuples * args.steps / wall, ms_per_step=wall / args.steps * 1e3,
                                  search_ms=search_ms, finalize_ms=fin_ms, tuples=tuples, tune=ctx.tune_get()[:2],
                                  blocks=tk[0]["grid_blocks"])
            print("bench: %s %.4g evals/s, search %.3f ms, finalize %.3f ms" %
                  (store, results[store]["value"], search_ms, fin_ms), flush=True)
            g.close()
        p50, p99 = (None, None) if args.no_latency else panel_latency(sb, ctx, cdp_off, geom, samples)
    if "f32" not in results or "f16" not in results:
        print(json.dumps({k: v for k, v in results.items()}), flush=True)
        return
    f32 = results["f32"]
    achieved = f32["tuples"] * FLOP_PER_TUPLE_CMP / (f32["search_ms"] * 1e-3) / 1e12
    traffic = None
    tf = os.path.join(ROOT, "profiles", "traffic_k_search_fast_cmp_2000cdp.json")
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
                   "launch": {"variant": f32["tune"][0], "pchunk": f32["tune"][1], "blocks": f32["blocks"]}},
        "fp16_store": {"value": results["f16"]["value"], "ms_per_step": results["f16"]["ms_per_step"],
                       "speedup_vs_f32": results["f16"]["value"] / f32["value"]},
        "panel_latency_ms": {"p50": p50, "p99": p99, "what": "1-CDP CMP panel, upload+search+download, sequential"},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_VECTOR_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP32_VECTOR_TFLOPS, "traffic": traffic,
                     "kernel": "k_search_fast (CMP, W=3)", "kernel_ms": f32["search_ms"],
                     "flop_per_eval": FLOP_PER_TUPLE_CMP},
    }
    if not args.no_cpu:
        print("bench: CPU baseline", flush=True)
        line["cpu_baseline"] = cpu_baseline(sb, cdp_off, geom, samples)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
