"""Writes the golden fixtures: small seeded SU inputs and the outputs of the
reference's own OpenMP programs on them.

    python tests/golden/make_golden.py

Needs oracle/_ref/{cmp-omp2,crs-omp2}, built from the reference sources by
`make -C oracle ref` (only possible where the reference tree is mounted).
Each case lands in tests/golden/<case>/ with input.su, the reference output
files, and an entry in tests/golden/manifest.json (arguments, the reference's
INFO line).  The inputs come from the library's seeded generator; some cases
rewrite coordinates to other scalco conventions or permute traces.
"""
import importlib.util
import json
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
GOLD = os.path.dirname(os.path.abspath(__file__))
REF = os.path.join(ROOT, "oracle", "_ref")


def _pkg():
    spec = importlib.util.spec_from_file_location(
        "sbamd", os.path.join(ROOT, "ipdps-crs-cmp-code_amd", "__init__.py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


sb = _pkg()

CMP_PARAMS = dict(c0="1.98e-7", c1="1.77e-6")
CRS_PARAMS = dict(a0="-0.7e-3", a1="0.7e-3", b0="-1e-7", b1="1e-7", c0="1.975e-7", c1="1.384e-6")


def rescale(geom, scalco):
    """Rewrites cm coordinates (scalco -100) to another scalco convention."""
    g = geom.copy()
    for k in ("sx", "sy", "gx", "gy"):
        v = g[k].astype(np.float64) / 100.0  # metres
        if scalco == 0 or scalco == 1:
            g[k] = np.round(v).astype(np.int32)
        elif scalco > 0:
            g[k] = np.round(v / scalco).astype(np.int32)
        else:
            g[k] = np.round(v * -scalco).astype(np.int32)
    g["scalco"] = scalco
    return g


def run_ref(binary, args, su_path, outdir):
    with tempfile.TemporaryDirectory() as td:
        cmd = [os.path.join(REF, binary)] + args + ["-i", os.path.abspath(su_path), "-v", "2"]
        p = subprocess.run(cmd, cwd=td, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True,
                           env=dict(os.environ, OMP_NUM_THREADS="4"))
        if p.returncode != 0:
            raise RuntimeError("%s failed: %s %s" % (binary, p.stdout, p.stderr))
        outs = []
        for f in sorted(os.listdir(td)):
            if f.endswith(".su"):
                shutil.copy(os.path.join(td, f), os.path.join(outdir, f))
                outs.append(f)
        info = [l for l in p.stdout.splitlines() if "INFO" in l]
        return outs, (info[0].split("]: ", 1)[1].split(": ", 2)[-1] if info else "")


def cmp_case(name, ncdp, fold, ns, dt_us, seed, nc, tau, aph, scalco=-100, regular=True, h_step=100.0,
             h_max=1100.0, h_jitter=0.0, shuffle=False, nevents=3, snr=5.0):
    cdp_off, geom, samples = sb.synth(0, ncdp, fold, ns, dt_us, seed, h_max=h_max, h_jitter=h_jitter,
                                      regular_offsets=regular, h_step=h_step, nevents=nevents, snr=snr)
    if scalco != -100:
        geom = rescale(geom, scalco)
    order = np.arange(len(geom))
    if shuffle:
        order = np.random.default_rng(seed).permutation(len(geom))
    d = os.path.join(GOLD, name)
    os.makedirs(d, exist_ok=True)
    sb.write_traces(os.path.join(d, "input.su"), geom[order], samples[order], dt_us)
    args = ["-c0", CMP_PARAMS["c0"], "-c1", CMP_PARAMS["c1"], "-nc", str(nc), "-aph", str(aph), "-tau", str(tau)]
    outs, info = run_ref("cmp-omp2", args, os.path.join(d, "input.su"), d)
    return dict(method="cmp", args=args, outputs=outs, info=info)


def crs_case(name, ncdp, fold, ns, dt_us, seed, n, tau, aph, apm, h_max=800.0, mid=30.0, scalco=-100,
             nevents=3, snr=3.0, spacing=12.5):
    cdp_off, geom, samples = sb.synth(1, ncdp, fold, ns, dt_us, seed, h_max=h_max, mid_half_width=mid,
                                      nevents=nevents, snr=snr, cdp_spacing=spacing)
    if scalco != -100:
        geom = rescale(geom, scalco)
    d = os.path.join(GOLD, name)
    os.makedirs(d, exist_ok=True)
    sb.write_traces(os.path.join(d, "input.su"), geom, samples, dt_us)
    na, nb, nc = n
    args = ["-a0", CRS_PARAMS["a0"], "-a1", CRS_PARAMS["a1"], "-na", str(na), "-b0", CRS_PARAMS["b0"], "-b1",
            CRS_PARAMS["b1"], "-nb", str(nb), "-c0", CRS_PARAMS["c0"], "-c1", CRS_PARAMS["c1"], "-nc", str(nc),
            "-aph", str(aph), "-apm", str(apm), "-tau", str(tau)]
    outs, info = run_ref("crs-omp2", args, os.path.join(d, "input.su"), d)
    return dict(method="crs", args=args, outputs=outs, info=info)


def main():
    for b in ("cmp-omp2", "crs-omp2"):
        if not os.path.exists(os.path.join(REF, b)):
            sys.exit("missing oracle/_ref/%s: run make -C oracle ref" % b)
    cases = {}
    # Regular line, w = 3 (tau 0.004 s at dt 2 ms).
    cases["cmp_w3"] = cmp_case("cmp_w3", 8, 12, 251, 2000, 11, 16, "0.004", 2000)
    # Uneven fold with repeated folds (pins the std::sort CDP order), w = 5.
    cases["cmp_uneven_w5"] = cmp_case("cmp_uneven_w5", 14, (6, 11), 201, 4000, 12, 12, "0.008", 2000,
                                      regular=False, h_max=900.0, h_jitter=5.0)
    # aph drops far traces; traces written in shuffled order; scalco 0.
    cases["cmp_aph_shuffled"] = cmp_case("cmp_aph_shuffled", 6, 10, 151, 4000, 13, 9, "0.004", 650,
                                         scalco=0, shuffle=True)
    # Positive scalco (decametres), tau 0 (w = 1), pure noise rows.
    cases["cmp_scalco_pos_w1"] = cmp_case("cmp_scalco_pos_w1", 5, 8, 121, 4000, 14, 7, "0.001", 2000,
                                          scalco=10, nevents=2)
    # CRS: small supergathers with gaps in the neighbour lists, w = 3.
    cases["crs_w3"] = crs_case("crs_w3", 10, 14, 151, 4000, 21, (3, 3, 3), "0.004", 500, 30)
    # CRS: wider window, more CDPs per supergather, scalco -10 (decimetres).
    cases["crs_w5"] = crs_case("crs_w5", 8, 10, 121, 4000, 22, (2, 3, 4), "0.008", 700, 60, scalco=-10)
    with open(os.path.join(GOLD, "manifest.json"), "w") as f:
        json.dump(cases, f, indent=1, sort_keys=True)
    print(json.dumps(cases, indent=1))


if __name__ == "__main__":
    main()
