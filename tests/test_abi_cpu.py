"""C ABI surface and host-side behaviour (no GPU needed)."""
import os
import re

import numpy as np
import pytest

from conftest import GOLD, ROOT, has_gpu


def header_functions():
    with open(os.path.join(ROOT, "include", "semblance.h")) as f:
        text = f.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(sb_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol(sb):
    declared = header_functions()
    assert sorted(sb.EXPORTS) == declared
    L = sb.lib()
    for name in declared:
        assert hasattr(L, name), name
    assert b"gfx950" in L.sb_version()


def test_library_has_gfx950_code_object():
    so = os.path.join(ROOT, "ipdps-crs-cmp-code_amd", "lib", "libsemblance.so")
    with open(so, "rb") as f:
        blob = f.read()
    assert b"gfx950" in blob


def test_error_strings(sb):
    L = sb.lib()
    for code in range(0, 10):
        assert L.sb_strerror(code) and L.sb_strerror(code) != b"unknown error", code
    assert L.sb_strerror(9) == b"admission deadline passed"


@pytest.mark.skipif(has_gpu(), reason="checks the no-device error path")
def test_ctx_without_device_reports_error(sb):
    with pytest.raises(sb.SbError, match="HIP"):
        sb.Context(0)


def test_su_load_errors(sb, tmp_path):
    with pytest.raises(sb.SbError, match="i/o"):
        sb.SuGather(str(tmp_path / "missing.su"), sb.METHOD_CMP, 1000)
    src = os.path.join(GOLD, "cmp_w3", "input.su")
    raw = open(src, "rb").read()
    # Truncated inside a trace's samples: the reference fails with "Bad input data".
    bad = tmp_path / "trunc.su"
    bad.write_bytes(raw[:-10])
    with pytest.raises(sb.SbError, match="Bad input data"):
        sb.SuGather(str(bad), sb.METHOD_CMP, 2000)
    # A partial header at the end only ends the read (su_trace::fgettr).
    part = tmp_path / "part.su"
    part.write_bytes(raw + raw[:100])
    G = sb.SuGather(str(part), sb.METHOD_CMP, 2000)
    G0 = sb.SuGather(src, sb.METHOD_CMP, 2000)
    assert G.ntr == G0.ntr
    # Mixed ns among kept traces.
    cdp_off, geom, s1 = sb.synth(0, 2, 3, 50, 4000, seed=3)
    _, _, s2 = sb.synth(0, 2, 3, 60, 4000, seed=3)
    sb.write_traces(str(tmp_path / "a.su"), geom, s1, 4000)
    sb.write_traces(str(tmp_path / "b.su"), geom, s2, 4000)
    (tmp_path / "mixed.su").write_bytes((tmp_path / "a.su").read_bytes() + (tmp_path / "b.su").read_bytes())
    with pytest.raises(sb.SbError, match="format"):
        sb.SuGather(str(tmp_path / "mixed.su"), sb.METHOD_CMP, 5000)


def test_synth_is_deterministic_and_binned(sb):
    a = sb.synth(1, 5, (3, 9), 101, 4000, seed=42, h_max=800.0, aph=500)
    b = sb.synth(1, 5, (3, 9), 101, 4000, seed=42, h_max=800.0, aph=500)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    cdp_off, geom, samples = a
    for i in range(5):
        assert np.all(geom["cdp"][cdp_off[i]:cdp_off[i + 1]] == i + 1)
    # every kept trace passes the CRS aperture test
    L = sb.lib()
    for g in geom:
        arr = np.array([g], sb.GEOM_DTYPE)
        assert L.sb_keep_trace(arr.ctypes.data, sb.METHOD_CRS, 500) == 1


def test_keep_trace_matches_float_and_double_rules(sb):
    """CMP squares and sums the half-offsets in float, CRS in double
    (CMP/OpenMP/src/su_gather.cpp:44-47 with real = float; the same lines of
    CRS/OpenMP/src/su_gather.cpp with real = double).  Constructed case, scalco
    0 (factor 1), aph 5000: hx = 5000, hy = 1 gives hx^2 + hy^2 = 25 000 001,
    above aph^2 = 25 000 000 exactly, but in float the sum lies halfway between
    the neighbours 25 000 000 and 25 000 002 (ulp 2 above 2^24) and rounds to
    the even one, 25 000 000, which is not above aph^2: CMP keeps the trace,
    CRS drops it."""
    L = sb.lib()

    def keep(gx, gy, method):
        g = np.array([(1, 0, 0, 0, gx, gy)], sb.GEOM_DTYPE)  # cdp, scalco, sx, sy, gx, gy
        return L.sb_keep_trace(g.ctypes.data, method, 5000)

    # half-offsets (gx - sx) / 2 = 5000 and (gy - sy) / 2 = 1
    assert keep(10000, 2, sb.METHOD_CMP) == 1
    assert keep(10000, 2, sb.METHOD_CRS) == 0
    # on the aperture exactly (not above): both keep; hy = 2 (sum 25 000 004,
    # exact in float): both drop
    assert keep(10000, 0, sb.METHOD_CMP) == 1 and keep(10000, 0, sb.METHOD_CRS) == 1
    assert keep(10000, 4, sb.METHOD_CMP) == 0 and keep(10000, 4, sb.METHOD_CRS) == 0


def test_neighbors_match_bruteforce(sb):
    cdp_off, geom, _ = sb.synth(1, 12, 4, 51, 4000, seed=5, mid_half_width=40.0, h_max=300.0)
    first = geom[cdp_off[:-1]]
    off, nbr = sb.neighbors(first, 45)
    f = np.float32
    s = np.float32(1.0) / np.float32(-100.0)
    mx = np.array([f(f(s * f(g["gx"] + g["sx"])) * 0.5) for g in first], np.float32)
    my = np.array([f(f(s * f(g["gy"] + g["sy"])) * 0.5) for g in first], np.float32)
    for i in range(12):
        dx = mx - mx[i]
        dy = my - my[i]
        want = np.nonzero(dx * dx + dy * dy <= np.float32(45 * 45))[0]
        assert np.array_equal(nbr[off[i]:off[i + 1]], want)
