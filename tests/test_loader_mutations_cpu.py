"""Malformed inputs to the SU and SEG-Y readers (sb_su_load, sb_segy_load):
seeded mutations of valid files -- truncation at any byte, flipped bytes,
corrupted ns / dt / format / extended-header fields, trailing garbage -- must
either load or fail with a library status, never crash, and whatever loads
must be self-consistent (offsets monotone and covering every trace, one ns,
finite geometry).  The SEG-Y bases are written by tests/segy_util.py,
independent of the library, in every sample format."""
import os

import numpy as np
import pytest

from conftest import GOLD

from segy_util import read_su, write_segy


def _check_loaded(sb, g):
    assert g.ncdp >= 1 and g.ntr >= g.ncdp and g.ns >= 1
    assert g.cdp_off[0] == 0 and g.cdp_off[-1] == g.ntr and np.all(np.diff(g.cdp_off) >= 1)
    assert g.samples.shape == (g.ntr, g.ns)
    if g.nbr_off is not None:
        assert g.nbr_off[0] == 0 and np.all(np.diff(g.nbr_off) >= 1)
        assert np.all((g.nbr >= 0) & (g.nbr < g.ncdp))


def _mutate(rng, b, header_at):
    n = len(b)
    k = rng.integers(0, 5)
    if k == 0:
        return b[:rng.integers(0, n)]
    if k == 1:
        for _ in range(rng.integers(1, 40)):
            b[rng.integers(0, n)] = rng.integers(0, 256)
        return b
    if k == 2:
        off = header_at[rng.integers(0, len(header_at))]
        b[off:off + 10] = bytes(rng.integers(0, 256, 10).astype(np.uint8))
        return b
    if k == 3:
        return b + bytes(rng.integers(0, 256, rng.integers(1, 700)).astype(np.uint8))
    return b[:n // 2] + b[n // 2 + rng.integers(1, 241):]  # bytes missing mid-file


def _run(sb, rng, path, segy, n):
    loaded = rejected = 0
    for _ in range(n):
        for method in (sb.METHOD_CMP, sb.METHOD_CRS):
            try:
                g = sb.SuGather(path, method, int(rng.integers(0, 3000)), int(rng.integers(0, 200)), segy=segy)
            except sb.SbError:
                rejected += 1
                continue
            _check_loaded(sb, g)
            g.close()
            loaded += 1
    return loaded, rejected


def test_su_mutations(sb, tmp_path):
    rng = np.random.default_rng(11)
    src = open(os.path.join(GOLD, "cmp_w3", "input.su"), "rb").read()
    ns = int.from_bytes(src[114:116], "little")
    stride = 240 + 4 * ns
    heads = [i * stride + 110 for i in range(len(src) // stride)]  # around ns (@114) and dt (@116)
    path = str(tmp_path / "m.su")
    total = [0, 0]
    for _ in range(150):
        open(path, "wb").write(bytes(_mutate(rng, bytearray(src), heads)))
        lo, rj = _run(sb, rng, path, False, 1)
        total[0] += lo
        total[1] += rj
    print("SU mutations: %d loaded, %d rejected" % tuple(total))
    assert total[0] > 0 and total[1] > 0


@pytest.mark.parametrize("fmt", [1, 2, 3, 5, 8])
def test_segy_mutations(sb, tmp_path, fmt):
    rng = np.random.default_rng(20 + fmt)
    traces = read_su(os.path.join(GOLD, "cmp_w3", "input.su"))[:40]
    base = str(tmp_path / "b.sgy")
    scale = 100.0 if fmt in (2, 3, 8) else 1.0
    write_segy(base, [(h, d * scale) for h, d in traces], fmt, 2000, ext_text=int(fmt == 8))
    src = open(base, "rb").read()
    first = 3600 + 3200 * int(fmt == 8)
    heads = [3200 + 12, 3200 + 16, 3200 + 300, first + 110]  # binary header dt/ns/format, ext count, trace ns/dt
    path = str(tmp_path / "m.sgy")
    total = [0, 0]
    for _ in range(40):
        open(path, "wb").write(bytes(_mutate(rng, bytearray(src), heads)))
        lo, rj = _run(sb, rng, path, True, 1)
        total[0] += lo
        total[1] += rj
    print("SEG-Y format %d mutations: %d loaded, %d rejected" % (fmt, total[0], total[1]))
    assert total[0] + total[1] == 80
