"""SEG-Y import (sb_segy_load): a big-endian SEG-Y file carrying the same
traces as an SU fixture must bin, order and write exactly as the SU file does
(the reference programs' own outputs are then reproduced byte for byte from
the SEG-Y input), for IEEE and integer sample formats; IBM floats decode to
the values the IBM System/360 format defines; malformed files fail cleanly.
The SEG-Y writer (tests/segy_util.py) is independent of the library."""
import filecmp
import os
import struct

import numpy as np
import pytest

from conftest import GOLD, cmp_args, crs_args

from segy_util import float_to_ibm, read_su, swap_header, write_segy


def _same(a, b):
    assert a.ncdp == b.ncdp and a.ntr == b.ntr and a.ns == b.ns and a.dt_us == b.dt_us
    assert np.array_equal(a.cdp_off, b.cdp_off) and np.array_equal(a.geom, b.geom)
    assert np.array_equal(a.samples, b.samples)
    if a.nbr_off is not None:
        assert np.array_equal(a.nbr_off, b.nbr_off) and np.array_equal(a.nbr, b.nbr)


@pytest.mark.parametrize("case", ["cmp_w3", "cmp_uneven_w5", "cmp_aph_shuffled", "cmp_scalco_pos_w1"])
def test_segy_ieee_reproduces_cmp_reference_outputs(sb, orc, manifest, tmp_path, case):
    p = cmp_args(manifest[case])
    su = os.path.join(GOLD, case, "input.su")
    sg = str(tmp_path / "input.sgy")
    write_segy(sg, read_su(su), 5, 0, ext_text=1)
    A = sb.SuGather(su, sb.METHOD_CMP, p["aph"])
    B = sb.SuGather(sg, sb.METHOD_CMP, p["aph"], segy=True)
    _same(A, B)
    out = orc.cmp(B.cdp_off, B.geom, B.samples, B.dt_us, p["c0"], p["c1"], p["nc"], p["tau"])
    for key, fname in (("c", "cmp.c.su"), ("coher", "cmp.coher.su"), ("stack", "cmp.stack.su")):
        B.write_section(str(tmp_path / fname), out[key], key == "c")
        assert filecmp.cmp(str(tmp_path / fname), os.path.join(GOLD, case, fname), shallow=False), (case, fname)


@pytest.mark.parametrize("case", ["crs_w3", "crs_w5"])
def test_segy_ieee_reproduces_crs_binning(sb, manifest, tmp_path, case):
    p = crs_args(manifest[case])
    su = os.path.join(GOLD, case, "input.su")
    sg = str(tmp_path / "input.sgy")
    write_segy(sg, read_su(su), 5, 0)
    _same(sb.SuGather(su, sb.METHOD_CRS, p["aph"], p["apm"]),
          sb.SuGather(sg, sb.METHOD_CRS, p["aph"], p["apm"], segy=True))


@pytest.mark.parametrize("fmt", [1, 2, 3, 8])
def test_segy_sample_formats(sb, tmp_path, fmt):
    """Integer-valued and dyadic samples survive every format exactly; a
    trace ns/dt of 0 takes the binary header's."""
    su = os.path.join(GOLD, "cmp_w3", "input.su")
    traces = read_su(su)
    rng = np.random.default_rng(fmt)
    lim = {1: 1000, 2: 1 << 20, 3: 30000, 8: 100}[fmt]
    new = []
    for hdr, data in traces:
        v = rng.integers(-lim, lim, len(data)).astype(np.float32)
        if fmt == 1:
            v = v / np.float32(64.0)  # dyadic, <= 17 significant bits
        new.append((hdr, v))
    su2 = str(tmp_path / "in.su")
    with open(su2, "wb") as f:
        for hdr, v in new:
            f.write(hdr)
            f.write(v.astype("<f4").tobytes())
    sg = str(tmp_path / "in.sgy")
    dt = struct.unpack_from("<H", traces[0][0], 116)[0]
    write_segy(sg, new, fmt, dt, zero_trace_ns=True)
    A = sb.SuGather(su2, sb.METHOD_CMP, 2000)
    B = sb.SuGather(sg, sb.METHOD_CMP, 2000, segy=True)
    _same(A, B)


def test_ibm_known_values(sb, tmp_path):
    """IBM hex-float encodings of 1.0, 100.0, -118.625, 0.15625, 2^-64,
    zeros and 16^-6 (IBM System/360: (-1)^s * 0.F * 16^(e-64)) decode to
    those values."""
    words = [0x41100000, 0x42640000, 0xC276A000, 0x40280000, 0x31100000, 0x00000000, 0x80000000, 0x3B100000]
    want = np.array([1.0, 100.0, -118.625, 0.15625, 2.0 ** -64, 0.0, -0.0, 16.0 ** -6], np.float32)
    su = os.path.join(GOLD, "cmp_w3", "input.su")
    hdr = read_su(su)[0][0]
    hdr = bytearray(hdr)
    struct.pack_into("<H", hdr, 114, len(words))
    sg = str(tmp_path / "ibm.sgy")
    binh = bytearray(400)
    struct.pack_into(">H", binh, 16, 4000)
    struct.pack_into(">H", binh, 20, len(words))
    struct.pack_into(">H", binh, 24, 1)
    with open(sg, "wb") as f:
        f.write(b"\x40" * 3200 + bytes(binh) + swap_header(bytes(hdr)))
        f.write(b"".join(struct.pack(">I", w) for w in words))
    G = sb.SuGather(sg, sb.METHOD_CMP, 100000, segy=True)
    assert np.array_equal(G.samples[0], want)
    assert float_to_ibm(-118.625) == 0xC276A000 and float_to_ibm(1.0) == 0x41100000


def test_segy_errors(sb, tmp_path):
    bad = str(tmp_path / "short.sgy")
    open(bad, "wb").write(b"\x40" * 3000)
    with pytest.raises(sb.SbError):
        sb.SuGather(bad, sb.METHOD_CMP, 1000, segy=True)
    su = os.path.join(GOLD, "cmp_w3", "input.su")
    fmt4 = str(tmp_path / "fmt4.sgy")
    write_segy(fmt4, read_su(su)[:3], 5, 0)
    raw = bytearray(open(fmt4, "rb").read())
    struct.pack_into(">H", raw, 3224, 4)  # fixed-point-with-gain: not supported
    open(fmt4, "wb").write(bytes(raw))
    with pytest.raises(sb.SbError):
        sb.SuGather(fmt4, sb.METHOD_CMP, 1000, segy=True)
    trunc = str(tmp_path / "trunc.sgy")
    write_segy(trunc, read_su(su)[:3], 5, 0)
    raw = open(trunc, "rb").read()
    open(trunc, "wb").write(raw[:-100])
    with pytest.raises(sb.SbError):
        sb.SuGather(trunc, sb.METHOD_CMP, 1000, segy=True)
