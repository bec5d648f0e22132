"""An SU -> SEG-Y writer for the tests, independent of the library: the
240-byte trace header is byte-swapped field by field from the SEG-Y rev 1
layout; samples in IEEE, IBM or integer formats."""
import struct

import numpy as np

# SEG-Y rev 1 trace header: (byte offset, width) of every field.
FIELDS = ([(4 * i, 4) for i in range(7)] + [(28 + 2 * i, 2) for i in range(4)] + [(36 + 4 * i, 4) for i in range(8)] +
          [(68, 2), (70, 2)] + [(72 + 4 * i, 4) for i in range(4)] + [(88 + 2 * i, 2) for i in range(46)] +
          [(180 + 4 * i, 4) for i in range(5)] + [(200, 2), (202, 2), (204, 4)] +
          [(208 + 2 * i, 2) for i in range(8)] + [(224, 4), (228, 2), (230, 2), (232, 4), (236, 4)])


def read_su(path):
    raw = open(path, "rb").read()
    out, pos = [], 0
    while pos + 240 <= len(raw):
        ns = struct.unpack_from("<H", raw, pos + 114)[0]
        hdr = raw[pos:pos + 240]
        data = np.frombuffer(raw, "<f4", ns, pos + 240).copy()
        out.append((hdr, data))
        pos += 240 + 4 * ns
    return out


def swap_header(hdr):
    be = bytearray(240)
    for off, w in FIELDS:
        be[off:off + w] = hdr[off:off + w][::-1]
    return bytes(be)


def float_to_ibm(v):
    """IEEE value -> IBM single (exact for values with <= 21 significant bits)."""
    if v == 0:
        return 0
    sign = 0x80000000 if v < 0 else 0
    m, e2 = np.frexp(abs(float(v)))  # |v| = m * 2^e2, m in [0.5, 1)
    e16 = -(-int(e2) // 4)  # ceil(e2 / 4): |v| = F * 16^e16, F in [1/16, 1)
    frac = int(round(abs(float(v)) / 16.0 ** e16 * (1 << 24)))
    if frac >= 1 << 24:
        frac >>= 4
        e16 += 1
    return sign | ((e16 + 64) << 24) | frac


def write_segy(path, traces, fmt, dt_us, zero_trace_ns=False, ext_text=0):
    ns = len(traces[0][1])
    binh = bytearray(400)
    struct.pack_into(">H", binh, 16, dt_us)
    struct.pack_into(">H", binh, 20, ns)
    struct.pack_into(">H", binh, 24, fmt)
    struct.pack_into(">h", binh, 304, ext_text)
    with open(path, "wb") as f:
        f.write(b"\x40" * 3200)
        f.write(bytes(binh))
        f.write(b"\x40" * (3200 * ext_text))
        for hdr, data in traces:
            be = bytearray(swap_header(hdr))
            if zero_trace_ns:
                be[114:118] = b"\0\0\0\0"
            f.write(bytes(be))
            if fmt == 5:
                f.write(data.astype(">f4").tobytes())
            elif fmt == 1:
                f.write(b"".join(struct.pack(">I", float_to_ibm(x)) for x in data))
            elif fmt == 2:
                f.write(data.astype(">i4").tobytes())
            elif fmt == 3:
                f.write(data.astype(">i2").tobytes())
            elif fmt == 8:
                f.write(data.astype("i1").tobytes())
