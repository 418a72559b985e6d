"""A small PNG reader written from the PNG specification (ISO/IEC 15948),
used to check the library's writer: 8-bit RGBA, non-interlaced; every chunk
CRC is verified and every filter type is undone."""
import struct
import zlib

import numpy as np

SIGNATURE = b"\x89PNG\r\n\x1a\n"


def chunks(data):
    assert data[:8] == SIGNATURE, "bad signature"
    pos = 8
    while pos < len(data):
        (n,) = struct.unpack(">I", data[pos:pos + 4])
        ctype = data[pos + 4:pos + 8]
        body = data[pos + 8:pos + 8 + n]
        (crc,) = struct.unpack(">I", data[pos + 8 + n:pos + 12 + n])
        assert zlib.crc32(ctype + body) & 0xFFFFFFFF == crc, f"CRC mismatch in {ctype!r}"
        yield ctype, body
        pos += 12 + n


def _paeth(a, b, c):
    p = a + b - c
    pa, pb, pc = abs(p - a), abs(p - b), abs(p - c)
    if pa <= pb and pa <= pc:
        return a
    return b if pb <= pc else c


def decode(data):
    """-> ((H, W, 4) uint8 array with row 0 = first row in the file, filter types used)."""
    ihdr, idat, seen_end = None, b"", False
    for ctype, body in chunks(data):
        if ctype == b"IHDR":
            ihdr = struct.unpack(">IIBBBBB", body)
        elif ctype == b"IDAT":
            idat += body
        elif ctype == b"IEND":
            seen_end = True
    assert ihdr is not None and seen_end
    w, h, depth, ctype, comp, filt, interlace = ihdr
    assert (depth, ctype, comp, filt, interlace) == (8, 6, 0, 0, 0)
    raw = zlib.decompress(idat)
    stride = 4 * w
    assert len(raw) == h * (stride + 1)
    out = np.zeros((h, stride), np.uint8)
    types = []
    prev = [0] * stride
    for r in range(h):
        f = raw[r * (stride + 1)]
        types.append(f)
        line = list(raw[r * (stride + 1) + 1:(r + 1) * (stride + 1)])
        cur = [0] * stride
        for i in range(stride):
            a = cur[i - 4] if i >= 4 else 0
            b = prev[i]
            c = prev[i - 4] if i >= 4 else 0
            pred = {0: 0, 1: a, 2: b, 3: (a + b) // 2, 4: _paeth(a, b, c)}[f]
            cur[i] = (line[i] + pred) & 0xFF
        out[r] = cur
        prev = cur
    return out.reshape(h, w, 4), types
