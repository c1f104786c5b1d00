"""Decoder for 8-bit greyscale, non-interlaced PNG files (the reconstruction previews)."""
import struct
import zlib

import numpy as np


def read_png_gray8(path):
    b = open(path, "rb").read()
    assert b[:8] == b"\x89PNG\r\n\x1a\n", "not a PNG file"
    pos, idat, w = 8, b"", None
    while pos < len(b):
        n, = struct.unpack(">I", b[pos:pos + 4])
        typ, data = b[pos + 4:pos + 8], b[pos + 8:pos + 8 + n]
        pos += 12 + n
        if typ == b"IHDR":
            w, h, depth, ctype, _, _, interlace = struct.unpack(">IIBBBBB", data)
            assert depth == 8 and ctype == 0 and interlace == 0, "only 8-bit grey, non-interlaced"
        elif typ == b"IDAT":
            idat += data
        elif typ == b"IEND":
            break
    raw = zlib.decompress(idat)
    out = np.zeros((h, w), dtype=np.int64)
    prev = np.zeros(w, dtype=np.int64)
    for y in range(h):
        f = raw[y * (w + 1)]
        line = np.frombuffer(raw[y * (w + 1) + 1:(y + 1) * (w + 1)], dtype=np.uint8).astype(np.int64)
        cur = np.zeros(w, dtype=np.int64)
        for x in range(w):
            a = cur[x - 1] if x > 0 else 0
            c = prev[x - 1] if x > 0 else 0
            up = prev[x]
            if f == 0:
                p = 0
            elif f == 1:
                p = a
            elif f == 2:
                p = up
            elif f == 3:
                p = (a + up) // 2
            else:  # Paeth
                pa, pb, pc = abs(up - c), abs(a - c), abs(a + up - 2 * c)
                p = a if (pa <= pb and pa <= pc) else (up if pb <= pc else c)
            cur[x] = (line[x] + p) & 0xff
        out[y] = cur
        prev = cur
    return out.astype(np.uint8)
