"""Write tests/golden/formats.npz: small PNG, TIFF, DM3/DM4 and MetaImage files together with the
arrays the reference's loader (qc::ScalarArray<QC_2D>(file), through oracle/_ref/ref_harness "load",
built against the system libpng and libtiff) reads from them, or the fact that it refused them.

The files come from the encoders below (every PNG colour type, bit depth, filter and Adam7; TIFF in
both byte orders, classic and BigTIFF, 8-64-bit int/uint/float samples, none/LZW/Deflate/PackBits,
both predictors, several strips, FillOrder 2; DM3/DM4 tag trees around a thumbnail and an image;
MetaImage headers) and from PIL's writers (libtiff's LZW and Deflate among them), so the reference's
decoders judge each encoding. Runs in the build container only.
Usage: python tests/golden/make_formats.py
"""
import io
import os
import struct
import subprocess
import sys
import tempfile
import zlib

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
HARNESS = os.path.join(ROOT, "oracle", "_ref", "ref_harness")


# ------------------------------------------------------------------------------------------------
# PNG
# ------------------------------------------------------------------------------------------------

def png_chunk(typ, data, bad_crc=False):
    crc = zlib.crc32(typ + data) & 0xffffffff
    return struct.pack(">I", len(data)) + typ + data + struct.pack(">I", crc ^ (1 if bad_crc else 0))


def png_filter(ftype, line, prev, bpp):
    out = bytearray(len(line))
    for i in range(len(line)):
        a = line[i - bpp] if i >= bpp else 0
        b = prev[i]
        c = prev[i - bpp] if i >= bpp else 0
        if ftype == 0:
            p = 0
        elif ftype == 1:
            p = a
        elif ftype == 2:
            p = b
        elif ftype == 3:
            p = (a + b) // 2
        else:
            pp = a + b - c
            pa, pb, pc = abs(pp - a), abs(pp - b), abs(pp - c)
            p = a if (pa <= pb and pa <= pc) else (b if pb <= pc else c)
        out[i] = (line[i] - p) & 0xff
    return bytes(out)


def png_pack_row(samples, depth):
    """samples: 1-D int array (channels interleaved) -> packed scanline bytes"""
    if depth == 8:
        return bytes(np.asarray(samples, dtype=np.uint8))
    if depth == 16:
        return np.asarray(samples, dtype=">u2").tobytes()
    bits = "".join(format(int(s), f"0{depth}b") for s in samples)
    bits += "0" * (-len(bits) % 8)
    return bytes(int(bits[i:i + 8], 2) for i in range(0, len(bits), 8))


def make_png(img, depth, ctype, interlace=False, palette=None, trns=None, idat_pieces=1, extra=(), filt_seed=0):
    """img: (h, w) or (h, w, c) integer samples of the file's colour type"""
    h, w = img.shape[:2]
    ch = {0: 1, 2: 3, 3: 1, 4: 2, 6: 4}[ctype]
    img = img.reshape(h, w, ch)
    bpp = max(1, ch * depth // 8)
    rng = np.random.default_rng(filt_seed)
    passes = [(0, 0, 8, 8), (4, 0, 8, 8), (0, 4, 4, 8), (2, 0, 4, 4), (0, 2, 2, 4), (1, 0, 2, 2), (0, 1, 1, 2)] \
        if interlace else [(0, 0, 1, 1)]
    raw = b""
    for x0, y0, dx, dy in passes:
        sub = img[y0::dy, x0::dx]
        if sub.size == 0:
            continue
        rowbytes = len(png_pack_row(sub[0].ravel(), depth))
        prev = bytes(rowbytes)
        for r in range(sub.shape[0]):
            line = png_pack_row(sub[r].ravel(), depth)
            ft = int(rng.integers(0, 5))
            raw += bytes([ft]) + png_filter(ft, line, prev, bpp)
            prev = line
    z = zlib.compress(raw, 9)
    out = b"\x89PNG\r\n\x1a\n" + png_chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, depth, ctype, 0, 0, int(interlace)))
    for typ, data, bad in extra:
        out += png_chunk(typ, data, bad)
    if palette is not None:
        out += png_chunk(b"PLTE", bytes(np.asarray(palette, dtype=np.uint8).ravel()))
    if trns is not None:
        out += png_chunk(b"tRNS", trns)
    k = max(1, len(z) // idat_pieces + 1)
    for i in range(0, len(z), k):
        out += png_chunk(b"IDAT", z[i:i + k])
    return out + png_chunk(b"IEND", b"")


def png_cases(rng):
    cases = {}
    for depth in (1, 2, 4, 8):
        img = rng.integers(0, 1 << depth, (7, 13))
        cases[f"png_gray{depth}"] = make_png(img, depth, 0, filt_seed=depth)
        cases[f"png_gray{depth}_adam7"] = make_png(rng.integers(0, 1 << depth, (11, 13)), depth, 0, interlace=True,
                                                   filt_seed=depth + 10)
    cases["png_gray8_adam7_tiny"] = make_png(rng.integers(0, 256, (2, 3)), 8, 0, interlace=True)
    cases["png_gray8_trns_pieces"] = make_png(rng.integers(0, 256, (9, 10)), 8, 0, trns=b"\x00\x07", idat_pieces=4,
                                              extra=[(b"tEXt", b"Comment\x00test", False),
                                                     (b"zTXt", b"bad\x00\x00xx", True)])
    cases["png_rgb8"] = make_png(rng.integers(0, 256, (7, 13, 3)), 8, 2, filt_seed=3)
    cases["png_rgb8_adam7_trns"] = make_png(rng.integers(0, 256, (10, 9, 3)), 8, 2, interlace=True,
                                            trns=b"\x00\x01\x00\x02\x00\x03")
    cases["png_rgba8"] = make_png(rng.integers(0, 256, (7, 13, 4)), 8, 6, filt_seed=4)
    cases["png_rgba8_adam7"] = make_png(rng.integers(0, 256, (9, 12, 4)), 8, 6, interlace=True, filt_seed=5)
    cases["png_ga8"] = make_png(rng.integers(0, 256, (7, 13, 2)), 8, 4, filt_seed=6)
    for depth in (1, 2, 4, 8):
        npal = min(1 << depth, 5 if depth >= 4 else 1 << depth)
        pal = rng.integers(0, 256, (npal, 3))
        img = rng.integers(0, 1 << depth, (7, 13))  # 4- and 8-bit: indices beyond the palette too
        trns = bytes(rng.integers(0, 256, min(npal, 3)).astype(np.uint8)) if depth != 2 else None
        cases[f"png_pal{depth}"] = make_png(img, depth, 3, palette=pal, trns=trns, filt_seed=20 + depth)
    cases["png_pal4_adam7"] = make_png(rng.integers(0, 16, (11, 9)), 4, 3, interlace=True,
                                       palette=rng.integers(0, 256, (16, 3)))
    cases["png_gray16"] = make_png(rng.integers(0, 65536, (5, 6)), 16, 0)
    cases["png_rgb16"] = make_png(rng.integers(0, 65536, (5, 6, 3)), 16, 2)
    from PIL import Image
    for mode, arr in (("L", rng.integers(0, 256, (9, 14)).astype(np.uint8)),
                      ("RGB", rng.integers(0, 256, (9, 14, 3)).astype(np.uint8)),
                      ("LA", rng.integers(0, 256, (9, 14, 2)).astype(np.uint8))):
        b = io.BytesIO()
        Image.fromarray(arr).save(b, format="PNG", optimize=True)
        cases[f"png_pil_{mode}"] = b.getvalue()
    b = io.BytesIO()
    Image.fromarray(rng.integers(0, 256, (9, 14, 3)).astype(np.uint8)).quantize(7).save(b, format="PNG")
    cases["png_pil_P"] = b.getvalue()
    return {k + ".png": v for k, v in cases.items()}


# ------------------------------------------------------------------------------------------------
# TIFF
# ------------------------------------------------------------------------------------------------

def lzw_encode(data):
    """TIFF LZW: MSB-first codes, clear first, code width growing one code early, clear at 4094."""
    out, acc, nacc = bytearray(), 0, 0
    nbits = 9

    def put(code):
        nonlocal acc, nacc
        acc = acc << nbits | code
        nacc += nbits
        while nacc >= 8:
            nacc -= 8
            out.append((acc >> nacc) & 0xff)
        acc &= (1 << nacc) - 1

    def reset():
        return {bytes([i]): i for i in range(256)}, 258

    table, free = reset()
    put(256)
    w = b""
    for c in data:
        wc = w + bytes([c])
        if wc in table:
            w = wc
            continue
        put(table[w])
        table[wc] = free
        free += 1
        if free == 4094:
            put(256)
            nbits = 9
            table, free = reset()
        elif free > (1 << nbits) - 1:
            nbits += 1
        w = bytes([c])
    if w:
        put(table[w])
        free += 1
        if free == 4094:
            put(256)
            nbits = 9
        elif free > (1 << nbits) - 1:
            nbits += 1
    put(257)
    if nacc:
        out.append((acc << (8 - nacc)) & 0xff)
    return bytes(out)


def packbits_encode(data):
    out, i, n = bytearray(), 0, len(data)
    while i < n:
        j = i
        while j + 1 < n and data[j + 1] == data[i] and j - i < 127:
            j += 1
        if j - i >= 2:
            out += bytes([(1 - (j - i + 1)) & 0xff, data[i]])
            i = j + 1
            continue
        k = i
        while k < n and k - i < 128 and not (k + 2 < n and data[k] == data[k + 1] == data[k + 2]):
            k += 1
        out += bytes([k - i - 1]) + data[i:k]
        i = k
    return bytes(out)


def make_tiff(img, dtype, order="<", big=False, compression=1, predictor=1, rows_per_strip=None, fill_order=None,
              sample_format=True, spp=1):
    """img: (h, w) array; dtype: numpy sample type"""
    h, w = img.shape[0], img.shape[1] // spp
    a = np.asarray(img).astype(dtype).reshape(h, w * spp)
    bits = a.dtype.itemsize * 8
    fmt = 3 if a.dtype.kind == "f" else (2 if a.dtype.kind == "i" else 1)
    rps = rows_per_strip or h
    strips = []
    for s in range(0, h, rps):
        rows = a[s:s + rps]
        if predictor == 2:
            u = rows.view(f"u{a.dtype.itemsize}").astype(np.uint64)
            d = u.copy()
            d[:, spp:] = (u[:, spp:] - u[:, :-spp]) & ((1 << bits) - 1 if bits < 64 else 0xffffffffffffffff)
            raw = d.astype(f"{order}u{a.dtype.itemsize}").tobytes()
        elif predictor == 3:
            raw = b""
            for r in rows:
                be = np.asarray(r, dtype=">" + a.dtype.str[1:]).tobytes()
                nb = a.dtype.itemsize
                planes = bytearray(nb * len(r))
                for j in range(len(r)):
                    for k in range(nb):
                        planes[k * len(r) + j] = be[j * nb + k]
                for i in range(len(planes) - 1, 0, -1):
                    planes[i] = (planes[i] - planes[i - 1]) & 0xff
                raw += bytes(planes)
        else:
            raw = rows.astype(order + a.dtype.str[1:]).tobytes()
        if compression == 5:
            raw = lzw_encode(raw)
        elif compression in (8, 32946):
            raw = zlib.compress(raw)
        elif compression == 32773:  # PackBits runs stop at the end of each row (TIFF 6.0 section 9)
            rb = len(raw) // len(rows)
            raw = b"".join(packbits_encode(raw[i:i + rb]) for i in range(0, len(raw), rb))
        if fill_order == 2:
            raw = bytes(int(format(c, "08b")[::-1], 2) for c in raw)
        strips.append(raw)
    tags = [(256, 4, [w]), (257, 4, [h]), (258, 3, [bits] * spp), (259, 3, [compression]), (262, 3, [1])]
    if fill_order:
        tags.append((266, 3, [fill_order]))
    tags += [(273, 16 if big else 4, None), (277, 3, [spp]), (278, 4, [rps]),
             (279, 16 if big else 4, [len(s) for s in strips])]
    if predictor != 1:
        tags.append((317, 3, [predictor]))
    if sample_format:
        tags.append((339, 3, [fmt] * spp))
    tags.sort()
    hdr_len = 16 if big else 8
    data = b"".join(strips)
    offsets, p = [], hdr_len
    for s in strips:
        offsets.append(p)
        p += len(s)
    ifd_pos = hdr_len + len(data)
    ifd_pos += ifd_pos & 1
    esz, cnt_fmt, off_fmt = (20, "Q", "Q") if big else (12, "H", "I")
    n_inline = 8 if big else 4
    ifd = struct.pack(order + cnt_fmt, len(tags))
    extra, extra_pos = b"", ifd_pos + len(ifd) + esz * len(tags) + (8 if big else 4)
    tsize = {3: 2, 4: 4, 16: 8}
    for tag, typ, vals in tags:
        if vals is None:
            vals = offsets
        vfmt = {3: "H", 4: "I", 16: "Q"}[typ]
        payload = struct.pack(order + vfmt * len(vals), *vals)
        if len(payload) <= n_inline:
            field = payload + bytes(n_inline - len(payload))
        else:
            field = struct.pack(order + off_fmt, extra_pos + len(extra))
            extra += payload
            extra += bytes(len(extra) & 1)
        ifd += struct.pack(order + "HH" + ("Q" if big else "I"), tag, typ, len(vals)) + field
    ifd += bytes(8 if big else 4)
    mark = b"II" if order == "<" else b"MM"
    hdr = mark + (struct.pack(order + "HHHQ", 43, 8, 0, ifd_pos) if big else struct.pack(order + "HI", 42, ifd_pos))
    body = hdr + data
    body += bytes(ifd_pos - len(body))
    return body + ifd + extra


def tiff_cases(rng):
    cases = {}

    def sample(dt, shape=(7, 13)):
        dt = np.dtype(dt)
        if dt.kind == "f":
            return (rng.standard_normal(shape) * 1e3).astype(dt)
        info = np.iinfo(dt)
        lo, hi = max(info.min, -2 ** 62), min(info.max, 2 ** 62)
        return rng.integers(lo, hi, shape, endpoint=True).astype(dt)

    for dt in ("u1", "i1", "u2", "i2", "u4", "i4", "u8", "i8", "f4", "f8"):
        cases[f"tif_{dt}_raw"] = make_tiff(sample(dt), dt)
        cases[f"tif_{dt}_mm_lzw_pred"] = make_tiff(sample(dt), dt, order=">", compression=5,
                                                   predictor=3 if dt[0] == "f" else 2, rows_per_strip=3)
    cases["tif_u2_deflate_strips"] = make_tiff(sample("u2", (9, 10)), "u2", compression=8, rows_per_strip=4)
    cases["tif_f4_adobe_deflate_fp"] = make_tiff(sample("f4", (9, 10)), "f4", compression=32946, predictor=3)
    cases["tif_f8_mm_deflate_fp"] = make_tiff(sample("f8", (6, 5)), "f8", order=">", compression=8, predictor=3)
    cases["tif_i2_packbits"] = make_tiff(np.repeat(sample("i2", (9, 5)), 3, axis=1), "i2", compression=32773,
                                         rows_per_strip=2)
    cases["tif_u1_packbits_runs"] = make_tiff(np.repeat(sample("u1", (6, 4)), 40, axis=1), "u1", compression=32773)
    cases["tif_u2_lzw_clear"] = make_tiff(sample("u2", (64, 48)), "u2", compression=5)
    cases["tif_u1_lzw_runs"] = make_tiff(np.repeat(sample("u1", (20, 10)), 30, axis=1), "u1", compression=5,
                                         rows_per_strip=7)
    cases["tif_f4_bigtiff"] = make_tiff(sample("f4"), "f4", big=True, rows_per_strip=2)
    cases["tif_u2_bigtiff_mm_lzw"] = make_tiff(sample("u2"), "u2", order=">", big=True, compression=5, predictor=2)
    cases["tif_u2_fillorder2"] = make_tiff(sample("u2"), "u2", fill_order=2)
    cases["tif_u1_fillorder2_lzw"] = make_tiff(sample("u1"), "u1", compression=5, fill_order=2)
    cases["tif_u2_no_sampleformat"] = make_tiff(sample("u2"), "u2", sample_format=False)
    cases["tif_u1_two_samples"] = make_tiff(sample("u1", (5, 16)), "u1", spp=2)
    cases["tif_f2_raw"] = make_tiff(np.ones((3, 4)), "f2")
    from PIL import Image
    for mode, arr in (("L", sample("u1", (11, 17))), ("I;16", sample("u2", (11, 17))),
                      ("I", sample("i4", (11, 17))), ("F", sample("f4", (11, 17)))):
        for comp in ("raw", "tiff_lzw", "tiff_deflate", "packbits"):
            b = io.BytesIO()
            Image.fromarray(arr).save(b, format="TIFF", compression=comp)
            cases[f"tif_pil_{mode.replace(';', '')}_{comp}"] = b.getvalue()
    return {k + ".tif": v for k, v in cases.items()}


# ------------------------------------------------------------------------------------------------
# Gatan DigitalMicrograph 3 / 4
# ------------------------------------------------------------------------------------------------

DM_CODES = {"i2": 2, "i4": 3, "u2": 4, "u4": 5, "f4": 6, "f8": 7, "b1": 8, "u1": 10, "i8": 11, "u8": 12}


def dm_file(version, root):
    """root: list of tags; a tag is (name, value) with value a dict-like list (group) or a leaf
    ('simple', code, bytes) / ('array', code, count, bytes) / ('struct', [codes], bytes) /
    ('structarray', [codes], count, bytes)."""
    cnt = ">i" if version == 3 else ">q"

    def group(tags):
        b = bytes([0, 0]) + struct.pack(cnt, len(tags))
        for name, val in tags:
            b += entry(name, val)
        return b

    def entry(name, val):
        nm = name.encode()
        if isinstance(val, list):
            ident, body = 20, group(val)
        else:
            ident = 21
            kind = val[0]
            if kind == "simple":
                info = [val[1]]
                payload = val[2]
            elif kind == "array":
                info = [20, val[1], val[2]]
                payload = val[3]
            elif kind == "struct":
                info = [15, 0, len(val[1])] + sum(([0, c] for c in val[1]), [])
                payload = val[2]
            else:
                info = [20, 15, 0, len(val[1])] + sum(([0, c] for c in val[1]), []) + [val[2]]
                payload = val[3]
            body = b"%%%%" + struct.pack(cnt, len(info)) + b"".join(struct.pack(cnt, x) for x in info) + payload
        size = struct.pack(">q", len(body)) if version == 4 else b""
        return bytes([ident]) + struct.pack(">h", len(nm)) + nm + size + body

    body = group(root)
    flen = struct.pack(">i", len(body) + 4) if version == 3 else struct.pack(">q", len(body) + 4)
    return struct.pack(">i", version) + flen + struct.pack(">i", 1) + body + bytes(8)


def dm_image(arr, code, dims=None):
    arr = np.asarray(arr)
    dims = dims or [arr.shape[-1], arr.shape[-2]] + ([arr.shape[0]] if arr.ndim == 3 else [])
    return [("Calibrations", [("Brightness", [("Origin", ("simple", 6, struct.pack("<f", 0.5)))])]),
            ("Data", ("array", code, arr.size, arr.astype("<" + arr.dtype.str[1:]).tobytes())),
            ("DataType", ("simple", 5, struct.pack("<I", code))),
            ("Dimensions", [("", ("simple", 5, struct.pack("<I", d))) for d in dims]),
            ("PixelDepth", ("simple", 5, struct.pack("<I", arr.dtype.itemsize)))]


def dm_cases(rng):
    cases = {}
    name = ("array", 4, 5, np.frombuffer("Frame".encode("utf-16-le"), "<u2").tobytes())
    for version in (3, 4):
        for dt in ("i2", "i4", "u2", "u4", "f4", "u1", "f8"):
            if dt[0] == "f":
                img = (rng.standard_normal((7, 13)) * 100).astype(dt)
            else:
                info = np.iinfo(dt)
                img = rng.integers(max(info.min, -10 ** 6), min(info.max, 10 ** 6), (7, 13)).astype(dt)
            thumb = rng.integers(0, 255, (4, 6)).astype("u1")
            root = [("ApplicationBounds", ("array", 3, 4, struct.pack("<4i", 0, 0, 800, 600))),
                    ("DocumentObjectList", [("", [("AnnotationType", ("simple", 3, struct.pack("<i", 20))),
                                                  ("Rectangle", ("struct", [6, 6, 6, 6], struct.pack("<4f", 1, 2, 3, 4))),
                                                  ("Flag", ("simple", 8, b"\x01")),
                                                  ("Points", ("structarray", [7, 7], 3, np.arange(6.0).tobytes())),
                                                  ("Id", ("simple", 12, struct.pack("<Q", 7)))])]),
                    ("ImageList", [("", [("ImageData", dm_image(thumb, 10)), ("Name", name)]),
                                   ("", [("ImageData", dm_image(img, DM_CODES[dt])), ("Name", name),
                                         ("UniqueID", ("struct", [5, 5, 5, 5], struct.pack("<4I", 1, 2, 3, 4)))])]),
                    ("InImageMode", ("simple", 10, b"\x01"))]
            cases[f"dm{version}_{dt}"] = dm_file(version, root)
        stack = rng.integers(0, 4000, (3, 5, 6)).astype("u2")
        cases[f"dm{version}_stack"] = dm_file(version, [("ImageList", [("", [("ImageData", dm_image(stack, 4))])])])
    return {k + (".dm3" if k.startswith("dm3") else ".dm4"): v for k, v in cases.items()}


# ------------------------------------------------------------------------------------------------
# MetaImage
# ------------------------------------------------------------------------------------------------

def mhd_cases(rng):
    cases = {}
    for met, dt in (("MET_FLOAT", "<f4"), ("MET_USHORT", "<u2"), ("MET_UCHAR", "u1")):
        arr = (rng.standard_normal((7, 13)) * 100).astype(dt) if dt == "<f4" else rng.integers(0, 250, (7, 13)).astype(dt)
        hdr = (f"NDims = 2\nDimSize = 13 7\nElementType = {met}\nElementSpacing = 1.0 1.0 1.0\n"
               f"ElementByteOrderMSB = False\nElementDataFile = mhd_{met}.raw\n").encode()
        cases[f"mhd_{met}.mhd"] = hdr
        cases[f"mhd_{met}.raw"] = arr.tobytes()
    return cases


def main():
    if not os.path.exists(HARNESS):
        sys.exit("oracle/_ref is not built (make -C oracle ref)")
    rng = np.random.default_rng(2024)
    files = {}
    files.update(png_cases(rng))
    files.update(tiff_cases(rng))
    files.update(dm_cases(rng))
    files.update(mhd_cases(rng))
    out = {}
    with tempfile.TemporaryDirectory() as tmp:
        for name, data in files.items():
            with open(os.path.join(tmp, name), "wb") as f:
                f.write(data)
        fin, fout = os.path.join(tmp, "in.bin"), os.path.join(tmp, "out.bin")
        np.zeros(1).tofile(fin)
        for name, data in sorted(files.items()):
            out[f"file:{name}"] = np.frombuffer(data, dtype=np.uint8)
            if name.endswith(".raw"):
                continue
            if os.path.exists(fout):
                os.remove(fout)
            r = subprocess.run([HARNESS, "load", "2", os.path.join(tmp, name), fin, fout], capture_output=True,
                               text=True)
            # libtiff reports a strip it cannot decode and leaves the scanline buffer undefined: such
            # a file counts as refused, like one the loader throws on (warnings do not count)
            errors = [ln for ln in r.stderr.splitlines() if ln.strip() and "warning" not in ln.lower()]
            if r.returncode != 0 or errors:
                out[f"error:{name}"] = np.array(1)
                print(f"{name}: refused by the reference loader", errors[:1])
                continue
            o = np.fromfile(fout, dtype="<f8")
            nx, ny = int(o[0]), int(o[1])
            out[f"array:{name}"] = o[2:2 + nx * ny].reshape(ny, nx)
            print(f"{name}: {nx}x{ny}")
    np.savez_compressed(os.path.join(HERE, "formats.npz"), **out)
    print("wrote", sum(k.startswith("file:") for k in out), "files")


if __name__ == "__main__":
    main()
