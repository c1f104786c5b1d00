#include "image_formats.h"

#include <zlib.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <sstream>
#include <vector>

#include "params.h"

namespace mshost {

namespace {

typedef std::vector<uint8_t> Bytes;

// The reference indexes its arrays with int: larger images cannot be loaded there either.
void checkPixels(uint64_t nx, uint64_t ny, const std::string &name) {
  if (nx > 0x7fffffffu || ny > 0x7fffffffu || nx * ny > 0x7fffffffu)
    MSH_THROW(strprintf("%s: image size %llux%llu exceeds the array size limit", name.c_str(), (unsigned long long)nx,
                        (unsigned long long)ny));
}

// Upper bounds of the decoded size of n compressed bytes: Deflate expands at most 1032:1, a 9-bit LZW
// code at most to 4096 bytes, a PackBits pair to 128 bytes. Checked before any buffer is sized.
size_t maxInflated(size_t n) { return n * 1032 + 64; }

uint32_t be32(const uint8_t *p) { return (uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | p[3]; }

// A zlib stream inflated into exactly `want` bytes. A stream that ends early is an error (libpng's
// "Not enough image data", libtiff's "Not enough data"); data beyond `want` is ignored.
Bytes inflateExact(const uint8_t *src, size_t n, size_t want, const std::string &what) {
  Bytes out(want);
  if (want == 0) return out;
  z_stream z;
  memset(&z, 0, sizeof(z));
  if (inflateInit(&z) != Z_OK) MSH_THROW("zlib initialisation failed");
  z.next_in = const_cast<Bytef *>(src);
  z.next_out = out.data();
  size_t inLeft = n, outLeft = want;
  int rc = Z_OK;
  while (outLeft > 0) {
    const uInt ci = (uInt)std::min<size_t>(inLeft, 1u << 30), co = (uInt)std::min<size_t>(outLeft, 1u << 30);
    z.avail_in = ci;
    z.avail_out = co;
    rc = inflate(&z, Z_NO_FLUSH);
    inLeft -= ci - z.avail_in;
    outLeft -= co - z.avail_out;
    if (rc == Z_STREAM_END) break;
    if (rc != Z_OK && !(rc == Z_BUF_ERROR && inLeft > 0)) break;
  }
  inflateEnd(&z);
  if (outLeft > 0)
    MSH_THROW(strprintf("%s: %s", what.c_str(),
                        rc == Z_STREAM_END || rc == Z_BUF_ERROR ? "not enough compressed data" : "corrupt zlib data"));
  return out;
}

// ------------------------------------------------------------------------------------------------
// PNG
// ------------------------------------------------------------------------------------------------

int pngChannels(int colorType) {
  switch (colorType) {
    case 0: return 1;
    case 2: return 3;
    case 3: return 1;
    case 4: return 2;
    case 6: return 4;
  }
  return 0;
}

// Reverses the adaptive filter of one scanline in place (PNG spec. 9.2-9.4); `prev` is the
// reconstructed previous line of the same pass, or null for the first.
void pngUnfilter(int type, uint8_t *cur, const uint8_t *prev, size_t len, size_t bpp, const std::string &name) {
  switch (type) {
    case 0:
      return;
    case 1:
      for (size_t i = bpp; i < len; ++i) cur[i] = (uint8_t)(cur[i] + cur[i - bpp]);
      return;
    case 2:
      if (prev)
        for (size_t i = 0; i < len; ++i) cur[i] = (uint8_t)(cur[i] + prev[i]);
      return;
    case 3:
      for (size_t i = 0; i < len; ++i) {
        const unsigned a = i >= bpp ? cur[i - bpp] : 0, b = prev ? prev[i] : 0;
        cur[i] = (uint8_t)(cur[i] + ((a + b) >> 1));
      }
      return;
    case 4:
      for (size_t i = 0; i < len; ++i) {
        const int a = i >= bpp ? cur[i - bpp] : 0, b = prev ? prev[i] : 0, c = (i >= bpp && prev) ? prev[i - bpp] : 0;
        const int p = a + b - c, pa = abs(p - a), pb = abs(p - b), pc = abs(p - c);
        const int pred = (pa <= pb && pa <= pc) ? a : (pb <= pc ? b : c);
        cur[i] = (uint8_t)(cur[i] + pred);
      }
      return;
  }
  MSH_THROW(strprintf("%s: bad PNG filter type %d", name.c_str(), type));
}

}  // namespace

Image decodePNG(const std::string &bytes, const std::string &name) {
  const uint8_t *b = reinterpret_cast<const uint8_t *>(bytes.data());
  const size_t n = bytes.size();
  static const uint8_t sig[8] = {137, 80, 78, 71, 13, 10, 26, 10};
  if (n < 4 || memcmp(b, sig, 4) != 0) MSH_THROW(strprintf("%s: File is not a png", name.c_str()));
  if (n < 8 || memcmp(b, sig, 8) != 0) MSH_THROW(strprintf("%s: corrupt PNG signature", name.c_str()));

  uint32_t w = 0, h = 0;
  int depth = 0, colorType = -1, interlace = 0;
  std::vector<uint8_t> palette;  // RGB triples
  Bytes idat;
  bool haveHeader = false, end = false;
  size_t pos = 8;
  while (!end) {
    if (pos + 12 > n) MSH_THROW(strprintf("%s: PNG ends before its IEND chunk", name.c_str()));
    const uint32_t len = be32(b + pos);
    if (len > 0x7fffffffu || pos + 12 + (size_t)len > n) MSH_THROW(strprintf("%s: truncated PNG chunk", name.c_str()));
    const uint8_t *type = b + pos + 4, *data = b + pos + 8;
    const std::string t(reinterpret_cast<const char *>(type), 4);
    const bool critical = !(type[0] & 0x20);
    const bool crcOk = (uint32_t)crc32(0L, type, 4 + len) == be32(data + len);
    pos += 12 + (size_t)len;
    if (!crcOk) {
      if (critical) MSH_THROW(strprintf("%s: CRC error in PNG chunk %s", name.c_str(), t.c_str()));
      continue;  // a damaged ancillary chunk is dropped, as libpng does by default
    }
    if (!haveHeader && t != "IHDR") MSH_THROW(strprintf("%s: PNG does not start with IHDR", name.c_str()));
    if (t == "IHDR") {
      if (haveHeader || len != 13) MSH_THROW(strprintf("%s: invalid IHDR chunk", name.c_str()));
      w = be32(data);
      h = be32(data + 4);
      depth = data[8];
      colorType = data[9];
      interlace = data[12];
      const bool depthOk =
          (colorType == 0 && (depth == 1 || depth == 2 || depth == 4 || depth == 8 || depth == 16)) ||
          (colorType == 3 && (depth == 1 || depth == 2 || depth == 4 || depth == 8)) ||
          ((colorType == 2 || colorType == 4 || colorType == 6) && (depth == 8 || depth == 16));
      if (!depthOk || data[10] != 0 || data[11] != 0 || interlace > 1)
        MSH_THROW(strprintf("%s: invalid PNG header (bit depth %d, colour type %d)", name.c_str(), depth, colorType));
      // libpng's default user limits, 1000000 pixels either way
      if (w == 0 || h == 0 || w > 1000000u || h > 1000000u)
        MSH_THROW(strprintf("%s: invalid PNG image size %ux%u", name.c_str(), w, h));
      checkPixels(w, h, name);
      haveHeader = true;
    } else if (t == "PLTE") {
      if (len == 0 || len % 3 != 0 || len / 3 > 256) MSH_THROW(strprintf("%s: invalid PLTE chunk", name.c_str()));
      if (colorType == 3) palette.assign(data, data + len);
    } else if (t == "IDAT") {
      idat.insert(idat.end(), data, data + len);
    } else if (t == "IEND") {
      end = true;
    } else if (critical) {
      MSH_THROW(strprintf("%s: unknown critical PNG chunk %s", name.c_str(), t.c_str()));
    }
  }
  if (colorType == 3 && palette.empty()) MSH_THROW(strprintf("%s: palette PNG without PLTE", name.c_str()));
  // The reference reads with PNG_TRANSFORM_EXPAND, which leaves 16-bit channels at 16 bits, and then
  // refuses them.
  if (depth == 16)
    MSH_THROW("qc::PNGLoadInterface::PNGLoadInterface: Reading 16-bit per channel PNG data not implemented");

  const int ch = pngChannels(colorType);
  const size_t bitsPerPixel = (size_t)ch * depth, bpp = std::max<size_t>(1, bitsPerPixel / 8);
  // Adam7 passes (start x, start y, step x, step y); a single pass covering the image otherwise
  static const int ax[7] = {0, 4, 0, 2, 0, 1, 0}, ay[7] = {0, 0, 4, 0, 2, 0, 1};
  static const int adx[7] = {8, 8, 4, 4, 2, 2, 1}, ady[7] = {8, 8, 8, 4, 4, 2, 2};
  const int npass = interlace ? 7 : 1;
  struct Pass { size_t x0, y0, dx, dy, pw, ph, rowBytes; };
  std::vector<Pass> passes;
  size_t total = 0;
  for (int p = 0; p < npass; ++p) {
    Pass q;
    q.x0 = interlace ? ax[p] : 0;
    q.y0 = interlace ? ay[p] : 0;
    q.dx = interlace ? adx[p] : 1;
    q.dy = interlace ? ady[p] : 1;
    q.pw = w > q.x0 ? (w - q.x0 + q.dx - 1) / q.dx : 0;
    q.ph = h > q.y0 ? (h - q.y0 + q.dy - 1) / q.dy : 0;
    q.rowBytes = (q.pw * bitsPerPixel + 7) / 8;
    if (q.pw > 0 && q.ph > 0) total += q.ph * (1 + q.rowBytes);
    passes.push_back(q);
  }
  if (total > maxInflated(idat.size())) MSH_THROW(strprintf("%s: PNG image data: not enough compressed data", name.c_str()));
  Bytes raw = inflateExact(idat.data(), idat.size(), total, name + ": PNG image data");

  Image im;
  im.nx = (int)w;
  im.ny = (int)h;
  im.data.assign((size_t)w * h, 0.0);
  const unsigned maxSample = (1u << depth) - 1;
  const unsigned scale8 = depth < 8 ? 255 / maxSample : 1;  // bit replication to 8 bits
  const size_t npal = palette.size() / 3;
  size_t off = 0;
  for (const Pass &q : passes) {
    if (q.pw == 0 || q.ph == 0) continue;
    const uint8_t *prev = nullptr;
    for (size_t r = 0; r < q.ph; ++r) {
      const int ft = raw[off];
      uint8_t *cur = raw.data() + off + 1;
      pngUnfilter(ft, cur, prev, q.rowBytes, bpp, name);
      double *dst = im.data.data() + (q.y0 + r * q.dy) * w;
      for (size_t i = 0; i < q.pw; ++i) {
        double v;
        if (depth < 8 || colorType == 0 || colorType == 3) {
          unsigned s;
          if (depth == 8) {
            s = cur[i];
          } else {
            const size_t bit = i * depth;
            s = (cur[bit >> 3] >> (8 - depth - (bit & 7))) & maxSample;
          }
          if (colorType == 3) {
            // out-of-range indices read the zero entries of libpng's 256-entry palette
            v = s < npal ? (double)(palette[3 * s] + palette[3 * s + 1] + palette[3 * s + 2]) / 3.0 : 0.0;
          } else {
            v = (double)(s * scale8);
          }
        } else if (colorType == 4) {
          v = cur[2 * i];
        } else {
          const uint8_t *px = cur + i * ch;  // RGB or RGBA: alpha stripped, mean of r, g, b
          v = (double)(px[0] + px[1] + px[2]) / 3.0;
        }
        dst[q.x0 + i * q.dx] = v;
      }
      prev = cur;
      off += 1 + q.rowBytes;
    }
  }
  return im;
}

// ------------------------------------------------------------------------------------------------
// TIFF
// ------------------------------------------------------------------------------------------------

namespace {

struct TiffReader {
  const uint8_t *b;
  size_t n;
  bool bigEndian = false, big = false;
  std::string name;

  void need(uint64_t pos, uint64_t len) const {
    if (pos > n || len > n - pos) MSH_THROW(strprintf("%s: TIFF structure points beyond the file", name.c_str()));
  }
  uint64_t uint(uint64_t pos, int bytes) const {
    need(pos, bytes);
    uint64_t v = 0;
    for (int i = 0; i < bytes; ++i) {
      const int k = bigEndian ? i : bytes - 1 - i;
      v = v << 8 | b[pos + k];
    }
    return v;
  }
  static int typeSize(int type) {
    switch (type) {
      case 1: case 2: case 6: case 7: return 1;
      case 3: case 8: return 2;
      case 4: case 9: case 11: case 13: return 4;
      case 5: case 10: case 12: case 16: case 17: case 18: return 8;
    }
    return 0;
  }

  struct Entry { int type = 0; uint64_t count = 0, valuePos = 0; };

  // First integer value of an entry, `dflt` for an entry without values
  uint64_t first(const Entry &e, uint64_t dflt) const {
    const std::vector<uint64_t> v = ints(e);
    return v.empty() ? dflt : v[0];
  }
  // Integer values of an entry (BYTE, SHORT, LONG, LONG8 and their signed kinds)
  std::vector<uint64_t> ints(const Entry &e) const {
    const int s = typeSize(e.type);
    if (s == 0 || e.type == 2 || e.type == 5 || e.type == 10 || e.type == 11 || e.type == 12)
      MSH_THROW(strprintf("%s: TIFF tag of type %d where integers are expected", name.c_str(), e.type));
    if (e.count > n) MSH_THROW(strprintf("%s: TIFF tag count out of range", name.c_str()));
    const uint64_t bytes = e.count * s;
    const uint64_t at = bytes <= (big ? 8u : 4u) ? e.valuePos : uint(e.valuePos, big ? 8 : 4);
    need(at, bytes);
    std::vector<uint64_t> v(e.count);
    for (uint64_t i = 0; i < e.count; ++i) {
      uint64_t x = uint(at + i * s, s);
      if (e.type == 6 || e.type == 8 || e.type == 9) {  // sign-extend
        const int sh = 64 - 8 * s;
        x = (uint64_t)((int64_t)(x << sh) >> sh);
      }
      v[i] = x;
    }
    return v;
  }
};

// TIFF LZW (TIFF 6.0 section 13): MSB-first codes of 9-12 bits, 256 = clear, 257 = end, the code
// width growing one code early.
Bytes lzwDecode(const uint8_t *src, size_t n, size_t want, const std::string &name) {
  if (n >= 2 && src[0] == 0 && (src[1] & 1))
    MSH_THROW(strprintf("%s: old-style (pre-TIFF 6.0) LZW data is not supported", name.c_str()));
  Bytes out;
  out.reserve(want);
  std::vector<uint16_t> prefix(4096);
  std::vector<uint8_t> suffix(4096), first(4096);
  std::vector<uint16_t> length(4096);
  for (int i = 0; i < 256; ++i) {
    prefix[i] = 0xffff;
    suffix[i] = first[i] = (uint8_t)i;
    length[i] = 1;
  }
  size_t bitPos = 0;
  const size_t totalBits = n * 8;
  int width = 9, next = 258, prev = -1;
  auto emit = [&](int code) {
    const size_t len = length[code], start = out.size();
    out.resize(start + len);
    for (size_t k = len; k-- > 0;) {
      out[start + k] = suffix[code];
      code = prefix[code];
    }
  };
  while (out.size() < want) {
    if (bitPos + width > totalBits) break;
    int code = 0;
    for (int k = 0; k < width; ++k, ++bitPos) code = code << 1 | ((src[bitPos >> 3] >> (7 - (bitPos & 7))) & 1);
    if (code == 256) {
      width = 9;
      next = 258;
      prev = -1;
      continue;
    }
    if (code == 257) break;
    if (prev < 0) {
      if (code > 255) MSH_THROW(strprintf("%s: corrupt LZW data", name.c_str()));
      emit(code);
      prev = code;
      continue;
    }
    int firstByte;
    if (code < next) {
      firstByte = first[code];
      emit(code);
    } else if (code == next) {
      firstByte = first[prev];
      emit(prev);
      out.push_back((uint8_t)firstByte);
    } else {
      MSH_THROW(strprintf("%s: corrupt LZW data", name.c_str()));
    }
    if (next < 4096) {
      prefix[next] = (uint16_t)prev;
      suffix[next] = (uint8_t)firstByte;
      first[next] = first[prev];
      length[next] = (uint16_t)(length[prev] + 1);
      ++next;
      if (next >= (1 << width) - 1 && width < 12) ++width;
    }
    prev = code;
  }
  if (out.size() < want) MSH_THROW(strprintf("%s: not enough LZW data", name.c_str()));
  out.resize(want);
  return out;
}

Bytes packBitsDecode(const uint8_t *src, size_t n, size_t want, const std::string &name) {
  Bytes out;
  out.reserve(want);
  size_t i = 0;
  while (out.size() < want && i < n) {
    const int c = (int8_t)src[i++];
    if (c >= 0) {
      const size_t len = std::min<size_t>(c + 1, n - i);
      out.insert(out.end(), src + i, src + i + len);
      i += len;
    } else if (c != -128) {
      if (i >= n) break;
      out.insert(out.end(), (size_t)(1 - c), src[i++]);
    }
  }
  if (out.size() < want) MSH_THROW(strprintf("%s: not enough PackBits data", name.c_str()));
  out.resize(want);
  return out;
}

}  // namespace

Image decodeTIFF(const std::string &bytes, const std::string &name) {
  TiffReader t;
  t.b = reinterpret_cast<const uint8_t *>(bytes.data());
  t.n = bytes.size();
  t.name = name;
  if (t.n < 8 || !((t.b[0] == 'I' && t.b[1] == 'I') || (t.b[0] == 'M' && t.b[1] == 'M')))
    MSH_THROW(strprintf("Error opening TIFF file \"%s\".", name.c_str()));
  t.bigEndian = t.b[0] == 'M';
  const uint64_t version = t.uint(2, 2);
  uint64_t ifd;
  if (version == 42) {
    ifd = t.uint(4, 4);
  } else if (version == 43) {
    t.big = true;
    if (t.uint(4, 2) != 8) MSH_THROW(strprintf("Error opening TIFF file \"%s\".", name.c_str()));
    ifd = t.uint(8, 8);
  } else {
    MSH_THROW(strprintf("Error opening TIFF file \"%s\".", name.c_str()));
  }
  const uint64_t count = t.uint(ifd, t.big ? 8 : 2);
  const int entrySize = t.big ? 20 : 12;
  const uint64_t first = ifd + (t.big ? 8 : 2);
  if (count > t.n / entrySize) MSH_THROW(strprintf("%s: TIFF directory out of range", name.c_str()));
  t.need(first, count * entrySize);
  TiffReader::Entry tags[8];  // width, length, bits, compression, offsets, spp, rows/strip, byte counts
  bool have[8] = {false};
  uint64_t fillOrder = 1, predictor = 1, sampleFormat = 0;
  bool tiled = false;
  for (uint64_t i = 0; i < count; ++i) {
    const uint64_t p = first + i * entrySize;
    TiffReader::Entry e;
    const int tag = (int)t.uint(p, 2);
    e.type = (int)t.uint(p + 2, 2);
    e.count = t.uint(p + 4, t.big ? 8 : 4);
    e.valuePos = p + (t.big ? 12 : 8);
    int slot = -1;
    switch (tag) {
      case 256: slot = 0; break;  // ImageWidth
      case 257: slot = 1; break;  // ImageLength
      case 258: slot = 2; break;  // BitsPerSample
      case 259: slot = 3; break;  // Compression
      case 273: slot = 4; break;  // StripOffsets
      case 277: slot = 5; break;  // SamplesPerPixel
      case 278: slot = 6; break;  // RowsPerStrip
      case 279: slot = 7; break;  // StripByteCounts
      case 266: fillOrder = t.first(e, fillOrder); break;
      case 317: predictor = t.first(e, predictor); break;
      case 339: sampleFormat = t.first(e, sampleFormat); break;
      case 322: case 323: case 324: case 325: tiled = true; break;
    }
    if (slot >= 0 && TiffReader::typeSize(e.type) > 0 && e.count > 0) {
      tags[slot] = e;
      have[slot] = true;
    }
  }
  auto one = [&](int slot, uint64_t dflt) { return have[slot] ? t.ints(tags[slot]).at(0) : dflt; };
  if (!have[0] || !have[1]) MSH_THROW(strprintf("%s: TIFF directory lacks the image size", name.c_str()));
  const uint64_t width = one(0, 0), height = one(1, 0), bits = one(2, 0), spp = one(5, 1);
  const uint64_t compression = one(3, 1);
  if (spp > 1) MSH_THROW("Reading a colored TIFF into a ScalarArray is not supported.");
  checkPixels(width, height, name);
  Image im;
  im.nx = (int)width;
  im.ny = (int)height;
  if (width == 0 || height == 0) return im;

  const bool isInt = sampleFormat == 2, isFloat = sampleFormat == 3;
  const bool bitsOk = isFloat ? (bits == 32 || bits == 64) : (bits == 8 || bits == 16 || bits == 32 || bits == 64);
  if (!bitsOk) MSH_THROW(strprintf("%d bits uint per sample unimplemented.", (int)bits));
  if (tiled) MSH_THROW(strprintf("%s: Can not read scanlines from a tiled image", name.c_str()));
  if (compression != 1 && compression != 5 && compression != 8 && compression != 32946 && compression != 32773)
    MSH_THROW(strprintf("%s: TIFF compression scheme %d is not supported", name.c_str(), (int)compression));
  if (predictor != 1 && predictor != 2 && predictor != 3)
    MSH_THROW(strprintf("%s: TIFF predictor %d is not supported", name.c_str(), (int)predictor));
  if (predictor == 2 && bits == 64)  // as the libtiff the reference links refuses it
    MSH_THROW(strprintf("%s: Horizontal differencing \"Predictor\" not supported with 64-bit samples", name.c_str()));
  if (predictor == 3 && !isFloat)
    MSH_THROW(strprintf("%s: floating-point predictor on non-float TIFF samples", name.c_str()));
  if (!have[4]) MSH_THROW(strprintf("%s: TIFF directory lacks StripOffsets", name.c_str()));

  const size_t bytesPerSample = bits / 8, scanline = width * bytesPerSample;
  const uint64_t rps = std::min<uint64_t>(std::max<uint64_t>(one(6, height), 1), height);
  const uint64_t nstrips = (height + rps - 1) / rps;
  const std::vector<uint64_t> offsets = t.ints(tags[4]);
  std::vector<uint64_t> counts;
  if (have[7]) {
    counts = t.ints(tags[7]);
  } else if (compression == 1) {  // libtiff estimates missing byte counts of uncompressed strips
    for (uint64_t s = 0; s < nstrips; ++s) counts.push_back(std::min<uint64_t>(rps, height - s * rps) * scanline);
  } else {
    MSH_THROW(strprintf("%s: TIFF directory lacks StripByteCounts", name.c_str()));
  }
  if (offsets.size() < nstrips || counts.size() < nstrips)
    MSH_THROW(strprintf("%s: TIFF has fewer strips than its rows need", name.c_str()));
  for (uint64_t s = 0; s < nstrips; ++s) {
    const uint64_t want = std::min<uint64_t>(rps, height - s * rps) * scanline, c = counts[s];
    const bool fits = compression == 1 ? want <= c && offsets[s] <= t.n && want <= t.n - offsets[s]
                      : compression == 5 ? want <= c * 4096 + 4096
                      : compression == 32773 ? want <= c * 64 + 128
                      : want <= maxInflated(c);
    if (!fits) MSH_THROW(strprintf("%s: TIFF strip %llu holds too little data for its rows", name.c_str(),
                                   (unsigned long long)s));
  }
  im.data.assign(width * height, 0.0);

  for (uint64_t s = 0; s < nstrips; ++s) {
    const uint64_t rows = std::min<uint64_t>(rps, height - s * rps), want = rows * scanline;
    if (counts[s] == 0) MSH_THROW(strprintf("%s: invalid TIFF strip byte count 0 (strip %llu)", name.c_str(),
                                            (unsigned long long)s));
    uint64_t len = counts[s];
    if (compression == 1) len = std::min(len, (uint64_t)want);
    t.need(offsets[s], len);
    Bytes rawStrip(t.b + offsets[s], t.b + offsets[s] + len);
    if (fillOrder == 2)  // lsb-to-msb bit order: libtiff reverses the bits of every byte before decoding
      for (uint8_t &c : rawStrip) {
        uint8_t r = 0;
        for (int k = 0; k < 8; ++k) r |= ((c >> k) & 1) << (7 - k);
        c = r;
      }
    Bytes dec;
    if (compression == 1) {
      if (rawStrip.size() < want) MSH_THROW(strprintf("%s: TIFF strip %llu is short", name.c_str(), (unsigned long long)s));
      dec.swap(rawStrip);
    } else if (compression == 5) {
      dec = lzwDecode(rawStrip.data(), rawStrip.size(), want, name);
    } else if (compression == 32773) {
      dec = packBitsDecode(rawStrip.data(), rawStrip.size(), want, name);
    } else {
      dec = inflateExact(rawStrip.data(), rawStrip.size(), want, name + ": TIFF Deflate strip");
    }
    for (uint64_t r = 0; r < rows; ++r) {
      uint8_t *row = dec.data() + r * scanline;
      if (predictor == 3) {
        // floating-point predictor: bytes differenced along the row, stored as byte planes with
        // the most significant plane first; undone into native (little-endian) samples
        for (size_t i = 1; i < scanline; ++i) row[i] = (uint8_t)(row[i] + row[i - 1]);
        Bytes tmp(row, row + scanline);
        for (size_t j = 0; j < width; ++j)
          for (size_t k = 0; k < bytesPerSample; ++k) row[j * bytesPerSample + k] = tmp[(bytesPerSample - 1 - k) * width + j];
      } else {
        if (t.bigEndian && bytesPerSample > 1)
          for (size_t j = 0; j < width; ++j) std::reverse(row + j * bytesPerSample, row + (j + 1) * bytesPerSample);
        if (predictor == 2) {
          for (size_t j = 1; j < width; ++j) {
            uint8_t *cur = row + j * bytesPerSample;
            const uint8_t *left = cur - bytesPerSample;
            if (bytesPerSample == 1) {
              cur[0] = (uint8_t)(cur[0] + left[0]);
            } else if (bytesPerSample == 2) {
              uint16_t a, c;
              memcpy(&a, left, 2), memcpy(&c, cur, 2), c = (uint16_t)(c + a), memcpy(cur, &c, 2);
            } else if (bytesPerSample == 4) {
              uint32_t a, c;
              memcpy(&a, left, 4), memcpy(&c, cur, 4), c += a, memcpy(cur, &c, 4);
            } else {
              uint64_t a, c;
              memcpy(&a, left, 8), memcpy(&c, cur, 8), c += a, memcpy(cur, &c, 8);
            }
          }
        }
      }
      double *dst = im.data.data() + (s * rps + r) * width;
      for (size_t j = 0; j < width; ++j) {
        const uint8_t *p = row + j * bytesPerSample;
        double v;
        if (isFloat) {
          if (bits == 32) { float f; memcpy(&f, p, 4); v = f; }
          else { double d; memcpy(&d, p, 8); v = d; }
        } else if (isInt) {
          if (bits == 8) v = (int8_t)p[0];
          else if (bits == 16) { int16_t x; memcpy(&x, p, 2); v = x; }
          else if (bits == 32) { int32_t x; memcpy(&x, p, 4); v = x; }
          else { int64_t x; memcpy(&x, p, 8); v = (double)x; }
        } else {
          if (bits == 8) v = p[0];
          else if (bits == 16) { uint16_t x; memcpy(&x, p, 2); v = x; }
          else if (bits == 32) { uint32_t x; memcpy(&x, p, 4); v = x; }
          else { uint64_t x; memcpy(&x, p, 8); v = (double)x; }
        }
        dst[j] = v;
      }
    }
  }
  return im;
}

// ------------------------------------------------------------------------------------------------
// Gatan DigitalMicrograph 3 / 4
// ------------------------------------------------------------------------------------------------

namespace {

// Walks the DM tag tree (all counts and type codes big-endian, 32-bit in version 3 and 64-bit in
// version 4; values little-endian) and records, for every ImageList.<i>.ImageData group, where its
// Data array lies and its Dimensions.
struct DMWalker {
  struct ImageEntry {
    long num = -1;
    int64_t offset = -1, count = -1, nx = -1, ny = -1, nz = -1;
    int type = -1;
  };
  const uint8_t *b;
  size_t n, pos = 0;
  int version = 3;
  std::string name;
  std::vector<std::string> path;
  std::vector<ImageEntry> images;

  void need(size_t k) const {
    if (pos > n || k > n - pos) MSH_THROW(strprintf("%s: DM file ends inside its tag tree", name.c_str()));
  }
  int64_t be(int bytes) {
    need(bytes);
    uint64_t v = 0;
    for (int i = 0; i < bytes; ++i) v = v << 8 | b[pos + i];
    pos += bytes;
    const int sh = 64 - 8 * bytes;
    return (int64_t)(v << sh) >> sh;
  }
  uint32_t le32() {
    need(4);
    const uint32_t v = (uint32_t)b[pos] | (uint32_t)b[pos + 1] << 8 | (uint32_t)b[pos + 2] << 16 | (uint32_t)b[pos + 3] << 24;
    pos += 4;
    return v;
  }
  int64_t count() { return be(version == 3 ? 4 : 8); }
  // bytes of one value of a simple type; the reference's parser knows these and no others
  int valueSize(int64_t type) const {
    switch (type) {
      case 2: case 4: return 2;
      case 3: case 5: case 6: return 4;
      case 7: case 11: case 12: return 8;
      case 8: case 10: return 1;
    }
    MSH_THROW(strprintf("type %d not implemented ", (int)type));
  }
  void skip(int64_t k) {
    if (k < 0) MSH_THROW(strprintf("%s: negative size in DM tag", name.c_str()));
    need((size_t)k);
    pos += (size_t)k;
  }
  bool under(size_t depth, const char *leaf) const {
    return path.size() == depth && path[0] == "ImageList" && path[2] == "ImageData" && (!leaf || path[3] == leaf) &&
           !images.empty() && images.back().num == atol(path[1].c_str());
  }

  void name_(long index) {
    const int64_t len = be(2);
    if (len < 0) MSH_THROW("negative tag name length.");
    need((size_t)len);
    std::string s(reinterpret_cast<const char *>(b + pos), (size_t)len);
    pos += (size_t)len;
    // a tag with an empty name is named by its index in the enclosing group
    path.push_back(!s.empty() ? s : index >= 0 ? strprintf("%ld", index) : std::string("nameless"));
    if (path.size() == 3 && path[0] == "ImageList" && path[2] == "ImageData") {
      images.emplace_back();
      images.back().num = atol(path[1].c_str());
    }
  }

  void data() {
    for (int i = 0; i < 4; ++i) {
      need(1);
      if (b[pos++] != '%') MSH_THROW("Unexpected character read");
    }
    const int64_t ninfo = count();
    if (ninfo < 1) MSH_THROW("Info size must be positive");
    if ((uint64_t)ninfo > n) MSH_THROW(strprintf("%s: DM tag info out of range", name.c_str()));
    std::vector<int64_t> info((size_t)ninfo);
    for (auto &x : info) x = count();
    if (ninfo == 1) {
      if (info[0] == 5 && path.size() == 5 && under(5, "Dimensions")) {
        const int64_t v = le32();
        const long dir = atol(path[4].c_str());
        if (dir == 0) images.back().nx = v;
        else if (dir == 1) images.back().ny = v;
        else if (dir == 2) images.back().nz = v;
      } else if (info[0] == 5 && under(4, "DataType")) {
        le32();  // declared data type: read, not used
      } else {
        skip(valueSize(info[0]));
      }
    } else if (info[0] == 15) {  // struct: 0, field count, (0, type) per field
      if (info[1] != 0) MSH_THROW("length of groupname has to be zero");
      const int64_t nf = info[2];
      if ((int64_t)info.size() - 3 != 2 * nf) MSH_THROW("inconsistent sizes");
      int64_t bytes = 0;
      for (int64_t i = 0; i < nf; ++i) {
        if (info[3 + 2 * i] != 0) MSH_THROW("length of fieldname has to be zero");
        bytes += valueSize(info[4 + 2 * i]);
      }
      skip(bytes);
    } else if (info[0] == 20) {
      if (info[1] == 15) {  // array of structs: 15, 0, field count, (0, type) per field, length
        if (info[2] != 0) MSH_THROW("length of groupname has to be zero");
        const int64_t nf = info[3];
        if ((int64_t)info.size() - 5 != 2 * nf) MSH_THROW("inconsistent sizes");
        for (int64_t i = 0; i < nf; ++i)
          if (info[4 + 2 * i] != 0) MSH_THROW("length of fieldname has to be zero");
        const int64_t len = info[4 + 2 * nf];
        int64_t bytes = 0;
        if (len > 0)
          for (int64_t i = 0; i < nf; ++i) bytes += valueSize(info[5 + 2 * i]);
        if (len > 0 && bytes > 0) {
          if (len > (int64_t)n / bytes) MSH_THROW(strprintf("%s: DM array beyond the file", name.c_str()));
          skip(len * bytes);
        }
      } else {  // array of simple values: element type, length
        if (info.size() < 3) MSH_THROW("inconsistent sizes");
        if (under(4, "Data")) {
          images.back().offset = (int64_t)pos;
          images.back().type = (int)info[1];
          images.back().count = info[2];
        }
        if (info[2] > 0) {
          const int s = valueSize(info[1]);
          if (info[2] > (int64_t)n / s) MSH_THROW(strprintf("%s: DM array beyond the file", name.c_str()));
          skip(info[2] * s);
        }
      }
    } else {
      MSH_THROW(strprintf("numberType %d not implemented", (int)info[0]));
    }
    path.pop_back();
  }

  // One tag entry; returns its identifier (0 ends the list)
  int entry(long index) {
    need(1);
    const int id = b[pos++];
    if (id == 0) return 0;
    name_(index);
    if (version == 4) be(8);  // the entry's byte size
    if (id == 21) {
      data();
    } else if (id == 20) {
      group();
    } else {
      MSH_THROW(strprintf("Invalid tag identifier %d", id));
    }
    return id;
  }

  void group() {
    skip(2);  // sorted, open
    const int64_t k = count();
    for (int64_t i = 0; i < k; ++i) entry((long)i);
    path.pop_back();
  }

  void run() {
    need(4);
    version = (int)be(4);
    if (version != 3 && version != 4)
      MSH_THROW(strprintf("Unexpected version (got %d), expected 3 or 4", version));
    count();  // file length
    if (be(4) != 1) MSH_THROW("Reading of detected byte order not implemented");
    skip(2);  // sorted, open
    count();  // number of root tags: the root's list is read up to its terminating zero
    while (pos < n) {
      if (entry(-1) == 0) break;
      path.clear();
    }
  }
};

}  // namespace

Image decodeDM(const std::string &bytes, const std::string &name) {
  DMWalker w;
  w.b = reinterpret_cast<const uint8_t *>(bytes.data());
  w.n = bytes.size();
  w.name = name;
  w.run();
  if (w.images.empty()) MSH_THROW("No data entry found.");
  const DMWalker::ImageEntry &e = w.images.back();
  int size;
  switch (e.type) {
    case 2: case 4: size = 2; break;
    case 3: case 5: case 6: size = 4; break;
    case 10: size = 1; break;
    default: MSH_THROW(strprintf("Data type %d not implemented ", e.type));
  }
  if (e.nx < 1 || e.ny < 1 || e.nx > (1 << 30) || e.ny > (1 << 30) || e.offset < 0)
    MSH_THROW(strprintf("%s: DM image without valid dimensions", name.c_str()));
  checkPixels(e.nx, e.ny, name);
  const size_t np = (size_t)e.nx * e.ny;
  if ((size_t)e.offset > w.n || np > (w.n - e.offset) / size)
    MSH_THROW("aol::Vector<DataType>::loadRaw: Reading from istream failed");
  Image im;
  im.nx = (int)e.nx;
  im.ny = (int)e.ny;
  im.data.resize(np);
  const uint8_t *p = w.b + e.offset;
  for (size_t i = 0; i < np; ++i, p += size) {
    switch (e.type) {
      case 2: { int16_t x; memcpy(&x, p, 2); im.data[i] = x; break; }
      case 4: { uint16_t x; memcpy(&x, p, 2); im.data[i] = x; break; }
      case 3: { int32_t x; memcpy(&x, p, 4); im.data[i] = x; break; }
      case 5: { uint32_t x; memcpy(&x, p, 4); im.data[i] = x; break; }
      case 6: { float x; memcpy(&x, p, 4); im.data[i] = x; break; }
      default: im.data[i] = p[0]; break;
    }
  }
  return im;
}

void saveInDMContainer(const std::string &dmFile, const double *data, int nx, int ny, const std::string &outBase) {
  std::ifstream in(dmFile, std::ios::binary);
  if (!in) MSH_THROW(strprintf("Cannot open file %s for reading", dmFile.c_str()));
  std::stringstream ss;
  ss << in.rdbuf();
  std::string bytes = ss.str();
  DMWalker w;
  w.b = reinterpret_cast<const uint8_t *>(bytes.data());
  w.n = bytes.size();
  w.name = dmFile;
  w.run();
  if (w.images.empty()) MSH_THROW("No data entry found.");
  const DMWalker::ImageEntry e = w.images.back();
  int size;
  switch (e.type) {
    case 4: size = 2; break;
    case 3: case 5: case 6: size = 4; break;
    case 10: size = 1; break;
    default: MSH_THROW(strprintf("type %d not implemented ", e.type));
  }
  if (e.nx < 1 || e.ny < 1 || e.offset < 0 || e.count < 0 || e.count > e.nx * e.ny ||
      (size_t)e.offset + (size_t)e.count * size > bytes.size())
    MSH_THROW(strprintf("%s: DM image data do not fit the image size", dmFile.c_str()));
  // padFrom: the array centred into e.nx x e.ny (an odd difference shifts it by one), zeros around
  const int PX = (int)e.nx, PY = (int)e.ny;
  std::vector<double> pad((size_t)PX * PY, 0.0);
  const int dx = PX - nx, dy = PY - ny, px = dx / 2, py = dy / 2;
  const int tx0 = std::max(0, px), tx1 = PX - std::max(0, px + dx % 2);
  const int ty0 = std::max(0, py), ty1 = PY - std::max(0, py + dy % 2);
  const int sx0 = -std::min(0, px + dx % 2), sy0 = -std::min(0, py + dy % 2);
  for (int y = 0; y < ty1 - ty0; ++y)
    for (int x = 0; x < tx1 - tx0; ++x) pad[(size_t)(ty0 + y) * PX + tx0 + x] = data[(size_t)(sy0 + y) * nx + sx0 + x];
  char *dst = &bytes[(size_t)e.offset];
  for (int64_t i = 0; i < e.count; ++i, dst += size) {
    const double v = pad[(size_t)i];
    switch (e.type) {
      case 3: { const int32_t x = (int32_t)v; memcpy(dst, &x, 4); break; }
      case 4: { const uint16_t x = (uint16_t)v; memcpy(dst, &x, 2); break; }
      case 5: { const uint32_t x = (uint32_t)v; memcpy(dst, &x, 4); break; }
      case 6: { const float x = (float)v; memcpy(dst, &x, 4); break; }
      default: { const uint8_t x = (uint8_t)(v < 0 ? 0.0 : v > 255 ? 255.0 : v); memcpy(dst, &x, 1); break; }
    }
  }
  const std::string out = outBase + (w.version == 3 ? ".dm3" : ".dm4");
  std::ofstream os(out, std::ios::binary);
  if (!os)
    MSH_THROW(strprintf("qc::DM3Reader::saveQuocDataInDM3Container: Cannot open \"%s\" for writing", out.c_str()));
  os.write(bytes.data(), (std::streamsize)bytes.size());
}

// ------------------------------------------------------------------------------------------------
// MetaImage
// ------------------------------------------------------------------------------------------------

Image loadMetaImage(const std::string &fileName) {
  std::ifstream in(fileName, std::ios::binary);
  if (!in) MSH_THROW(strprintf("MetaImageInfo: Cannot open file %s for input.", fileName.c_str()));
  auto expectWord = [&](const char *w) {
    std::string s;
    in >> s;
    if (s != w) MSH_THROW(strprintf("Unknown header: Expected \n\"%s\"\n, got \n\"%s\".\n", w, s.c_str()));
  };
  auto expectLine = [&](const char *w) {
    std::string s;
    std::getline(in, s);
    if (s != w) MSH_THROW(strprintf("Unknown header: Expected \n\"%s\"\n, got \n\"%s\".\n", w, s.c_str()));
  };
  int dim = 0;
  expectWord("NDims");
  expectWord("=");
  in >> dim;
  expectLine("");
  expectWord("DimSize");
  expectWord("=");
  long size[3] = {1, 1, 1};
  for (int i = 0; i < dim; ++i) {
    long v = 1;
    in >> v;
    if (i < 3) size[i] = v;
  }
  expectLine("");
  expectWord("ElementType");
  expectWord("=");
  std::string type;
  in >> type;
  int bytesPer;
  if (type == "MET_FLOAT") bytesPer = 4;
  else if (type == "MET_USHORT") bytesPer = 2;
  else MSH_THROW(strprintf("MetaImageInfo: Data type %s not implemented ", type.c_str()));
  expectLine("");
  expectLine("ElementSpacing = 1.0 1.0 1.0");
  expectLine("ElementByteOrderMSB = False");
  expectWord("ElementDataFile");
  expectWord("=");
  std::string raw;
  in >> raw;
  if (dim != 2) fprintf(stderr, "Warning: Reading %dD mhd file \"%s\" into a 2D ScalarArray\n", dim, fileName.c_str());
  const size_t slash = fileName.rfind('/');
  const std::string path = (slash == std::string::npos ? std::string() : fileName.substr(0, slash + 1)) + raw;
  if (size[0] < 0 || size[1] < 0) MSH_THROW(strprintf("%s: MetaImage size out of range", fileName.c_str()));
  checkPixels(size[0], size[1], fileName);
  std::ifstream data(path, std::ios::binary | std::ios::ate);
  if (!data) MSH_THROW(strprintf("cannot open file %s for input.", path.c_str()));
  Image im;
  im.nx = (int)size[0];
  im.ny = (int)size[1];
  const size_t np = (size_t)im.nx * im.ny;
  if ((uint64_t)data.tellg() < (uint64_t)np * bytesPer)
    MSH_THROW("aol::Vector<DataType>::loadRaw: Reading from istream failed");
  data.seekg(0);
  std::vector<char> buf(np * bytesPer);
  data.read(buf.data(), (std::streamsize)buf.size());
  if ((size_t)data.gcount() != buf.size()) MSH_THROW("aol::Vector<DataType>::loadRaw: Reading from istream failed");
  im.data.resize(np);
  for (size_t i = 0; i < np; ++i) {
    if (bytesPer == 4) {
      float f;
      memcpy(&f, &buf[4 * i], 4);
      im.data[i] = f;
    } else {
      uint16_t u;
      memcpy(&u, &buf[2 * i], 2);
      im.data[i] = u;
    }
  }
  return im;
}

}  // namespace mshost
