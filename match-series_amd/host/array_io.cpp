#include "array_io.h"

#include <bzlib.h>
#include <zlib.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <ctime>
#include <fstream>
#include <sstream>

#include "image_formats.h"
#include "params.h"

namespace mshost {

bool hasBzipSuffix(const std::string &f) {
  auto ends = [&](const char *s) {
    const size_t n = strlen(s);
    return f.size() >= n && f.compare(f.size() - n, n, s) == 0;
  };
  return ends(".bz2") || ends(".q2bz");
}

namespace {

std::string readFile(const std::string &fileName) {
  std::ifstream in(fileName, std::ios::binary);
  if (!in) MSH_THROW(strprintf("Could not open \"%s\" for input.", fileName.c_str()));
  std::stringstream ss;
  ss << in.rdbuf();
  return ss.str();
}

std::string bunzip(const std::string &in, const std::string &name) {
  bz_stream s;
  memset(&s, 0, sizeof(s));
  if (BZ2_bzDecompressInit(&s, 0, 0) != BZ_OK) MSH_THROW("bzip2 initialisation failed");
  std::string out;
  std::vector<char> buf(1 << 20);
  s.next_in = const_cast<char *>(in.data());
  s.avail_in = (unsigned)in.size();
  int rc = BZ_OK;
  while (rc == BZ_OK) {
    s.next_out = buf.data();
    s.avail_out = (unsigned)buf.size();
    rc = BZ2_bzDecompress(&s);
    if (rc != BZ_OK && rc != BZ_STREAM_END) {
      BZ2_bzDecompressEnd(&s);
      MSH_THROW(strprintf("bzip2 decompression of \"%s\" failed (%d)", name.c_str(), rc));
    }
    out.append(buf.data(), buf.size() - s.avail_out);
    if (rc == BZ_OK && s.avail_in == 0 && s.avail_out != 0) break;
  }
  BZ2_bzDecompressEnd(&s);
  return out;
}

std::string bzip(const std::string &in) {
  unsigned cap = (unsigned)(in.size() + in.size() / 100 + 1024);
  std::string out(cap, '\0');
  if (BZ2_bzBuffToBuffCompress(&out[0], &cap, const_cast<char *>(in.data()), (unsigned)in.size(), 9, 0, 0) != BZ_OK)
    MSH_THROW("bzip2 compression failed");
  out.resize(cap);
  return out;
}

// next whitespace-separated header token, skipping '#' comment lines
std::string token(const std::string &b, size_t &pos) {
  while (pos < b.size()) {
    if (b[pos] == '#') {
      while (pos < b.size() && b[pos] != '\n') ++pos;
    } else if (isspace((unsigned char)b[pos])) {
      ++pos;
    } else {
      break;
    }
  }
  const size_t start = pos;
  while (pos < b.size() && !isspace((unsigned char)b[pos])) ++pos;
  return b.substr(start, pos - start);
}

}  // namespace

Image loadImage(const std::string &fileName) {
  // the reference's dispatch on the (case-sensitive) suffix, quoc/scalarArray.cpp:547-569
  auto ends = [&](const char *s) {
    const size_t n = strlen(s);
    return fileName.size() >= n && fileName.compare(fileName.size() - n, n, s) == 0;
  };
  if (ends(".png")) return decodePNG(readFile(fileName), fileName);
  if (ends(".dm3") || ends(".dm4")) return decodeDM(readFile(fileName), fileName);
  if (ends(".tif") || ends(".tiff")) return decodeTIFF(readFile(fileName), fileName);
  if (ends(".mhd")) return loadMetaImage(fileName);
  if (ends(".nc"))
    MSH_THROW(strprintf("%s: reading NetCDF files needs the netCDF library, which this build does not link",
                        fileName.c_str()));
  std::string b = readFile(fileName);
  if (hasBzipSuffix(fileName)) b = bunzip(b, fileName);
  size_t pos = 0;
  const std::string magic = token(b, pos);
  if (magic.size() != 2 || magic[0] != 'P')
    MSH_THROW(strprintf("%s: wrong file format (only the reference's P2/P5/P7/P8/P9/Pa/Pb arrays are read)",
                        fileName.c_str()));
  Image im;
  im.nx = atoi(token(b, pos).c_str());
  im.ny = atoi(token(b, pos).c_str());
  const double mx = atof(token(b, pos).c_str());
  ++pos;  // the single whitespace after the header
  const size_t n = (size_t)im.nx * im.ny;
  im.data.resize(n);
  const char t = magic[1];
  auto need = [&](size_t bytes) {
    if (pos + bytes > b.size()) MSH_THROW(strprintf("%s: file too short", fileName.c_str()));
  };
  switch (t) {
    case '2':
    case '7': {
      std::istringstream ss(b.substr(pos));
      for (size_t i = 0; i < n; ++i) ss >> im.data[i];
      break;
    }
    case '5':
      if (mx > 255) {
        need(2 * n);
        for (size_t i = 0; i < n; ++i)
          im.data[i] = (unsigned char)b[pos + 2 * i] * 256.0 + (unsigned char)b[pos + 2 * i + 1];
      } else {
        need(n);
        for (size_t i = 0; i < n; ++i) im.data[i] = (unsigned char)b[pos + i];
      }
      break;
    case '8':
      need(4 * n);
      for (size_t i = 0; i < n; ++i) {
        float v;
        memcpy(&v, &b[pos + 4 * i], 4);
        im.data[i] = v;
      }
      break;
    case '9':
      need(8 * n);
      memcpy(im.data.data(), &b[pos], 8 * n);
      break;
    case 'a':
      need(2 * n);
      for (size_t i = 0; i < n; ++i) {
        uint16_t v;
        memcpy(&v, &b[pos + 2 * i], 2);
        im.data[i] = v;
      }
      break;
    case 'b':
      need(2 * n);
      for (size_t i = 0; i < n; ++i) {
        int16_t v;
        memcpy(&v, &b[pos + 2 * i], 2);
        im.data[i] = v;
      }
      break;
    default:
      MSH_THROW(strprintf("%s: wrong magic number", fileName.c_str()));
  }
  return im;
}

void saveP9(const std::string &fileName, const double *data, int nx, int ny, const char *comment) {
  const size_t n = (size_t)nx * ny;
  double mx = n ? data[0] : 0;
  for (size_t i = 1; i < n; ++i) mx = std::max(mx, data[i]);
  std::ostringstream hdr;
  hdr << "P9\n";
  if (comment) hdr << (comment[0] == '#' ? "" : "#") << comment << "\n";
  else hdr << "# This is a QuOcMesh file of type 9 (=PGM_DOUBLE_BINARY) written by msgpu\n";
  hdr << nx << " " << ny << "\n" << (std::isfinite(mx) ? (long long)mx : 0) << "\n";
  std::string out = hdr.str();
  out.append(reinterpret_cast<const char *>(data), 8 * n);
  if (hasBzipSuffix(fileName)) out = bzip(out);
  std::ofstream f(fileName, std::ios::binary);
  if (!f) MSH_THROW(strprintf("Could not open \"%s\" for output.", fileName.c_str()));
  f.write(out.data(), (std::streamsize)out.size());
}

void loadField(const std::string &mask, int n, std::vector<double> &phi) {
  phi.assign((size_t)2 * n * n, 0.0);
  for (int c = 0; c < 2; ++c) {
    const Image im = loadImage(strprintf(mask.c_str(), c));
    if (im.nx != n || im.ny != n)
      MSH_THROW(strprintf("deformation %s has size %dx%d, expected %dx%d", strprintf(mask.c_str(), c).c_str(), im.nx,
                          im.ny, n, n));
    std::copy(im.data.begin(), im.data.end(), phi.begin() + (size_t)c * n * n);
  }
}

void saveField(const std::string &mask, const double *phi, int n) {
  for (int c = 0; c < 2; ++c) saveP9(strprintf(mask.c_str(), c), phi + (size_t)c * n * n, n, n);
}


namespace {
void putU32(std::string &b, uint32_t v) {
  for (int k = 3; k >= 0; --k) b.push_back((char)((v >> (8 * k)) & 0xff));
}
void pngChunk(std::string &out, const char *type, const std::string &data) {
  putU32(out, (uint32_t)data.size());
  std::string td(type, 4);
  td += data;
  out += td;
  putU32(out, (uint32_t)crc32(0L, reinterpret_cast<const Bytef *>(td.data()), (uInt)td.size()));
}
}  // namespace

void savePNGScaled(const std::string &fileName, const double *data, int nx, int ny) {
  if (nx <= 0 || ny <= 0) MSH_THROW("Cannot save ScalarArray of size 0 as PNG.");
  double mn = data[0], mx = data[0];
  const size_t cnt = (size_t)nx * ny;
  for (size_t i = 1; i < cnt; ++i) {
    mn = std::min(mn, data[i]);
    mx = std::max(mx, data[i]);
  }
  savePNGClipScaled(fileName, data, nx, ny, mn, mx);
}

void savePNGClipScaled(const std::string &fileName, const double *data, int nx, int ny, double mn, double mx) {
  if (nx <= 0 || ny <= 0) MSH_THROW("Cannot save ScalarArray of size 0 as PNG.");
  const size_t cnt = (size_t)nx * ny;
  const double range = mx - mn;
  std::string raw;
  raw.reserve(cnt + ny);
  for (int y = 0; y < ny; ++y) {
    raw.push_back(0);  // filter type 0 per scanline
    for (int x = 0; x < nx; ++x) {
      double v = data[(size_t)y * nx + x];
      v = v < mn ? mn : (v > mx ? mx : v);
      const double s = range > 0 ? (v - mn) / (1.0 * range) * 255 : 0.0;
      raw.push_back((char)static_cast<unsigned char>(s));
    }
  }
  uLongf zlen = compressBound((uLong)raw.size());
  std::string z(zlen, '\0');
  if (compress2(reinterpret_cast<Bytef *>(&z[0]), &zlen, reinterpret_cast<const Bytef *>(raw.data()), (uLong)raw.size(),
                Z_DEFAULT_COMPRESSION) != Z_OK)
    MSH_THROW("PNG compression failed");
  z.resize(zlen);
  std::string out("\x89PNG\r\n\x1a\n", 8);
  std::string ihdr;
  putU32(ihdr, (uint32_t)nx);
  putU32(ihdr, (uint32_t)ny);
  ihdr += std::string("\x08\x00\x00\x00\x00", 5);  // 8-bit grey, deflate, no filter, no interlace
  pngChunk(out, "IHDR", ihdr);
  pngChunk(out, "IDAT", z);
  pngChunk(out, "IEND", std::string());
  std::ofstream f(fileName, std::ios::binary);
  if (!f) MSH_THROW("cannot open " + fileName + " for writing");
  f.write(out.data(), (std::streamsize)out.size());
  if (!f) MSH_THROW("error writing " + fileName);
}

}  // namespace mshost
