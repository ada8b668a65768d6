// png.cpp -- the frame image file: write_image saves a 24-bit PNG through
// FreeImage (src/main.cpp:672, ./img/NNNN.png from gen_image, :293-300).
// FreeImage is not part of this build; the file is written here directly:
// signature, IHDR (8-bit RGB, no interlace), one IDAT of the zlib-
// compressed rows (filter byte 0 each), IEND, each chunk with its CRC-32.
#include <zlib.h>

#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../../include/nbs_host.h"

namespace {

void put32(std::vector<unsigned char>& b, uint32_t v) {
  b.push_back((unsigned char)(v >> 24));
  b.push_back((unsigned char)(v >> 16));
  b.push_back((unsigned char)(v >> 8));
  b.push_back((unsigned char)v);
}

void chunk(std::vector<unsigned char>& out, const char* type, const unsigned char* data,
           size_t len) {
  put32(out, (uint32_t)len);
  const size_t at = out.size();
  out.insert(out.end(), type, type + 4);
  if (len) out.insert(out.end(), data, data + len);
  const uLong crc = crc32(0L, out.data() + at, (uInt)(len + 4));
  put32(out, (uint32_t)crc);
}

}  // namespace

extern "C" int nbs_write_png(const char* path, uint32_t w, uint32_t h, const uint8_t* rgb) {
  if (!path || !rgb || w == 0 || h == 0) return NBS_ERR_ARG;
  const size_t row = 3 * (size_t)w;
  std::vector<unsigned char> raw((row + 1) * h);
  for (uint32_t r = 0; r < h; r++) {
    raw[r * (row + 1)] = 0;
    std::memcpy(&raw[r * (row + 1) + 1], rgb + r * row, row);
  }
  uLongf zlen = compressBound((uLong)raw.size());
  std::vector<unsigned char> z(zlen);
  if (compress2(z.data(), &zlen, raw.data(), (uLong)raw.size(), 6) != Z_OK) return NBS_ERR_IO;
  std::vector<unsigned char> out = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1a, '\n'};
  std::vector<unsigned char> ihdr;
  put32(ihdr, w);
  put32(ihdr, h);
  ihdr.insert(ihdr.end(), {8, 2, 0, 0, 0});  // 8-bit, RGB, deflate, filter 0, no interlace
  chunk(out, "IHDR", ihdr.data(), ihdr.size());
  chunk(out, "IDAT", z.data(), zlen);
  chunk(out, "IEND", nullptr, 0);
  std::string tmp = std::string(path) + ".tmp";
  FILE* f = std::fopen(tmp.c_str(), "wb");
  if (!f) return NBS_ERR_IO;
  const bool ok = std::fwrite(out.data(), 1, out.size(), f) == out.size();
  if (std::fclose(f) != 0 || !ok || std::rename(tmp.c_str(), path) != 0) {
    std::remove(tmp.c_str());
    return NBS_ERR_IO;
  }
  return NBS_OK;
}
