// tiff.cpp -- multi-page grayscale TIFF I/O for the command-line front end
// (src/PGURE-SVT.cpp:118-174 reads, 204-240 writes, through libtiff).
//
// Baseline TIFF only, which is what the front end reads and writes:
// uncompressed, one sample per pixel, 8 or 16 bits, strips, either byte
// order.  A cube element (r, c, k) is pixel (row r, column c) of page k, the
// orientation the reference produces with its transpose of each scanline
// buffer (src/PGURE-SVT.cpp:148-150).
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "tiff.hpp"

namespace pgsvt {
namespace {

struct Reader {
  std::vector<uint8_t> buf;
  bool be = false;  // big-endian file
  uint16_t u16(size_t at) const {
    if (at + 2 > buf.size()) throw std::runtime_error("TIFF: truncated file");
    return be ? (uint16_t)(buf[at] << 8 | buf[at + 1]) : (uint16_t)(buf[at] | buf[at + 1] << 8);
  }
  uint32_t u32(size_t at) const {
    if (at + 4 > buf.size()) throw std::runtime_error("TIFF: truncated file");
    return be ? ((uint32_t)buf[at] << 24 | (uint32_t)buf[at + 1] << 16 | (uint32_t)buf[at + 2] << 8 | buf[at + 3])
              : ((uint32_t)buf[at] | (uint32_t)buf[at + 1] << 8 | (uint32_t)buf[at + 2] << 16 |
                 (uint32_t)buf[at + 3] << 24);
  }
};

struct Page {
  uint32_t width = 0, height = 0, bits = 1, spp = 1, compression = 1, rows_per_strip = 0xffffffffu;
  std::vector<uint32_t> offsets, counts;
};

// value(s) of an IFD entry of type SHORT (3) or LONG (4)
std::vector<uint32_t> entry_values(const Reader& R, size_t e) {
  const uint16_t type = R.u16(e + 2);
  const uint32_t n = R.u32(e + 4);
  const size_t size = (type == 3) ? 2 : (type == 4) ? 4 : 0;
  if (size == 0) throw std::runtime_error("TIFF: unsupported field type");
  // every value must lie inside the file before anything is allocated
  if (n == 0) throw std::runtime_error("TIFF: empty field");
  if (size * (size_t)n > R.buf.size()) throw std::runtime_error("TIFF: field larger than the file");
  const size_t at = (size * n <= 4) ? e + 8 : R.u32(e + 8);
  if (at + size * (size_t)n > R.buf.size()) throw std::runtime_error("TIFF: field outside the file");
  std::vector<uint32_t> v(n);
  for (uint32_t k = 0; k < n; ++k) v[k] = (size == 2) ? R.u16(at + 2 * k) : R.u32(at + 4 * k);
  return v;
}

std::vector<Page> parse(const Reader& R) {
  std::vector<Page> pages;
  size_t ifd = R.u32(4);
  while (ifd != 0) {
    if (pages.size() > 1000000) throw std::runtime_error("TIFF: too many pages");
    Page p;
    const uint16_t n = R.u16(ifd);
    for (uint16_t i = 0; i < n; ++i) {
      const size_t e = ifd + 2 + 12 * (size_t)i;
      const uint16_t tag = R.u16(e);
      switch (tag) {
        case 256: p.width = entry_values(R, e)[0]; break;
        case 257: p.height = entry_values(R, e)[0]; break;
        case 258: p.bits = entry_values(R, e)[0]; break;
        case 259: p.compression = entry_values(R, e)[0]; break;
        case 273: p.offsets = entry_values(R, e); break;
        case 277: p.spp = entry_values(R, e)[0]; break;
        case 278: p.rows_per_strip = entry_values(R, e)[0]; break;
        case 279: p.counts = entry_values(R, e); break;
        default: break;
      }
    }
    pages.push_back(p);
    ifd = R.u32(ifd + 2 + 12 * (size_t)n);
  }
  return pages;
}

Reader load(const char* path) {
  Reader R;
  FILE* f = std::fopen(path, "rb");
  if (!f) throw std::runtime_error(std::string("File ") + path + " not found");
  std::fseek(f, 0, SEEK_END);
  const long n = std::ftell(f);
  std::fseek(f, 0, SEEK_SET);
  R.buf.resize(n > 0 ? (size_t)n : 0);
  const size_t got = R.buf.empty() ? 0 : std::fread(R.buf.data(), 1, R.buf.size(), f);
  std::fclose(f);
  if (got != R.buf.size() || R.buf.size() < 8) throw std::runtime_error("TIFF: cannot read file");
  if (R.buf[0] == 'M' && R.buf[1] == 'M') R.be = true;
  else if (!(R.buf[0] == 'I' && R.buf[1] == 'I')) throw std::runtime_error("TIFF: bad byte-order mark");
  if (R.u16(2) != 42) throw std::runtime_error("TIFF: not a classic TIFF file");
  return R;
}

}  // namespace

TiffInfo tiff_info(const char* path) {
  const Reader R = load(path);
  const auto pages = parse(R);
  if (pages.empty()) throw std::runtime_error("TIFF: no pages");
  TiffInfo i;
  i.width = pages[0].width;
  i.height = pages[0].height;
  i.bits = pages[0].bits;
  i.pages = (uint32_t)pages.size();
  return i;
}

void tiff_read_u16(const char* path, uint32_t first, uint32_t count, uint16_t* out) {
  const Reader R = load(path);
  const auto pages = parse(R);
  if ((size_t)first + count > pages.size()) throw std::runtime_error("TIFF: page range out of the file");
  const uint32_t W = pages[0].width, H = pages[0].height;
  for (uint32_t k = 0; k < count; ++k) {
    const Page& p = pages[first + k];
    if (p.width != W || p.height != H) throw std::runtime_error("TIFF: pages differ in size");
    if (p.compression != 1) throw std::runtime_error("TIFF: compressed pages are not supported");
    if (p.spp != 1) throw std::runtime_error("TIFF: only one sample per pixel is supported");
    if (p.bits != 8 && p.bits != 16) throw std::runtime_error("TIFF: only 8-bit and 16-bit pages are supported");
    if (p.offsets.size() != p.counts.size()) throw std::runtime_error("TIFF: bad strip tables");
    if (p.rows_per_strip == 0) throw std::runtime_error("TIFF: RowsPerStrip is 0");
    if (W == 0 || H == 0) throw std::runtime_error("TIFF: empty page");
    const uint32_t bpp = p.bits / 8, rps = std::min(p.rows_per_strip, H);
    uint16_t* F = out + (size_t)k * W * H;
    for (uint32_t row = 0; row < H; ++row) {
      const uint32_t s = row / rps;
      if (s >= p.offsets.size()) throw std::runtime_error("TIFF: missing strip");
      const size_t at = (size_t)p.offsets[s] + (size_t)(row % rps) * W * bpp;
      if (at + (size_t)W * bpp > R.buf.size()) throw std::runtime_error("TIFF: truncated strip");
      for (uint32_t col = 0; col < W; ++col)
        F[row + (size_t)H * col] = (bpp == 1) ? R.buf[at + col] : R.u16(at + 2 * (size_t)col);
    }
  }
}

void tiff_write_u16(const char* path, const uint16_t* cube, uint32_t width, uint32_t height, uint32_t pages) {
  // little-endian, page k: header/IFD chain, then one strip of height rows
  std::vector<uint8_t> b;
  auto put16 = [&](uint16_t v) {
    b.push_back(v & 0xff);
    b.push_back(v >> 8);
  };
  auto put32 = [&](uint32_t v) {
    for (int i = 0; i < 4; ++i) b.push_back((v >> (8 * i)) & 0xff);
  };
  auto set32 = [&](size_t at, uint32_t v) {
    for (int i = 0; i < 4; ++i) b[at + i] = (v >> (8 * i)) & 0xff;
  };
  b.insert(b.end(), {'I', 'I'});
  put16(42);
  size_t link = b.size();  // where the offset of the next IFD goes
  put32(0);
  const uint32_t strip = width * height * 2;
  for (uint32_t k = 0; k < pages; ++k) {
    const size_t data = b.size();
    for (uint32_t row = 0; row < height; ++row)
      for (uint32_t col = 0; col < width; ++col) put16(cube[(size_t)k * width * height + row + (size_t)height * col]);
    if (b.size() % 2) b.push_back(0);
    const size_t ifd = b.size();
    set32(link, (uint32_t)ifd);
    // the tags the reference sets for every page (src/PGURE-SVT.cpp:218-226)
    // plus the strip layout, in ascending tag order
    struct E {
      uint16_t tag, type;
      uint32_t count, value;
    };
    const E es[] = {
        {254, 4, 1, 2},                  // NewSubfileType: page of a multi-page image
        {256, 4, 1, width},              // ImageWidth
        {257, 4, 1, height},             // ImageLength
        {258, 3, 1, 16},                 // BitsPerSample
        {259, 3, 1, 1},                  // Compression: none
        {262, 3, 1, 1},                  // Photometric: min-is-black
        {273, 4, 1, (uint32_t)data},     // StripOffsets
        {274, 3, 1, 1},                  // Orientation: top-left
        {277, 3, 1, 1},                  // SamplesPerPixel
        {278, 4, 1, height},             // RowsPerStrip
        {279, 4, 1, strip},              // StripByteCounts
        {284, 3, 1, 1},                  // PlanarConfiguration: contiguous
        {297, 3, 2, k | (pages << 16)},  // PageNumber (page, total)
    };
    put16((uint16_t)(sizeof(es) / sizeof(es[0])));
    for (const E& e : es) {
      put16(e.tag);
      put16(e.type);
      put32(e.count);
      put32(e.value);  // SHORT values sit left-justified: low half first
    }
    link = b.size();
    put32(0);
  }
  FILE* f = std::fopen(path, "wb");
  if (!f) throw std::runtime_error(std::string("File ") + path + " could not be written");
  const size_t w = std::fwrite(b.data(), 1, b.size(), f);
  const int c = std::fclose(f);
  if (w != b.size() || c != 0) throw std::runtime_error(std::string("File ") + path + " could not be written");
}

}  // namespace pgsvt
