// tiff.hpp -- multi-page grayscale TIFF I/O of the command-line front end.
#pragma once

#include <cstdint>

namespace pgsvt {

struct TiffInfo {
  uint32_t width, height, bits, pages;
};

// Size, depth and page count (from the first page).  Throws on files that
// are not classic TIFF.
TiffInfo tiff_info(const char* path);

// Pages first .. first+count-1 (8- or 16-bit, uncompressed) into a uint16
// cube height x width x count, column-major: out[r + height*(c + width*k)]
// is row r, column c of page first+k.
void tiff_read_u16(const char* path, uint32_t first, uint32_t count, uint16_t* out);

// Write a uint16 cube (same layout) as a multi-page 16-bit TIFF.
void tiff_write_u16(const char* path, const uint16_t* cube, uint32_t width, uint32_t height, uint32_t pages);

}  // namespace pgsvt
