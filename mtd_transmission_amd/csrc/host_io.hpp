// Host-side file formats of the reference's program interface: KEY value
// config files, PGM holograms and masks, headerless raw arrays, float32 TIFF.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace mtd {

// extract_val (pretraitement_hors_axe_CPU/src/IO_fonctions.cpp:81-118): the
// last "KEY value" line wins, '#' lines are skipped, the value goes through
// atof into a C float. Missing key -> def (has_default) or 0. *found tells
// whether a non-empty value was present; *readable whether the file opened.
float cfg_val(const std::string& path, const std::string& token, bool has_default,
              double def, bool* found = nullptr, bool* readable = nullptr);
// extract_string (IO_fonctions.cpp:44-79): last value, "" when absent.
std::string cfg_string(const std::string& path, const std::string& token,
                       bool* readable = nullptr);

// PGM (P5 8/16-bit big-endian, or P2). Samples widened to uint16.
bool read_pgm(const std::string& path, std::vector<uint16_t>& px, int& w, int& h,
              int& maxval, std::string* err = nullptr);
bool write_pgm8(const std::string& path, const uint8_t* px, int w, int h);

long long file_size(const std::string& path);
bool read_raw(const std::string& path, void* dst, size_t bytes);
bool write_raw(const std::string& path, const void* src, size_t bytes, bool append);

// Multi-page float32 TIFF, one strip per row, uncompressed; resolution
// 0.01/pixel_m pixels per centimetre; orientation 1 (top-left) or 4
// (bottom-left, the reference's 3-D writer). A file whose offsets would not
// fit 32 bits (over 4 GiB, e.g. a 1024^3 volume) is written as BigTIFF.
bool write_tiff_f32(const std::string& path, const float* data, int w, int h, int pages,
                    double pixel_m, int orientation);
// format 0: classic TIFF, BigTIFF when it would exceed 4 GiB; 1: classic
// only (-2, nothing written, when it would exceed 4 GiB); 2: BigTIFF.
// Returns 0 (classic written), 1 (BigTIFF written) or < 0 on failure.
int write_tiff_f32_fmt(const std::string& path, const float* data, int w, int h, int pages,
                       double pixel_m, int orientation, int format);
// file size of the layout write_tiff_f32_fmt produces
uint64_t tiff_file_bytes(bool big, int w, int h, int pages);

}  // namespace mtd
