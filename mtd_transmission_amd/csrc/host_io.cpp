#include "host_io.hpp"

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>

#include "mtd.h"

namespace mtd {

namespace {

// Lines as the reference tokenises them: split on '\n', drop lines whose
// first character is '#' (IO_fonctions.cpp:85-95).
bool config_lines(const std::string& path, std::vector<std::string>& out) {
  std::ifstream f(path, std::ios::binary);
  if (!f) return false;
  std::stringstream ss;
  ss << f.rdbuf();
  const std::string text = ss.str();
  size_t start = 0;
  while (true) {
    const size_t nl = text.find('\n', start);
    const std::string ln = text.substr(start, nl == std::string::npos ? std::string::npos : nl - start);
    if (ln.empty() || ln[0] != '#') out.push_back(ln);
    if (nl == std::string::npos) break;
    start = nl + 1;
  }
  return true;
}

// Last value string for token; a line without a space is its own key and
// its own value (IO_fonctions.cpp:97-112).
bool lookup(const std::vector<std::string>& lines, const std::string& token, std::string& value) {
  bool hit = false;
  for (const std::string& ln : lines) {
    if (ln.empty()) continue;
    const size_t pos = ln.find(' ');
    const std::string key = pos == std::string::npos ? ln : ln.substr(0, pos);
    if (key == token) {
      value = pos == std::string::npos ? ln : ln.substr(pos + 1);
      hit = true;
    }
  }
  return hit;
}

}  // namespace

float cfg_val(const std::string& path, const std::string& token, bool has_default, double def,
              bool* found, bool* readable) {
  std::vector<std::string> lines;
  const bool ok = config_lines(path, lines);
  if (readable) *readable = ok;
  std::string v;
  const bool hit = ok && lookup(lines, token, v);
  if (!hit || v.empty()) {
    if (found) *found = false;
    return has_default ? (float)def : 0.0f;
  }
  if (found) *found = true;
  return (float)std::atof(v.c_str());
}

std::string cfg_string(const std::string& path, const std::string& token, bool* readable) {
  std::vector<std::string> lines;
  const bool ok = config_lines(path, lines);
  if (readable) *readable = ok;
  std::string v;
  if (ok && lookup(lines, token, v)) return v;
  return std::string();
}

static bool pgm_token(FILE* f, std::string& tok) {
  tok.clear();
  int c;
  while ((c = fgetc(f)) != EOF) {
    if (c == '#') {
      while ((c = fgetc(f)) != EOF && c != '\n') {}
      continue;
    }
    if (!isspace(c)) break;
  }
  if (c == EOF) return false;
  tok.push_back((char)c);
  while ((c = fgetc(f)) != EOF && !isspace(c)) tok.push_back((char)c);
  return true;
}

bool read_pgm(const std::string& path, std::vector<uint16_t>& px, int& w, int& h, int& maxval,
              std::string* err) {
  FILE* f = fopen(path.c_str(), "rb");
  if (!f) {
    if (err) *err = "cannot open " + path;
    return false;
  }
  std::string magic, sw, sh, sm;
  if (!pgm_token(f, magic) || (magic != "P5" && magic != "P2") || !pgm_token(f, sw) ||
      !pgm_token(f, sh) || !pgm_token(f, sm)) {
    if (err) *err = "not a PGM file: " + path;
    fclose(f);
    return false;
  }
  w = atoi(sw.c_str());
  h = atoi(sh.c_str());
  maxval = atoi(sm.c_str());
  if (w <= 0 || h <= 0 || maxval <= 0 || maxval > 65535) {
    if (err) *err = "bad PGM header: " + path;
    fclose(f);
    return false;
  }
  const size_t n = (size_t)w * h;
  px.resize(n);
  bool ok = true;
  if (magic == "P5") {
    if (maxval < 256) {
      std::vector<uint8_t> b(n);
      ok = fread(b.data(), 1, n, f) == n;
      for (size_t i = 0; i < n; ++i) px[i] = b[i];
    } else {
      std::vector<uint8_t> b(2 * n);
      ok = fread(b.data(), 1, 2 * n, f) == 2 * n;
      for (size_t i = 0; i < n; ++i) px[i] = (uint16_t)((b[2 * i] << 8) | b[2 * i + 1]);
    }
  } else {
    std::string tok;
    for (size_t i = 0; i < n && ok; ++i) {
      ok = pgm_token(f, tok);
      if (ok) px[i] = (uint16_t)atoi(tok.c_str());
    }
  }
  fclose(f);
  if (!ok && err) *err = "truncated PGM: " + path;
  return ok;
}

bool write_pgm8(const std::string& path, const uint8_t* px, int w, int h) {
  FILE* f = fopen(path.c_str(), "wb");
  if (!f) return false;
  fprintf(f, "P5\n%d %d\n255\n", w, h);
  const bool ok = fwrite(px, 1, (size_t)w * h, f) == (size_t)w * h;
  fclose(f);
  return ok;
}

long long file_size(const std::string& path) {
  FILE* f = fopen(path.c_str(), "rb");
  if (!f) return -1;
  fseek(f, 0, SEEK_END);
  const long long n = ftell(f);
  fclose(f);
  return n;
}

bool read_raw(const std::string& path, void* dst, size_t bytes) {
  FILE* f = fopen(path.c_str(), "rb");
  if (!f) return false;
  const bool ok = fread(dst, 1, bytes, f) == bytes;
  fclose(f);
  return ok;
}

bool write_raw(const std::string& path, const void* src, size_t bytes, bool append) {
  FILE* f = fopen(path.c_str(), append ? "ab" : "wb");
  if (!f) return false;
  const bool ok = fwrite(src, 1, bytes, f) == bytes;
  fclose(f);
  return ok;
}

namespace {
// Classic TIFF (32-bit offsets) or BigTIFF (64-bit offsets, magic 43). An
// IFD entry's value is stored in place when it fits the entry's value field
// (4 bytes classic, 8 bytes BigTIFF), else the field holds its offset.
struct TiffOut {
  FILE* f;
  bool big;
  void u16(uint16_t v) { fwrite(&v, 2, 1, f); }
  void u32(uint32_t v) { fwrite(&v, 4, 1, f); }
  void u64(uint64_t v) { fwrite(&v, 8, 1, f); }
  void off(uint64_t v) {
    if (big) u64(v); else u32((uint32_t)v);
  }
  // value: the in-place bytes (little-endian, low bytes first) or an offset
  void entry(uint16_t tag, uint16_t type, uint64_t count, uint64_t value) {
    u16(tag);
    u16(type);
    off(count);
    if (type == 3 && count == 1) {  // SHORT in the low bytes of the field
      u16((uint16_t)value);
      u16(0);
      if (big) u32(0);
    } else {
      off(value);
    }
  }
};

constexpr int kTiffEntries = 16;

// Bytes of one page's IFD, its out-of-place tables and its pixels.
uint64_t tiff_page_bytes(bool big, int w, int h) {
  const uint64_t ifd = big ? 8 + kTiffEntries * 20 + 8 : 2 + kTiffEntries * 12 + 4;
  const uint64_t offsz = big ? 8 : 4;
  const uint64_t field = big ? 8 : 4;
  uint64_t extra = 0;
  if (!big) extra += 16;                          // two RATIONALs
  if ((uint64_t)h * offsz > field) extra += (uint64_t)h * offsz;  // strip offsets
  if ((uint64_t)h * 4 > field) extra += (uint64_t)h * 4;          // strip byte counts
  return ifd + extra + (uint64_t)w * h * 4;
}
}  // namespace

uint64_t tiff_file_bytes(bool big, int w, int h, int pages) {
  return (big ? 16 : 8) + (uint64_t)pages * tiff_page_bytes(big, w, h);
}

int write_tiff_f32_fmt(const std::string& path, const float* data, int w, int h, int pages,
                       double pixel_m, int orientation, int format) {
  if (w <= 0 || h <= 0 || pages <= 0) return -1;
  const bool fits = tiff_file_bytes(false, w, h, pages) <= 0xFFFFFFFFull;
  if (format == 1 && !fits) return -2;  // classic TIFF asked for, offsets would wrap
  const bool big = format == 2 || (format == 0 && !fits);
  FILE* f = fopen(path.c_str(), "wb");
  if (!f) return -1;
  TiffOut t{f, big};
  fwrite("II", 1, 2, f);
  if (big) {
    t.u16(43);
    t.u16(8);  // offset size
    t.u16(0);
    t.u64(16);  // first IFD
  } else {
    t.u16(42);
    t.u32(8);
  }
  const uint64_t row_bytes = (uint64_t)w * 4;
  const float resf = (float)(0.01 / pixel_m);
  // resolution as the rational num/den
  const uint32_t den = 1000;
  const uint32_t num = (uint32_t)(resf * den + 0.5f);
  const uint64_t rational = (uint64_t)num | ((uint64_t)den << 32);
  const uint64_t field = big ? 8 : 4, offsz = big ? 8 : 4;
  const uint16_t off_type = big ? 16 : 4;  // LONG8 or LONG strip offsets
  uint64_t pos = big ? 16 : 8;
  for (int z = 0; z < pages; ++z) {
    const uint64_t ifd = pos;
    const uint64_t ifd_size = big ? 8 + kTiffEntries * 20 + 8 : 2 + kTiffEntries * 12 + 4;
    uint64_t extra = ifd + ifd_size;  // out-of-place values, then the pixels
    uint64_t xres = rational, yres = rational;
    if (!big) {
      xres = extra;
      yres = extra + 8;
      extra += 16;
    }
    const bool offs_inline = (uint64_t)h * offsz <= field;
    const bool cnts_inline = (uint64_t)h * 4 <= field;
    const uint64_t offs_at = extra;
    if (!offs_inline) extra += (uint64_t)h * offsz;
    const uint64_t cnts_at = extra;
    if (!cnts_inline) extra += (uint64_t)h * 4;
    const uint64_t data_at = extra;
    const uint64_t next = (z + 1 < pages) ? data_at + row_bytes * h : 0;
    uint64_t offs_val = offs_at, cnts_val = cnts_at;
    if (offs_inline) {  // h == 1 (both formats)
      offs_val = data_at;
    }
    if (cnts_inline) {  // h == 1, or h == 2 in BigTIFF
      cnts_val = 0;
      for (int y = 0; y < h; ++y) cnts_val |= row_bytes << (32 * y);
    }
    if (big) t.u64(kTiffEntries); else t.u16(kTiffEntries);
    t.entry(254, 4, 1, pages > 1 ? 2 : 0);    // NewSubfileType (page)
    t.entry(256, 4, 1, (uint32_t)w);          // ImageWidth
    t.entry(257, 4, 1, (uint32_t)h);          // ImageLength
    t.entry(258, 3, 1, 32);                   // BitsPerSample
    t.entry(259, 3, 1, 1);                    // Compression none
    t.entry(262, 3, 1, 1);                    // Photometric min-is-black
    t.entry(273, off_type, (uint64_t)h, offs_val);           // StripOffsets
    t.entry(274, 3, 1, (uint32_t)orientation);                // Orientation
    t.entry(277, 3, 1, 1);                                    // SamplesPerPixel
    t.entry(278, 4, 1, 1);                                    // RowsPerStrip
    t.entry(279, 4, (uint64_t)h, cnts_val);                   // StripByteCounts
    t.entry(282, 5, 1, xres);                                 // XResolution
    t.entry(283, 5, 1, yres);                                 // YResolution
    t.entry(284, 3, 1, 1);                                    // PlanarConfig contig
    t.entry(296, 3, 1, 3);                                    // ResolutionUnit cm
    t.entry(339, 3, 1, 3);                                    // SampleFormat IEEE float
    t.off(next);
    if (!big) {
      t.u32(num); t.u32(den);
      t.u32(num); t.u32(den);
    }
    if (!offs_inline)
      for (int y = 0; y < h; ++y) t.off(data_at + row_bytes * y);
    if (!cnts_inline)
      for (int y = 0; y < h; ++y) t.u32((uint32_t)row_bytes);
    if (fwrite(data + (size_t)z * w * h, 4, (size_t)w * h, f) != (size_t)w * h) {
      fclose(f);
      return -1;
    }
    pos = data_at + row_bytes * h;
  }
  const bool ok = ferror(f) == 0;
  if (fclose(f) != 0 || !ok) return -1;
  return big ? 1 : 0;
}

bool write_tiff_f32(const std::string& path, const float* data, int w, int h, int pages,
                    double pixel_m, int orientation) {
  return write_tiff_f32_fmt(path, data, w, h, pages, pixel_m, orientation, 0) >= 0;
}

}  // namespace mtd

extern "C" int mtd_write_tiff_f32(const char* path, const float* data, int w, int h,
                                  int pages, double pixel_m, int orientation, int format) {
  if (!path || !data || format < 0 || format > 2) return -1;
  return mtd::write_tiff_f32_fmt(path, data, w, h, pages, pixel_m, orientation, format);
}

extern "C" int mtd_config_extract_val(const char* path, const char* token, int has_default,
                                      double default_value, float* out) {
  if (!path || !token || !out) return -1;
  bool found = false, readable = false;
  *out = mtd::cfg_val(path, token, has_default != 0, default_value, &found, &readable);
  if (!readable) return -1;
  return found ? 1 : 0;
}

extern "C" int mtd_config_extract_string(const char* path, const char* token, char* out,
                                         size_t cap) {
  if (!path || !token || !out || cap == 0) return -1;
  bool readable = false;
  const std::string v = mtd::cfg_string(path, token, &readable);
  if (!readable) {
    out[0] = 0;
    return -1;
  }
  std::strncpy(out, v.c_str(), cap - 1);
  out[cap - 1] = 0;
  return v.empty() ? 0 : 1;
}
