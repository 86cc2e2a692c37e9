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
struct TiffOut {
  FILE* f;
  void u16(uint16_t v) { fwrite(&v, 2, 1, f); }
  void u32(uint32_t v) { fwrite(&v, 4, 1, f); }
  void entry(uint16_t tag, uint16_t type, uint32_t count, uint32_t value) {
    u16(tag); u16(type); u32(count);
    if (type == 3 && count == 1) { u16((uint16_t)value); u16(0); }
    else u32(value);
  }
};
}  // namespace

bool write_tiff_f32(const std::string& path, const float* data, int w, int h, int pages,
                    double pixel_m, int orientation) {
  FILE* f = fopen(path.c_str(), "wb");
  if (!f) return false;
  TiffOut t{f};
  // little-endian header; first IFD offset patched per page
  fwrite("II", 1, 2, f);
  t.u16(42);
  t.u32(8);
  const uint32_t row_bytes = (uint32_t)w * 4;
  const float resf = (float)(0.01 / pixel_m);
  // rational: resolution as num/den
  const uint32_t den = 1000;
  const uint32_t num = (uint32_t)(resf * den + 0.5f);
  const int nent = 16;
  for (int z = 0; z < pages; ++z) {
    const long ifd = ftell(f);
    const uint32_t ifd_size = 2 + nent * 12 + 4;
    const uint32_t extra = ifd + ifd_size;  // rationals, strip tables, data
    const uint32_t xres_off = extra, yres_off = extra + 8;
    const uint32_t offs_off = extra + 16;
    const uint32_t cnts_off = offs_off + 4 * h;
    const uint32_t data_off = cnts_off + 4 * h;
    const uint32_t next = (z + 1 < pages) ? data_off + row_bytes * h : 0;
    t.u16(nent);
    t.entry(254, 4, 1, pages > 1 ? 2 : 0);    // NewSubfileType (page)
    t.entry(256, 4, 1, (uint32_t)w);          // ImageWidth
    t.entry(257, 4, 1, (uint32_t)h);          // ImageLength
    t.entry(258, 3, 1, 32);                   // BitsPerSample
    t.entry(259, 3, 1, 1);                    // Compression none
    t.entry(262, 3, 1, 1);                    // Photometric min-is-black
    t.entry(273, 4, (uint32_t)h, h == 1 ? data_off : offs_off);  // StripOffsets
    t.entry(274, 3, 1, (uint32_t)orientation);                    // Orientation
    t.entry(277, 3, 1, 1);                                        // SamplesPerPixel
    t.entry(278, 4, 1, 1);                                        // RowsPerStrip
    t.entry(279, 4, (uint32_t)h, h == 1 ? row_bytes : cnts_off);  // StripByteCounts
    t.entry(282, 5, 1, xres_off);                                 // XResolution
    t.entry(283, 5, 1, yres_off);                                 // YResolution
    t.entry(284, 3, 1, 1);                                        // PlanarConfig contig
    t.entry(296, 3, 1, 3);                                        // ResolutionUnit cm
    t.entry(339, 3, 1, 3);                                        // SampleFormat IEEE float
    t.u32(next);
    t.u32(num); t.u32(den);
    t.u32(num); t.u32(den);
    for (int y = 0; y < h; ++y) t.u32(data_off + row_bytes * y);
    for (int y = 0; y < h; ++y) t.u32(row_bytes);
    if (fwrite(data + (size_t)z * w * h, 4, (size_t)w * h, f) != (size_t)w * h) {
      fclose(f);
      return false;
    }
  }
  fclose(f);
  return true;
}

}  // namespace mtd

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
