// Event and image file formats of the strip tools.
//  - binary events: float32 triples (z_u, z_d, dl), no header
//    (reconstruction.h:156-167 reading a util::ibstream);
//  - text events: one "z_u z_d dl" line each ("i j z_u z_d dl" with
//    --3d-response), parsed with the same digit arithmetic as the
//    reference's fast parser (src/util/text_parser.h:140-175) so both read the
//    same float values;
//  - images: raw float32 row-major, NRRD header (src/util/nrrd_writer.h:19-28),
//    full-matrix text (pixel_map.h:134-147), sparse "x y value" tuples
//    (reconstruction.h:203-213), 8-bit grey PNG (src/util/png_writer.h:32-54;
//    written here without libpng, stored deflate blocks).
#pragma once

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace strip_io {

// ---------------------------------------------------------------------------
// text numbers, digit by digit
// ---------------------------------------------------------------------------

class TextNumbers {
 public:
  explicit TextNumbers(const char* p) : p_(p) {}

  int next_int() {
    skip_space();
    const int sign = parse_sign();
    int n = 0;
    return sign * parse_digits(n);
  }

  // whole part and fraction accumulated as int, fraction scaled by a cached
  // power of ten, exponent applied by cached powers, sign last
  float next_float() {
    skip_space();
    const int sign = parse_sign();
    int n_whole = 0;
    const int whole = parse_digits(n_whole);
    float val = (float)whole;
    if (*p_ == '.') {
      ++p_;
      int n_frac = 0;
      const int frac = parse_digits(n_frac);
      if (n_whole == 0 && n_frac == 0)
        throw std::runtime_error("period with no digits either before or after");
      if (n_frac > 0) val += (float)frac * pow10_negative(n_frac);
    } else if (n_whole == 0) {
      throw std::runtime_error("neither whole part nor period");
    }
    if (*p_ == 'e' || *p_ == 'E') {
      ++p_;
      const int esign = parse_sign();
      int n_exp = 0;
      const int exponent = esign * parse_digits(n_exp);
      if (n_exp == 0) throw std::runtime_error("found exponent char but no exponent");
      if (exponent > kMaxExponent || exponent < -kMaxExponent)
        throw std::runtime_error("exponent out of range");
      if (exponent > 0) val *= pow10_positive(exponent);
      else if (exponent < 0) val *= pow10_negative(-exponent);
    }
    return val * (float)sign;
  }

 private:
  static constexpr int kMaxExponent = 300;

  // The reference's caches: 10^-e as 0.1f * 10^-(e-1) in float, and its
  // "positive" power as 10 * 10^-(e-1) (text_parser.h:83-100).
  static float pow10_negative(int e) {
    static float cache[kMaxExponent + 1] = {1.0f};
    if (e < 0 || e > kMaxExponent) return 0.0f;
    if (cache[e] == 0.0f) cache[e] = 0.1f * pow10_negative(e - 1);
    return cache[e];
  }
  static float pow10_positive(int e) {
    static float cache[kMaxExponent + 1] = {1.0f};
    if (e < 0 || e > kMaxExponent) return 0.0f;
    if (cache[e] == 0.0f) cache[e] = 10.0f * pow10_negative(e - 1);
    return cache[e];
  }

  void skip_space() {
    while (*p_ == ' ') ++p_;
  }
  int parse_sign() {
    if (*p_ == '-') {
      ++p_;
      return -1;
    }
    if (*p_ == '+') ++p_;
    return 1;
  }
  int parse_digits(int& n) {
    n = 0;
    int sum = 0;
    while (*p_ >= '0' && *p_ <= '9') {
      sum = sum * 10 + (*p_ - '0');
      ++p_;
      ++n;
    }
    return sum;
  }

  const char* p_;
};

inline void append_text_events(const std::string& fn, bool is_3d,
                                std::vector<float>& out) {
  std::ifstream in(fn, std::ios::binary);
  if (!in) throw std::runtime_error("cannot open input file: " + fn);
  std::string line;
  while (std::getline(in, line)) {
    // like text_parser::read_lines, a last line without '\n' is not read
    if (in.eof()) break;
    TextNumbers t(line.c_str());
    try {
      if (is_3d) {
        t.next_int();
        t.next_int();
      }
      const float z_u = t.next_float();
      const float z_d = t.next_float();
      const float dl = t.next_float();
      out.push_back(z_u);
      out.push_back(z_d);
      out.push_back(dl);
    } catch (const std::exception&) {
      std::cerr << "error line: " << line << std::endl;
      throw;
    }
  }
}

inline void append_binary_events(const std::string& fn,
                                  std::vector<float>& out) {
  std::ifstream in(fn, std::ios::binary);
  if (!in) throw std::runtime_error("cannot open phantom responses: " + fn);
  float v[3];
  while (in.read(reinterpret_cast<char*>(v), sizeof(v)))
    out.insert(out.end(), v, v + 3);
}

// ---------------------------------------------------------------------------
// images
// ---------------------------------------------------------------------------

inline void write_raw(const std::string& fn, const float* img, size_t n) {
  std::ofstream out(fn, std::ios::binary);
  out.write(reinterpret_cast<const char*>(img), sizeof(float) * n);
}

inline void write_nrrd(const std::string& fn, const std::string& raw_fn,
                       int width, int height, bool textual) {
  std::ofstream out(fn);
  out << "NRRD0001" << std::endl
      << "type: float" << std::endl
      << "endian: little" << std::endl
      << "dimension: 2" << std::endl
      << "sizes: " << width << " " << height << std::endl
      << "data file: ./" << raw_fn << std::endl
      << "encoding: " << (textual ? "ascii" : "raw") << std::endl;
}

inline void write_matrix_text(const std::string& fn, const float* img,
                              int width, int height) {
  std::ofstream out(fn);
  const size_t n = (size_t)width * height;
  for (size_t i = 0; i < n; ++i) {
    out << img[i];
    out << (((i + 1) % width == 0) ? "\n" : " ");
  }
}

inline void write_tuples_text(const std::string& fn, const float* img,
                              int width, int height) {
  std::ofstream out(fn);
  out.precision(12);
  out << std::fixed;
  for (int y = 0; y < height; ++y)
    for (int x = 0; x < width; ++x) {
      const float v = img[(size_t)y * width + x];
      if (v >= (float)0.000000000001) out << x << ' ' << y << ' ' << v << std::endl;
    }
}

/// Start image for --rho (reconstruction_cmd.cpp:145-165).  Binary files
/// hold n floats, or n int32 counts that become count + 1000 with
/// --rho-is-int.  A ".txt" file is read as the reference reads it: its
/// PixelMap text extractor streams every value into a copy of the pixel
/// (`for (auto voxel : map) in >> voxel`, pixel_map.h:127-132), so the image
/// keeps its values (rho = 1) and a missing file is no error.
inline void read_image(const std::string& fn, bool is_int, float* img,
                       size_t n) {
  const bool txt = fn.size() >= 4 && fn.compare(fn.size() - 4, 4, ".txt") == 0;
  if (txt) {
    std::ifstream in(fn);
    for (size_t i = 0; i < n; ++i) {
      float copy = img[i];
      in >> copy;
    }
    return;
  }
  std::ifstream in(fn, std::ios::binary);
  if (!in) throw std::runtime_error("cannot open start image: " + fn);
  if (is_int) {
    std::vector<int32_t> v(n);
    in.read(reinterpret_cast<char*>(v.data()), sizeof(int32_t) * n);
    for (size_t i = 0; i < n; ++i) img[i] = (float)(v[i] + 1000);
  } else {
    in.read(reinterpret_cast<char*>(img), sizeof(float) * n);
  }
}

namespace png_detail {
inline uint32_t crc(const unsigned char* p, size_t n, uint32_t c = 0) {
  static uint32_t table[256];
  static bool init = false;
  if (!init) {
    for (uint32_t i = 0; i < 256; ++i) {
      uint32_t v = i;
      for (int k = 0; k < 8; ++k) v = (v & 1) ? 0xEDB88320u ^ (v >> 1) : v >> 1;
      table[i] = v;
    }
    init = true;
  }
  c = ~c;
  for (size_t i = 0; i < n; ++i) c = table[(c ^ p[i]) & 0xff] ^ (c >> 8);
  return ~c;
}
inline void put32(std::vector<unsigned char>& v, uint32_t x) {
  v.push_back(x >> 24);
  v.push_back(x >> 16);
  v.push_back(x >> 8);
  v.push_back(x);
}
inline void chunk(std::ofstream& out, const char* type,
                  const std::vector<unsigned char>& data) {
  std::vector<unsigned char> buf;
  put32(buf, (uint32_t)data.size());
  buf.insert(buf.end(), type, type + 4);
  buf.insert(buf.end(), data.begin(), data.end());
  const uint32_t c = crc(buf.data() + 4, buf.size() - 4);
  put32(buf, c);
  out.write(reinterpret_cast<const char*>(buf.data()), buf.size());
}
}  // namespace png_detail

/// 8-bit grey PNG: value v maps to max(0, 255 - v * 255 / max) with max the
/// image maximum, or scale_max when given (png_writer.h:32-54).
inline void write_png(const std::string& fn, const float* img, int width,
                      int height, double scale_max = 0) {
  using namespace png_detail;
  float output_max = (float)scale_max;
  const size_t n = (size_t)width * height;
  if (output_max == 0)
    for (size_t i = 0; i < n; ++i) output_max = std::max(output_max, img[i]);
  const double gain = 255.0 / output_max;
  std::vector<unsigned char> raw;
  raw.reserve((size_t)(width + 1) * height);
  for (int y = 0; y < height; ++y) {
    raw.push_back(0);  // filter: none
    for (int x = 0; x < width; ++x)
      raw.push_back((unsigned char)std::max(
          0.0, 255.0 - gain * img[(size_t)y * width + x]));
  }
  // zlib stream of stored deflate blocks
  std::vector<unsigned char> z = {0x78, 0x01};
  size_t pos = 0;
  do {
    const size_t len = std::min<size_t>(65535, raw.size() - pos);
    const bool last = pos + len == raw.size();
    z.push_back(last ? 1 : 0);
    z.push_back(len & 0xff);
    z.push_back(len >> 8);
    z.push_back(~len & 0xff);
    z.push_back((~len >> 8) & 0xff);
    z.insert(z.end(), raw.begin() + pos, raw.begin() + pos + len);
    pos += len;
  } while (pos < raw.size());
  uint32_t a = 1, b = 0;
  for (unsigned char c : raw) {
    a = (a + c) % 65521;
    b = (b + a) % 65521;
  }
  put32(z, (b << 16) | a);
  std::ofstream out(fn, std::ios::binary);
  if (!out) throw std::runtime_error("cannot create output file " + fn);
  static const unsigned char sig[8] = {137, 80, 78, 71, 13, 10, 26, 10};
  out.write(reinterpret_cast<const char*>(sig), 8);
  std::vector<unsigned char> ihdr;
  put32(ihdr, (uint32_t)width);
  put32(ihdr, (uint32_t)height);
  ihdr.push_back(8);  // bit depth
  ihdr.push_back(0);  // grey
  ihdr.push_back(0);
  ihdr.push_back(0);
  ihdr.push_back(0);
  chunk(out, "IHDR", ihdr);
  chunk(out, "IDAT", z);
  chunk(out, "IEND", {});
}

}  // namespace strip_io
