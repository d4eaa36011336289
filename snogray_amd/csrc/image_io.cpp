// Image output and input in the two formats the reference always
// supports without external libraries, and recovery of a partially
// written output image (the reference's --continue).
//
//   PFM  ("Portable Float Map", image/image-pfm.cc): "PF\n", "W H\n",
//        scale line (negative: little-endian floats), then RGB float
//        rows, bottom row first.
//   RGBE (Radiance .hdr / .pic, image/image-rgbe.cc, image/image-rgbe.h):
//        "#?RGBE" header, blank line, "-Y H +X W", then top-to-bottom
//        rows of shared-exponent RGBE pixels, each row in the "new"
//        run-length form (2, 2, width) with the four byte planes
//        run-length coded separately; "old"-format rows are read too.
//
// Pixels written are the alpha-scaled colours of the finished image
// (ImageRow::alpha_scaled_color), as the reference's sinks write them.
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "sgr_internal.h"

namespace {

using sgr::fail;

enum Fmt { FMT_PFM, FMT_RGBE };

// ImageIo::find_format (image/image-io.cc): an explicit format, else the
// file name's extension.
int find_format(const char *path, const char *format, Fmt &out) {
  std::string f = format ? format : "";
  if (f.empty()) {
    const char *dot = std::strrchr(path, '.');
    if (dot) f = dot + 1;
  }
  for (auto &c : f) c = (char)std::tolower((unsigned char)c);
  if (f == "pfm") { out = FMT_PFM; return SGR_OK; }
  if (f == "rgbe" || f == "hdr" || f == "pic") { out = FMT_RGBE; return SGR_OK; }
  return fail(SGR_ERR_INVALID, "image: unknown or unsupported image format '" + f + "'");
}

struct File {
  FILE *f = nullptr;
  explicit File(FILE *p) : f(p) {}
  ~File() { if (f) std::fclose(f); }
};

bool little_endian() {
  const uint32_t x = 1;
  unsigned char c;
  std::memcpy(&c, &x, 1);
  return c == 1;
}

// ---- RGBE pixel (image/image-rgbe.h: RgbeColor) --------------------------
struct Rgbe {
  unsigned char r = 0, g = 0, b = 0, e = 0;
};

Rgbe to_rgbe(float r, float g, float b) {
  Rgbe o;
  const float m = std::fmax(r, std::fmax(g, b));
  if (m > 1e-32f) {
    int ex;
    const float adj = std::frexp(m, &ex) * 255.9999f / m;
    o.r = (unsigned char)(adj * std::fmax(r, 0.f));
    o.g = (unsigned char)(adj * std::fmax(g, 0.f));
    o.b = (unsigned char)(adj * std::fmax(b, 0.f));
    o.e = (unsigned char)(ex + 128);
  }
  return o;
}

void from_rgbe(const Rgbe &p, float *rgb) {
  if (p.e == 0) {
    rgb[0] = rgb[1] = rgb[2] = 0.f;
    return;
  }
  const float s = (float)std::ldexp(1.0, (int)p.e - (128 + 8));
  rgb[0] = s * (p.r + 0.5f);
  rgb[1] = s * (p.g + 0.5f);
  rgb[2] = s * (p.b + 0.5f);
}

// One byte plane of a row, run-length coded: runs of >= 4 equal bytes as
// (128 + n, byte), n <= 127; other bytes in literal groups (n, bytes...),
// n <= 128 (RgbeImageSink::write_rle_component).
void write_plane(std::vector<unsigned char> &out, const unsigned char *v, unsigned n) {
  const unsigned kMinRun = 4;
  unsigned i = 0;
  while (i < n) {
    unsigned run_start = i, run_len = 0;
    for (; run_start < n; run_start += run_len) {
      run_len = 1;
      while (run_len < 127 && run_start + run_len < n && v[run_start + run_len] == v[run_start])
        run_len++;
      if (run_len >= kMinRun) break;
    }
    if (run_start - i > 1 && run_start - i < kMinRun) {
      // a short run of equal bytes just before the long run
      bool same = true;
      for (unsigned j = i + 1; j < run_start; j++) same = same && v[j] == v[i];
      if (same) {
        out.push_back((unsigned char)(128 + run_start - i));
        out.push_back(v[i]);
        i = run_start;
      }
    }
    while (i < run_start) {
      unsigned len = run_start - i;
      if (len > 128) len = 128;
      out.push_back((unsigned char)len);
      while (len--) out.push_back(v[i++]);
    }
    if (run_len >= kMinRun) {
      out.push_back((unsigned char)(128 + run_len));
      out.push_back(v[run_start]);
      i += run_len;
    }
  }
}

int write_pfm(const char *path, const float *rgba, uint32_t w, uint32_t h) {
  File f(std::fopen(path, "wb"));
  if (!f.f) return fail(SGR_ERR_INVALID, std::string("image: cannot open ") + path + ": " +
                                             std::strerror(errno));
  std::fprintf(f.f, "PF\n%u %u\n%s\n", w, h, little_endian() ? "-1" : "1");
  std::vector<float> row((size_t)w * 3);
  for (uint32_t k = 0; k < h; k++) {
    const uint32_t y = h - 1 - k;  // bottom row first
    for (uint32_t x = 0; x < w; x++) {
      const float *p = rgba + 4 * ((size_t)y * w + x);
      for (int c = 0; c < 3; c++) row[3 * x + c] = p[c] * p[3];
    }
    if (std::fwrite(row.data(), 4, row.size(), f.f) != row.size())
      return fail(SGR_ERR_INVALID, std::string("image: write failed: ") + path);
  }
  if (std::fflush(f.f) != 0) return fail(SGR_ERR_INVALID, std::string("image: write failed: ") + path);
  return SGR_OK;
}

int write_rgbe(const char *path, const float *rgba, uint32_t w, uint32_t h) {
  if (w >= (1u << 16)) return fail(SGR_ERR_LIMIT, "image: RGBE rows hold at most 65535 pixels");
  File f(std::fopen(path, "wb"));
  if (!f.f) return fail(SGR_ERR_INVALID, std::string("image: cannot open ") + path + ": " +
                                             std::strerror(errno));
  std::fprintf(f.f, "#?RGBE\n# Written by snogray\n\n-Y %u +X %u\n", h, w);
  std::vector<unsigned char> planes[4], out;
  for (auto &p : planes) p.resize(w);
  for (uint32_t y = 0; y < h; y++) {
    for (uint32_t x = 0; x < w; x++) {
      const float *p = rgba + 4 * ((size_t)y * w + x);
      const Rgbe e = to_rgbe(p[0] * p[3], p[1] * p[3], p[2] * p[3]);
      planes[0][x] = e.r; planes[1][x] = e.g; planes[2][x] = e.b; planes[3][x] = e.e;
    }
    out.clear();
    out.push_back(2);
    out.push_back(2);
    out.push_back((unsigned char)((w >> 8) & 0xff));
    out.push_back((unsigned char)(w & 0xff));
    for (auto &p : planes) write_plane(out, p.data(), w);
    if (std::fwrite(out.data(), 1, out.size(), f.f) != out.size())
      return fail(SGR_ERR_INVALID, std::string("image: write failed: ") + path);
  }
  if (std::fflush(f.f) != 0) return fail(SGR_ERR_INVALID, std::string("image: write failed: ") + path);
  return SGR_OK;
}

// ---- readers: header, then rows one at a time (false: the file ends
// or is malformed before the row is complete) -------------------------------
struct Reader {
  virtual ~Reader() {}
  uint32_t w = 0, h = 0;
  virtual bool row(uint32_t y, float *rgb) = 0;  // rows in top-to-bottom order
};

struct PfmReader : Reader {
  FILE *f;
  long data_off = 0;
  bool swap = false;
  std::vector<float> buf;
  explicit PfmReader(FILE *fp) : f(fp) {}
  bool open() {
    char magic[3];
    if (std::fread(magic, 1, 3, f) != 3 || magic[0] != 'P' || magic[1] != 'F' || magic[2] != '\n')
      return false;
    float scale;
    if (std::fscanf(f, "%u %u", &w, &h) != 2) return false;
    std::fgetc(f);
    if (std::fscanf(f, "%f", &scale) != 1) return false;
    std::fgetc(f);
    swap = (scale < 0) != little_endian();
    data_off = std::ftell(f);
    buf.resize((size_t)w * 3);
    return w > 0 && h > 0;
  }
  bool row(uint32_t y, float *rgb) override {
    // bottom row first in the file
    const long off = data_off + (long)((size_t)(h - 1 - y) * w * 12);
    if (std::fseek(f, off, SEEK_SET) != 0) return false;
    if (std::fread(buf.data(), 4, buf.size(), f) != buf.size()) return false;
    for (size_t i = 0; i < buf.size(); i++) {
      float v = buf[i];
      if (swap) {
        unsigned char b[4];
        std::memcpy(b, &v, 4);
        std::swap(b[0], b[3]);
        std::swap(b[1], b[2]);
        std::memcpy(&v, b, 4);
      }
      rgb[i] = v;
    }
    return true;
  }
};

struct RgbeReader : Reader {
  FILE *f;
  std::vector<Rgbe> px;
  explicit RgbeReader(FILE *fp) : f(fp) {}
  bool open() {
    char line[512];
    if (!std::fgets(line, sizeof line, f)) return false;
    if (std::strcmp(line, "#?RGBE\n") != 0 && std::strcmp(line, "#?RADIANCE\n") != 0) return false;
    for (;;) {  // header lines up to a blank one
      if (!std::fgets(line, sizeof line, f)) return false;
      if (line[0] == '\n') break;
    }
    if (std::fscanf(f, "-Y %u +X %u", &h, &w) != 2) return false;
    if (std::fgetc(f) != '\n') return false;
    px.resize(w);
    return w > 0 && h > 0;
  }
  bool plane(unsigned char Rgbe::*c) {
    uint32_t i = 0;
    while (i < w) {
      int len = std::fgetc(f);
      if (len == EOF) return false;
      if (len > 128) {
        int v = std::fgetc(f);
        if (v == EOF) return false;
        len -= 128;
        if (i + (uint32_t)len > w) return false;
        while (len-- > 0) px[i++].*c = (unsigned char)v;
      } else {
        if (len == 0 || i + (uint32_t)len > w) return false;
        while (len-- > 0) {
          int v = std::fgetc(f);
          if (v == EOF) return false;
          px[i++].*c = (unsigned char)v;
        }
      }
    }
    return true;
  }
  bool row(uint32_t, float *rgb) override {
    int b[4];
    for (int &x : b)
      if ((x = std::fgetc(f)) == EOF) return false;
    if (b[0] == 2 && b[1] == 2 && ((uint32_t)b[2] << 8 | (uint32_t)b[3]) == w) {
      if (!plane(&Rgbe::r) || !plane(&Rgbe::g) || !plane(&Rgbe::b) || !plane(&Rgbe::e))
        return false;
    } else {
      // "old" format: plain pixels, where (1, 1, 1, n) repeats the
      // previous pixel n << shift times (shift grows by 8 for each
      // consecutive repeat pixel)
      px[0] = Rgbe{(unsigned char)b[0], (unsigned char)b[1], (unsigned char)b[2],
                   (unsigned char)b[3]};
      uint32_t i = 1;
      unsigned shift = 0;
      while (i < w) {
        int c[4];
        for (int &x : c)
          if ((x = std::fgetc(f)) == EOF) return false;
        if (c[0] == 1 && c[1] == 1 && c[2] == 1) {
          const uint64_t n = (uint64_t)c[3] << shift;
          if (n > w - i) return false;
          for (uint64_t k = 0; k < n; k++, i++) px[i] = px[i - 1];
          shift += 8;
        } else {
          px[i++] = Rgbe{(unsigned char)c[0], (unsigned char)c[1], (unsigned char)c[2],
                         (unsigned char)c[3]};
          shift = 0;
        }
      }
    }
    for (uint32_t x = 0; x < w; x++) from_rgbe(px[x], rgb + 3 * x);
    return true;
  }
};

int open_reader(const char *path, const char *format, File &file,
                std::unique_ptr<Reader> &rd) {
  Fmt fmt;
  int rc = find_format(path, format, fmt);
  if (rc) return rc;
  file.f = std::fopen(path, "rb");
  if (!file.f)
    return fail(SGR_ERR_INVALID, std::string("image: cannot open ") + path + ": " +
                                     std::strerror(errno));
  if (fmt == FMT_PFM) {
    auto *r = new PfmReader(file.f);
    rd.reset(r);
    if (!r->open()) return fail(SGR_ERR_INVALID, std::string("image: not a PFM file: ") + path);
  } else {
    auto *r = new RgbeReader(file.f);
    rd.reset(r);
    if (!r->open())
      return fail(SGR_ERR_INVALID, std::string("image: not a Radiance RGBE file: ") + path);
  }
  return SGR_OK;
}

}  // namespace

extern "C" {

int sgr_image_write(const char *path, const char *format, const float *rgba, uint32_t width,
                    uint32_t height) {
  return sgr::guarded("sgr_image_write", [&] {
    if (!path || !rgba || width == 0 || height == 0)
      return fail(SGR_ERR_INVALID, "sgr_image_write: null argument or empty image");
    Fmt fmt;
    int rc = find_format(path, format, fmt);
    if (rc) return rc;
    return fmt == FMT_PFM ? write_pfm(path, rgba, width, height)
                          : write_rgbe(path, rgba, width, height);
  });
}

int sgr_image_read(const char *path, const char *format, float *rgb, uint32_t *width,
                   uint32_t *height) {
  return sgr::guarded("sgr_image_read", [&] {
    if (!path || !width || !height) return fail(SGR_ERR_INVALID, "sgr_image_read: null argument");
    File file(nullptr);
    std::unique_ptr<Reader> rd;
    int rc = open_reader(path, format, file, rd);
    if (rc) return rc;
    if (!rgb) {
      *width = rd->w;
      *height = rd->h;
      return SGR_OK;
    }
    if (*width != rd->w || *height != rd->h)
      return fail(SGR_ERR_INVALID, "sgr_image_read: buffer size does not match the image");
    for (uint32_t y = 0; y < rd->h; y++)
      if (!rd->row(y, rgb + (size_t)3 * y * rd->w))
        return fail(SGR_ERR_INVALID, "sgr_image_read: truncated or malformed row " +
                                         std::to_string(y) + " of " + path);
    return SGR_OK;
  });
}

int sgr_image_recover(const char *path, const char *format, float *accum, uint32_t width,
                      uint32_t height, uint32_t *rows) {
  return sgr::guarded("sgr_image_recover", [&] {
    if (!path || !accum || !rows) return fail(SGR_ERR_INVALID, "sgr_image_recover: null argument");
    *rows = 0;
    File file(nullptr);
    std::unique_ptr<Reader> rd;
    if (open_reader(path, format, file, rd) != SGR_OK) {
      sgr::g_err.clear();  // an unreadable partial image recovers nothing
      return SGR_OK;
    }
    if (rd->w != width)
      return fail(SGR_ERR_INVALID, "sgr_image_recover: partial image is " + std::to_string(rd->w) +
                                       " pixels wide, the render " + std::to_string(width));
    std::vector<float> rgb((size_t)3 * width);
    uint32_t got = 0;
    bool failed = false;
    const uint32_t limit = rd->h < height ? rd->h : height;
    while (got < limit) {
      if (!rd->row(got, rgb.data())) {
        failed = true;
        break;
      }
      float *a = accum + (size_t)5 * got * width;
      for (uint32_t x = 0; x < width; x++) {
        a[5 * x] = rgb[3 * x];
        a[5 * x + 1] = rgb[3 * x + 1];
        a[5 * x + 2] = rgb[3 * x + 2];
        a[5 * x + 3] = 1.0f;
        a[5 * x + 4] = 1.0f;
      }
      got++;
    }
    if (failed || got < height) {
      // a partial file's last rows may be garbage: drop up to
      // SGR_RECOVER_DISCARD_ROWS of them (image/recover-image.cc:94-103)
      const uint32_t drop = failed ? (got < SGR_RECOVER_DISCARD_ROWS ? got : SGR_RECOVER_DISCARD_ROWS) : 0;
      for (uint32_t k = 0; k < drop; k++) {
        got--;
        std::memset(accum + (size_t)5 * got * width, 0, (size_t)5 * width * sizeof(float));
      }
    }
    *rows = got;
    return SGR_OK;
  });
}

}  // extern "C"
