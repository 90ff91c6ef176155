// bh_io.cpp — image files on the host: the PNG decoder that stands in for the
// reference's texture input (cv::imread(path, IMREAD_COLOR), config.h:35-37),
// and the PPM / PNG / JPEG frame writer and Motion-JPEG AVI writer that stand
// in for its cv::VideoWriter (blackhole_solution_test.cc:69-72,330).
//
// The decoder follows the PNG specification (W3C PNG, 2nd edition): chunk
// layout and CRC, IHDR, PLTE, IDAT as one zlib stream, the five scanline
// filters and Adam7 interlacing.  Output conversion follows IMREAD_COLOR:
// three 8-bit channels in B, G, R order; alpha and tRNS dropped without
// compositing; grey replicated to three channels; grey below 8 bits scaled to
// 0..255 (x * 255 / (2^d - 1)); 16-bit samples keep their high byte;
// palette entries looked up (an index beyond the palette reads black); gAMA,
// cHRM, sRGB and iCCP are not applied.
#include <zlib.h>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <exception>
#include <new>
#include <string>
#include <utility>
#include <vector>

#include "bh_api.h"

namespace {

uint32_t be32(const uint8_t* p) {
  return ((uint32_t)p[0] << 24) | ((uint32_t)p[1] << 16) | ((uint32_t)p[2] << 8) | p[3];
}

void put_be32(std::vector<uint8_t>& v, uint32_t x) {
  v.push_back((uint8_t)(x >> 24));
  v.push_back((uint8_t)(x >> 16));
  v.push_back((uint8_t)(x >> 8));
  v.push_back((uint8_t)x);
}

const uint8_t kSig[8] = {137, 80, 78, 71, 13, 10, 26, 10};

// Paeth predictor (PNG spec, section 9.4).
inline uint8_t paeth(int a, int b, int c) {
  const int p = a + b - c;
  const int pa = p > a ? p - a : a - p;
  const int pb = p > b ? p - b : b - p;
  const int pc = p > c ? p - c : c - p;
  if (pa <= pb && pa <= pc) return (uint8_t)a;
  if (pb <= pc) return (uint8_t)b;
  return (uint8_t)c;
}

// Reverse the filter of one scanline in place; prev is the unfiltered
// previous line of the same pass (nullptr on its first line).
bool unfilter(int type, uint8_t* line, const uint8_t* prev, size_t len, size_t bpp) {
  switch (type) {
    case 0:
      return true;
    case 1:
      for (size_t i = bpp; i < len; ++i) line[i] = (uint8_t)(line[i] + line[i - bpp]);
      return true;
    case 2:
      if (prev)
        for (size_t i = 0; i < len; ++i) line[i] = (uint8_t)(line[i] + prev[i]);
      return true;
    case 3:
      for (size_t i = 0; i < len; ++i) {
        const int a = i >= bpp ? line[i - bpp] : 0;
        const int b = prev ? prev[i] : 0;
        line[i] = (uint8_t)(line[i] + ((a + b) >> 1));
      }
      return true;
    case 4:
      for (size_t i = 0; i < len; ++i) {
        const int a = i >= bpp ? line[i - bpp] : 0;
        const int b = prev ? prev[i] : 0;
        const int c = (prev && i >= bpp) ? prev[i - bpp] : 0;
        line[i] = (uint8_t)(line[i] + paeth(a, b, c));
      }
      return true;
    default:
      return false;
  }
}

struct Png {
  uint32_t w = 0, h = 0;
  int depth = 0, ctype = 0, interlace = 0;
  int channels = 0;  // samples per pixel
  uint8_t pal[256][3] = {};
  int npal = 0;
  std::vector<uint8_t> idat;
};

int parse(const uint8_t* p, size_t n, Png& g, std::string& err) {
  if (!p || n < 8 || std::memcmp(p, kSig, 8) != 0) {
    err = "not a PNG file (signature)";
    return BH_ERR_INVALID_ARG;
  }
  size_t off = 8;
  bool have_ihdr = false, have_iend = false;
  while (off + 12 <= n) {
    const uint32_t len = be32(p + off);
    if (len > 0x7fffffffu || (size_t)len > n - off - 12) {
      err = "truncated chunk";
      return BH_ERR_INVALID_ARG;
    }
    const uint8_t* type = p + off + 4;
    const uint8_t* data = p + off + 8;
    const uint32_t crc = be32(data + len);
    if ((uint32_t)crc32(crc32(0L, Z_NULL, 0), type, len + 4) != crc) {
      err = "chunk CRC mismatch";
      return BH_ERR_INVALID_ARG;
    }
    const std::string t((const char*)type, 4);
    if (!have_ihdr && t != "IHDR") {
      err = "first chunk is not IHDR";
      return BH_ERR_INVALID_ARG;
    }
    if (t == "IHDR") {
      if (have_ihdr || len != 13) {
        err = "bad IHDR";
        return BH_ERR_INVALID_ARG;
      }
      g.w = be32(data);
      g.h = be32(data + 4);
      g.depth = data[8];
      g.ctype = data[9];
      g.interlace = data[12];
      if (g.w == 0 || g.h == 0 || g.w > 0x7fffffffu || g.h > 0x7fffffffu || data[10] != 0 ||
          data[11] != 0 || g.interlace > 1) {
        err = "bad IHDR fields";
        return BH_ERR_INVALID_ARG;
      }
      const int d = g.depth;
      bool ok = false;
      switch (g.ctype) {
        case 0: ok = d == 1 || d == 2 || d == 4 || d == 8 || d == 16; g.channels = 1; break;
        case 2: ok = d == 8 || d == 16; g.channels = 3; break;
        case 3: ok = d == 1 || d == 2 || d == 4 || d == 8; g.channels = 1; break;
        case 4: ok = d == 8 || d == 16; g.channels = 2; break;
        case 6: ok = d == 8 || d == 16; g.channels = 4; break;
        default: ok = false;
      }
      if (!ok) {
        err = "bad colour type / bit depth";
        return BH_ERR_INVALID_ARG;
      }
      if ((uint64_t)g.w * g.h > ((uint64_t)1 << 28)) {
        err = "image too large";
        return BH_ERR_UNSUPPORTED;
      }
      have_ihdr = true;
    } else if (t == "PLTE") {
      if (len % 3 != 0 || len == 0 || len > 768) {
        err = "bad PLTE";
        return BH_ERR_INVALID_ARG;
      }
      g.npal = (int)(len / 3);
      for (int i = 0; i < g.npal; ++i)
        for (int k = 0; k < 3; ++k) g.pal[i][k] = data[3 * i + k];
    } else if (t == "IDAT") {
      g.idat.insert(g.idat.end(), data, data + len);
    } else if (t == "IEND") {
      have_iend = true;
      break;
    } else if (!(type[0] & 0x20)) {
      err = "unknown critical chunk " + t;
      return BH_ERR_UNSUPPORTED;
    }
    off += 12 + (size_t)len;
  }
  if (!have_ihdr || !have_iend) {
    err = "missing IHDR or IEND";
    return BH_ERR_INVALID_ARG;
  }
  if (g.ctype == 3 && g.npal == 0) {
    err = "palette image without PLTE";
    return BH_ERR_INVALID_ARG;
  }
  if (g.idat.empty()) {
    err = "no IDAT";
    return BH_ERR_INVALID_ARG;
  }
  return BH_OK;
}

// Bytes of one filtered line of `width` pixels (without the filter byte).
size_t line_bytes(const Png& g, uint32_t width) {
  return ((size_t)width * g.channels * g.depth + 7) / 8;
}

// Sample k (0-based over the line) of an unfiltered line.
inline uint32_t sample(const uint8_t* line, size_t k, int depth) {
  switch (depth) {
    case 16: return ((uint32_t)line[2 * k] << 8) | line[2 * k + 1];
    case 8: return line[k];
    default: {
      const size_t bit = k * depth;
      const int shift = 8 - depth - (int)(bit & 7);
      return (line[bit >> 3] >> shift) & ((1u << depth) - 1);
    }
  }
}

// One pixel of an unfiltered line to B, G, R.
inline void to_bgr(const Png& g, const uint8_t* line, uint32_t x, uint8_t* out) {
  const int d = g.depth;
  const size_t s0 = (size_t)x * g.channels;
  auto eight = [d](uint32_t v) -> uint8_t {
    if (d == 16) return (uint8_t)(v >> 8);
    if (d == 8) return (uint8_t)v;
    return (uint8_t)(v * 255u / ((1u << d) - 1));
  };
  switch (g.ctype) {
    case 0:
    case 4: {
      const uint8_t v = eight(sample(line, s0, d));
      out[0] = out[1] = out[2] = v;
      break;
    }
    case 3: {
      const uint32_t i = sample(line, s0, d);
      if ((int)i < g.npal) {
        out[0] = g.pal[i][2];
        out[1] = g.pal[i][1];
        out[2] = g.pal[i][0];
      } else {
        out[0] = out[1] = out[2] = 0;
      }
      break;
    }
    default: {  // 2, 6: R, G, B(, A)
      out[2] = eight(sample(line, s0, d));
      out[1] = eight(sample(line, s0 + 1, d));
      out[0] = eight(sample(line, s0 + 2, d));
    }
  }
}

int decode(const Png& g, uint8_t* dst, std::string& err) {
  // Adam7 passes: x0, y0, dx, dy; a non-interlaced image is one pass.
  static const int kAdam7[7][4] = {{0, 0, 8, 8}, {4, 0, 8, 8}, {0, 4, 4, 8}, {2, 0, 4, 4},
                                   {0, 2, 2, 4}, {1, 0, 2, 2}, {0, 1, 1, 2}};
  static const int kFull[1][4] = {{0, 0, 1, 1}};
  const int (*passes)[4] = g.interlace ? kAdam7 : kFull;
  const int npass = g.interlace ? 7 : 1;
  size_t total = 0;
  for (int p = 0; p < npass; ++p) {
    const uint32_t pw = g.w > (uint32_t)passes[p][0]
                            ? (g.w - passes[p][0] + passes[p][2] - 1) / passes[p][2] : 0;
    const uint32_t ph = g.h > (uint32_t)passes[p][1]
                            ? (g.h - passes[p][1] + passes[p][3] - 1) / passes[p][3] : 0;
    if (pw && ph) total += (size_t)ph * (line_bytes(g, pw) + 1);
  }
  std::vector<uint8_t> raw(total);
  z_stream zs;
  std::memset(&zs, 0, sizeof zs);
  if (inflateInit(&zs) != Z_OK) {
    err = "inflateInit failed";
    return BH_ERR_INVALID_ARG;
  }
  zs.next_in = const_cast<Bytef*>(g.idat.data());
  zs.avail_in = (uInt)g.idat.size();
  zs.next_out = raw.data();
  zs.avail_out = (uInt)raw.size();
  const int zr = inflate(&zs, Z_FINISH);
  const size_t got = raw.size() - zs.avail_out;
  inflateEnd(&zs);
  // Data beyond the image (a full buffer before the stream end) is ignored.
  const bool full = zs.avail_out == 0 && (zr == Z_OK || zr == Z_BUF_ERROR);
  if ((zr != Z_STREAM_END && !full) || got != total) {
    err = "IDAT stream is corrupt or short";
    return BH_ERR_INVALID_ARG;
  }
  const size_t bpp = ((size_t)g.channels * g.depth + 7) / 8;
  size_t off = 0;
  for (int p = 0; p < npass; ++p) {
    const int x0 = passes[p][0], y0 = passes[p][1], dx = passes[p][2], dy = passes[p][3];
    const uint32_t pw = g.w > (uint32_t)x0 ? (g.w - x0 + dx - 1) / dx : 0;
    const uint32_t ph = g.h > (uint32_t)y0 ? (g.h - y0 + dy - 1) / dy : 0;
    if (!pw || !ph) continue;
    const size_t lb = line_bytes(g, pw);
    const uint8_t* prev = nullptr;
    for (uint32_t j = 0; j < ph; ++j) {
      uint8_t* rec = raw.data() + off;
      if (!unfilter(rec[0], rec + 1, prev, lb, bpp)) {
        err = "bad filter type";
        return BH_ERR_INVALID_ARG;
      }
      const size_t y = (size_t)y0 + (size_t)j * dy;
      for (uint32_t i = 0; i < pw; ++i) {
        const size_t x = (size_t)x0 + (size_t)i * dx;
        to_bgr(g, rec + 1, i, dst + (y * g.w + x) * 3);
      }
      prev = rec + 1;
      off += lb + 1;
    }
  }
  return BH_OK;
}

thread_local std::string t_err;

int io_fail(int code, const std::string& msg) {
  t_err = msg;
  return code;
}

// Body of a C entry point: no exception crosses the C boundary.
template <class F>
int guarded(F&& body) {
  try {
    return body();
  } catch (const std::bad_alloc&) {
    return io_fail(BH_ERR_OUT_OF_MEMORY, "out of host memory");
  } catch (const std::exception& e) {
    return io_fail(BH_ERR_INVALID_ARG, e.what());
  }
}

bool read_file(const char* path, std::vector<uint8_t>& out) {
  FILE* f = std::fopen(path, "rb");
  if (!f) return false;
  uint8_t buf[1 << 16];
  size_t n;
  while ((n = std::fread(buf, 1, sizeof buf, f)) > 0) out.insert(out.end(), buf, buf + n);
  const bool ok = !std::ferror(f);
  std::fclose(f);
  return ok;
}

// Binary PPM (P6, maxval 255) in R, G, B order, returned as B, G, R.
int decode_ppm(const std::vector<uint8_t>& f, uint8_t* dst, size_t cap, int* rows, int* cols) {
  size_t i = 2;
  long vals[3];
  for (int k = 0; k < 3; ++k) {
    for (;;) {  // whitespace and comments
      while (i < f.size() && (f[i] == ' ' || f[i] == '\t' || f[i] == '\n' || f[i] == '\r')) ++i;
      if (i < f.size() && f[i] == '#') {
        while (i < f.size() && f[i] != '\n') ++i;
        continue;
      }
      break;
    }
    long v = 0;
    int nd = 0;
    while (i < f.size() && f[i] >= '0' && f[i] <= '9' && nd < 9) v = v * 10 + (f[i++] - '0'), ++nd;
    if (nd == 0) return io_fail(BH_ERR_INVALID_ARG, "bad PPM header");
    vals[k] = v;
  }
  if (i >= f.size() || vals[0] <= 0 || vals[1] <= 0 || vals[2] != 255)
    return io_fail(BH_ERR_UNSUPPORTED, "PPM must be binary 8-bit (P6, maxval 255)");
  ++i;  // the single whitespace byte before the raster
  const size_t n = (size_t)vals[0] * vals[1];
  if (f.size() - i < n * 3) return io_fail(BH_ERR_INVALID_ARG, "truncated PPM raster");
  *cols = (int)vals[0];
  *rows = (int)vals[1];
  if (!dst) return BH_OK;
  if (cap < n * 3) return io_fail(BH_ERR_INVALID_ARG, "dst is too small");
  for (size_t k = 0; k < n; ++k) {
    dst[3 * k] = f[i + 3 * k + 2];
    dst[3 * k + 1] = f[i + 3 * k + 1];
    dst[3 * k + 2] = f[i + 3 * k];
  }
  return BH_OK;
}

void png_chunk(std::vector<uint8_t>& out, const char* type, const uint8_t* data, size_t len) {
  put_be32(out, (uint32_t)len);
  const size_t t0 = out.size();
  out.insert(out.end(), type, type + 4);
  if (len) out.insert(out.end(), data, data + len);
  put_be32(out, (uint32_t)crc32(crc32(0L, Z_NULL, 0), out.data() + t0, (uInt)(len + 4)));
}

}  // namespace

extern "C" {

int bh_jpeg_decode(const uint8_t* data, size_t size, uint8_t* dst, size_t dst_size, int* rows,
                   int* cols);
const char* bh_jpeg_last_error(void);

const char* bh_io_last_error(void) { return t_err.c_str(); }

}  // extern "C"

namespace {

bool ends_with(const std::string& p, const char* suf) {
  const size_t n = std::strlen(suf);
  return p.size() >= n && p.compare(p.size() - n, n, suf) == 0;
}

void put_le32(uint8_t* p, uint32_t x) {
  for (int i = 0; i < 4; ++i) p[i] = (uint8_t)(x >> (8 * i));
}

// Header layout of the AVI file written below (byte offsets):
//   0   RIFF <size> 'AVI '
//   12  LIST <size> 'hdrl'
//   24  'avih' 56: us/frame, max bytes/s, padding, flags (HASINDEX),
//       total frames (@48), initial frames, streams, buffer size, W, H, 4 reserved
//   88  LIST <size> 'strl'
//   100 'strh' 56: 'vids' 'MJPG', flags, priority, language, initial frames,
//       scale, rate, start, length (@140), buffer size, quality, sample size,
//       rcFrame (4 x int16)
//   164 'strf' 40: BITMAPINFOHEADER (24-bit, compression 'MJPG')
//   212 LIST <size> 'movi'   (frames follow at 224)
constexpr size_t kHdr = 224, kMoviList = 212;
constexpr uint64_t kMaxAvi = (uint64_t)1 << 30;

}  // namespace

struct bh_video {
  FILE* f = nullptr;
  int rows = 0, cols = 0, quality = 95;
  double fps = 29;
  uint64_t pos = kHdr;           // file size so far
  uint32_t max_frame = 0;
  std::vector<uint8_t> idx;      // idx1 entries, 16 bytes each
  int err = BH_OK;
  std::string msg;
  uint8_t hdr[kHdr] = {};
};

namespace {

int video_fail(bh_video* v, int code, const std::string& m) {
  if (v->err == BH_OK) {
    v->err = code;
    v->msg = m;
  }
  return io_fail(code, m);
}

// Header for `frames` frames written so far; the RIFF and movi sizes come
// from the file position and the index built so far.
void fill_header(bh_video* v, uint32_t frames) {
  uint8_t* h = v->hdr;
  const double fps = v->fps;
  auto tag = [&](size_t at, const char* t) { std::memcpy(h + at, t, 4); };
  // frame rate as rate / scale: whole fps exactly, else in 1/1000 frames/s
  const bool whole = fps == (double)(uint32_t)fps;
  const uint32_t scale = whole ? 1 : 1000, rate = whole ? (uint32_t)fps : (uint32_t)(fps * 1000 + 0.5);
  const uint32_t frame_bytes = (uint32_t)v->rows * v->cols * 3;
  tag(0, "RIFF");
  put_le32(h + 4, (uint32_t)(v->pos + 8 + v->idx.size() - 8));  // file size - 8
  tag(8, "AVI ");
  tag(12, "LIST");
  put_le32(h + 16, 212 - 20);
  tag(20, "hdrl");
  tag(24, "avih");
  put_le32(h + 28, 56);
  put_le32(h + 32, (uint32_t)(1e6 / fps + 0.5));
  put_le32(h + 36, (uint32_t)(v->max_frame * fps + 0.5));
  put_le32(h + 44, 0x10);  // AVIF_HASINDEX
  put_le32(h + 48, frames);
  put_le32(h + 56, 1);
  put_le32(h + 60, v->max_frame);
  put_le32(h + 64, (uint32_t)v->cols);
  put_le32(h + 68, (uint32_t)v->rows);
  tag(88, "LIST");
  put_le32(h + 92, 212 - 96);
  tag(96, "strl");
  tag(100, "strh");
  put_le32(h + 104, 56);
  tag(108, "vids");
  tag(112, "MJPG");
  put_le32(h + 128, scale);
  put_le32(h + 132, rate);
  put_le32(h + 140, frames);
  put_le32(h + 144, v->max_frame);
  put_le32(h + 148, 0xFFFFFFFFu);  // default quality
  h[160] = (uint8_t)(v->cols & 0xFF);
  h[161] = (uint8_t)(v->cols >> 8);
  h[162] = (uint8_t)(v->rows & 0xFF);
  h[163] = (uint8_t)(v->rows >> 8);
  tag(164, "strf");
  put_le32(h + 168, 40);
  put_le32(h + 172, 40);
  put_le32(h + 176, (uint32_t)v->cols);
  put_le32(h + 180, (uint32_t)v->rows);
  h[184] = 1;   // planes
  h[186] = 24;  // bits per pixel
  tag(188, "MJPG");
  put_le32(h + 192, frame_bytes);
  tag(kMoviList, "LIST");
  put_le32(h + kMoviList + 4, (uint32_t)(v->pos - kMoviList - 8));
  tag(kMoviList + 8, "movi");
}

}  // namespace

extern "C" {

int bh_video_open(const char* path, int rows, int cols, double fps, int quality, bh_video** out) {
  return guarded([&]() -> int {
    if (!out) return io_fail(BH_ERR_INVALID_ARG, "out is NULL");
    *out = nullptr;
    if (!path || rows <= 0 || cols <= 0 || rows > 65535 || cols > 65535)
      return io_fail(BH_ERR_INVALID_ARG, "bad video size");
    if (!(fps > 0 && fps <= 1000)) return io_fail(BH_ERR_INVALID_ARG, "fps must be in (0, 1000]");
    if (quality < 1 || quality > 100) return io_fail(BH_ERR_INVALID_ARG, "quality must be 1..100");
    bh_video* v = new (std::nothrow) bh_video;
    if (!v) return io_fail(BH_ERR_OUT_OF_MEMORY, "out of host memory");
    v->rows = rows;
    v->cols = cols;
    v->quality = quality;
    v->f = std::fopen(path, "wb");
    if (!v->f) {
      delete v;
      return io_fail(BH_ERR_INVALID_ARG, std::string("cannot write ") + path);
    }
    v->fps = fps;
    fill_header(v, 0);  // rewritten by bh_video_close with the counts
    if (std::fwrite(v->hdr, 1, kHdr, v->f) != kHdr) {
      std::fclose(v->f);
      delete v;
      return io_fail(BH_ERR_INVALID_ARG, std::string("write failed: ") + path);
    }
    *out = v;
    return BH_OK;
  });
}

int bh_video_write(bh_video* v, const uint8_t* bgr) {
  return guarded([&]() -> int {
    if (!v || !bgr) return io_fail(BH_ERR_INVALID_ARG, "video/frame is NULL");
    if (v->err != BH_OK) return io_fail(v->err, v->msg);
    uint8_t* jpg = nullptr;
    size_t n = 0;
    int rc = bh_jpeg_encode(bgr, v->rows, v->cols, v->quality, &jpg, &n);
    if (rc != BH_OK) return video_fail(v, rc, bh_jpeg_last_error());
    const uint64_t chunk = 8 + n + (n & 1);
    if (v->pos + chunk + v->idx.size() + 16 + 8 > kMaxAvi) {
      bh_free(jpg);
      return video_fail(v, BH_ERR_UNSUPPORTED, "AVI file would pass 1 GiB");
    }
    try {
      v->idx.reserve(v->idx.size() + 16);  // the index entry cannot fail once the chunk is out
    } catch (...) {
      bh_free(jpg);
      throw;
    }
    uint8_t ch[8] = {'0', '0', 'd', 'c'};
    put_le32(ch + 4, (uint32_t)n);
    const uint8_t pad = 0;
    bool ok = std::fwrite(ch, 1, 8, v->f) == 8 && std::fwrite(jpg, 1, n, v->f) == n &&
              (!(n & 1) || std::fwrite(&pad, 1, 1, v->f) == 1);
    bh_free(jpg);
    if (!ok) return video_fail(v, BH_ERR_INVALID_ARG, "video write failed");
    uint8_t e[16] = {'0', '0', 'd', 'c'};
    put_le32(e + 4, 0x10);  // AVIIF_KEYFRAME
    put_le32(e + 8, (uint32_t)(v->pos - (kMoviList + 8)));  // from the 'movi' tag
    put_le32(e + 12, (uint32_t)n);
    v->idx.insert(v->idx.end(), e, e + 16);
    v->pos += chunk;
    if (n > v->max_frame) v->max_frame = (uint32_t)n;
    return BH_OK;
  });
}

int bh_video_close(bh_video* v) {
  return guarded([&]() -> int {
    if (!v) return io_fail(BH_ERR_INVALID_ARG, "video is NULL");
    int rc = v->err;
    std::string msg = std::move(v->msg);
    uint8_t ih[8] = {'i', 'd', 'x', '1'};
    put_le32(ih + 4, (uint32_t)v->idx.size());
    bool ok = std::fwrite(ih, 1, 8, v->f) == 8 &&
              std::fwrite(v->idx.data(), 1, v->idx.size(), v->f) == v->idx.size();
    fill_header(v, (uint32_t)(v->idx.size() / 16));
    ok = ok && std::fseek(v->f, 0, SEEK_SET) == 0 &&
         std::fwrite(v->hdr, 1, kHdr, v->f) == kHdr;
    ok = (std::fclose(v->f) == 0) && ok;
    delete v;
    if (rc == BH_OK && !ok) {
      rc = BH_ERR_INVALID_ARG;
      msg = "video write failed";
    }
    return rc == BH_OK ? BH_OK : io_fail(rc, msg);
  });
}

int bh_png_decode(const uint8_t* data, size_t size, uint8_t* dst, size_t dst_size, int* rows,
                  int* cols) {
  return guarded([&]() -> int {
    if (!rows || !cols) return io_fail(BH_ERR_INVALID_ARG, "rows/cols are NULL");
    Png g;
    std::string err;
    int rc = parse(data, size, g, err);
    if (rc) return io_fail(rc, err);
    *rows = (int)g.h;
    *cols = (int)g.w;
    if (!dst) return BH_OK;
    if (dst_size < (size_t)g.w * g.h * 3) return io_fail(BH_ERR_INVALID_ARG, "dst is too small");
    rc = decode(g, dst, err);
    return rc ? io_fail(rc, err) : BH_OK;
  });
}

int bh_image_load(const char* path, uint8_t* dst, size_t dst_size, int* rows, int* cols) {
  return guarded([&]() -> int {
    if (!path || !rows || !cols) return io_fail(BH_ERR_INVALID_ARG, "path/rows/cols are NULL");
    std::vector<uint8_t> f;
    if (!read_file(path, f)) return io_fail(BH_ERR_INVALID_ARG, std::string("cannot read ") + path);
    if (f.size() >= 8 && std::memcmp(f.data(), kSig, 8) == 0)
      return bh_png_decode(f.data(), f.size(), dst, dst_size, rows, cols);
    if (f.size() >= 2 && f[0] == 'P' && f[1] == '6') return decode_ppm(f, dst, dst_size, rows, cols);
    if (f.size() >= 3 && f[0] == 0xFF && f[1] == 0xD8 && f[2] == 0xFF) {
      const int rc = bh_jpeg_decode(f.data(), f.size(), dst, dst_size, rows, cols);
      return rc ? io_fail(rc, bh_jpeg_last_error()) : BH_OK;
    }
    return io_fail(BH_ERR_UNSUPPORTED, "only PNG, JPEG and binary PPM files are read");
  });
}

int bh_image_write(const char* path, const uint8_t* bgr, int rows, int cols) {
  return guarded([&]() -> int {
    if (!path || !bgr || rows <= 0 || cols <= 0) return io_fail(BH_ERR_INVALID_ARG, "bad image");
    const std::string p(path);
    const bool png = ends_with(p, ".png"), ppm = ends_with(p, ".ppm");
    const bool jpg = ends_with(p, ".jpg") || ends_with(p, ".jpeg");
    if (!png && !ppm && !jpg)
      return io_fail(BH_ERR_UNSUPPORTED, "path must end in .ppm, .png, .jpg or .jpeg");
    const size_t n = (size_t)rows * cols;
    std::vector<uint8_t> out;
    if (jpg) {
      uint8_t* d = nullptr;
      size_t dn = 0;
      const int rc = bh_jpeg_encode(bgr, rows, cols, 95, &d, &dn);
      if (rc != BH_OK) return io_fail(rc, bh_jpeg_last_error());
      out.assign(d, d + dn);
      bh_free(d);
    } else if (ppm) {
      char hdr[64];
      const int hl = std::snprintf(hdr, sizeof hdr, "P6\n%d %d\n255\n", cols, rows);
      out.assign(hdr, hdr + hl);
      out.resize(hl + n * 3);
      uint8_t* o = out.data() + hl;
      for (size_t k = 0; k < n; ++k) {
        o[3 * k] = bgr[3 * k + 2];
        o[3 * k + 1] = bgr[3 * k + 1];
        o[3 * k + 2] = bgr[3 * k];
      }
    } else {
      // RGB 8-bit, filter 0 on every line, one zlib stream at speed level 1.
      const size_t lb = (size_t)cols * 3 + 1;
      std::vector<uint8_t> raw(lb * rows);
      for (int y = 0; y < rows; ++y) {
        uint8_t* l = raw.data() + y * lb;
        l[0] = 0;
        const uint8_t* s = bgr + (size_t)y * cols * 3;
        for (int x = 0; x < cols; ++x) {
          l[1 + 3 * x] = s[3 * x + 2];
          l[2 + 3 * x] = s[3 * x + 1];
          l[3 + 3 * x] = s[3 * x];
        }
      }
      uLongf zlen = compressBound((uLong)raw.size());
      std::vector<uint8_t> z(zlen);
      if (compress2(z.data(), &zlen, raw.data(), (uLong)raw.size(), 1) != Z_OK)
        return io_fail(BH_ERR_INVALID_ARG, "deflate failed");
      out.assign(kSig, kSig + 8);
      std::vector<uint8_t> ihdr;
      put_be32(ihdr, (uint32_t)cols);
      put_be32(ihdr, (uint32_t)rows);
      const uint8_t tail[5] = {8, 2, 0, 0, 0};
      ihdr.insert(ihdr.end(), tail, tail + 5);
      png_chunk(out, "IHDR", ihdr.data(), ihdr.size());
      png_chunk(out, "IDAT", z.data(), zlen);
      png_chunk(out, "IEND", nullptr, 0);
    }
    FILE* f = std::fopen(path, "wb");
    if (!f) return io_fail(BH_ERR_INVALID_ARG, std::string("cannot write ") + path);
    const bool ok = std::fwrite(out.data(), 1, out.size(), f) == out.size();
    if (std::fclose(f) != 0 || !ok) return io_fail(BH_ERR_INVALID_ARG, std::string("write failed: ") + path);
    return BH_OK;
  });
}

}  // extern "C"
