// bh_jpeg.cpp — baseline JPEG decoder (host code), the JPEG half of the
// texture input that stands in for cv::imread(path, IMREAD_COLOR)
// (config.h:35-37); the reference's flat scene textures mooni.jpeg,
// karina.jpeg and winter.jpg are baseline 4:2:0 JFIF files.
//
// Follows ITU-T T.81 (baseline and extended sequential Huffman, 8-bit):
// DQT, DHT, SOF0/SOF1, SOS (interleaved and single-component scans), DRI and
// restart markers.  The output is meant to equal what libjpeg's default
// decompression path gives (the path imread takes): the integer
// Loeffler-Ligtenberg-Moschytz IDCT with 13-bit constants and 2 extra bits
// in the first pass, "fancy" triangle upsampling of 2:1 chroma (3/4 and 1/4
// weights with alternating rounding), and the fixed-point YCbCr -> RGB of
// JFIF with 16-bit scaling; three 8-bit channels in B, G, R order.
// Progressive, arithmetic-coded, 12-bit and 4-component (CMYK/YCCK) files
// are refused with BH_ERR_UNSUPPORTED.
//
// The second half is a baseline encoder for frame output (the JPEG frames of
// the MJPG video the reference writes with cv::VideoWriter,
// blackhole_solution_test.cc:71-72,330): JFIF YCbCr, 2x2-averaged chroma
// (4:2:0), the Annex K quantization tables scaled by quality, and Huffman
// tables made for each image from its own symbol counts (Annex K.2).
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <algorithm>
#include <array>
#include <exception>
#include <string>
#include <system_error>
#include <thread>
#include <vector>

#include "bh_api.h"

namespace {

const int kZigzag[64] = {0,  1,  8,  16, 9,  2,  3,  10, 17, 24, 32, 25, 18, 11, 4,  5,
                         12, 19, 26, 33, 40, 48, 41, 34, 27, 20, 13, 6,  7,  14, 21, 28,
                         35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23, 30, 37, 44, 51,
                         58, 59, 52, 45, 38, 31, 39, 46, 53, 60, 61, 54, 47, 55, 62, 63};

struct Huff {
  bool present = false;
  uint8_t vals[256] = {};
  int mincode[17] = {}, maxcode[18] = {}, valptr[17] = {};
};

struct Comp {
  int id = 0, h = 1, v = 1, tq = 0, td = 0, ta = 0;
  int bw = 0, bh = 0;     // blocks per row / column stored (padded to whole MCUs)
  int dw = 0, dh = 0;     // downsampled width / height (real samples)
  int dc = 0;             // DC predictor
  std::vector<uint8_t> px;  // bw*8 x bh*8 samples
};

struct Jpeg {
  int W = 0, H = 0, n = 0, hmax = 1, vmax = 1, mcux = 0, mcuy = 0, restart = 0;
  bool frame = false, adobe = false, jfif = false;
  int adobe_transform = -1;
  Comp c[4];
  uint16_t q[4][64] = {};
  bool qset[4] = {};
  Huff dc[4], ac[4];
};

struct Fail {
  int code;
  std::string msg;
};

[[noreturn]] void fail(int code, const std::string& m) { throw Fail{code, m}; }

void build_huff(Huff& t, const uint8_t* counts, const uint8_t* vals, int nvals) {
  t.present = true;
  std::memcpy(t.vals, vals, (size_t)nvals);
  int code = 0, k = 0;
  for (int l = 1; l <= 16; ++l) {
    t.valptr[l] = k;
    t.mincode[l] = code;
    code += counts[l - 1];
    k += counts[l - 1];
    t.maxcode[l] = counts[l - 1] ? code - 1 : -1;
    code <<= 1;
  }
  t.maxcode[17] = 0x7fffffff;
}

// Entropy-coded segment reader: byte stuffing (FF 00) removed; a marker ends
// the data (zeros are shifted in past it, as decoders do for truncated data).
struct Bits {
  const uint8_t* p;
  const uint8_t* end;
  uint32_t acc = 0;
  int n = 0;
  bool marker = false;
  int bit() {
    if (n == 0) {
      uint32_t b = 0;
      if (!marker && p < end) {
        b = *p;
        if (b == 0xFF) {
          const uint8_t nx = p + 1 < end ? p[1] : 0;
          if (nx == 0x00) {
            p += 2;
          } else {
            marker = true;  // leave the marker for the parser
            b = 0;
          }
        } else {
          ++p;
        }
      }
      acc = b;
      n = 8;
    }
    --n;
    return (int)((acc >> n) & 1u);
  }
  int bits(int k) {
    int v = 0;
    for (int i = 0; i < k; ++i) v = (v << 1) | bit();
    return v;
  }
  void align() { n = 0; }
};

int decode_huff(Bits& b, const Huff& t) {
  int code = b.bit();
  int l = 1;
  while (code > t.maxcode[l]) {
    code = (code << 1) | b.bit();
    if (++l > 16) fail(BH_ERR_INVALID_ARG, "bad Huffman code");
  }
  return t.vals[t.valptr[l] + code - t.mincode[l]];
}

inline int extend(int v, int s) { return (s && v < (1 << (s - 1))) ? v - (1 << s) + 1 : v; }

// ---- IDCT: integer LL&M, 13-bit constants, 2 extra bits in pass 1 --------
constexpr int kConstBits = 13, kPass1Bits = 2;
constexpr int32_t fix(double x) { return (int32_t)(x * (1 << kConstBits) + 0.5); }
constexpr int32_t F0298 = fix(0.298631336), F0390 = fix(0.390180644), F0541 = fix(0.541196100),
                  F0765 = fix(0.765366865), F0899 = fix(0.899976223), F1175 = fix(1.175875602),
                  F1501 = fix(1.501321110), F1847 = fix(1.847759065), F1961 = fix(1.961570560),
                  F2053 = fix(2.053119869), F2562 = fix(2.562915447), F3072 = fix(3.072711026);

inline int32_t descale(int64_t x, int n) { return (int32_t)((x + ((int64_t)1 << (n - 1))) >> n); }

// One 1-D pass over 8 values v (already dequantized / from pass 1) into o,
// scaled down by `shift` bits after the rotation.
inline void idct_1d(const int32_t* v, int64_t* o) {
  // even part
  int64_t z2 = v[2], z3 = v[6];
  int64_t z1 = (z2 + z3) * F0541;
  int64_t t2 = z1 - z3 * F1847;
  int64_t t3 = z1 + z2 * F0765;
  int64_t t0 = ((int64_t)v[0] + v[4]) * (1 << kConstBits);
  int64_t t1 = ((int64_t)v[0] - v[4]) * (1 << kConstBits);
  const int64_t t10 = t0 + t3, t13 = t0 - t3, t11 = t1 + t2, t12 = t1 - t2;
  // odd part
  t0 = v[7];
  t1 = v[5];
  t2 = v[3];
  t3 = v[1];
  z1 = t0 + t3;
  z2 = t1 + t2;
  z3 = t0 + t2;
  int64_t z4 = t1 + t3;
  const int64_t z5 = (z3 + z4) * F1175;
  t0 *= F0298;
  t1 *= F2053;
  t2 *= F3072;
  t3 *= F1501;
  z1 *= -F0899;
  z2 *= -F2562;
  z3 *= -F1961;
  z4 *= -F0390;
  z3 += z5;
  z4 += z5;
  t0 += z1 + z3;
  t1 += z2 + z4;
  t2 += z2 + z3;
  t3 += z1 + z4;
  o[0] = t10 + t3;
  o[7] = t10 - t3;
  o[1] = t11 + t2;
  o[6] = t11 - t2;
  o[2] = t12 + t1;
  o[5] = t12 - t1;
  o[3] = t13 + t0;
  o[4] = t13 - t0;
}

void idct_block(const int32_t* coef, uint8_t* out, int stride) {
  int32_t ws[64];
  for (int x = 0; x < 8; ++x) {  // pass 1: columns
    int32_t v[8];
    for (int k = 0; k < 8; ++k) v[k] = coef[k * 8 + x];
    int64_t o[8];
    idct_1d(v, o);
    for (int k = 0; k < 8; ++k) ws[k * 8 + x] = descale(o[k], kConstBits - kPass1Bits);
  }
  for (int y = 0; y < 8; ++y) {  // pass 2: rows, then +128 and clamp
    int64_t o[8];
    idct_1d(ws + y * 8, o);
    for (int k = 0; k < 8; ++k) {
      int v = descale(o[k], kConstBits + kPass1Bits + 3) + 128;
      out[y * stride + k] = (uint8_t)(v < 0 ? 0 : (v > 255 ? 255 : v));
    }
  }
}

// ---- upsampling to full resolution -----------------------------------------
// Row r of a component plane, with the edge rows repeated above and below.
inline const uint8_t* crow(const Comp& c, int r) {
  if (r < 0) r = 0;
  if (r >= c.dh) r = c.dh - 1;
  return c.px.data() + (size_t)r * c.bw * 8;
}

// Horizontal 2:1 triangle filter of one row of w >= 2 samples into 2w.
void h2_row(const int* cs, int w, uint8_t* out, int bias_even, int bias_odd, int scale) {
  // cs: column sums (already weighted vertically, or the samples x scale)
  out[0] = (uint8_t)((cs[0] * 4 + bias_even) >> scale);
  out[1] = (uint8_t)((cs[0] * 3 + cs[1] + bias_odd) >> scale);
  for (int j = 1; j < w - 1; ++j) {
    out[2 * j] = (uint8_t)((cs[j] * 3 + cs[j - 1] + bias_even) >> scale);
    out[2 * j + 1] = (uint8_t)((cs[j] * 3 + cs[j + 1] + bias_odd) >> scale);
  }
  out[2 * w - 2] = (uint8_t)((cs[w - 1] * 3 + cs[w - 2] + bias_even) >> scale);
  out[2 * w - 1] = (uint8_t)((cs[w - 1] * 4 + bias_odd) >> scale);
}

// Full-resolution plane (W x H) of component c.  As in libjpeg, the
// horizontal 2:1 cases (h2v1, h2v2) use the triangle filter only for
// components more than 2 samples wide; narrower ones are replicated.
std::vector<uint8_t> upsample(const Jpeg& j, const Comp& c) {
  const int W = j.W, H = j.H;
  std::vector<uint8_t> out((size_t)W * H);
  int fx = j.hmax / c.h;
  const int fy = j.vmax / c.v;
  const bool exact_x = c.h * fx == j.hmax, exact_y = c.v * fy == j.vmax;
  if (fx == 2 && c.dw <= 2) fx = 0;  // to the replication branch
  std::vector<int> cs((size_t)c.dw);
  std::vector<uint8_t> tmp((size_t)c.dw * 2 + 2);
  for (int y = 0; y < H; ++y) {
    uint8_t* o = out.data() + (size_t)y * W;
    if (fx == 1 && fy == 1 && exact_x && exact_y) {
      std::memcpy(o, crow(c, y), (size_t)W);
    } else if (fx == 2 && fy == 1 && exact_x && exact_y) {
      // h2v1: 3/4 nearer + 1/4 farther horizontally, bias 1 / 2
      const uint8_t* r = crow(c, y);
      for (int i = 0; i < c.dw; ++i) cs[i] = r[i];
      h2_row(cs.data(), c.dw, tmp.data(), 1, 2, 2);
      std::memcpy(o, tmp.data(), (size_t)W);
    } else if (fx == 2 && fy == 2 && exact_x && exact_y) {
      // h2v2: vertical 3/4 + 1/4 column sums, then horizontal, bias 8 / 7
      const int ry = y >> 1;
      const uint8_t* r0 = crow(c, ry);
      const uint8_t* r1 = crow(c, (y & 1) ? ry + 1 : ry - 1);
      for (int i = 0; i < c.dw; ++i) cs[i] = r0[i] * 3 + r1[i];
      h2_row(cs.data(), c.dw, tmp.data(), 8, 7, 4);
      std::memcpy(o, tmp.data(), (size_t)W);
    } else if (fx == 1 && fy == 2 && exact_x && exact_y) {
      // h1v2: vertical 3/4 + 1/4, bias 1 (upper output row) / 2 (lower)
      const int ry = y >> 1;
      const uint8_t* r0 = crow(c, ry);
      const uint8_t* r1 = crow(c, (y & 1) ? ry + 1 : ry - 1);
      const int bias = (y & 1) ? 2 : 1;
      for (int x = 0; x < W; ++x) o[x] = (uint8_t)((r0[x] * 3 + r1[x] + bias) >> 2);
    } else {
      // other ratios: replication
      const uint8_t* r = crow(c, y * c.v / j.vmax);
      for (int x = 0; x < W; ++x) o[x] = r[x * c.h / j.hmax];
    }
  }
  return out;
}

inline uint8_t clamp8(int v) { return (uint8_t)(v < 0 ? 0 : (v > 255 ? 255 : v)); }

void color_convert(const Jpeg& j, const std::vector<uint8_t>* pl, uint8_t* dst) {
  const size_t n = (size_t)j.W * j.H;
  if (j.n == 1) {
    for (size_t i = 0; i < n; ++i) dst[3 * i] = dst[3 * i + 1] = dst[3 * i + 2] = pl[0][i];
    return;
  }
  const bool rgb = (j.adobe && j.adobe_transform == 0) ||
                   (!j.jfif && !j.adobe && j.c[0].id == 'R' && j.c[1].id == 'G' &&
                    j.c[2].id == 'B');
  if (rgb) {
    for (size_t i = 0; i < n; ++i) {
      dst[3 * i] = pl[2][i];
      dst[3 * i + 1] = pl[1][i];
      dst[3 * i + 2] = pl[0][i];
    }
    return;
  }
  // JFIF YCbCr -> RGB with 16-bit fixed-point coefficients
  constexpr int kS = 16;
  constexpr int64_t kHalf = (int64_t)1 << (kS - 1);
  auto F = [](double x) { return (int64_t)(x * (1 << kS) + 0.5); };
  int cr_r[256], cb_b[256];
  int64_t cr_g[256], cb_g[256];
  for (int i = 0; i < 256; ++i) {
    const int64_t x = i - 128;
    cr_r[i] = (int)((F(1.40200) * x + kHalf) >> kS);
    cb_b[i] = (int)((F(1.77200) * x + kHalf) >> kS);
    cr_g[i] = -F(0.71414) * x;
    cb_g[i] = -F(0.34414) * x + kHalf;
  }
  for (size_t i = 0; i < n; ++i) {
    const int y = pl[0][i], cb = pl[1][i], cr = pl[2][i];
    dst[3 * i + 2] = clamp8(y + cr_r[cr]);
    dst[3 * i + 1] = clamp8(y + (int)((cb_g[cb] + cr_g[cr]) >> kS));
    dst[3 * i] = clamp8(y + cb_b[cb]);
  }
}

uint16_t be16(const uint8_t* p) { return (uint16_t)((p[0] << 8) | p[1]); }

void decode_scan(Jpeg& j, const uint8_t* p, const uint8_t* end, int ns, const int* ci,
                 const uint8_t** after) {
  Bits b{p, end};
  for (int k = 0; k < ns; ++k) j.c[ci[k]].dc = 0;
  int32_t coef[64];
  auto block = [&](Comp& c, int bx, int by) {
    std::memset(coef, 0, sizeof coef);
    const Huff& hd = j.dc[c.td];
    const Huff& ha = j.ac[c.ta];
    if (!hd.present || !ha.present) fail(BH_ERR_INVALID_ARG, "missing Huffman table");
    const uint16_t* q = j.q[c.tq];
    const int t = decode_huff(b, hd);
    if (t > 11) fail(BH_ERR_INVALID_ARG, "bad DC magnitude");
    c.dc += extend(b.bits(t), t);
    coef[0] = c.dc * q[0];
    for (int k = 1; k < 64;) {
      const int rs = decode_huff(b, ha);
      const int r = rs >> 4, s = rs & 15;
      if (s == 0) {
        if (r != 15) break;
        k += 16;
        continue;
      }
      k += r;
      if (k > 63) fail(BH_ERR_INVALID_ARG, "AC run past the block");
      coef[kZigzag[k]] = extend(b.bits(s), s) * q[k];
      ++k;
    }
    if (bx < c.bw && by < c.bh)
      idct_block(coef, c.px.data() + (size_t)by * 8 * c.bw * 8 + bx * 8, c.bw * 8);
  };
  int todo = j.restart;
  auto restart = [&]() {
    if (!j.restart) return;
    if (todo-- > 0) return;
    // expect RSTn: realign, skip the marker, reset predictors
    b.align();
    if (b.marker && b.p + 1 < end && b.p[1] >= 0xD0 && b.p[1] <= 0xD7) {
      b.p += 2;
      b.marker = false;
    } else {
      // search forward for the marker (tolerate garbage before it)
      while (b.p + 1 < end && !(b.p[0] == 0xFF && b.p[1] >= 0xD0 && b.p[1] <= 0xD7)) ++b.p;
      if (b.p + 1 < end) b.p += 2;
      b.marker = false;
    }
    for (int k = 0; k < ns; ++k) j.c[ci[k]].dc = 0;
    todo = j.restart - 1;
  };
  if (ns == 1) {
    Comp& c = j.c[ci[0]];
    const int nbx = (c.dw + 7) / 8, nby = (c.dh + 7) / 8;
    for (int by = 0; by < nby; ++by)
      for (int bx = 0; bx < nbx; ++bx) {
        restart();
        block(c, bx, by);
      }
  } else {
    for (int my = 0; my < j.mcuy; ++my)
      for (int mx = 0; mx < j.mcux; ++mx) {
        restart();
        for (int k = 0; k < ns; ++k) {
          Comp& c = j.c[ci[k]];
          for (int v = 0; v < c.v; ++v)
            for (int h = 0; h < c.h; ++h) block(c, mx * c.h + h, my * c.v + v);
        }
      }
  }
  // resume parsing at the next marker
  const uint8_t* q = b.p;
  while (q + 1 < end && !(q[0] == 0xFF && q[1] != 0x00 && !(q[1] >= 0xD0 && q[1] <= 0xD7)))
    ++q;
  *after = q;
}

void decode(const uint8_t* data, size_t size, Jpeg& j, bool header_only) {
  if (size < 4 || data[0] != 0xFF || data[1] != 0xD8) fail(BH_ERR_INVALID_ARG, "not a JPEG (SOI)");
  const uint8_t* p = data + 2;
  const uint8_t* end = data + size;
  bool scanned = false;
  while (p < end) {
    if (p[0] != 0xFF) fail(BH_ERR_INVALID_ARG, "marker expected");
    while (p < end && p[0] == 0xFF) ++p;  // fill bytes
    if (p >= end) break;
    const int m = *p++;
    if (m == 0xD9) break;  // EOI
    if (m >= 0xD0 && m <= 0xD7) continue;
    if (p + 2 > end) fail(BH_ERR_INVALID_ARG, "truncated segment");
    const int len = be16(p);
    if (len < 2 || p + len > end) fail(BH_ERR_INVALID_ARG, "truncated segment");
    const uint8_t* s = p + 2;
    const uint8_t* se = p + len;
    switch (m) {
      case 0xDB:  // DQT
        while (s < se) {
          const int pq = s[0] >> 4, tq = s[0] & 15;
          if (tq > 3 || pq > 1) fail(BH_ERR_INVALID_ARG, "bad DQT");
          ++s;
          for (int k = 0; k < 64; ++k) {
            if (s + (pq ? 2 : 1) > se) fail(BH_ERR_INVALID_ARG, "truncated DQT");
            j.q[tq][k] = pq ? be16(s) : s[0];
            s += pq ? 2 : 1;
          }
          j.qset[tq] = true;
        }
        break;
      case 0xC4:  // DHT
        while (s < se) {
          if (s + 17 > se) fail(BH_ERR_INVALID_ARG, "truncated DHT");
          const int tc = s[0] >> 4, th = s[0] & 15;
          if (tc > 1 || th > 3) fail(BH_ERR_INVALID_ARG, "bad DHT");
          int nv = 0;
          for (int k = 0; k < 16; ++k) nv += s[1 + k];
          if (nv > 256 || s + 17 + nv > se) fail(BH_ERR_INVALID_ARG, "bad DHT");
          build_huff(tc ? j.ac[th] : j.dc[th], s + 1, s + 17, nv);
          s += 17 + nv;
        }
        break;
      case 0xDD:  // DRI
        if (len != 4) fail(BH_ERR_INVALID_ARG, "bad DRI");
        j.restart = be16(s);
        break;
      case 0xC0:
      case 0xC1: {  // SOF0 / SOF1
        if (j.frame) fail(BH_ERR_INVALID_ARG, "second frame header");
        if (len < 8 || s[0] != 8) fail(BH_ERR_UNSUPPORTED, "only 8-bit JPEG is read");
        j.H = be16(s + 1);
        j.W = be16(s + 3);
        j.n = s[5];
        if (j.W == 0 || j.H == 0) fail(BH_ERR_UNSUPPORTED, "JPEG height from DNL is not read");
        if ((uint64_t)j.W * j.H > ((uint64_t)1 << 28)) fail(BH_ERR_UNSUPPORTED, "image too large");
        if (j.n != 1 && j.n != 3) fail(BH_ERR_UNSUPPORTED, "only 1- or 3-component JPEG is read");
        if (len != 8 + 3 * j.n) fail(BH_ERR_INVALID_ARG, "bad SOF");
        j.hmax = j.vmax = 1;
        for (int k = 0; k < j.n; ++k) {
          Comp& c = j.c[k];
          c.id = s[6 + 3 * k];
          c.h = s[7 + 3 * k] >> 4;
          c.v = s[7 + 3 * k] & 15;
          c.tq = s[8 + 3 * k];
          if (c.h < 1 || c.h > 4 || c.v < 1 || c.v > 4 || c.tq > 3)
            fail(BH_ERR_INVALID_ARG, "bad component");
          j.hmax = c.h > j.hmax ? c.h : j.hmax;
          j.vmax = c.v > j.vmax ? c.v : j.vmax;
        }
        j.mcux = (j.W + 8 * j.hmax - 1) / (8 * j.hmax);
        j.mcuy = (j.H + 8 * j.vmax - 1) / (8 * j.vmax);
        for (int k = 0; k < j.n; ++k) {
          Comp& c = j.c[k];
          c.dw = (j.W * c.h + j.hmax - 1) / j.hmax;
          c.dh = (j.H * c.v + j.vmax - 1) / j.vmax;
          c.bw = j.mcux * c.h;
          c.bh = j.mcuy * c.v;
        }
        j.frame = true;
        if (header_only) return;
        for (int k = 0; k < j.n; ++k) j.c[k].px.assign((size_t)j.c[k].bw * 8 * j.c[k].bh * 8, 0);
        break;
      }
      case 0xC2:
      case 0xC3:
      case 0xC5:
      case 0xC6:
      case 0xC7:
      case 0xC9:
      case 0xCA:
      case 0xCB:
      case 0xCD:
      case 0xCE:
      case 0xCF:
        fail(BH_ERR_UNSUPPORTED, "only baseline / extended sequential Huffman JPEG is read");
      case 0xDA: {  // SOS
        if (!j.frame) fail(BH_ERR_INVALID_ARG, "scan before frame header");
        const int ns = s[0];
        if (ns < 1 || ns > j.n || len != 6 + 2 * ns) fail(BH_ERR_INVALID_ARG, "bad SOS");
        int ci[4];
        for (int k = 0; k < ns; ++k) {
          const int id = s[1 + 2 * k];
          int found = -1;
          for (int c = 0; c < j.n; ++c)
            if (j.c[c].id == id) found = c;
          if (found < 0) fail(BH_ERR_INVALID_ARG, "scan names an unknown component");
          ci[k] = found;
          j.c[found].td = s[2 + 2 * k] >> 4;
          j.c[found].ta = s[2 + 2 * k] & 15;
          if (j.c[found].td > 3 || j.c[found].ta > 3) fail(BH_ERR_INVALID_ARG, "bad table id");
          if (!j.qset[j.c[found].tq]) fail(BH_ERR_INVALID_ARG, "missing quantization table");
        }
        const uint8_t* after = nullptr;
        decode_scan(j, se, end, ns, ci, &after);
        scanned = true;
        p = after;
        continue;
      }
      case 0xE0:  // APP0: JFIF
        if (len >= 7 && std::memcmp(s, "JFIF\0", 5) == 0) j.jfif = true;
        break;
      case 0xEE:  // APP14: Adobe
        if (len >= 14 && std::memcmp(s, "Adobe", 5) == 0) {
          j.adobe = true;
          j.adobe_transform = s[11];
        }
        break;
      default:
        break;  // other APPn, COM, DNL...: skipped
    }
    p += len;
  }
  if (!j.frame) fail(BH_ERR_INVALID_ARG, "no frame header");
  if (!header_only && !scanned) fail(BH_ERR_INVALID_ARG, "no scan");
}

thread_local std::string t_jerr;

}  // namespace

extern "C" {

const char* bh_jpeg_last_error(void) { return t_jerr.c_str(); }

int bh_jpeg_decode(const uint8_t* data, size_t size, uint8_t* dst, size_t dst_size, int* rows,
                   int* cols) {
  if (!data || !rows || !cols) {
    t_jerr = "data/rows/cols are NULL";
    return BH_ERR_INVALID_ARG;
  }
  try {
    Jpeg j;
    decode(data, size, j, dst == nullptr);
    *rows = j.H;
    *cols = j.W;
    if (!dst) return BH_OK;
    if (dst_size < (size_t)j.W * j.H * 3) fail(BH_ERR_INVALID_ARG, "dst is too small");
    std::vector<uint8_t> pl[3];
    for (int k = 0; k < j.n; ++k) pl[k] = upsample(j, j.c[k]);
    color_convert(j, pl, dst);
    return BH_OK;
  } catch (const Fail& f) {
    t_jerr = f.msg;
    return f.code;
  } catch (const std::bad_alloc&) {
    t_jerr = "out of host memory";
    return BH_ERR_OUT_OF_MEMORY;
  }
}

}  // extern "C"

// ===========================================================================
// Encoder: baseline JPEG (4:2:0, JFIF) for frame output and MJPG video.
// ===========================================================================
namespace {

// Quantization tables of ITU-T T.81 Annex K.1 (natural order), scaled by
// quality the usual way: s = q < 50 ? 5000 / q : 200 - 2 q, t = (b s + 50) / 100.
const uint8_t kLumQ[64] = {16, 11, 10, 16, 24,  40,  51,  61,  12, 12, 14, 19, 26,  58,  60,  55,
                           14, 13, 16, 24, 40,  57,  69,  56,  14, 17, 22, 29, 51,  87,  80,  62,
                           18, 22, 37, 56, 68,  109, 103, 77,  24, 35, 55, 64, 81,  104, 113, 92,
                           49, 64, 78, 87, 103, 121, 120, 101, 72, 92, 95, 98, 112, 100, 103, 99};
const uint8_t kChrQ[64] = {17, 18, 24, 47, 99, 99, 99, 99, 18, 21, 26, 66, 99, 99, 99, 99,
                           24, 26, 56, 99, 99, 99, 99, 99, 47, 66, 99, 99, 99, 99, 99, 99,
                           99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99,
                           99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99};

// Entropy-coded data writer: bits MSB first, FF followed by a stuffed 00.
// Whole 32-bit words go out at once unless one of their bytes is FF; `out`
// grows ahead of the write position and is cut to length by flush().
struct BitWriter {
  std::vector<uint8_t>& out;
  size_t len;        // bytes written
  uint64_t acc = 0;  // the low n bits are pending, n < 32 between calls
  int n = 0;
  explicit BitWriter(std::vector<uint8_t>& o) : out(o), len(o.size()) {}
  void room() {
    if (len + 8 > out.size()) out.resize(out.size() * 2 + 4096);
  }
  void byte(uint8_t b) {
    out[len++] = b;
    if (b == 0xFF) out[len++] = 0x00;
  }
  void put(uint32_t bits, int nb) {  // nb <= 32, bits < 2^nb
    acc = (acc << nb) | bits;
    n += nb;
    if (n >= 32) {
      n -= 32;
      const uint32_t w = (uint32_t)(acc >> n);
      room();
      if (((~w - 0x01010101u) & w & 0x80808080u) == 0) {  // no FF byte
        const uint32_t be = __builtin_bswap32(w);
        std::memcpy(out.data() + len, &be, 4);
        len += 4;
      } else {
        for (int k = 24; k >= 0; k -= 8) byte((uint8_t)(w >> k));
      }
    }
  }
  void align() {  // whole bytes out, the last padded with 1 bits
    const int pad = (8 - (n & 7)) & 7;
    acc = (acc << pad) | ((1u << pad) - 1);
    n += pad;
    room();
    while (n > 0) {
      n -= 8;
      byte((uint8_t)(acc >> n));
    }
  }
  void marker(uint8_t m) {  // after align()
    room();
    out[len++] = 0xFF;
    out[len++] = m;
  }
};

// Code lengths (<= 16) of a Huffman code for the symbol frequencies, with one
// reserved all-ones code point, by merging least-frequent pairs and then
// shortening codes longer than 16 bits (T.81 Annex K.2); returns BITS[1..16]
// and HUFFVAL in code order.
void optimal_table(const uint32_t* freq_in, uint8_t* bits, std::vector<uint8_t>& vals) {
  std::vector<int64_t> freq(257, 0);
  for (int i = 0; i < 256; ++i) freq[i] = freq_in[i];
  freq[256] = 1;  // reserved: no code may be all ones
  std::vector<int> codesize(257, 0), others(257, -1);
  for (;;) {
    int c1 = -1, c2 = -1;
    for (int i = 0; i <= 256; ++i) {  // least frequency, ties to the larger index
      if (freq[i] == 0) continue;
      if (c1 < 0 || freq[i] <= freq[c1]) c1 = i;
    }
    for (int i = 0; i <= 256; ++i) {
      if (freq[i] == 0 || i == c1) continue;
      if (c2 < 0 || freq[i] <= freq[c2]) c2 = i;
    }
    if (c2 < 0) break;
    freq[c1] += freq[c2];
    freq[c2] = 0;
    ++codesize[c1];
    while (others[c1] >= 0) {
      c1 = others[c1];
      ++codesize[c1];
    }
    others[c1] = c2;
    ++codesize[c2];
    while (others[c2] >= 0) {
      c2 = others[c2];
      ++codesize[c2];
    }
  }
  int count[33] = {0};
  for (int i = 0; i <= 256; ++i)
    if (codesize[i]) ++count[codesize[i] > 32 ? 32 : codesize[i]];
  for (int i = 32; i > 16; --i) {  // limit to 16 bits
    while (count[i] > 0) {
      int j = i - 2;
      while (count[j] == 0) --j;
      count[i] -= 2;
      count[i - 1] += 1;
      count[j + 1] += 2;
      count[j] -= 1;
    }
  }
  int i = 16;
  while (count[i] == 0) --i;
  count[i] -= 1;  // drop the reserved code
  for (int l = 1; l <= 16; ++l) bits[l - 1] = (uint8_t)count[l];
  vals.clear();
  for (int l = 1; l <= 32; ++l)
    for (int s = 0; s < 256; ++s)
      if (codesize[s] == l) vals.push_back((uint8_t)s);
}

struct Code {
  uint16_t code[256];
  uint8_t len[256];
};

void make_codes(const uint8_t* bits, const std::vector<uint8_t>& vals, Code& c) {
  std::memset(&c, 0, sizeof c);
  int k = 0;
  uint32_t code = 0;
  for (int l = 1; l <= 16; ++l) {
    for (int i = 0; i < bits[l - 1]; ++i, ++k) {
      c.code[vals[(size_t)k]] = (uint16_t)code;
      c.len[vals[(size_t)k]] = (uint8_t)l;
      ++code;
    }
    code <<= 1;
  }
}

inline int nbits(int v) {  // magnitude category: bits of |v|
  const unsigned a = (unsigned)(v < 0 ? -v : v);
  return a ? 32 - __builtin_clz(a) : 0;
}

// One 8-point forward DCT pass by the Arai-Agui-Nakajima factorisation
// (Arai, Agui, Nakajima, Trans. IEICE E71(11), 1988; 5 multiplies): d[k *
// stride] becomes 8 s(k) F(k) with the T.81 normalisation of F and
// s(0) = 1, s(k) = sqrt 2 cos(k pi / 16) for k > 0.
inline void aan_1d(float* d, int st) {
  const float t0 = d[0] + d[7 * st], t7 = d[0] - d[7 * st];
  const float t1 = d[st] + d[6 * st], t6 = d[st] - d[6 * st];
  const float t2 = d[2 * st] + d[5 * st], t5 = d[2 * st] - d[5 * st];
  const float t3 = d[3 * st] + d[4 * st], t4 = d[3 * st] - d[4 * st];
  // even part
  const float e10 = t0 + t3, e13 = t0 - t3, e11 = t1 + t2, e12 = t1 - t2;
  d[0] = e10 + e11;
  d[4 * st] = e10 - e11;
  const float z1 = (e12 + e13) * 0.707106781f;  // cos(pi / 4)
  d[2 * st] = e13 + z1;
  d[6 * st] = e13 - z1;
  // odd part
  const float o10 = t4 + t5, o11 = t5 + t6, o12 = t6 + t7;
  const float z5 = (o10 - o12) * 0.382683433f;  // cos(3 pi / 8)
  const float z2 = 0.541196100f * o10 + z5;     // cos(pi / 8) - cos(3 pi / 8)
  const float z4 = 1.306562965f * o12 + z5;     // cos(pi / 8) + cos(3 pi / 8)
  const float z3 = o11 * 0.707106781f;
  const float z11 = t7 + z3, z13 = t7 - z3;
  d[5 * st] = z13 + z2;
  d[3 * st] = z13 - z2;
  d[st] = z11 + z4;
  d[7 * st] = z11 - z4;
}

// 2-D DCT of a level-shifted block (rows, then columns), quantized by
// `div` = 1 / (8 s(u) s(v) Q(v, u)), rounded to nearest, natural order
// (row = vertical frequency), clamped to baseline's 10-bit AC categories.
inline void fdct_quant(float* blk, const float* div, int16_t* out) {
  for (int y = 0; y < 8; ++y) aan_1d(blk + 8 * y, 1);
  for (int x = 0; x < 8; ++x) aan_1d(blk + x, 8);
  for (int k = 0; k < 64; ++k) {
    const float c = std::nearbyint(blk[k] * div[k]);
    out[k] = (int16_t)(c < -1023.f ? -1023 : (c > 1023.f ? 1023 : (int)c));
  }
}

// Transforms MCU rows [j0, j1) of the image into blocks Y0 Y1 Y2 Y3 Cb Cr per
// 16x16 MCU (edge pixels replicated past the image); chroma is the mean of
// each 2x2 pixel group.
void mcu_rows(const uint8_t* bgr, int H, int W, int mx, int j0, int j1, const float (*div)[64],
              int16_t* coefs) {
  float y[4][64], cb[64], cr[64];
  float rc[16][16], bc[16][16];
  for (int j = j0; j < j1; ++j)
    for (int i = 0; i < mx; ++i) {
      for (int r = 0; r < 16; ++r) {
        const int sy = j * 16 + r < H ? j * 16 + r : H - 1;
        const uint8_t* row = bgr + (size_t)sy * W * 3;
        for (int c = 0; c < 16; ++c) {
          const int sx = i * 16 + c < W ? i * 16 + c : W - 1;
          const float b = row[3 * sx], g = row[3 * sx + 1], rr = row[3 * sx + 2];
          y[(r >> 3) * 2 + (c >> 3)][(r & 7) * 8 + (c & 7)] =
              0.299f * rr + 0.587f * g + 0.114f * b - 128.f;
          bc[r][c] = -0.168736f * rr - 0.331264f * g + 0.5f * b;
          rc[r][c] = 0.5f * rr - 0.418688f * g - 0.081312f * b;
        }
      }
      for (int r = 0; r < 8; ++r)
        for (int c = 0; c < 8; ++c) {
          cb[r * 8 + c] = 0.25f * (bc[2 * r][2 * c] + bc[2 * r][2 * c + 1] +
                                   bc[2 * r + 1][2 * c] + bc[2 * r + 1][2 * c + 1]);
          cr[r * 8 + c] = 0.25f * (rc[2 * r][2 * c] + rc[2 * r][2 * c + 1] +
                                   rc[2 * r + 1][2 * c] + rc[2 * r + 1][2 * c + 1]);
        }
      int16_t* o = coefs + ((size_t)j * mx + i) * 6 * 64;
      for (int b = 0; b < 4; ++b) fdct_quant(y[b], div[0], o + b * 64);
      fdct_quant(cb, div[1], o + 4 * 64);
      fdct_quant(cr, div[1], o + 5 * 64);
    }
}

}  // namespace

namespace {

// Calls emit(table, symbol, value) for every Huffman symbol of MCU row j
// (table 0/1: luminance DC/AC, 2/3: chrominance DC/AC), DC predictors
// starting from zero as they do after a restart marker.
template <class F>
void walk_row(const int16_t* coefs, int mx, int j, F&& emit) {
  int pred[3] = {0, 0, 0};
  const int16_t* blk = coefs + (size_t)j * mx * 6 * 64;
  for (int b = 0; b < mx * 6; ++b, blk += 64) {
    const int comp = b % 6 < 4 ? 0 : b % 6 - 3;
    const int t = comp ? 2 : 0;
    const int diff = blk[0] - pred[comp];
    pred[comp] = blk[0];
    emit(t, nbits(diff), diff);
    int run = 0;
    for (int k = 1; k < 64; ++k) {
      const int v = blk[kZigzag[k]];
      if (v == 0) {
        ++run;
        continue;
      }
      for (; run > 15; run -= 16) emit(t + 1, 0xF0, 0);
      emit(t + 1, (run << 4) | nbits(v), v);
      run = 0;
    }
    if (run) emit(t + 1, 0x00, 0);
  }
}

// Runs fn(part, j0, j1) over `parts` contiguous ranges of [0, n) on host
// threads (range 0 on the calling thread; inline where no thread starts).
// An exception from any part is rethrown here once every thread has joined.
template <class F>
void parallel_ranges(int parts, int n, F&& fn) {
  std::vector<std::exception_ptr> err((size_t)parts);
  auto run = [&](int t, int j0, int j1) {
    try {
      fn(t, j0, j1);
    } catch (...) {
      err[(size_t)t] = std::current_exception();
    }
  };
  std::vector<std::thread> th;
  for (int t = 1; t < parts; ++t) {
    const int j0 = (int)((int64_t)n * t / parts), j1 = (int)((int64_t)n * (t + 1) / parts);
    try {
      th.emplace_back(run, t, j0, j1);
    } catch (const std::system_error&) {
      run(t, j0, j1);
    }
  }
  run(0, 0, (int)((int64_t)n / parts));
  for (auto& t : th) t.join();
  for (const auto& e : err)
    if (e) std::rethrow_exception(e);
}

// Encode rows x cols x 3 B,G,R bytes; appends the file to `out`.  Each MCU
// row is one restart interval, so host threads code their rows on their own
// and the segments are joined with RST0..RST7 markers between them.
void encode_jpeg(const uint8_t* bgr, int H, int W, int quality, std::vector<uint8_t>& out) {
  quality = quality < 1 ? 1 : (quality > 100 ? 100 : quality);
  const int scale = quality < 50 ? 5000 / quality : 200 - 2 * quality;
  uint16_t q[2][64];
  float div[2][64];
  for (int k = 0; k < 64; ++k) {
    int a = (kLumQ[k] * scale + 50) / 100, b = (kChrQ[k] * scale + 50) / 100;
    q[0][k] = (uint16_t)(a < 1 ? 1 : (a > 255 ? 255 : a));
    q[1][k] = (uint16_t)(b < 1 ? 1 : (b > 255 ? 255 : b));
  }
  auto aan_scale = [](int k) {
    return k ? std::sqrt(2.0) * std::cos(k * 3.14159265358979323846 / 16) : 1.0;
  };
  for (int v = 0; v < 8; ++v)
    for (int u = 0; u < 8; ++u) {
      const double f = 8.0 * aan_scale(u) * aan_scale(v);
      div[0][v * 8 + u] = (float)(1.0 / (f * q[0][v * 8 + u]));
      div[1][v * 8 + u] = (float)(1.0 / (f * q[1][v * 8 + u]));
    }
  const int mx = (W + 15) / 16, my = (H + 15) / 16;
  std::vector<int16_t> coefs((size_t)mx * my * 6 * 64);
  const unsigned hw = std::thread::hardware_concurrency();
  const int nt = std::max(1, std::min({(int)(hw ? hw : 1), 8, my / 2}));
  // blocks and symbol counts per thread
  std::vector<std::array<std::array<uint32_t, 256>, 4>> freq((size_t)nt);
  parallel_ranges(nt, my, [&](int t, int j0, int j1) {
    mcu_rows(bgr, H, W, mx, j0, j1, div, coefs.data());
    auto& f = freq[(size_t)t];
    for (auto& a : f) a.fill(0);
    for (int j = j0; j < j1; ++j) walk_row(coefs.data(), mx, j, [&](int k, int sym, int) { ++f[k][sym]; });
  });
  uint32_t tot[4][256] = {};
  for (const auto& f : freq)
    for (int k = 0; k < 4; ++k)
      for (int s = 0; s < 256; ++s) tot[k][s] += f[k][s];
  uint8_t bits[4][16];
  std::vector<uint8_t> vals[4];
  Code codes[4];
  for (int k = 0; k < 4; ++k) {
    optimal_table(tot[k], bits[k], vals[k]);
    make_codes(bits[k], vals[k], codes[k]);
  }
  // entropy-coded segments, one buffer per thread
  std::vector<std::vector<uint8_t>> seg((size_t)nt);
  parallel_ranges(nt, my, [&](int t, int j0, int j1) {
    auto& o = seg[(size_t)t];
    o.resize((size_t)(j1 - j0) * mx * 6 * 24 + 4096);
    BitWriter w(o);
    w.len = 0;
    for (int j = j0; j < j1; ++j) {
      walk_row(coefs.data(), mx, j, [&](int k, int sym, int v) {
        const Code& c = codes[k];
        const int nx = (k & 1) ? (sym & 15) : sym;  // code, then nx extra bits (<= 16 + 11)
        const uint32_t extra = (uint32_t)(v < 0 ? v - 1 : v) & ((1u << nx) - 1);
        w.put(((uint32_t)c.code[sym] << nx) | extra, c.len[sym] + nx);
      });
      w.align();
      if (j + 1 < my) w.marker((uint8_t)(0xD0 + (j & 7)));
    }
    o.resize(w.len);
  });
  // headers
  auto u8 = [&](int v) { out.push_back((uint8_t)v); };
  auto u16 = [&](int v) {
    u8(v >> 8);
    u8(v & 0xFF);
  };
  u16(0xFFD8);
  u16(0xFFE0);  // JFIF 1.01, no units, 1:1, no thumbnail
  u16(16);
  for (char ch : {'J', 'F', 'I', 'F', '\0'}) u8(ch);
  u8(1);
  u8(1);
  u8(0);
  u16(1);
  u16(1);
  u8(0);
  u8(0);
  for (int t = 0; t < 2; ++t) {
    u16(0xFFDB);
    u16(67);
    u8(t);
    for (int k = 0; k < 64; ++k) u8(q[t][kZigzag[k]]);
  }
  u16(0xFFC0);  // 8-bit, 3 components: Y 2x2 (table 0), Cb and Cr 1x1 (table 1)
  u16(17);
  u8(8);
  u16(H);
  u16(W);
  u8(3);
  u8(1), u8(0x22), u8(0);
  u8(2), u8(0x11), u8(1);
  u8(3), u8(0x11), u8(1);
  for (int k = 0; k < 4; ++k) {
    u16(0xFFC4);
    u16(3 + 16 + (int)vals[k].size());
    u8(((k & 1) << 4) | (k >> 1));  // class, id: DC0 AC0 DC1 AC1
    for (int l = 0; l < 16; ++l) u8(bits[k][l]);
    for (uint8_t v : vals[k]) u8(v);
  }
  u16(0xFFDD);  // restart interval: one MCU row
  u16(4);
  u16(mx);
  u16(0xFFDA);
  u16(12);
  u8(3);
  u8(1), u8(0x00);
  u8(2), u8(0x11);
  u8(3), u8(0x11);
  u8(0);
  u8(63);
  u8(0);
  for (const auto& o : seg) out.insert(out.end(), o.begin(), o.end());
  u16(0xFFD9);
}

}  // namespace

extern "C" {

int bh_jpeg_encode(const uint8_t* bgr, int rows, int cols, int quality, uint8_t** out_data,
                   size_t* out_size) {
  if (!bgr || rows <= 0 || cols <= 0 || rows > 65535 || cols > 65535 || !out_data || !out_size) {
    t_jerr = "bad image";
    return BH_ERR_INVALID_ARG;
  }
  try {
    std::vector<uint8_t> v;
    encode_jpeg(bgr, rows, cols, quality, v);
    uint8_t* p = (uint8_t*)std::malloc(v.size());
    if (!p) {
      t_jerr = "out of host memory";
      return BH_ERR_OUT_OF_MEMORY;
    }
    std::memcpy(p, v.data(), v.size());
    *out_data = p;
    *out_size = v.size();
    return BH_OK;
  } catch (const std::bad_alloc&) {
    t_jerr = "out of host memory";
    return BH_ERR_OUT_OF_MEMORY;
  } catch (const std::exception& e) {  // e.g. a failed thread join
    t_jerr = e.what();
    return BH_ERR_INVALID_ARG;
  }
}

void bh_free(void* p) { std::free(p); }

}  // extern "C"
