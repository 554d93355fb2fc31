// Block decoders (BC1-BC7), one block per lane, all-integer and bit-exact
// with the reference's decoders (src/decode/block.rs, a port of bcdec v0.98).
//
// BC7 and BC6H are parsed from per-mode field descriptions (the format's
// layouts) rather than one hand-written routine per mode; BC6H reuses the
// encoder's field table (bc6h_block.hpp), so encoder and decoder share one
// statement of the bit layout, and the tests check both against the
// reference's routines.
#pragma once
#include "bc6h_block.hpp"
#include "bcn_common.hpp"

namespace bcn {

struct BitReader {
  uint64_t lo, hi;
  uint32_t pos = 0;
  BCN_HD BitReader(const uint32_t w[4])
      : lo((uint64_t)w[0] | ((uint64_t)w[1] << 32)), hi((uint64_t)w[2] | ((uint64_t)w[3] << 32)) {}
  BCN_HD uint32_t read(uint32_t n) {
    if (n == 0) return 0;
    uint64_t v;
    if (pos >= 64) v = hi >> (pos - 64);
    else if (pos == 0) v = lo;
    else v = (lo >> pos) | (hi << (64 - pos));
    pos += n;
    return (uint32_t)(v & ((1ull << n) - 1));
  }
};

// (a * (64 - w) + b * w + 32) >> 6 (block.rs:1256-1258)
BCN_HD int interp(int a, int b, int w) { return (a * (64 - w) + b * w + 32) >> 6; }

// ---- BC1 colour, BC2/BC3 alpha, BC4/BC5 channel (block.rs:122-258) ----------

// 16 RGBA8 texels of a BC1 colour block; opaque = BC2/BC3's always-4-colour.
BCN_HD void decode_color(uint32_t w0, uint32_t w1, bool opaque, uint32_t px[16]) {
  uint32_t c0 = w0 & 0xFFFFu, c1 = w0 >> 16;
  uint32_t r0 = (c0 >> 11) & 31, g0 = (c0 >> 5) & 63, b0 = c0 & 31;
  uint32_t r1 = (c1 >> 11) & 31, g1 = (c1 >> 5) & 63, b1 = c1 & 31;
  uint32_t ref[4];
  ref[0] = 0xFF000000u | (((b0 * 527 + 23) >> 6) << 16) | (((g0 * 259 + 33) >> 6) << 8) | ((r0 * 527 + 23) >> 6);
  ref[1] = 0xFF000000u | (((b1 * 527 + 23) >> 6) << 16) | (((g1 * 259 + 33) >> 6) << 8) | ((r1 * 527 + 23) >> 6);
  if (c0 > c1 || opaque) {
    ref[2] = 0xFF000000u | ((((2 * b0 + b1) * 351 + 61) >> 7) << 16) |
             ((((2 * g0 + g1) * 2763 + 1039) >> 11) << 8) | (((2 * r0 + r1) * 351 + 61) >> 7);
    ref[3] = 0xFF000000u | ((((b0 + 2 * b1) * 351 + 61) >> 7) << 16) |
             ((((g0 + 2 * g1) * 2763 + 1039) >> 11) << 8) | (((r0 + 2 * r1) * 351 + 61) >> 7);
  } else {
    ref[2] = 0xFF000000u | ((((b0 + b1) * 1053 + 125) >> 8) << 16) |
             ((((g0 + g1) * 4145 + 1019) >> 11) << 8) | (((r0 + r1) * 1053 + 125) >> 8);
    ref[3] = 0;
  }
#pragma unroll
  for (int k = 0; k < 16; k++) {
    uint32_t i = (w1 >> (2 * k)) & 3u;
    px[k] = i == 0 ? ref[0] : (i == 1 ? ref[1] : (i == 2 ? ref[2] : ref[3]));
  }
}

// 16 values of an 8-level alpha / BC4 channel block (block.rs:219-258).
BCN_HD void decode_channel(uint32_t w0, uint32_t w1, uint8_t v[16]) {
  uint32_t a0 = w0 & 0xFFu, a1 = (w0 >> 8) & 0xFFu;
  uint32_t a[8];
  a[0] = a0;
  a[1] = a1;
  if (a0 > a1) {
    for (int i = 1; i < 7; i++) a[1 + i] = ((7 - i) * a0 + i * a1) / 7;
  } else {
    for (int i = 1; i < 5; i++) a[1 + i] = ((5 - i) * a0 + i * a1) / 5;
    a[6] = 0;
    a[7] = 255;
  }
  uint64_t idx = ((uint64_t)w0 >> 16) | ((uint64_t)w1 << 16);
#pragma unroll
  for (int k = 0; k < 16; k++) {
    uint32_t i = (uint32_t)(idx >> (3 * k)) & 7u;
    uint32_t r = a[0];
#pragma unroll
    for (int t = 1; t < 8; t++) r = i == (uint32_t)t ? a[t] : r;
    v[k] = (uint8_t)r;
  }
}

// ---- BC7 (block.rs:801-1252) -------------------------------------------------

// Per mode: subsets, partition bits, rotation bits, index-selector bits,
// colour bits, alpha bits, p-bit kind (0 none, 1 per endpoint, 2 shared per
// subset), index bits, secondary index bits.
struct Bc7ModeInfo {
  uint8_t ns, pb, rb, isb, cb, ab, pbit, ib, ib2;
};
BCN_TABLE Bc7ModeInfo kBc7Modes[8] = {
    {3, 4, 0, 0, 4, 0, 1, 3, 0}, {2, 6, 0, 0, 6, 0, 2, 3, 0}, {3, 6, 0, 0, 5, 0, 0, 2, 0},
    {2, 6, 0, 0, 7, 0, 1, 2, 0}, {1, 0, 2, 1, 5, 6, 0, 2, 3}, {1, 0, 2, 0, 7, 8, 0, 2, 2},
    {1, 0, 0, 0, 7, 7, 1, 4, 0}, {2, 6, 0, 0, 5, 5, 1, 2, 0}};

BCN_HD int expand_bits(int v, int bits) {
  // Left-align to 8 bits and replicate the MSBs (block.rs:1078-1092).
  v <<= 8 - bits;
  return v | (v >> bits);
}

template <int IB>
BCN_HD int bptc_weight(int i) {
  return IB == 2 ? weight2(i) : (IB == 3 ? weight3(i) : weight4(i));
}
BCN_HD int bptc_weight_rt(int bits, int i) {
  return bits == 2 ? weight2(i) : (bits == 3 ? weight3(i) : weight4(i));
}

BCN_HD void decode_bc7(const uint32_t in[4], uint32_t px[16]) {
  uint32_t first = in[0] & 0xFFu;
  if (first == 0) {
#pragma unroll
    for (int k = 0; k < 16; k++) px[k] = 0;  // invalid mode: transparent black
    return;
  }
  int mode = __builtin_ctz(first);
  const Bc7ModeInfo mi = kBc7Modes[mode];
  BitReader br(in);
  br.pos = (uint32_t)mode + 1;
  int part = (int)br.read(mi.pb);
  int rotation = (int)br.read(mi.rb);
  int isb = (int)br.read(mi.isb);
  const int ne = mi.ns * 2;
  // Endpoint arrays are indexed with compile-time indices only (unrolled,
  // guarded loops and selects) so they stay in registers.
  int ep[6][4];
#pragma unroll
  for (int c = 0; c < 3; c++)
#pragma unroll
    for (int e = 0; e < 6; e++) ep[e][c] = e < ne ? (int)br.read(mi.cb) : 0;
#pragma unroll
  for (int e = 0; e < 6; e++) ep[e][3] = (e < ne && mi.ab) ? (int)br.read(mi.ab) : 0;
  int cbits = mi.cb, abits = mi.ab;
  if (mi.pbit == 1) {
#pragma unroll
    for (int e = 0; e < 6; e++) {
      if (e < ne) {
        int b = (int)br.read(1);
#pragma unroll
        for (int c = 0; c < 4; c++) ep[e][c] = (ep[e][c] << 1) | b;
      }
    }
    cbits++;
    abits++;
  } else if (mi.pbit == 2) {
#pragma unroll
    for (int sb = 0; sb < 2; sb++) {
      int b = (int)br.read(1);
#pragma unroll
      for (int c = 0; c < 3; c++) {
        ep[2 * sb][c] = (ep[2 * sb][c] << 1) | b;
        ep[2 * sb + 1][c] = (ep[2 * sb + 1][c] << 1) | b;
      }
    }
    cbits++;
  }
#pragma unroll
  for (int e = 0; e < 6; e++) {
#pragma unroll
    for (int c = 0; c < 3; c++) ep[e][c] = expand_bits(ep[e][c], cbits);
    ep[e][3] = mi.ab ? expand_bits(ep[e][3], abits) : 255;
  }
  uint32_t pattern = mi.ns == 1 ? 0u : kPart.subset[mi.ns == 3 ? 64 + part : part];
  int a1 = mi.ns >= 2 ? kPart.anchor1[mi.ns == 3 ? 64 + part : part] : -1;
  int a2 = mi.ns == 3 ? kPart.anchor2[64 + part] : -1;
  int idx[16];
#pragma unroll
  for (int k = 0; k < 16; k++) {
    bool anchor = k == 0 || k == a1 || k == a2;
    idx[k] = (int)br.read(anchor ? mi.ib - 1 : mi.ib);
  }
#pragma unroll
  for (int k = 0; k < 16; k++) {
    int s = (int)((pattern >> (2 * k)) & 3u);
    int w = bptc_weight_rt(mi.ib, idx[k]);
    int lo[4], hi[4];
#pragma unroll
    for (int ch = 0; ch < 4; ch++) {
      lo[ch] = s == 0 ? ep[0][ch] : (s == 1 ? ep[2][ch] : ep[4][ch]);
      hi[ch] = s == 0 ? ep[1][ch] : (s == 1 ? ep[3][ch] : ep[5][ch]);
    }
    int c[4];
    if (mi.ib2 == 0) {
#pragma unroll
      for (int ch = 0; ch < 4; ch++) c[ch] = interp(lo[ch], hi[ch], w);
    } else {
      int i2 = (int)br.read(k == 0 ? mi.ib2 - 1 : mi.ib2);
      int w2 = bptc_weight_rt(mi.ib2, i2);
      int wc = isb ? w2 : w, wa = isb ? w : w2;
#pragma unroll
      for (int ch = 0; ch < 3; ch++) c[ch] = interp(lo[ch], hi[ch], wc);
      c[3] = interp(lo[3], hi[3], wa);
    }
    int r0 = rotation == 1 ? c[3] : c[0];
    int g0 = rotation == 2 ? c[3] : c[1];
    int b0 = rotation == 3 ? c[3] : c[2];
    int a0 = rotation == 1 ? c[0] : (rotation == 2 ? c[1] : (rotation == 3 ? c[2] : c[3]));
    px[k] = (uint32_t)r0 | ((uint32_t)g0 << 8) | ((uint32_t)b0 << 16) | ((uint32_t)a0 << 24);
  }
}

// ---- BC6H (block.rs:262-769) --------------------------------------------------

BCN_HD int extend_sign(int v, int bits) { return (int)((uint32_t)v << (32 - bits)) >> (32 - bits); }

BCN_HD int bc6h_unquantize(int v, int bits, bool is_signed) {
  if (!is_signed) {
    if (bits >= 15) return v;
    if (v == 0) return 0;
    if (v == (1 << bits) - 1) return 0xFFFF;
    return ((v << 16) + 0x8000) >> bits;
  }
  if (bits >= 16) return v;
  bool s = v < 0;
  int m = s ? -v : v;
  int u;
  if (m == 0) u = 0;
  else if (m >= (1 << (bits - 1)) - 1) u = 0x7FFF;
  else u = ((m << 15) + 0x4000) >> (bits - 1);
  return s ? -u : u;
}

BCN_HD uint16_t bc6h_finish(int v, bool is_signed) {
  if (!is_signed) return (uint16_t)((v * 31) >> 6);
  int sc = v < 0 ? -(((-v) * 31) >> 5) : (v * 31) >> 5;
  return sc < 0 ? (uint16_t)(0x8000 | (uint32_t)(-sc)) : (uint16_t)sc;
}

// Base-endpoint bits per mode (the decoder's W precision).
BCN_TABLE uint8_t kBc6hBaseBits[14] = {10, 7, 11, 11, 11, 9, 8, 8, 8, 6, 10, 11, 12, 16};

// RGB half-float bit patterns of the 16 texels: out[k*3 + c].
BCN_HD void decode_bc6h(const uint32_t in[4], bool is_signed, uint16_t out[48]) {
  BitReader br(in);
  uint32_t mb = br.read(2);
  if (mb > 1) mb |= br.read(3) << 2;
  int m = -1;
#pragma unroll
  for (int i = 0; i < 14; i++) m = (kBc6hPrefix[i] == mb) ? i : m;
  if (m < 0) {
    for (int i = 0; i < 48; i++) out[i] = 0;  // reserved mode: RGB zero
    return;
  }
  // Field value of (endpoint e, channel c) lives in f[e*3 + c]; fields come
  // with runtime (e, c), so they are merged through selects, keeping f in
  // registers.
  int f[12];
#pragma unroll
  for (int i = 0; i < 12; i++) f[i] = 0;
  for (int i = 0; i < kBc6hFieldCount[m]; i++) {
    uint32_t t = kBc6hFields[m][i];
    int c = (int)(t & 3u), e = (int)((t >> 2) & 3u), first = (int)((t >> 4) & 31u);
    int n = (int)((t >> 9) & 63u);
    int v;
    if (t >> 15) {
      v = 0;
      for (int k = 0; k < n; k++) v |= (int)br.read(1) << (first + n - 1 - k);
    } else {
      v = (int)br.read((uint32_t)n) << first;
    }
    int slot = e * 3 + c;
#pragma unroll
    for (int j = 0; j < 12; j++) f[j] |= j == slot ? v : 0;
  }
  const bool two = m < 10;
  int part = two ? (int)br.read(5) : 0;
  const int ne = two ? 4 : 2;
  const int b0 = kBc6hBaseBits[m];
  if (is_signed)
#pragma unroll
    for (int c = 0; c < 3; c++) f[c] = extend_sign(f[c], b0);
  if ((m != 9 && m != 10) || is_signed)
#pragma unroll
    for (int e = 1; e < 4; e++)
#pragma unroll
      for (int c = 0; c < 3; c++)
        if (e < ne) f[e * 3 + c] = extend_sign(f[e * 3 + c], kBc6hDeltaBits[c][m]);
  if (m != 9 && m != 10)
#pragma unroll
    for (int e = 1; e < 4; e++)
#pragma unroll
      for (int c = 0; c < 3; c++) {
        int t = (f[e * 3 + c] + f[c]) & ((1 << b0) - 1);
        if (e < ne) f[e * 3 + c] = is_signed ? extend_sign(t, b0) : t;
      }
#pragma unroll
  for (int i = 0; i < 12; i++) f[i] = bc6h_unquantize(f[i], b0, is_signed);
  uint32_t pattern = two ? kPart.subset[part] : 0u;
  int a1 = two ? kPart.anchor1[part] : -1;
  const int ib = two ? 3 : 4;
#pragma unroll
  for (int k = 0; k < 16; k++) {
    bool anchor = k == 0 || k == a1;
    int i = (int)br.read(anchor ? ib - 1 : ib);
    int s = (int)((pattern >> (2 * k)) & 1u);
    int w = ib == 3 ? weight3(i) : weight4(i);
#pragma unroll
    for (int c = 0; c < 3; c++) {
      int lo = s ? f[6 + c] : f[c], hi = s ? f[9 + c] : f[3 + c];
      out[k * 3 + c] = bc6h_finish(interp(lo, hi, w), is_signed);
    }
  }
}

#if defined(__HIPCC__)
// sRGB8 code thresholds of linear_to_srgb (src/decode.rs:151-159) over the
// non-negative half floats: t[k] is the smallest half bit pattern whose code
// is >= k (t[0] = 0).  Non-negative halves order like their bit patterns,
// so code(h) = #{k >= 1 : h >= t[k]}.
struct Srgb8Table {
  uint16_t t[256];
};

__device__ __forceinline__ float half_to_float(uint16_t h) {
  return (float)__builtin_bit_cast(_Float16, h);
}

// sRGB8 code estimate of a non-negative half from the hardware log2/exp2;
// within one code of linear_to_srgb on every input.
__device__ __forceinline__ int srgb8_estimate(uint16_t h) {
  float lin = half_to_float(h);
  float v = lin <= 0.0031308f
                ? lin * 12.92f
                : 1.055f * __builtin_amdgcn_exp2f(__builtin_amdgcn_logf(lin) * (1.0f / 2.4f)) - 0.055f;
  return (int)(fminf(fmaxf(v, 0.0f), 1.0f) * 255.0f + 0.5f);
}

// linear_to_srgb of an unsigned BC6H texel value, exactly: the estimate is
// corrected by one comparison against each neighbouring threshold.
__device__ __forceinline__ uint32_t srgb8_code(uint16_t h, const uint16_t* lt) {
  int k = srgb8_estimate(h);
  if (k > 0 && h < lt[k])
    k--;
  else if (k < 255 && h >= lt[k + 1])
    k++;
  return (uint32_t)k;
}
#endif

}  // namespace bcn
