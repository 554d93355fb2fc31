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

// (hi:lo) >> n, low 32 bits, for n in 0..31 (v_alignbit_b32 on the GPU).
BCN_HD uint32_t funnel_shr(uint32_t hi, uint32_t lo, uint32_t n) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_alignbit(hi, lo, n);
#else
  return (uint32_t)((((uint64_t)hi << 32) | lo) >> n);
#endif
}

// LSB-first reader of a 128-bit block that consumes its bits from the bottom:
// a read takes the low n bits (n <= 31) and shifts the four words right by n
// with three funnel shifts.  There is no position-dependent branch, so reads
// whose width differs per lane (mode-dependent fields, anchor indices) cost
// the same as fixed ones, and a read of 0 bits is a no-op.
struct BitReader {
  uint32_t w0, w1, w2, w3;
  BCN_HD BitReader(const uint32_t w[4]) : w0(w[0]), w1(w[1]), w2(w[2]), w3(w[3]) {}
  BCN_HD uint32_t read(uint32_t n) {
    uint32_t v = w0 & ((1u << n) - 1u);
    w0 = funnel_shr(w1, w0, n);
    w1 = funnel_shr(w2, w1, n);
    w2 = funnel_shr(w3, w2, n);
    w3 >>= n;
    return v;
  }
  // Drop n bits (n <= 31).
  BCN_HD void skip(uint32_t n) { (void)read(n); }
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
// The BPTC index weights are round(64 i / m) with m = 2^bits - 1, and for
// m = 3, 7, 15 they equal (i * (4095 / m) + 32) >> 6 at every index (checked
// at compile time against the format's tables, weight_mad_holds in
// bcn_common.hpp): one multiply instead of a per-lane choice among three
// tables when `bits` differs by lane.
BCN_HD int bptc_weight_mul(int bits) { return bits == 2 ? 1365 : (bits == 3 ? 585 : 273); }
BCN_HD int bptc_weight_rt(int bits, int i) { return (mul24(i, bptc_weight_mul(bits)) + 32) >> 6; }


// Two 16-bit lanes of (a * (64 - w) + b * w + 32) >> 6 (block.rs:1256-1258):
// lo/hi/w hold one channel per half-word.  Every term stays below 2^16
// (255 * 64 + 32), so packed 16-bit integer arithmetic is exact.
BCN_HD uint32_t interp2(uint32_t lo, uint32_t hi, uint32_t w) {
#if defined(__HIP_DEVICE_COMPILE__)
  typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));
  const u16x2 a = __builtin_bit_cast(u16x2, lo), b = __builtin_bit_cast(u16x2, hi);
  const u16x2 ww = __builtin_bit_cast(u16x2, w);
  const u16x2 c64 = {64, 64}, c32 = {32, 32}, c6 = {6, 6};
  const u16x2 r = (a * (c64 - ww) + b * ww + c32) >> c6;
  return __builtin_bit_cast(uint32_t, r);
#else
  uint32_t r = 0;
  for (int h = 0; h < 2; h++) {
    uint32_t a = (lo >> (16 * h)) & 0xFFFFu, b = (hi >> (16 * h)) & 0xFFFFu;
    uint32_t ww = (w >> (16 * h)) & 0xFFFFu;
    r |= (((a * (64 - ww) + b * ww + 32) >> 6) & 0xFFFFu) << (16 * h);
  }
  return r;
#endif
}

// Every read below has a per-lane width (0 where the mode lacks the field)
// rather than a per-lane branch, so a wave of mixed modes runs one path.
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
  br.skip((uint32_t)mode + 1);
  int part = (int)br.read(mi.pb);
  int rotation = (int)br.read(mi.rb);
  int isb = (int)br.read(mi.isb);
  const int ne = mi.ns * 2;
  // Endpoint arrays are indexed with compile-time indices only (unrolled
  // loops and selects) so they stay in registers.
  int ep[6][4];
#pragma unroll
  for (int c = 0; c < 3; c++)
#pragma unroll
    for (int e = 0; e < 6; e++) ep[e][c] = (int)br.read(e < ne ? mi.cb : 0u);
#pragma unroll
  for (int e = 0; e < 6; e++) ep[e][3] = (int)br.read(e < ne ? mi.ab : 0u);
  // P-bits: one per endpoint (kind 1) or one per subset shared by its two
  // endpoints' colour channels (kind 2); at most one of the two loops reads.
  int pb[6];
#pragma unroll
  for (int e = 0; e < 6; e++) pb[e] = (int)br.read(mi.pbit == 1 && e < ne ? 1u : 0u);
#pragma unroll
  for (int sb = 0; sb < 2; sb++) {
    int b = (int)br.read(mi.pbit == 2 ? 1u : 0u);
    pb[2 * sb] |= b;
    pb[2 * sb + 1] |= b;
  }
  const int has_p = mi.pbit != 0, alpha_p = mi.pbit == 1;
  const int cbits = mi.cb + has_p, abits = mi.ab + alpha_p;
  // Expand, apply the rotation to the endpoints (output channel rotation-1
  // takes the alpha endpoints and alpha the channel's, block.rs:1228-1240),
  // and pack each subset's endpoints as (R | G << 16), (B | A << 16).
  uint32_t lrg[3], lba[3], hrg[3], hba[3];
#pragma unroll
  for (int sb = 0; sb < 3; sb++) {
    int v[2][4];
#pragma unroll
    for (int h = 0; h < 2; h++) {
      const int e = 2 * sb + h;
      int c[4];
#pragma unroll
      for (int ch = 0; ch < 3; ch++) c[ch] = expand_bits(has_p ? (ep[e][ch] << 1) | pb[e] : ep[e][ch], cbits);
      c[3] = mi.ab ? expand_bits(alpha_p ? (ep[e][3] << 1) | pb[e] : ep[e][3], abits) : 255;
      v[h][0] = rotation == 1 ? c[3] : c[0];
      v[h][1] = rotation == 2 ? c[3] : c[1];
      v[h][2] = rotation == 3 ? c[3] : c[2];
      v[h][3] = rotation == 1 ? c[0] : (rotation == 2 ? c[1] : (rotation == 3 ? c[2] : c[3]));
    }
    lrg[sb] = (uint32_t)v[0][0] | ((uint32_t)v[0][1] << 16);
    lba[sb] = (uint32_t)v[0][2] | ((uint32_t)v[0][3] << 16);
    hrg[sb] = (uint32_t)v[1][0] | ((uint32_t)v[1][1] << 16);
    hba[sb] = (uint32_t)v[1][2] | ((uint32_t)v[1][3] << 16);
  }
  // Half-words interpolated with the alpha-set weight: the output channel
  // that holds the alpha endpoints.
  const uint32_t mrg = (rotation == 1 ? 0x0000FFFFu : 0u) | (rotation == 2 ? 0xFFFF0000u : 0u);
  const uint32_t mba = (rotation == 3 ? 0x0000FFFFu : 0u) | (rotation == 0 ? 0xFFFF0000u : 0u);
  uint32_t pattern = mi.ns == 1 ? 0u : kPart.subset[mi.ns == 3 ? 64 + part : part];
  int a1 = mi.ns >= 2 ? kPart.anchor1[mi.ns == 3 ? 64 + part : part] : -1;
  int a2 = mi.ns == 3 ? kPart.anchor2[64 + part] : -1;
  const int k1 = bptc_weight_mul(mi.ib), k2 = bptc_weight_mul(mi.ib2);
  int w1[16];
#pragma unroll
  for (int k = 0; k < 16; k++) {
    bool anchor = k == 0 || k == a1 || k == a2;
    w1[k] = (mul24((int)br.read(anchor ? mi.ib - 1u : mi.ib), k1) + 32) >> 6;
  }
#pragma unroll
  for (int k = 0; k < 16; k++) {
    int s = (int)((pattern >> (2 * k)) & 3u);
    // Second index set (modes 4 and 5), read after all of the first.
    int i2 = (int)br.read(mi.ib2 == 0 ? 0u : (k == 0 ? mi.ib2 - 1u : mi.ib2));
    int w2 = mi.ib2 == 0 ? w1[k] : (mul24(i2, k2) + 32) >> 6;
    uint32_t wc = (uint32_t)(isb ? w2 : w1[k]), wa = (uint32_t)(isb ? w1[k] : w2);
    wc |= wc << 16;
    wa |= wa << 16;
    uint32_t rg = interp2(s == 0 ? lrg[0] : (s == 1 ? lrg[1] : lrg[2]),
                          s == 0 ? hrg[0] : (s == 1 ? hrg[1] : hrg[2]), (wc & ~mrg) | (wa & mrg));
    uint32_t ba = interp2(s == 0 ? lba[0] : (s == 1 ? lba[1] : lba[2]),
                          s == 0 ? hba[0] : (s == 1 ? hba[1] : hba[2]), (wc & ~mba) | (wa & mba));
    px[k] = (rg & 0xFFu) | ((rg >> 8) & 0xFF00u) | ((ba & 0xFFu) << 16) | ((ba >> 16) << 24);
  }
}

// ---- BC6H (block.rs:262-769) --------------------------------------------------

BCN_HD int extend_sign(int v, int bits) { return (int)((uint32_t)v << (32 - bits)) >> (32 - bits); }

// unquantize (block.rs:1369-1389), written as selects.
template <bool SIGNED>
BCN_HD int bc6h_unquantize(int v, int bits) {
  if (!SIGNED) {
    int u = ((v << 16) + 0x8000) >> bits;
    u = v == (1 << bits) - 1 ? 0xFFFF : u;
    u = v == 0 ? 0 : u;
    return bits >= 15 ? v : u;
  }
  int m = v < 0 ? -v : v;
  int u = ((m << 15) + 0x4000) >> (bits - 1);
  u = m >= (1 << (bits - 1)) - 1 ? 0x7FFF : u;
  u = m == 0 ? 0 : u;
  u = v < 0 ? -u : u;
  return bits >= 16 ? v : u;
}

// finish_unquantize (block.rs:1391-1397): the half bit pattern.  A negative
// value whose magnitude scales to 0 gives +0.
template <bool SIGNED>
BCN_HD uint16_t bc6h_finish(int v) {
  if (!SIGNED) return (uint16_t)((v * 31) >> 6);
  uint32_t mag = (uint32_t)(((v < 0 ? -v : v) * 31) >> 5);
  return (uint16_t)((v < 0 && mag != 0) ? (0x8000u | mag) : mag);
}

// Base-endpoint bits per mode (the decoder's W precision).
BCN_TABLE uint8_t kBc6hBaseBits[14] = {10, 7, 11, 11, 11, 9, 8, 8, 8, 6, 10, 11, 12, 16};

// The low n bits of v in reverse order (n in 1..31).
BCN_HD uint32_t reverse_low_bits(uint32_t v, uint32_t n) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_bitreverse32(v) >> (32u - n);
#else
  uint32_t r = 0;
  for (uint32_t k = 0; k < n; k++) r |= ((v >> k) & 1u) << (n - 1 - k);
  return r;
#endif
}

// RGB half-float bit patterns of the 16 texels: out[k*3 + c].  As in
// decode_bc7, per-lane differences (mode, field widths) are read widths and
// selects, not branches.
template <bool SIGNED>
BCN_HD void decode_bc6h_t(const uint32_t in[4], uint16_t out[48]) {
  BitReader br(in);
  uint32_t mb = br.read(2);
  mb |= br.read(mb > 1 ? 3u : 0u) << 2;
  int m = -1;
#pragma unroll
  for (int i = 0; i < 14; i++) m = (kBc6hPrefix[i] == mb) ? i : m;
  if (m < 0) {
    for (int i = 0; i < 48; i++) out[i] = 0;  // reserved mode: RGB zero
    return;
  }
  // Field value of (endpoint e, channel c) goes to slot e*3 + c; slots are
  // kept as 16-bit halves of six words (every field value is below 2^16
  // before sign extension) and merged through selects, so they stay in
  // registers although the slot differs by lane.  Every mode runs the
  // longest field list: the table's entries past a mode's count are 0, a
  // read of 0 bits into slot 0.
  uint32_t f2[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
  for (int i = 0; i < 23; i++) {
    uint32_t t = kBc6hFields[m][i];
    uint32_t c = t & 3u, e = (t >> 2) & 3u, first = (t >> 4) & 31u, n = (t >> 9) & 63u;
    uint32_t raw = br.read(n);
    // A reversed field stores its bits from first+n-1 down to first.
    uint32_t v = ((t >> 15) ? reverse_low_bits(raw, n) : raw) << first;
    uint32_t slot = e * 3 + c;
    v <<= (slot & 1u) * 16;
#pragma unroll
    for (uint32_t w = 0; w < 6; w++) f2[w] |= w == (slot >> 1) ? v : 0u;
  }
  int f[12];
#pragma unroll
  for (int i = 0; i < 12; i++) f[i] = (int)((f2[i >> 1] >> (16 * (i & 1))) & 0xFFFFu);
  const bool two = m < 10;
  int part = (int)br.read(two ? 5u : 0u);
  const int ne = two ? 4 : 2;
  const int b0 = kBc6hBaseBits[m];
  const bool transformed = m != 9 && m != 10;
  if (SIGNED)
#pragma unroll
    for (int c = 0; c < 3; c++) f[c] = extend_sign(f[c], b0);
#pragma unroll
  for (int e = 1; e < 4; e++)
#pragma unroll
    for (int c = 0; c < 3; c++) {
      int v = f[e * 3 + c];
      if (SIGNED) {
        v = extend_sign(v, kBc6hDeltaBits[c][m]);
      } else {
        v = transformed ? extend_sign(v, kBc6hDeltaBits[c][m]) : v;
      }
      int t = (v + f[c]) & ((1 << b0) - 1);
      v = transformed ? (SIGNED ? extend_sign(t, b0) : t) : v;
      f[e * 3 + c] = e < ne ? v : f[e * 3 + c];
    }
#pragma unroll
  for (int i = 0; i < 12; i++) f[i] = bc6h_unquantize<SIGNED>(f[i], b0);
  uint32_t pattern = two ? kPart.subset[part] : 0u;
  int a1 = two ? kPart.anchor1[part] : -1;
  const int ib = two ? 3 : 4;
  const int kw = bptc_weight_mul(ib);
#pragma unroll
  for (int k = 0; k < 16; k++) {
    bool anchor = k == 0 || k == a1;
    int w = (mul24i((int)br.read(anchor ? ib - 1 : ib), kw) + 32) >> 6;
    int s = (int)((pattern >> (2 * k)) & 1u);
#pragma unroll
    for (int c = 0; c < 3; c++) {
      int lo = s ? f[6 + c] : f[c], hi = s ? f[9 + c] : f[3 + c];
      // (lo*(64-w) + hi*w + 32) >> 6 == lo + ((w*(hi-lo) + 32) >> 6): 64*lo is
      // a multiple of 64 and >> floors; |w*(hi-lo)| <= 64 * 65535 < 2^23.
      out[k * 3 + c] = bc6h_finish<SIGNED>(lo + ((mul24i(w, hi - lo) + 32) >> 6));
    }
  }
}

BCN_HD void decode_bc6h(const uint32_t in[4], bool is_signed, uint16_t out[48]) {
  if (is_signed)
    decode_bc6h_t<true>(in, out);
  else
    decode_bc6h_t<false>(in, out);
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
  // Both branches are computed and selected: no per-lane branch.
  float curve = 1.055f * __builtin_amdgcn_exp2f(__builtin_amdgcn_logf(lin) * (1.0f / 2.4f)) - 0.055f;
  float v = lin <= 0.0031308f ? lin * 12.92f : curve;
  return (int)(fminf(fmaxf(v, 0.0f), 1.0f) * 255.0f + 0.5f);
}

// linear_to_srgb of an unsigned BC6H texel value, exactly: the estimate is
// corrected by one comparison against each neighbouring threshold (both
// thresholds are read and the correction is a select).
__device__ __forceinline__ uint32_t srgb8_code(uint16_t h, const uint16_t* lt) {
  int k = srgb8_estimate(h);
  const uint32_t below = lt[k], above = lt[k < 255 ? k + 1 : 255];
  // Bitwise, not short-circuit, conditions: both reads stay unconditional.
  const int dn = (int)((k > 0) & (h < below)), up = (int)((k < 255) & (h >= above));
  return (uint32_t)(k - dn + (up & (dn ^ 1)));
}
#endif

}  // namespace bcn
