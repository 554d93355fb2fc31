// Block decoders (BC1-BC7), one block per lane, all-integer and bit-exact
// with the reference's decoders (src/decode/block.rs, a port of bcdec v0.98).
//
// BC7 and BC6H are parsed from per-mode field descriptions (the format's
// layouts) rather than one hand-written routine per mode; BC6H reuses the
// encoder's field table (bc6h_block.hpp), so encoder and decoder share one
// statement of the bit layout, and the tests check both against the
// reference's routines.
#pragma once
#include <cstddef>

#include "bc6h_block.hpp"
#include "bcn_common.hpp"

namespace bcn {

// (hi:lo) >> n, low 32 bits, for n modulo 32 (v_alignbit_b32 on the GPU).
BCN_HD uint32_t funnel_shr(uint32_t hi, uint32_t lo, uint32_t n) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_alignbit(hi, lo, n);
#else
  return (uint32_t)((((uint64_t)hi << 32) | lo) >> (n & 31u));
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
    n &= 31u;  // the shift instructions' own rule: a caller may pass a word whose low bits are n
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

// Byte 2 of x and of y in bytes 0 and 1 (the rest 0).
BCN_HD uint32_t bytes2(uint32_t x, uint32_t y) { return perm_b32(y, x, 0x0C0C0602u); }

// The 16 values of an 8-level alpha / BC4 channel block (block.rs:219-258)
// as four words of bytes: byte t of v4[j] is texel 4 j + t.
//   * The palette is two words of bytes (levels 0..3 and 4..7), and one
//     v_perm_b32 looks up four texels: its selector bytes are their 3-bit
//     indices.
//   * The interpolated levels ((n - i) a0 + i a1) / n, n = 7 or 5, equal
//     (v * ceil(2^16 / n)) >> 16 for v <= 255 n (the rounding-up error,
//     v * 5 / (7 * 2^16) or v * 4 / (5 * 2^16), stays below 1/n): byte 2 of
//     a 24-bit multiply-add, which the packing perms take directly.
BCN_HD void decode_channel4(uint32_t w0, uint32_t w1, uint32_t v4[4]) {
  const int a0 = (int)(w0 & 0xFFu), a1 = (int)((w0 >> 8) & 0xFFu);
  constexpr int k7 = 9363, k5 = 13108;  // ceil(2^16 / 7), ceil(2^16 / 5)
  auto q7 = [&](int i) { return (uint32_t)(mul24(a0, (7 - i) * k7) + mul24(a1, i * k7)); };
  auto q5 = [&](int i) { return (uint32_t)(mul24(a0, (5 - i) * k5) + mul24(a1, i * k5)); };
  // Levels 0..3 = a0, a1, q(1), q(2); 4..7 = q(3), q(4) and q(5), q(6) or 0, 255.
  const uint32_t lo7 = perm_b32(bytes2(q7(1), q7(2)), w0, 0x05040100u);
  const uint32_t hi7 = perm_b32(bytes2(q7(5), q7(6)), bytes2(q7(3), q7(4)), 0x05040100u);
  const uint32_t lo5 = perm_b32(bytes2(q5(1), q5(2)), w0, 0x05040100u);
  const uint32_t hi5 = bytes2(q5(3), q5(4)) | 0xFF000000u;
  const bool six = a0 > a1;
  const uint32_t lo = six ? lo7 : lo5, hi = six ? hi7 : hi5;
  // 48 index bits from bit 16 of the block; four 3-bit fields (12 bits) per
  // word, spread to one field per byte in two steps (6-bit halves to 16-bit
  // lanes, then 3-bit fields to bytes).
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const uint32_t x = j == 0 ? w0 >> 16 : (j == 1 ? funnel_shr(w1, w0, 28) : w1 >> (12 * j - 16));
    const uint32_t y = bfe_u32(x, 0, 6) | (bfe_u32(x, 6, 6) << 16);
    const uint32_t sel = (y & 0x00070007u) | ((y << 5) & 0x07000700u);
    v4[j] = perm_b32(hi, lo, sel);
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


// Per-mode fields as nibble tables, mode m in bits 4m..4m+3 (kBc7Modes in
// that order): one v_bfe_u32 per field instead of a per-lane table load.
constexpr uint32_t bc7_nibbles(uint8_t Bc7ModeInfo::*f) {
  uint32_t k = 0;
  const Bc7ModeInfo t[8] = {{3, 4, 0, 0, 4, 0, 1, 3, 0}, {2, 6, 0, 0, 6, 0, 2, 3, 0}, {3, 6, 0, 0, 5, 0, 0, 2, 0},
                            {2, 6, 0, 0, 7, 0, 1, 2, 0}, {1, 0, 2, 1, 5, 6, 0, 2, 3}, {1, 0, 2, 0, 7, 8, 0, 2, 2},
                            {1, 0, 0, 0, 7, 7, 1, 4, 0}, {2, 6, 0, 0, 5, 5, 1, 2, 0}};
  for (int m = 0; m < 8; m++) k |= (uint32_t)(t[m].*f) << (4 * m);
  return k;
}

// Per-lane subset data of the BC7 decode: for subsets 0..2, c = 256 a + 128
// and d = 4 (b - a) for endpoints a, b, in 16-bit lanes (R | G << 16 and
// B | A << 16).  Texel k of subset s is then c + w d modulo 2^16 per lane,
// which is 4 ((64 - w) a + w b + 32) < 2^16: the interpolated channel
// (block.rs:1256-1258) is the lane's high byte.
struct alignas(16) Bc7SubsetData {
  uint32_t crg, cba, drg, dba;
};

// The record of subset s for lane l (l < 64: one wave per workgroup) is
// element s * kBc7SubRows + l of the record array: byte address
// s * 1024 + 16 l, whose low byte comes from the lane alone and whose second
// byte is s << 2 | l >> 4.
constexpr uint32_t kBc7SubRows = 64;

// Per-block constants of the BC7 decode, one 64-byte row per form:
//   rows 0-3: modes 0-3; row 8: a first byte of 0 (every constant 0, so the
//   block decodes to transparent black); rows 9-16: mode 4 by rotation and
//   index selector; rows 17-20: mode 5 by rotation; rows 21, 22: modes 6, 7.
// A block's row is mode + kBc7Adj[mode] + (pr & kBc7Vmask[mode]), pr being
// the bits that follow the mode bits (the partition, or the rotation and
// index selector).
struct alignas(16) Bc7DecForm {
  // Endpoint expansion in 16-bit lanes (R, G) and (B, A).  A field v of w
  // bits with p-bit p (t = w + 1 bits) or without (t = w) expands as
  // expand_bits (block.rs:1159-1163) to x (2^8 + 2^(8-t)) >> t for x = 2v + p
  // or v (the two bit groups of the expansion do not overlap, so the OR is a
  // sum): v * mul + p * k, then >> bt.  c_ba adds 255 << 8 to the alpha lane
  // of modes without alpha (alpha 255).
  uint32_t mul_rg, mul_ba, k_rg, k_ba;
  // lens bytes: header bits, colour chunk bits (ne fields), alpha chunk bits,
  // alpha chunk + p-bits.  wid bytes: colour and alpha field bits, index bits
  // of set 1 and set 2.
  uint32_t c_ba, bt_rg, bt_ba, lens;
  // kk: the 4095 / (2^bits - 1) weight multipliers of the colour (low) and
  // alpha (high) lane; ipsel: v_perm selector that places the set 1 and set
  // 2 indices in those lanes (the index selector of mode 4 swaps them);
  // out_sel: selector packing the texel from the lanes' high bytes, with the
  // rotated channel and alpha exchanged (block.rs:1228-1240).
  uint32_t wid, kk, ipsel, out_sel;
  // prow: first row of the block's partition kind in kBc7Dec.part (the
  // decoder takes it from kBc7Prow16; a first byte of 0 then reads mode 0's
  // rows, harmless with every constant 0); s2sh: where index set 2 starts in
  // the block's top 64 bits; ins1, ins2: low words of the insertion masks
  // (insert_zero64) of texel 0's missing index bit in set 1 and set 2 (their
  // high words are all ones).
  uint32_t prow, s2sh, ins1, ins2;
};

// Per partition kind (one subset; two or three subsets at 2 or 3 index bits)
// and partition: the subset of texel 4 q + t in byte t of sub4[q] (times 4,
// its part of the record address's second byte), and the masks ~((1 << P) - 1) (64 bits,
// low word first) that insert the anchors' missing index bits at their
// positions P = (a + 1) ib - 1 of the set 1 stream, in ascending order (0
// inserts nothing; insert_zero64).
struct alignas(16) Bc7DecPart {
  uint32_t sub4[4];
  uint32_t ins[4];
};

struct Bc7DecTables {
  Bc7DecForm form[23];
  Bc7DecPart part[272];  // 1 subset (rows 0-15), 2 subsets ib 2 (16-79), ib 3 (80-143), 3 subsets ib 2 (144-207), ib 3 (208-271)
};

constexpr uint32_t kBc7Pbw = 0x60236664u;  // bits after the mode bits: pb + rb + isb, by mode (nibbles)
constexpr uint32_t kBc7Adj = 0xFFC50000u;  // row - mode
constexpr uint32_t kBc7Vmask = 0x00370000u;  // pr bits that select the row (modes 4, 5)
constexpr uint32_t kBc7Prow16 = 0x1000195Du;  // first partition row / 16, by mode (a first byte of 0 reads mode 0's)

constexpr uint32_t bc7_part_row(int ns, int ib) {
  return ns == 1 ? 0u : (ns == 2 ? (ib == 2 ? 16u : 80u) : (ib == 2 ? 144u : 208u));
}

// ~((1 << p) - 1) as two words; p = 64 gives 0.
constexpr void bc7_ins_mask(int p, uint32_t& lo, uint32_t& hi) {
  const uint64_t m = p >= 64 ? ~0ull : (1ull << p) - 1ull;
  lo = (uint32_t)~m;
  hi = (uint32_t)(~m >> 32);
}

constexpr Bc7DecForm bc7_dec_form(int m, int rot, int isb) {
  constexpr Bc7ModeInfo t[8] = {{3, 4, 0, 0, 4, 0, 1, 3, 0}, {2, 6, 0, 0, 6, 0, 2, 3, 0}, {3, 6, 0, 0, 5, 0, 0, 2, 0},
                                {2, 6, 0, 0, 7, 0, 1, 2, 0}, {1, 0, 2, 1, 5, 6, 0, 2, 3}, {1, 0, 2, 0, 7, 8, 0, 2, 2},
                                {1, 0, 0, 0, 7, 7, 1, 4, 0}, {2, 6, 0, 0, 5, 5, 1, 2, 0}};
  const Bc7ModeInfo& i = t[m];
  Bc7DecForm f{};
  const int ne = 2 * i.ns, hp = i.pbit != 0 ? 1 : 0, ap = i.pbit == 1 ? 1 : 0;
  const int tc = i.cb + hp, ta = i.ab ? i.ab + ap : 8;
  const uint32_t mc = 256u + (1u << (8 - tc)), ma = 256u + (1u << (8 - ta));
  const uint32_t mul_c = mc << hp, k_c = hp ? mc : 0u;
  const uint32_t mul_a = i.ab ? ma << ap : 0u, k_a = (i.ab && ap) ? ma : 0u;
  f.mul_rg = mul_c * 0x10001u;
  f.mul_ba = mul_c | (mul_a << 16);
  f.k_rg = k_c * 0x10001u;
  f.k_ba = k_c | (k_a << 16);
  f.c_ba = i.ab ? 0u : 0xFF000000u;
  f.bt_rg = (uint32_t)tc * 0x10001u;
  f.bt_ba = (uint32_t)tc | ((uint32_t)ta << 16);
  const int np = i.pbit == 1 ? ne : (i.pbit == 2 ? 2 : 0);
  const int h = m + 1 + i.pb + i.rb + i.isb;
  f.lens = (uint32_t)h | ((uint32_t)(ne * i.cb) << 8) | ((uint32_t)(ne * i.ab) << 16) |
           ((uint32_t)(ne * i.ab + np) << 24);
  f.wid = (uint32_t)i.cb | ((uint32_t)i.ab << 8) | ((uint32_t)i.ib << 16) | ((uint32_t)i.ib2 << 24);
  const uint32_t k1 = i.ib == 2 ? 1365u : (i.ib == 3 ? 585u : 273u);
  const uint32_t k2 = i.ib2 == 2 ? 1365u : (i.ib2 == 3 ? 585u : 273u);
  f.kk = i.ib2 == 0 ? k1 * 0x10001u : (isb ? k2 | (k1 << 16) : k1 | (k2 << 16));
  f.ipsel = i.ib2 == 0 ? 0x0C000C00u : (isb ? 0x0C000C04u : 0x0C040C00u);
  f.out_sel = rot == 1 ? 0x01050307u : (rot == 2 ? 0x03050701u : (rot == 3 ? 0x05070301u : 0x07050301u));
  f.prow = bc7_part_row(i.ns, i.ib);
  f.s2sh = i.ib2 ? (uint32_t)(65 - 16 * i.ib2) : 0u;
  f.ins1 = ~((1u << (i.ib - 1)) - 1u);
  f.ins2 = i.ib2 ? ~((1u << (i.ib2 - 1)) - 1u) : 0u;
  return f;
}

constexpr Bc7DecTables make_bc7_dec_tables() {
  constexpr PartitionTables pt = make_partition_tables();
  Bc7DecTables d{};
  for (int m = 0; m < 8; m++) {
    const int base = m + (int)((kBc7Adj >> (4 * m)) & 15u), nv = m == 4 ? 8 : (m == 5 ? 4 : 1);
    for (int v = 0; v < nv; v++) d.form[base + v] = bc7_dec_form(m, m == 4 ? (v & 3) : (m == 5 ? v : 0), m == 4 ? v >> 2 : 0);
  }
  d.form[8] = Bc7DecForm{};  // a first byte of 0
  d.form[8].out_sel = 0x07050301u;
  for (int kind = 0; kind < 4; kind++) {
    const int ns = kind < 2 ? 2 : 3, ib = (kind & 1) ? 3 : 2;
    for (int p = 0; p < 64; p++) {
      Bc7DecPart& r = d.part[16 + 64 * kind + p];
      const int pi = ns == 3 ? 64 + p : p;
      for (int k = 0; k < 16; k++) r.sub4[k >> 2] |= ((kShapes[pi] >> (2 * k)) & 3u) << (8 * (k & 3) + 2);
      int a = (pt.anchor1[pi] + 1) * ib - 1, b = ns == 3 ? (pt.anchor2[pi] + 1) * ib - 1 : 64;
      if (b < a) { const int s = a; a = b; b = s; }
      bc7_ins_mask(a, r.ins[0], r.ins[1]);
      bc7_ins_mask(b, r.ins[2], r.ins[3]);
    }
  }
  return d;
}

BCN_TABLE Bc7DecTables kBc7Dec = make_bc7_dec_tables();

constexpr bool bc7_prow16_holds() {
  constexpr Bc7DecTables d = make_bc7_dec_tables();
  for (int m = 0; m < 8; m++)
    if (d.form[m + (int)((kBc7Adj >> (4 * m)) & 15u)].prow != ((kBc7Prow16 >> (4 * m)) & 15u) << 4) return false;
  return true;
}
static_assert(bc7_prow16_holds(), "kBc7Prow16 disagrees with the form rows");

// All ones if bit `off` (modulo 32) of x is set, else 0 (v_bfe_i32).
BCN_HD uint32_t bit_mask(uint32_t x, uint32_t off) {
#if defined(__HIP_DEVICE_COMPILE__)
  return (uint32_t)__builtin_amdgcn_sbfe((int)x, off, 1);
#else
  return 0u - ((x >> (off & 31u)) & 1u);
#endif
}

// x with a 0 bit inserted at bit P, for im = ~((1 << P) - 1): the bits below
// P stay and those from P move up one, (x & ~im) + 2 (x & im) = x + (x & im).
BCN_HD uint64_t insert_zero64(uint64_t x, uint64_t im) { return x + (x & im); }

// decode_block_bc7 (block.rs:801-1252) for one block, every mode in one
// branch-free path.  `sub` is the record array of the workgroup (or of the
// caller, with lane 0): this lane's subset records are written and read back
// at the addresses kBc7SubRows describes.
//   * Per-block constants come from one row of kBc7Dec.form, chosen by the
//     mode and, for modes 4 and 5, the rotation and index selector; the
//     partition's texel subsets and anchor positions from one row of
//     kBc7Dec.part.
//   * The header, then the colour chunks (each channel's ne fields), are
//     consumed by four skips of a 128-bit reader; the alpha chunk, the p-bits
//     and index set 1 follow within the next 96 bits.  Index set 2 (modes 4
//     and 5) ends the block.
//   * Each index set is made uniform before the texel loop: a 0 bit is
//     inserted at every anchor index's missing top bit, so index k sits at bit
//     k * bits (k < 8) of the stream or (k - 8) * bits of the stream shifted by
//     8 * bits -- one v_bfe_u32 per index.
//   * Weights are round(64 i / (2^bits - 1)) = (i * (4095 / m) + 32) >> 6 in
//     16-bit lanes: the colour weight in the low lane, the alpha weight in the
//     high lane.  R and G take the low lane, B the low and A the high lane.
//   * A texel's record address is one v_perm_b32 of the partition row's byte
//     and the lane's base.
//   * Each row of four texels goes to `emit(y, t0, t1, t2, t3)` as soon as it
//     is decoded (the image kernel stores it then).
template <typename RowSink>
BCN_HD void decode_bc7_rows(const uint32_t in[4], Bc7SubsetData* sub, uint32_t lane, RowSink&& emit) {
  const uint32_t mode = (uint32_t)__builtin_ctz(in[0] | 0x100u), m4 = 4u * mode;
  const uint32_t pr = bfe_u32(in[0], mode + 1u, bfe_u32(kBc7Pbw, m4, 4));
  const uint32_t row = mode + bfe_u32(kBc7Adj, m4, 4) + (pr & bfe_u32(kBc7Vmask, m4, 4));
  // Both rows through one 32-bit byte offset from the table's base, each
  // indexed by the mode and the bits after it alone (the partition row's
  // base comes from the mode, not from the form row), so the two loads go
  // out together.
  const uint32_t prow = bfe_u32(kBc7Prow16, m4, 4) << 4;
  const char* tab = reinterpret_cast<const char*>(&kBc7Dec);
  const Bc7DecForm f = *reinterpret_cast<const Bc7DecForm*>(tab + row * (uint32_t)sizeof(Bc7DecForm));
  const Bc7DecPart pt = *reinterpret_cast<const Bc7DecPart*>(
      tab + offsetof(Bc7DecTables, part) + (prow + pr) * (uint32_t)sizeof(Bc7DecPart));
  // Chunks: header, then R, G, B (L bits each), then alpha, p-bits, set 1.
  BitReader br(in);
  br.skip(f.lens);
  const uint32_t L = f.lens >> 8, La = f.lens >> 16;
  const uint32_t chr = br.w0;
  br.skip(L);
  const uint32_t chg = br.w0;
  br.skip(L);
  const uint32_t chb = br.w0;
  br.skip(L);
  const uint32_t cha = br.w0;
  const uint32_t pbits = funnel_shr(br.w1, br.w0, La);
  const uint32_t s1o = f.lens >> 24;
  uint64_t s1 = ((uint64_t)funnel_shr(br.w2, br.w1, s1o) << 32) | funnel_shr(br.w1, br.w0, s1o);
  uint64_t s2 = (((uint64_t)in[3] << 32) | in[2]) >> (f.s2sh & 63u);
  // Texel 0's missing bit in both sets, then the anchors of set 1.
  s1 = insert_zero64(s1, 0xFFFFFFFF00000000ull | f.ins1);
  s1 = insert_zero64(s1, ((uint64_t)pt.ins[1] << 32) | pt.ins[0]);
  s1 = insert_zero64(s1, ((uint64_t)pt.ins[3] << 32) | pt.ins[2]);
  s2 = insert_zero64(s2, 0xFFFFFFFF00000000ull | f.ins2);
  // Endpoints: field e of a channel at bit e * width of its chunk.
  const uint32_t cbx = f.wid, abx = f.wid >> 8, ibx = f.wid >> 16, ib2x = f.wid >> 24;
  // P-bit number of endpoint e: e, or e >> 1 for the shared p-bits of mode 1.
  const uint32_t pks = bfe_u32(2u, mode, 1);
  // Field offsets e * width as running sums (in_vgpr keeps them additions:
  // the compiler would otherwise fold them into slower multiplies and shifts).
  uint32_t oc = 0u, oa = 0u;
  // 256 in both lanes, in a register: with the constant in view the compiler
  // splits each multiply-add into a shift and an OR.
  const uint32_t k256 = in_vgpr(0x01000100u);
  // Subsets 1 and 2 only where some block of the wave has them.
  const bool two = wave_any(prow >= 16u), three = wave_any(prow >= 144u);
#pragma unroll
  for (int sb = 0; sb < 3; sb++) {
    if ((sb == 1 && !two) || (sb == 2 && !three)) continue;
    uint32_t rg[2], ba[2];
#pragma unroll
    for (int h = 0; h < 2; h++) {
      const uint32_t e = 2u * (uint32_t)sb + (uint32_t)h;
      if (e == 1) {
        oc = cbx;
        oa = abx;
      } else if (e > 1) {
        oc = in_vgpr(oc + cbx);
        oa = in_vgpr(oa + abx);
      }
      const uint32_t vrg = bfe_u32(chr, oc, cbx) | (bfe_u32(chg, oc, cbx) << 16);
      const uint32_t vba = bfe_u32(chb, oc, cbx) | (bfe_u32(cha, oa, abx) << 16);
      const uint32_t pm = bit_mask(pbits, e >> pks);
      rg[h] = pk_shr16(pk_mad16(vrg, f.mul_rg, pm & f.k_rg), f.bt_rg);
      ba[h] = pk_shr16(pk_mad16(vba, f.mul_ba, (pm & f.k_ba) | f.c_ba), f.bt_ba);
    }
    Bc7SubsetData d;
    d.crg = pk_mad16(rg[0], k256, 0x00800080u);
    d.cba = pk_mad16(ba[0], k256, 0x00800080u);
    d.drg = pk_shl16(pk_sub16(rg[1], rg[0]), 0x00020002u);
    d.dba = pk_shl16(pk_sub16(ba[1], ba[0]), 0x00020002u);
    sub[(uint32_t)sb * kBc7SubRows + lane] = d;
  }
  // Index k < 8 at bit k * bits of the stream, k >= 8 at (k - 8) * bits of
  // the stream shifted by 8 * bits.
  const uint32_t s1lo = (uint32_t)s1, s1mid = (uint32_t)(s1 >> ((ibx << 3) & 63u));
  const uint32_t s2lo = (uint32_t)s2, s2mid = (uint32_t)(s2 >> ((ib2x << 3) & 63u));
  const uint32_t lanehi = (lane >> 4) * 0x01010101u, laneb = lane << 4;
  const char* base = reinterpret_cast<const char*>(sub);
  uint32_t o1[8], o2[8];
  o1[0] = 0u;
  o2[0] = 0u;
  o1[1] = ibx;
  o2[1] = ib2x;
#pragma unroll
  for (int j = 2; j < 8; j++) {
    o1[j] = in_vgpr(o1[j - 1] + ibx);
    o2[j] = in_vgpr(o2[j - 1] + ib2x);
  }
  // A row at a time: its four record reads go out together.
#pragma unroll
  for (int y = 0; y < 4; y++) {
    const uint32_t row4 = pt.sub4[y] | lanehi;
    Bc7SubsetData d[4];
#pragma unroll
    for (int x = 0; x < 4; x++)
      d[x] = *reinterpret_cast<const Bc7SubsetData*>(base + perm_b32(row4, laneb, 0x0C0C0400u + ((uint32_t)x << 8)));
    uint32_t px[4];
#pragma unroll
    for (int x = 0; x < 4; x++) {
      const int k = 4 * y + x;
      const uint32_t x1 = bfe_u32(k < 8 ? s1lo : s1mid, o1[k & 7], ibx);
      const uint32_t x2 = bfe_u32(k < 8 ? s2lo : s2mid, o2[k & 7], ib2x);
      // Both weights: i * (4095 / m) + 32 <= 4127 per lane, then >> 6.
      const uint32_t wp = pk_mad_shr6(perm_b32(x2, x1, f.ipsel), f.kk, 0x00200020u);
      const uint32_t rg = pk_mad16_lo(d[x].drg, wp, d[x].crg);
      const uint32_t ba = pk_mad16(d[x].dba, wp, d[x].cba);
      px[x] = perm_b32(ba, rg, f.out_sel);
    }
    emit(y, px[0], px[1], px[2], px[3]);
  }
}

// decode_bc7_rows into px[16].
BCN_HD void decode_bc7(const uint32_t in[4], uint32_t px[16], Bc7SubsetData* sub, uint32_t lane) {
  decode_bc7_rows(in, sub, lane, [&](int y, uint32_t a, uint32_t b, uint32_t c, uint32_t e) {
    px[4 * y] = a;
    px[4 * y + 1] = b;
    px[4 * y + 2] = c;
    px[4 * y + 3] = e;
  });
}

// The RGBA8 texels of one block of format FMT (bcn_variant BC1..BC5, BC7) as
// decompress_blocks_as_rgba8 writes them (src/decode.rs:336-387): BC4 ->
// (R, 0, 0, 0), BC5 -> (R, G, 0, 0).  `in` holds the block's 8 or 16 bytes;
// `sub`/`lane` are BC7's subset records (decode_bc7).
template <int FMT>
BCN_HD void decode_rgba8_block(const uint32_t in[4], uint32_t px[16], Bc7SubsetData* sub, uint32_t lane) {
  if (FMT == 0) {  // BC1
    decode_color(in[0], in[1], false, px);
  } else if (FMT == 1 || FMT == 2) {  // BC2, BC3: colour, then alpha in byte 3
    decode_color(in[2], in[3], true, px);
    if (FMT == 1) {
      // decode_sharp_alpha_block (block.rs:200-217): 4-bit alpha x 17
#pragma unroll
      for (int k = 0; k < 16; k++) {
        uint32_t a = ((k < 8 ? in[0] : in[1]) >> (4 * (k & 7))) & 15u;
        px[k] = (px[k] & 0x00FFFFFFu) | ((a * 17u) << 24);
      }
    } else {
      uint32_t a4[4];
      decode_channel4(in[0], in[1], a4);
#pragma unroll
      for (int k = 0; k < 16; k++) px[k] = perm_b32(a4[k >> 2], px[k], 0x00020100u | ((4u + (k & 3)) << 24));
    }
  } else if (FMT == 3) {  // BC4 (decode.rs:97-118)
    uint32_t r4[4];
    decode_channel4(in[0], in[1], r4);
#pragma unroll
    for (int k = 0; k < 16; k++) px[k] = bfe_u32(r4[k >> 2], 8u * (k & 3), 8);
  } else if (FMT == 4) {  // BC5
    uint32_t r4[4], g4[4];
    decode_channel4(in[0], in[1], r4);
    decode_channel4(in[2], in[3], g4);
#pragma unroll
    for (int k = 0; k < 16; k++) px[k] = perm_b32(g4[k >> 2], r4[k >> 2], 0x0C0C0400u | ((k & 3) * 0x101u));
  } else {  // BC7
    decode_bc7(in, px, sub, lane);
  }
}

// ---- BC6H (block.rs:262-769) --------------------------------------------------

// The low `bits` bits of v (1..32) as a signed value (v_bfe_i32 on the GPU).
BCN_HD int extend_sign(int v, int bits) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_sbfe(v, 0, bits);
#else
  return (int)((uint32_t)v << (32 - bits)) >> (32 - bits);
#endif
}

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
  // |v| < 2^17: the products are 24-bit multiplies.
  if (!SIGNED) return (uint16_t)(mul24i(v, 31) >> 6);
  uint32_t mag = (uint32_t)(mul24i(v < 0 ? -v : v, 31) >> 5);
  return (uint16_t)((v < 0 && mag != 0) ? (0x8000u | mag) : mag);
}

// The low n bits of v in reverse order (n in 0..31; 0 gives 0).
BCN_HD uint32_t reverse_low_bits(uint32_t v, uint32_t n) {
#if defined(__HIP_DEVICE_COMPILE__)
  return n ? __builtin_bitreverse32(v) >> (32u - n) : 0u;
#else
  uint32_t r = 0;
  for (uint32_t k = 0; k < n; k++) r |= ((v >> k) & 1u) << (n - 1 - k);
  return r;
#endif
}

// Decoder form of kBc6hFields: width | destination << 5 | reversed << 15,
// where the destination is the field's first bit in the 192-bit value that
// holds the 12 endpoint components as 16-bit slots (slot = endpoint * 3 +
// channel, bit = slot * 16 + first_bit).  Entries past a mode's count are 0:
// a read of no bits ORed into bit 0.  Row 14 is all 0, the row of the
// reserved mode prefixes: no fields, so every component is 0.
constexpr int kBc6hDecRows = 15;
// Rows of 24 entries (48 bytes, the last entry 0), so a lane loads its
// mode's row with three 16-byte reads.
struct alignas(16) Bc6hDecFields {
  uint16_t f[kBc6hDecRows][24];
};
constexpr Bc6hDecFields make_bc6h_dec_fields() {
  Bc6hDecFields t{};
  for (int m = 0; m < 14; m++)
    for (int i = 0; i < 23; i++) {
      const uint32_t e = kBc6hFieldSrc.f[m][i];
      const uint32_t c = e & 3u, ep = (e >> 2) & 3u, first = (e >> 4) & 31u, n = (e >> 9) & 63u;
      const uint32_t dst = (ep * 3 + c) * 16 + first;
      t.f[m][i] = (uint16_t)(n | (dst << 5) | ((e >> 15) << 15));
    }
  return t;
}
BCN_TABLE Bc6hDecFields kBc6hDec = make_bc6h_dec_fields();

// Every field fits its 16-bit slot, and the packed form round-trips.
constexpr bool bc6h_dec_fields_hold() {
  const Bc6hDecFields t = make_bc6h_dec_fields();
  for (int m = 0; m < 14; m++)
    for (int i = 0; i < 23; i++) {
      const uint32_t e = t.f[m][i], n = e & 31u, dst = (e >> 5) & 0xFFu;
      const uint32_t src = kBc6hFieldSrc.f[m][i];
      if (n != ((src >> 9) & 63u) || (dst & 15u) + n > 16) return false;
      if (dst >> 4 != ((src >> 2) & 3u) * 3 + (src & 3u) || (e >> 15) != (src >> 15u)) return false;
    }
  return true;
}
static_assert(bc6h_dec_fields_hold(), "BC6H decoder field table");

// Mode prefix (5 bits, or 2 for prefixes 0 and 1) -> mode index + 1 as a
// nibble table over the 32 five-bit values; 0 marks the reserved prefixes.
constexpr uint32_t bc6h_mode_nibbles(int word) {
  const uint8_t prefix[14] = {0, 1, 2, 6, 10, 14, 18, 22, 26, 30, 3, 7, 11, 15};
  uint32_t k = 0;
  for (int v = 8 * word; v < 8 * word + 8; v++) {
    int m = -1;
    for (int i = 0; i < 14; i++)
      if ((i < 2 ? (v & 3) : v) == prefix[i] && (i < 2 ? (v & 3) < 2 : (v & 3) >= 2)) m = i;
    k |= (uint32_t)(m + 1) << (4 * (v - 8 * word));
  }
  return k;
}

// Base bits (which = 0) or the delta bits of R, G, B (which = 1..3) of the 14
// modes, minus one, as nibbles (mode m at bits 4m of the low word for m < 8,
// 4(m-8) of the high word otherwise), from kBc6hPrecSrc.
constexpr uint64_t bc6h_prec_nibbles(int which) {
  uint64_t k = 0;
  for (int m = 0; m < 14; m++) {
    const int v = which == 0 ? kBc6hPrecSrc.base[m] : kBc6hPrecSrc.delta[which - 1][m];
    k |= (uint64_t)(v - 1) << (4 * (m < 8 ? m : m - 8) + (m < 8 ? 0 : 32));
  }
  return k;
}

// Per-lane data of the BC6H decode held in LDS on the GPU: words 0..5 the
// 12 endpoint components as 16-bit slots (ORed field by field), words 6..17
// the two subsets' interpolation records (c = 64 lo + 32, d = hi - lo per
// channel).  Rows are interleaved by lane at `stride` words.
struct Bc6hScratch {
  uint32_t* row;
  int stride;
  BCN_HD uint32_t& at(int w) const { return row[w * stride]; }
};

// Mode index of a BC6H block (0..13; 14 for the reserved prefixes) from its
// 2- or 5-bit prefix (block.rs:270-290).
BCN_HD int bc6h_mode(const uint32_t in[4]) {
  const uint32_t b = in[0] & 31u;
  const uint32_t mb = (b & 3u) > 1 ? b : (b & 3u);
  constexpr uint32_t kM0 = bc6h_mode_nibbles(0), kM1 = bc6h_mode_nibbles(1), kM2 = bc6h_mode_nibbles(2),
                     kM3 = bc6h_mode_nibbles(3);
  const uint32_t mw = mb >> 3;
  const uint32_t mk = mw == 0 ? kM0 : (mw == 1 ? kM1 : (mw == 2 ? kM2 : kM3));
  const int m1 = (int)bfe_u32(mk, 4u * (mb & 7u), 4);
  return m1 ? m1 - 1 : kBc6hDecRows - 1;
}

// Whether mode index m has one subset (modes 10..13, and the reserved row).
BCN_HD bool bc6h_one_subset(int m) { return m >= 10; }

// BC6H index state after the block's header: the remaining index bits and
// what the texels need to read them (decode_bc6h_texel reads them in order),
// and, for one-subset blocks (ONE), the subset's interpolation records.
struct Bc6hIdx {
  uint32_t lo, mid;  // the index stream with the anchors' missing bits inserted: bits 0.., 8 ib..
  uint32_t off[8];   // j * ib: index k at bit off[k & 7] of lo (k < 8) or mid
  uint32_t pattern;  // 2-bit subset ids of the 16 texels
  uint32_t ib;       // index bits
  int kw;            // weight multiplier: w = (i * kw + 32) >> 6
  uint32_t w2, w3;   // ONE: the block's last two words (the indices)
  int c[3], d[3];    // ONE: the records of the one subset
};

// decode_block_bc6h (block.rs:262-769), the header of a block of mode index
// m (bc6h_mode): endpoint fields, transforms and unquantize, leaving each
// subset's interpolation records in the lane's scratch row (words 6..17)
// and returning the index state.  Every mode runs one branch-free path:
// field widths, destinations and precisions come from tables indexed by the
// lane's mode, and a field the mode lacks is a read of width 0.
//   * Each endpoint field is read and ORed into its 16-bit slot in the lane's
//     scratch row (an LDS OR on the GPU), rather than merged through selects
//     over all 12 slots.  `fields` is kBc6hDec, 16-byte aligned; the
//     lane's row is read once into registers.
//     The field loop ends once no lane of the wave has a field left.
//   * A reserved mode prefix takes the field row of no fields: its
//     components are all 0 and so is every texel (block.rs:282-290).
//   * ONE: every lane of the wave holds a one-subset block (at most 9
//     fields, 2 endpoints, 4-bit indices); the records stay in registers.
//   * The indices end the block, so they are read from its top 64 bits;
//     anchor texels read one bit less, from an anchor mask.
//   * Interpolation lo + ((w (hi - lo) + 32) >> 6) is (c + w d) >> 6 with
//     c = 64 lo + 32 (64 lo is a multiple of 64): one 24-bit multiply-add.
template <bool SIGNED, bool ONE>
BCN_HD Bc6hIdx decode_bc6h_header(const uint32_t in[4], int m, Bc6hScratch sc, const uint16_t* fields) {
  BitReader br(in);
  br.skip((in[0] & 3u) > 1 ? 5u : 2u);
#pragma unroll
  for (int w = 0; w < 6; w++) sc.at(w) = 0;
  // The mode's row, loaded whole into registers (two entries a word).
  uint32_t rw[12];
  {
    const uint32_t* rp = static_cast<const uint32_t*>(__builtin_assume_aligned(fields + m * 24, 16));
#pragma unroll
    for (int q = 0; q < 12; q++) rw[q] = rp[q];
  }
  // Every field is merged as read; the reversed ones are put in order below.
#pragma unroll
  for (int i = 0; i < (ONE ? 9 : 23); i++) {
    const uint32_t t = (rw[i >> 1] >> (16 * (i & 1))) & 0xFFFFu;
    if (!ONE && !wave_any(t != 0u)) break;  // entries past a mode's count are 0
    const uint32_t n = t & 31u, dst = (t >> 5) & 0xFFu;
    const uint32_t v = br.read(n);
#if defined(__HIP_DEVICE_COMPILE__)
    __hip_atomic_fetch_or(&sc.at((int)(dst >> 5)), v << (dst & 31u), __ATOMIC_RELAXED,
                          __HIP_MEMORY_SCOPE_WORKGROUP);
#else
    sc.at((int)(dst >> 5)) |= v << (dst & 31u);
#endif
  }
  constexpr int kComp = ONE ? 6 : 12;  // components in use
  int f[12];
#pragma unroll
  for (int w = 0; w < kComp / 2; w++) {
    const uint32_t x = sc.at(w);
    f[2 * w] = (int)(x & 0xFFFFu);
    f[2 * w + 1] = (int)(x >> 16);
  }
  // The reversed fields are the top 2 (mode index 12) or 6 (13) bits of each
  // base component, from bit 10, stored last bit first (block.rs:656-676):
  // reverse them in place.
  const uint32_t nrev = m == 12 ? 2u : (m == 13 ? 6u : 0u);
#pragma unroll
  for (int c = 0; c < 3; c++) {
    const uint32_t x = (uint32_t)f[c];
    f[c] = nrev ? (int)((x & 0x3FFu) | (reverse_low_bits(x >> 10, nrev) << 10)) : f[c];
  }
  const bool two = !ONE && m < 10;
  const uint32_t part = ONE ? 0u : br.read(two ? 5u : 0u);
  const int ne = two ? 4 : 2;
  // Base bits (6..16) and the delta bits of R, G, B (kBc6hPrec), minus one,
  // as nibbles: modes 0..7 in the low word (the reserved row reads 0).
  constexpr uint64_t kBase = bc6h_prec_nibbles(0), kDR = bc6h_prec_nibbles(1), kDG = bc6h_prec_nibbles(2),
                     kDB = bc6h_prec_nibbles(3);
  const uint32_t mo = 4u * (uint32_t)(m & 7), mlo = m < 8;
  auto nib = [&](uint64_t k) { return (int)bfe_u32(mlo ? (uint32_t)k : (uint32_t)(k >> 32), mo, 4) + 1; };
  const int b0 = nib(kBase);
  const int db[3] = {nib(kDR), nib(kDG), nib(kDB)};
  const bool transformed = m != 9 && m != 10;
  if (SIGNED)
#pragma unroll
    for (int c = 0; c < 3; c++) f[c] = extend_sign(f[c], b0);
#pragma unroll
  for (int e = 1; e < kComp / 3; e++)
#pragma unroll
    for (int c = 0; c < 3; c++) {
      int v = f[e * 3 + c];
      if (SIGNED) {
        v = extend_sign(v, db[c]);
      } else {
        v = transformed ? extend_sign(v, db[c]) : v;
      }
      int t = (v + f[c]) & ((1 << b0) - 1);
      v = transformed ? (SIGNED ? extend_sign(t, b0) : t) : v;
      f[e * 3 + c] = e < ne ? v : f[e * 3 + c];
    }
#pragma unroll
  for (int i = 0; i < kComp; i++) f[i] = bc6h_unquantize<SIGNED>(f[i], b0);
  // Subset records: c = 64 lo + 32 and d = hi - lo per channel (|d| < 2^17,
  // 64 |d| + c < 2^23: one v_mad_i32_i24).
  Bc6hIdx st;
#pragma unroll
  for (int sb = 0; sb < kComp / 6; sb++)
#pragma unroll
    for (int c = 0; c < 3; c++) {
      const int lo = f[6 * sb + c], hi = f[6 * sb + 3 + c];
      if (ONE) {
        st.c[c] = 64 * lo + 32;
        st.d[c] = hi - lo;
      } else {
        sc.at(6 + 6 * sb + c) = (uint32_t)(64 * lo + 32);
        sc.at(6 + 6 * sb + 3 + c) = (uint32_t)(hi - lo);
      }
    }
  st.w2 = in[2];
  st.w3 = in[3];
  st.pattern = two ? kPart.subset[part] : 0u;
  const uint32_t a1 = two ? kPart.anchor1[part] : 0u;
  st.ib = two ? 3u : 4u;
  st.kw = two ? 585 : 273;
  // The indices: the block's last 16 ib - (1 + two) bits (46 or 63), made
  // uniform by inserting a 0 at each anchor's missing top bit -- texel 0's at
  // bit ib - 1, then subset 1's anchor a1 at (a1 + 1) ib - 1 -- so index k
  // sits at bit k ib (insert_zero64).
  const uint32_t sh = two ? 18u : 1u;
  uint64_t s = (((uint64_t)in[3] << 32) | in[2]) >> sh;
  s = insert_zero64(s, ~0ull << (st.ib - 1u));
  s = insert_zero64(s, two ? ~0ull << ((a1 + 1u) * 3u - 1u) : 0ull);
  st.lo = (uint32_t)s;
  st.mid = (uint32_t)(s >> (8u * st.ib));
  st.off[0] = 0u;
  st.off[1] = st.ib;
#pragma unroll
  for (int j = 2; j < 8; j++) st.off[j] = in_vgpr(st.off[j - 1] + st.ib);
  return st;
}

// Texel k's RGB half bit patterns; texels are read in order 0..15.  ONE:
// the 4-bit indices of texels 1..15 are the nibbles 1..15 of the block's
// last 64 bits (texel 0, the anchor, is bits 65..67).
template <bool SIGNED, bool ONE>
BCN_HD void decode_bc6h_texel(Bc6hIdx& st, int k, Bc6hScratch sc, uint16_t rgb[3]) {
  uint32_t x;
  int s = 0;
  if (ONE) {
    x = k == 0 ? bfe_u32(st.w2, 1, 3) : bfe_u32(k < 8 ? st.w2 : st.w3, 4u * (uint32_t)(k & 7), 4);
  } else {
    x = bfe_u32(k < 8 ? st.lo : st.mid, st.off[k & 7], st.ib);
    s = (int)bfe_u32(st.pattern, 2u * (uint32_t)k, 1);
  }
  const int w = (mul24((int)x, ONE ? 273 : st.kw) + 32) >> 6;
#pragma unroll
  for (int c = 0; c < 3; c++) {
    const int cc = ONE ? st.c[c] : (int)sc.at(6 + 6 * s + c);
    const int dd = ONE ? st.d[c] : (int)sc.at(6 + 6 * s + 3 + c);
    rgb[c] = bc6h_finish<SIGNED>((mul24i(w, dd) + cc) >> 6);
  }
}

// The whole block: out[k*3 + c] for texel k, channel c.  ONE may be set only
// for one-subset blocks (bc6h_one_subset); the general path takes any block.
template <bool SIGNED, bool ONE = false>
BCN_HD void decode_bc6h_t(const uint32_t in[4], uint16_t out[48], Bc6hScratch sc,
                          const uint16_t* fields) {
  const int m = bc6h_mode(in);
  Bc6hIdx st = decode_bc6h_header<SIGNED, ONE>(in, m, sc, fields);
#pragma unroll
  for (int k = 0; k < 16; k++) decode_bc6h_texel<SIGNED, ONE>(st, k, sc, out + 3 * k);
}

// Host form with private scratch (tests and the CPU kernel-check build).
// one_path: one-subset blocks take the ONE path (as in a wave of only such
// blocks); otherwise every block takes the general path.
BCN_HD void decode_bc6h(const uint32_t in[4], bool is_signed, uint16_t out[48], bool one_path = false) {
  uint32_t row[18];
  const Bc6hScratch sc{row, 1};
  const uint16_t* ft = &kBc6hDec.f[0][0];
  const bool one = one_path && bc6h_one_subset(bc6h_mode(in));
  if (is_signed)
    one ? decode_bc6h_t<true, true>(in, out, sc, ft) : decode_bc6h_t<true, false>(in, out, sc, ft);
  else
    one ? decode_bc6h_t<false, true>(in, out, sc, ft) : decode_bc6h_t<false, false>(in, out, sc, ft);
}

// ---- sRGB8 codes of BC6H texels (linear_to_srgb, src/decode.rs:151-159) -----
//
// Over the non-negative half floats the code is a step function of the bit
// pattern h (they order like their values): code(h) = #{k >= 1 : h >= t[k]}
// for the 255 thresholds t[k], the smallest pattern whose code is >= k.
// Within any bucket of 8 consecutive patterns the code rises at most once
// (the steepest stretch of the curve rises by ~0.11 codes per half ulp), so
// one 16-bit entry per bucket holds it:
//   entry = 256 * base + 32 * (8 - off),
// with base the code of the bucket's first pattern and off in 1..8 the
// position of the rise (8: none).  For j = h & 7,
//   entry + 32 j = 256 * base + 32 * (8 - off + j),  8 - off + j in [1, 14],
// so its high byte is base + (j >= off), the code of h.
constexpr uint32_t kSrgb8Buckets = 0x7C00u >> 3;  // every finite non-negative half

// A 16-bit word whose high byte is the sRGB8 code of the non-negative
// finite half h (h <= 0x7BFF).
BCN_HD uint32_t srgb8_word(uint32_t h, const uint16_t* buckets) {
  return (uint32_t)buckets[h >> 3] + ((h & 7u) << 5);
}

// Row and column of block `idx` in an image `bw` blocks wide, without a
// division: m = floor((2^32 - 1) / bw) gives q = mulhi(idx, m) in
// {idx / bw - 1, idx / bw} for idx < 2^31 (m bw > 2^32 - 1 - bw), and one
// compare of the remainder settles it.  bw and the row count below 2^24
// (the remainder product is a 24-bit multiply).
struct BlockRows {
  uint32_t bw, m;
};
BCN_HD BlockRows block_rows(uint32_t bw) { return BlockRows{bw, 0xFFFFFFFFu / bw}; }
BCN_HD void block_pos(BlockRows r, uint32_t idx, uint32_t& yy, uint32_t& xx) {
#if defined(__HIP_DEVICE_COMPILE__)
  const uint32_t q = __umulhi(idx, r.m);
  const uint32_t rem = idx - __umul24(q, r.bw);
#else
  const uint32_t q = (uint32_t)(((uint64_t)idx * r.m) >> 32);
  const uint32_t rem = idx - q * r.bw;
#endif
  const bool up = rem >= r.bw;
  yy = q + (up ? 1u : 0u);
  xx = up ? rem - r.bw : rem;
}

#if defined(__HIPCC__)
__device__ __forceinline__ float half_to_float(uint16_t h) {
  return (float)__builtin_bit_cast(_Float16, h);
}
#endif

}  // namespace bcn
