// TEST INFRASTRUCTURE ONLY -- the CPU oracle (see bc_oracle.hpp).
//
// Restatement of the BC6H (unsigned) block encoder in src/encode/bc6h.rs.
#include "bc_oracle.hpp"

namespace bco {

// IEEE-754 binary16 <- binary32, round to nearest even.  This is the
// published behaviour of the `half` crate's f16::from_f32 (used at
// src/encode/bc6h.rs:59-63); it is pinned in tests against numpy float16.
uint16_t f32_to_f16_bits(float x) {
  uint32_t u;
  std::memcpy(&u, &x, 4);
  uint32_t sign = (u >> 16) & 0x8000u;
  uint32_t exp = (u >> 23) & 0xFFu;
  uint32_t man = u & 0x7FFFFFu;
  if (exp == 0xFF) {  // inf / NaN
    if (man == 0) return (uint16_t)(sign | 0x7C00u);
    return (uint16_t)(sign | 0x7E00u | (man >> 13));
  }
  int32_t e = (int32_t)exp - 127 + 15;
  if (e >= 0x1F) return (uint16_t)(sign | 0x7C00u);  // overflow -> inf
  if (e <= 0) {                                       // subnormal or zero
    if (e < -10) return (uint16_t)sign;
    man |= 0x800000u;
    uint32_t shift = (uint32_t)(14 - e);
    uint32_t half_man = man >> shift;
    uint32_t rem = man & ((1u << shift) - 1);
    uint32_t halfway = 1u << (shift - 1);
    if (rem > halfway || (rem == halfway && (half_man & 1))) half_man++;
    return (uint16_t)(sign | half_man);
  }
  uint32_t half = sign | ((uint32_t)e << 10) | (man >> 13);
  uint32_t rem = man & 0x1FFFu;
  if (rem > 0x1000u || (rem == 0x1000u && (half & 1))) half++;  // may carry into exp: correct
  return (uint16_t)half;
}

// binary16 -> binary32 (exact).
float f16_bits_to_f32(uint16_t h) {
  uint32_t sign = ((uint32_t)h & 0x8000u) << 16;
  uint32_t exp = ((uint32_t)h >> 10) & 0x1Fu;
  uint32_t man = (uint32_t)h & 0x3FFu;
  uint32_t u;
  if (exp == 0) {
    if (man == 0) {
      u = sign;
    } else {
      int e = -1;
      do {
        e++;
        man <<= 1;
      } while ((man & 0x400u) == 0);
      u = sign | ((uint32_t)(127 - 15 - e) << 23) | ((man & 0x3FFu) << 13);
    }
  } else if (exp == 0x1F) {
    u = sign | 0x7F800000u | (man << 13);
  } else {
    u = sign | ((exp + 127 - 15) << 23) | (man << 13);
  }
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}

namespace {

// bc6h.rs:20-27.  powf is the platform libm's (parity unpinned for this
// function: it is not on any measured path).
float srgb_to_linear(uint8_t srgb) {
  float v = (float)srgb / 255.0f;
  if (v <= 0.04045f) return v / 12.92f;
  return std::pow((v + 0.055f) / 1.055f, 2.4f);
}

struct Bc6h {
  F block[64];
  uint32_t data[5];
  F best_err;
  F rgb_bounds[6];
  F max_span;
  size_t max_span_idx;
  size_t mode;
  uint32_t epb;
  int32_t qbounds[8];
  const BC6HSettings* settings;
  // Signed (SF16) encoding.  Not in the reference, whose encoder is unsigned
  // only; its decoder reads signed blocks (decode/block.rs:266, 697-719).
  // The same search over the signed value domain: a texel of half bits h is
  // +-(|h| / 31) * 32 (the inverse of the signed finish_unquantize,
  // block.rs:1391-1397), endpoints quantise to +-(2^(b-1) - 1) and
  // dequantise as the decoder's signed unquantize (block.rs:1369-1389), and
  // endpoint fields are two's complement in b bits.
  bool sgn = false;

  // bc6h.rs:131-135
  static uint32_t get_mode_prefix(size_t mode) {
    static const uint32_t T[14] = {0, 1, 2, 6, 10, 14, 18, 22, 26, 30, 3, 7, 11, 15};
    return T[mode];
  }

  // bc6h.rs:137-156 (constants evaluated in f32 as Rust const-evaluates them)
  static F get_span(size_t mode) {
    static const float T[14] = {
        0.9f * 65535.0f / 64.0f, 0.9f * 65535.0f / 4.0f, 0.8f * 65535.0f / 256.0f, -1.0f, -1.0f,
        0.9f * 65535.0f / 32.0f, 0.9f * 65535.0f / 16.0f, -1.0f, -1.0f, 65535.0f, 65535.0f,
        0.95f * 65535.0f / 8.0f, 0.95f * 65535.0f / 32.0f, 6.0f};
    return F(T[mode]);
  }

  // bc6h.rs:158-164
  static uint32_t get_mode_bits(size_t mode) {
    static const uint32_t T[14] = {10, 7, 11, 0xFFFFFFFFu, 0xFFFFFFFFu, 9, 8,
                                   0xFFFFFFFFu, 0xFFFFFFFFu, 6, 10, 11, 12, 16};
    return T[mode];
  }

  // Signed quantisation of one endpoint value: the magnitude rounds to
  // 0..2^(b-1)-1 over 0..32767 as the unsigned one rounds over 0..65535.
  static int32_t quant_signed(F e, uint32_t bits) {
    int32_t top = (1 << (bits - 1)) - 1;
    int32_t m = iclamp(f2i(fabs_(e) / F(32767.0f) * i2f(top) + F(0.5f)), 0, top);
    return e < F(0.0f) ? -m : m;
  }

  // Signed unquantize (block.rs:1369-1389).
  static int32_t unquant_signed(int32_t v, uint32_t bits) {
    if (bits >= 16) return v;
    int32_t m = v < 0 ? -v : v, u;
    if (m == 0)
      u = 0;
    else if (m >= (1 << (bits - 1)) - 1)
      u = 0x7FFF;
    else
      u = ((m << 15) + 0x4000) >> (bits - 1);
    return v < 0 ? -u : u;
  }

  // bc6h.rs:166-173
  void ep_quant_bc6h_8(const F ep[8], uint32_t bits, uint32_t pairs) {
    int32_t levels = 1 << bits;
    for (size_t i = 0; i < 8 * (size_t)pairs; i++) {
      if (sgn) {
        qbounds[i] = quant_signed(ep[i], bits);
        continue;
      }
      int32_t v = f2i(ep[i] / F(256.0f * 256.0f - 1.0f) * i2f(levels - 1) + F(0.5f));
      qbounds[i] = iclamp(v, 0, levels - 1);
    }
  }

  // bc6h.rs:175-185
  void compute_qbounds_core(const F rgb_span[3]) {
    F bounds[8];
    for (int i = 0; i < 8; i++) bounds[i] = 0.0f;
    for (int p = 0; p < 3; p++) {
      F middle = (rgb_bounds[p] + rgb_bounds[3 + p]) / F(2.0f);
      bounds[p] = middle - rgb_span[p] / F(2.0f);
      bounds[4 + p] = middle + rgb_span[p] / F(2.0f);
    }
    ep_quant_bc6h_8(bounds, epb, 1);
  }

  // bc6h.rs:187-197
  void compute_qbounds(F span) {
    F s[3] = {span, span, span};
    compute_qbounds_core(s);
  }
  void compute_qbounds2(F span, size_t idx) {
    F s[3] = {span, span, span};
    if (idx < 3) s[idx] *= F(2.0f);
    compute_qbounds_core(s);
  }

  // bc6h.rs:199-211
  static uint32_t unpack_to_uf16(uint32_t v, uint32_t bits) {
    if (bits >= 15) return v;
    if (v == 0) return 0;
    if (v == (1u << bits) - 1) return 0xFFFF;
    return (v * 2 + 1) << (15 - bits);
  }

  // bc6h.rs:213-220
  static void ep_quant_bc6h(int32_t qep[24], const F ep[24], uint32_t bits, size_t pairs,
                            bool sgn) {
    int32_t levels = 1 << bits;
    for (size_t i = 0; i < 8 * pairs; i++) {
      if (sgn) {
        qep[i] = quant_signed(ep[i], bits);
        continue;
      }
      int32_t v = f2i(ep[i] / F(256.0f * 256.0f - 1.0f) * i2f(levels - 1) + F(0.5f));
      qep[i] = iclamp(v, 0, levels - 1);
    }
  }

  // bc6h.rs:222-226
  static void ep_dequant_bc6h(F ep[24], const int32_t qep[24], uint32_t bits, size_t pairs,
                              bool sgn) {
    for (size_t i = 0; i < 8 * pairs; i++)
      ep[i] = sgn ? i2f(unquant_signed(qep[i], bits))
                  : u2f(unpack_to_uf16((uint32_t)qep[i], bits));
  }

  // bc6h.rs:228-239
  void ep_quant_dequant_bc6h(int32_t qep[24], F ep[24], size_t pairs) const {
    uint32_t bits = epb;
    ep_quant_bc6h(qep, ep, bits, pairs, sgn);
    for (size_t i = 0; i < 2 * pairs; i++)
      for (int p = 0; p < 3; p++)
        qep[i * 4 + p] = iclamp(qep[i * 4 + p], qbounds[p], qbounds[4 + p]);
    ep_dequant_bc6h(ep, qep, bits, pairs, sgn);
  }

  // Signed endpoints go to the packer as their low `epb` bits (two's
  // complement): every field it writes lies in those bits, and the deltas
  // modulo 2^delta_bits are unchanged.
  void mask_signed(int32_t qep[24], size_t pairs) const {
    if (!sgn) return;
    for (size_t i = 0; i < 2 * pairs; i++)
      for (int p = 0; p < 3; p++) qep[i * 4 + p] &= (1 << epb) - 1;
  }

  static uint32_t bit_at(int32_t v, uint32_t pos) { return (uint32_t)((v >> pos) & 1); }

  // bc6h.rs:452-463
  static uint32_t reverse_bits(uint32_t v, uint32_t bits) {
    if (bits == 2) return (v >> 1) + (v & 1) * 2;
    // bits == 6 (the only other call)
    uint32_t vv = (v & 0x5555) * 2 + ((v >> 1) & 0x5555);
    return (vv >> 4) + ((vv >> 2) & 3) * 4 + (vv & 3) * 16;
  }

  // bc6h.rs:465-791
  static void pack(uint32_t packed[4], const int32_t qep[24], size_t mode) {
    packed[0] = packed[1] = packed[2] = packed[3] = 0;
    if (mode == 0) {
      int32_t pr[16] = {0};
      for (int p = 0; p < 3; p++) {
        pr[p] = qep[p];
        pr[4 + p] = (qep[4 + p] - qep[p]) & 31;
        pr[8 + p] = (qep[8 + p] - qep[p]) & 31;
        pr[12 + p] = (qep[12 + p] - qep[p]) & 31;
      }
      uint32_t pq[10] = {0};
      pq[4] = (uint32_t)pr[4] + (uint32_t)(pr[8 + 1] & 15) * 64;
      pq[5] = (uint32_t)pr[5] + (uint32_t)(pr[12 + 1] & 15) * 64;
      pq[6] = (uint32_t)pr[6] + (uint32_t)(pr[8 + 2] & 15) * 64;
      pq[4] += bit_at(pr[12 + 1], 4) << 5;
      pq[5] += bit_at(pr[12 + 2], 0) << 5;
      pq[6] += bit_at(pr[12 + 2], 1) << 5;
      pq[8] = (uint32_t)pr[8] + bit_at(pr[12 + 2], 2) * 32;
      pq[9] = (uint32_t)pr[12] + bit_at(pr[12 + 2], 3) * 32;
      packed[0] = get_mode_prefix(0);
      packed[0] += bit_at(pr[8 + 1], 4) << 2;
      packed[0] += bit_at(pr[8 + 2], 4) << 3;
      packed[0] += bit_at(pr[12 + 2], 4) << 4;
      packed[1] = ((uint32_t)pr[2] << 20) + ((uint32_t)pr[1] << 10) + (uint32_t)pr[0];
      packed[2] = (pq[6] << 20) + (pq[5] << 10) + pq[4];
      packed[3] = (pq[9] << 6) + pq[8];
    } else if (mode == 1) {
      int32_t pr[16] = {0};
      for (int p = 0; p < 3; p++) {
        pr[p] = qep[p];
        pr[4 + p] = (qep[4 + p] - qep[p]) & 63;
        pr[8 + p] = (qep[8 + p] - qep[p]) & 63;
        pr[12 + p] = (qep[12 + p] - qep[p]) & 63;
      }
      uint32_t pq[8] = {0};
      pq[0] = (uint32_t)pr[0];
      pq[0] += bit_at(pr[12 + 2], 0) << 7;
      pq[0] += bit_at(pr[12 + 2], 1) << 8;
      pq[0] += bit_at(pr[8 + 2], 4) << 9;
      pq[1] = (uint32_t)pr[1];
      pq[1] += bit_at(pr[8 + 2], 5) << 7;
      pq[1] += bit_at(pr[12 + 2], 2) << 8;
      pq[1] += bit_at(pr[8 + 1], 4) << 9;
      pq[2] = (uint32_t)pr[2];
      pq[2] += bit_at(pr[12 + 2], 3) << 7;
      pq[2] += bit_at(pr[12 + 2], 5) << 8;
      pq[2] += bit_at(pr[12 + 2], 4) << 9;
      pq[4] = (uint32_t)pr[4] + (uint32_t)(pr[8 + 1] & 15) * 64;
      pq[5] = (uint32_t)pr[5] + (uint32_t)(pr[12 + 1] & 15) * 64;
      pq[6] = (uint32_t)pr[6] + (uint32_t)(pr[8 + 2] & 15) * 64;
      packed[0] = get_mode_prefix(1);
      packed[0] += bit_at(pr[8 + 1], 5) << 2;
      packed[0] += bit_at(pr[12 + 1], 4) << 3;
      packed[0] += bit_at(pr[12 + 1], 5) << 4;
      packed[1] = (pq[2] << 20) + (pq[1] << 10) + pq[0];
      packed[2] = (pq[6] << 20) + (pq[5] << 10) + pq[4];
      packed[3] = ((uint32_t)pr[12] << 6) + (uint32_t)pr[8];
    } else if (mode == 2 || mode == 3 || mode == 4) {
      int32_t dq[16] = {0};
      for (int p = 0; p < 3; p++) {
        int32_t mask = ((size_t)p == mode - 2) ? 31 : 15;
        dq[p] = qep[p];
        dq[4 + p] = (qep[4 + p] - qep[p]) & mask;
        dq[8 + p] = (qep[8 + p] - qep[p]) & mask;
        dq[12 + p] = (qep[12 + p] - qep[p]) & mask;
      }
      uint32_t pq[10] = {0};
      pq[0] = (uint32_t)(dq[0] & 1023);
      pq[1] = (uint32_t)(dq[1] & 1023);
      pq[2] = (uint32_t)(dq[2] & 1023);
      pq[4] = (uint32_t)dq[4] + (uint32_t)(dq[8 + 1] & 15) * 64;
      pq[5] = (uint32_t)dq[5] + (uint32_t)(dq[12 + 1] & 15) * 64;
      pq[6] = (uint32_t)dq[6] + (uint32_t)(dq[8 + 2] & 15) * 64;
      pq[8] = (uint32_t)dq[8];
      pq[9] = (uint32_t)dq[12];
      if (mode == 2) {
        packed[0] = get_mode_prefix(2);
        pq[5] += bit_at(dq[1], 10) << 4;
        pq[6] += bit_at(dq[2], 10) << 4;
        pq[4] += bit_at(dq[0], 10) << 5;
        pq[5] += bit_at(dq[12 + 2], 0) << 5;
        pq[6] += bit_at(dq[12 + 2], 1) << 5;
        pq[8] += bit_at(dq[12 + 2], 2) << 5;
        pq[9] += bit_at(dq[12 + 2], 3) << 5;
      } else if (mode == 3) {
        packed[0] = get_mode_prefix(3);
        pq[4] += bit_at(dq[0], 10) << 4;
        pq[6] += bit_at(dq[2], 10) << 4;
        pq[8] += bit_at(dq[12 + 2], 0) << 4;
        pq[9] += bit_at(dq[8 + 1], 4) << 4;
        pq[4] += bit_at(dq[12 + 1], 4) << 5;
        pq[5] += bit_at(dq[1], 10) << 5;
        pq[6] += bit_at(dq[12 + 2], 1) << 5;
        pq[8] += bit_at(dq[12 + 2], 2) << 5;
        pq[9] += bit_at(dq[12 + 2], 3) << 5;
      } else {
        packed[0] = get_mode_prefix(4);
        pq[4] += bit_at(dq[0], 10) << 4;
        pq[5] += bit_at(dq[1], 10) << 4;
        pq[8] += bit_at(dq[12 + 2], 1) << 4;
        pq[9] += bit_at(dq[12 + 2], 4) << 4;
        pq[4] += bit_at(dq[8 + 2], 4) << 5;
        pq[5] += bit_at(dq[12 + 2], 0) << 5;
        pq[6] += bit_at(dq[2], 10) << 5;
        pq[8] += bit_at(dq[12 + 2], 2) << 5;
        pq[9] += bit_at(dq[12 + 2], 3) << 5;
      }
      packed[1] = (pq[2] << 20) + (pq[1] << 10) + pq[0];
      packed[2] = (pq[6] << 20) + (pq[5] << 10) + pq[4];
      packed[3] = (pq[9] << 6) + pq[8];
    } else if (mode == 5) {
      int32_t dq[16] = {0};
      for (int p = 0; p < 3; p++) {
        dq[p] = qep[p];
        dq[4 + p] = (qep[4 + p] - qep[p]) & 31;
        dq[8 + p] = (qep[8 + p] - qep[p]) & 31;
        dq[12 + p] = (qep[12 + p] - qep[p]) & 31;
      }
      uint32_t pq[10] = {0};
      pq[0] = (uint32_t)dq[0];
      pq[1] = (uint32_t)dq[1];
      pq[2] = (uint32_t)dq[2];
      pq[4] = (uint32_t)dq[4] + (uint32_t)(dq[8 + 1] & 15) * 64;
      pq[5] = (uint32_t)dq[5] + (uint32_t)(dq[12 + 1] & 15) * 64;
      pq[6] = (uint32_t)dq[6] + (uint32_t)(dq[8 + 2] & 15) * 64;
      pq[8] = (uint32_t)dq[8];
      pq[9] = (uint32_t)dq[12];
      pq[0] += bit_at(dq[8 + 2], 4) << 9;
      pq[1] += bit_at(dq[8 + 1], 4) << 9;
      pq[2] += bit_at(dq[12 + 2], 4) << 9;
      pq[4] += bit_at(dq[12 + 1], 4) << 5;
      pq[5] += bit_at(dq[12 + 2], 0) << 5;
      pq[6] += bit_at(dq[12 + 2], 1) << 5;
      pq[8] += bit_at(dq[12 + 2], 2) << 5;
      pq[9] += bit_at(dq[12 + 2], 3) << 5;
      packed[0] = get_mode_prefix(5);
      packed[1] = (pq[2] << 20) + (pq[1] << 10) + pq[0];
      packed[2] = (pq[6] << 20) + (pq[5] << 10) + pq[4];
      packed[3] = (pq[9] << 6) + pq[8];
    } else if (mode == 6 || mode == 7 || mode == 8) {
      int32_t dq[16] = {0};
      for (int p = 0; p < 3; p++) {
        int32_t mask = ((size_t)p == mode - 6) ? 63 : 31;
        dq[p] = qep[p];
        dq[4 + p] = (qep[4 + p] - qep[p]) & mask;
        dq[8 + p] = (qep[8 + p] - qep[p]) & mask;
        dq[12 + p] = (qep[12 + p] - qep[p]) & mask;
      }
      uint32_t pq[10] = {0};
      pq[0] = (uint32_t)dq[0];
      pq[0] += bit_at(dq[8 + 2], 4) << 9;
      pq[1] = (uint32_t)dq[1];
      pq[1] += bit_at(dq[8 + 1], 4) << 9;
      pq[2] = (uint32_t)dq[2];
      pq[2] += bit_at(dq[12 + 2], 4) << 9;
      pq[4] = (uint32_t)dq[4] + (uint32_t)(dq[8 + 1] & 15) * 64;
      pq[5] = (uint32_t)dq[5] + (uint32_t)(dq[12 + 1] & 15) * 64;
      pq[6] = (uint32_t)dq[6] + (uint32_t)(dq[8 + 2] & 15) * 64;
      pq[8] = (uint32_t)dq[8];
      pq[9] = (uint32_t)dq[12];
      if (mode == 6) {
        packed[0] = get_mode_prefix(6);
        pq[0] += bit_at(dq[12 + 1], 4) << 8;
        pq[1] += bit_at(dq[12 + 2], 2) << 8;
        pq[2] += bit_at(dq[12 + 2], 3) << 8;
        pq[5] += bit_at(dq[12 + 2], 0) << 5;
        pq[6] += bit_at(dq[12 + 2], 1) << 5;
      } else if (mode == 7) {
        packed[0] = get_mode_prefix(7);
        pq[0] += bit_at(dq[12 + 2], 0) << 8;
        pq[1] += bit_at(dq[8 + 1], 5) << 8;
        pq[2] += bit_at(dq[12 + 1], 5) << 8;
        pq[4] += bit_at(dq[12 + 1], 4) << 5;
        pq[6] += bit_at(dq[12 + 2], 1) << 5;
        pq[8] += bit_at(dq[12 + 2], 2) << 5;
        pq[9] += bit_at(dq[12 + 2], 3) << 5;
      } else {
        packed[0] = get_mode_prefix(8);
        pq[0] += bit_at(dq[12 + 2], 1) << 8;
        pq[1] += bit_at(dq[8 + 2], 5) << 8;
        pq[2] += bit_at(dq[12 + 2], 5) << 8;
        pq[4] += bit_at(dq[12 + 1], 4) << 5;
        pq[5] += bit_at(dq[12 + 2], 0) << 5;
        pq[8] += bit_at(dq[12 + 2], 2) << 5;
        pq[9] += bit_at(dq[12 + 2], 3) << 5;
      }
      packed[1] = (pq[2] << 20) + (pq[1] << 10) + pq[0];
      packed[2] = (pq[6] << 20) + (pq[5] << 10) + pq[4];
      packed[3] = (pq[9] << 6) + pq[8];
    } else if (mode == 9) {
      uint32_t pq[10] = {0};
      pq[0] = (uint32_t)qep[0];
      pq[0] += bit_at(qep[12 + 1], 4) << 6;
      pq[0] += bit_at(qep[12 + 2], 0) << 7;
      pq[0] += bit_at(qep[12 + 2], 1) << 8;
      pq[0] += bit_at(qep[8 + 2], 4) << 9;
      pq[1] = (uint32_t)qep[1];
      pq[1] += bit_at(qep[8 + 1], 5) << 6;
      pq[1] += bit_at(qep[8 + 2], 5) << 7;
      pq[1] += bit_at(qep[12 + 2], 2) << 8;
      pq[1] += bit_at(qep[8 + 1], 4) << 9;
      pq[2] = (uint32_t)qep[2];
      pq[2] += bit_at(qep[12 + 1], 5) << 6;
      pq[2] += bit_at(qep[12 + 2], 3) << 7;
      pq[2] += bit_at(qep[12 + 2], 5) << 8;
      pq[2] += bit_at(qep[12 + 2], 4) << 9;
      pq[4] = (uint32_t)qep[4] + (uint32_t)(qep[8 + 1] & 15) * 64;
      pq[5] = (uint32_t)qep[5] + (uint32_t)(qep[12 + 1] & 15) * 64;
      pq[6] = (uint32_t)qep[6] + (uint32_t)(qep[8 + 2] & 15) * 64;
      packed[0] = get_mode_prefix(9);
      packed[1] = (pq[2] << 20) + (pq[1] << 10) + pq[0];
      packed[2] = (pq[6] << 20) + (pq[5] << 10) + pq[4];
      packed[3] = ((uint32_t)qep[12] << 6) + (uint32_t)qep[8];
    } else if (mode == 10) {
      packed[0] = get_mode_prefix(10);
      packed[1] = ((uint32_t)qep[2] << 20) + ((uint32_t)qep[1] << 10) + (uint32_t)qep[0];
      packed[2] = ((uint32_t)qep[6] << 20) + ((uint32_t)qep[5] << 10) + (uint32_t)qep[4];
    } else if (mode == 11) {
      int32_t dq[8] = {0};
      for (int p = 0; p < 3; p++) {
        dq[p] = qep[p];
        dq[4 + p] = (qep[4 + p] - qep[p]) & 511;
      }
      uint32_t pq[10] = {0};
      pq[0] = (uint32_t)(dq[0] & 1023);
      pq[1] = (uint32_t)(dq[1] & 1023);
      pq[2] = (uint32_t)(dq[2] & 1023);
      pq[4] = (uint32_t)dq[4] + (uint32_t)(dq[0] >> 10) * 512;
      pq[5] = (uint32_t)dq[5] + (uint32_t)(dq[1] >> 10) * 512;
      pq[6] = (uint32_t)dq[6] + (uint32_t)(dq[2] >> 10) * 512;
      packed[0] = get_mode_prefix(11);
      packed[1] = (pq[2] << 20) + (pq[1] << 10) + pq[0];
      packed[2] = (pq[6] << 20) + (pq[5] << 10) + pq[4];
    } else if (mode == 12) {
      int32_t dq[8] = {0};
      for (int p = 0; p < 3; p++) {
        dq[p] = qep[p];
        dq[4 + p] = (qep[4 + p] - qep[p]) & 255;
      }
      uint32_t pq[8] = {0};
      pq[0] = (uint32_t)(dq[0] & 1023);
      pq[1] = (uint32_t)(dq[1] & 1023);
      pq[2] = (uint32_t)(dq[2] & 1023);
      pq[4] = (uint32_t)dq[4] + reverse_bits((uint32_t)(dq[0] >> 10), 2) * 256;
      pq[5] = (uint32_t)dq[5] + reverse_bits((uint32_t)(dq[1] >> 10), 2) * 256;
      pq[6] = (uint32_t)dq[6] + reverse_bits((uint32_t)(dq[2] >> 10), 2) * 256;
      packed[0] = get_mode_prefix(12);
      packed[1] = (pq[2] << 20) + (pq[1] << 10) + pq[0];
      packed[2] = (pq[6] << 20) + (pq[5] << 10) + pq[4];
    } else if (mode == 13) {
      int32_t dq[8] = {0};
      for (int p = 0; p < 3; p++) {
        dq[p] = qep[p];
        dq[4 + p] = (qep[4 + p] - qep[p]) & 15;
      }
      uint32_t pq[8] = {0};
      pq[0] = (uint32_t)(dq[0] & 1023);
      pq[1] = (uint32_t)(dq[1] & 1023);
      pq[2] = (uint32_t)(dq[2] & 1023);
      pq[4] = (uint32_t)dq[4] + reverse_bits((uint32_t)(dq[0] >> 10), 6) * 16;
      pq[5] = (uint32_t)dq[5] + reverse_bits((uint32_t)(dq[1] >> 10), 6) * 16;
      pq[6] = (uint32_t)dq[6] + reverse_bits((uint32_t)(dq[2] >> 10), 6) * 16;
      packed[0] = get_mode_prefix(13);
      packed[1] = (pq[2] << 20) + (pq[1] << 10) + pq[0];
      packed[2] = (pq[6] << 20) + (pq[5] << 10) + pq[4];
    }
  }

  // bc6h.rs:241-266
  void code_2p(int32_t qep[24], const uint32_t qblock[2], int32_t part_id, size_t m) {
    const uint32_t bits = 3;
    uint32_t flips = bc7_code_apply_swap_mode01237(qep, qblock, 1, part_id);
    mask_signed(qep, 2);
    std::memset(data, 0, sizeof(data));
    uint32_t pos = 0;
    uint32_t packed[4];
    pack(packed, qep, m);
    put_bits(data, &pos, 5, packed[0]);
    put_bits(data, &pos, 30, packed[1]);
    put_bits(data, &pos, 30, packed[2]);
    put_bits(data, &pos, 12, packed[3]);
    put_bits(data, &pos, 5, (uint32_t)part_id);
    bc7_code_qblock(data, &pos, qblock, bits, flips);
    bc7_code_adjust_skip_mode01237(data, 1, part_id);
  }

  // bc6h.rs:268-286
  void code_1p(int32_t qep[24], uint32_t qblock[2], size_t m) {
    bc7_code_apply_swap_mode456(qep, 4, qblock, 4);
    mask_signed(qep, 1);
    std::memset(data, 0, sizeof(data));
    uint32_t pos = 0;
    uint32_t packed[4];
    pack(packed, qep, m);
    put_bits(data, &pos, 5, packed[0]);
    put_bits(data, &pos, 30, packed[1]);
    put_bits(data, &pos, 30, packed[2]);
    bc7_code_qblock(data, &pos, qblock, 4, 0);
  }

  // bc6h.rs:304-324
  F enc_2p_part_fast(int32_t qep[24], uint32_t qblock[2], int32_t part_id) const {
    uint32_t pattern = get_pattern(part_id);
    const uint32_t bits = 3;
    const size_t pairs = 2, channels = 3;
    F ep[24];
    for (int i = 0; i < 24; i++) ep[i] = 0.0f;
    for (size_t j = 0; j < pairs; j++) {
      uint32_t mask = get_pattern_mask(part_id, (uint32_t)j);
      block_segment_core(ep + j * 8, block, mask, channels);
    }
    ep_quant_dequant_bc6h(qep, ep, 2);
    return block_quant(qblock, block, bits, ep, pattern, channels, sgn);
  }

  // bc6h.rs:326-388
  void enc_2p_list(const int32_t part_list[32], uint32_t part_count) {
    if (part_count == 0) return;
    const uint32_t bits = 3;
    const size_t pairs = 2, channels = 3;
    int32_t best_qep[24] = {0};
    uint32_t best_qblock[2] = {0, 0};
    int32_t best_part_id = -1;
    F best = F(INFINITY);
    for (uint32_t part = 0; part < part_count; part++) {
      int32_t part_id = part_list[part] & 31;
      int32_t qep[24] = {0};
      uint32_t qblock[2] = {0, 0};
      F err = enc_2p_part_fast(qep, qblock, part_id);
      if (err < best) {
        std::memcpy(best_qep, qep, sizeof(int32_t) * 8 * pairs);
        best_qblock[0] = qblock[0];
        best_qblock[1] = qblock[1];
        best_part_id = part_id;
        best = err;
      }
    }
    for (uint32_t it = 0; it < settings->refine_iterations_2p; it++) {
      F ep[24];
      for (int i = 0; i < 24; i++) ep[i] = 0.0f;
      for (size_t j = 0; j < pairs; j++) {
        uint32_t mask = get_pattern_mask(best_part_id, (uint32_t)j);
        opt_endpoints(ep + j * 8, block, bits, best_qblock, mask, channels);
      }
      int32_t qep[24] = {0};
      uint32_t qblock[2] = {0, 0};
      ep_quant_dequant_bc6h(qep, ep, 2);
      uint32_t pattern = get_pattern(best_part_id);
      F err = block_quant(qblock, block, bits, ep, pattern, channels, sgn);
      if (err < best) {
        std::memcpy(best_qep, qep, sizeof(int32_t) * 8 * pairs);
        best_qblock[0] = qblock[0];
        best_qblock[1] = qblock[1];
        best = err;
      }
    }
    if (best < best_err) {
      best_err = best;
      code_2p(best_qep, best_qblock, best_part_id, mode);
    }
  }

  // bc6h.rs:288-302
  void enc_2p() {
    F full_stats[15];
    for (int i = 0; i < 15; i++) full_stats[i] = 0.0f;
    compute_stats_masked(full_stats, block, 0xFFFFFFFFu, 3);
    int32_t part_list[32];
    for (int32_t part = 0; part < 32; part++) {
      uint32_t mask = get_pattern_mask(part, 0);
      F bound12 = block_pca_bound_split(block, mask, full_stats, 3);
      int32_t bound = f2i(bound12);
      part_list[part] = part + bound * 64;
    }
    partial_sort_list(part_list, 32, settings->fast_skip_threshold);
    enc_2p_list(part_list, settings->fast_skip_threshold);
  }

  // bc6h.rs:390-412
  void enc_1p() {
    F ep[24];
    for (int i = 0; i < 24; i++) ep[i] = 0.0f;
    block_segment_core(ep, block, 0xFFFFFFFFu, 3);
    int32_t qep[24] = {0};
    ep_quant_dequant_bc6h(qep, ep, 1);
    uint32_t qblock[2] = {0, 0};
    F err = block_quant(qblock, block, 4, ep, 0, 3, sgn);
    for (uint32_t it = 0; it < settings->refine_iterations_1p; it++) {
      opt_endpoints(ep, block, 4, qblock, 0xFFFFFFFFu, 3);
      ep_quant_dequant_bc6h(qep, ep, 1);
      err = block_quant(qblock, block, 4, ep, 0, 3, sgn);
    }
    if (err < best_err) {
      best_err = err;
      code_1p(qep, qblock, mode);
    }
  }

  // bc6h.rs:414-446
  void test_mode(size_t m, bool enc, F margin) {
    uint32_t mode_bits = get_mode_bits(m);
    F span = get_span(m);
    if (max_span * margin > span) return;
    if (m >= 10) {
      epb = mode_bits;
      mode = m;
      compute_qbounds(span);
      if (enc) enc_1p();
    } else if (m <= 1 || m == 5 || m == 9) {
      epb = mode_bits;
      mode = m;
      compute_qbounds(span);
      if (enc) enc_2p();
    } else {
      epb = mode_bits;
      mode = m + max_span_idx;
      compute_qbounds2(span, max_span_idx);
      if (enc) enc_2p();
    }
  }

  // bc6h.rs:793-819
  void setup() {
    for (int p = 0; p < 3; p++) {
      rgb_bounds[p] = F(65535.0f);
      rgb_bounds[3 + p] = sgn ? F(-65535.0f) : F(0.0f);
    }
    for (int p = 0; p < 3; p++)
      for (int k = 0; k < 16; k++) {
        F value;
        if (sgn) {
          // Sign and magnitude of the half; -0 maps to -0.0.
          uint32_t h = (uint32_t)raw(block[p * 16 + k]);
          value = (u2f(h & 0x7FFFu) / F(31.0f)) * F(32.0f);
          if (h & 0x8000u) value = -value;
        } else {
          value = (block[p * 16 + k] / F(31.0f)) * F(64.0f);
        }
        block[p * 16 + k] = value;
        rgb_bounds[p] = fmin_(rgb_bounds[p], value);
        rgb_bounds[3 + p] = fmax_(rgb_bounds[3 + p], value);
      }
    max_span = 0.0f;
    max_span_idx = 0;
    for (int p = 0; p < 3; p++) {
      F span = rgb_bounds[3 + p] - rgb_bounds[p];
      if (span > max_span) {
        max_span_idx = (size_t)p;
        max_span = span;
      }
    }
  }

  // bc6h.rs:821-860
  void compress_core() {
    setup();
    if (settings->slow_mode != 0) {
      test_mode(0, true, F(0.0f));
      test_mode(1, true, F(0.0f));
      test_mode(2, true, F(0.0f));
      test_mode(5, true, F(0.0f));
      test_mode(6, true, F(0.0f));
      test_mode(9, true, F(0.0f));
      test_mode(10, true, F(0.0f));
      test_mode(11, true, F(0.0f));
      test_mode(12, true, F(0.0f));
      test_mode(13, true, F(0.0f));
    } else {
      if (settings->fast_skip_threshold > 0) {
        test_mode(9, false, F(0.0f));
        if (settings->fast_mode != 0) test_mode(1, false, F(1.0f));
        test_mode(6, false, F(1.0f / 1.2f));
        test_mode(5, false, F(1.0f / 1.2f));
        test_mode(0, false, F(1.0f / 1.2f));
        test_mode(2, false, F(1.0f));
        enc_2p();
        if (settings->fast_mode == 0) test_mode(1, true, F(0.0f));
      }
      test_mode(10, false, F(0.0f));
      test_mode(11, false, F(1.0f));
      test_mode(12, false, F(1.0f));
      test_mode(13, false, F(1.0f));
      enc_1p();
    }
  }

  void init(const BC6HSettings* s) {
    settings = s;
    for (int i = 0; i < 64; i++) block[i] = 0.0f;
    std::memset(data, 0, sizeof(data));
    best_err = F(INFINITY);
    for (int i = 0; i < 6; i++) rgb_bounds[i] = 0.0f;
    max_span = 0.0f;
    max_span_idx = 0;
    mode = 0;
    epb = 0;
    for (int i = 0; i < 8; i++) qbounds[i] = 0;
  }

  void store(uint8_t out[16]) const {
    for (int i = 0; i < 4; i++) {
      out[i * 4 + 0] = (uint8_t)data[i];
      out[i * 4 + 1] = (uint8_t)(data[i] >> 8);
      out[i * 4 + 2] = (uint8_t)(data[i] >> 16);
      out[i * 4 + 3] = (uint8_t)(data[i] >> 24);
    }
  }
};

}  // namespace

// bc6h.rs:73-97 then encode.rs:601-604; `is_signed` selects the signed
// (SF16) encoding described at Bc6h::sgn.
float bc6h_encode_block_16(const uint16_t* rgba16, size_t stride, const BC6HSettings& s,
                           uint8_t out[16], bool is_signed) {
  Bc6h st;
  st.init(&s);
  st.sgn = is_signed;
  for (int y = 0; y < 4; y++)
    for (int x = 0; x < 4; x++) {
      const uint16_t* px = rgba16 + y * stride + x * 4;
      st.block[y * 4 + x] = u2f(px[0]);
      st.block[16 + y * 4 + x] = u2f(px[1]);
      st.block[32 + y * 4 + x] = u2f(px[2]);
      st.block[48 + y * 4 + x] = 0.0f;
    }
  st.compress_core();
  st.store(out);
  return raw(st.best_err);
}

// bc6h.rs:45-71 then encode.rs:562-565
float bc6h_encode_block_8(const uint8_t* rgba, size_t stride, const BC6HSettings& s,
                          uint8_t out[16]) {
  Bc6h st;
  st.init(&s);
  for (int y = 0; y < 4; y++)
    for (int x = 0; x < 4; x++) {
      const uint8_t* px = rgba + y * stride + x * 4;
      for (int c = 0; c < 3; c++)
        st.block[c * 16 + y * 4 + x] = u2f(f32_to_f16_bits(srgb_to_linear(px[c])));
      st.block[48 + y * 4 + x] = 0.0f;
    }
  st.compress_core();
  st.store(out);
  return raw(st.best_err);
}

}  // namespace bco
