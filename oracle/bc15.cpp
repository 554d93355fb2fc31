// TEST INFRASTRUCTURE ONLY -- the CPU oracle (see bc_oracle.hpp).
//
// Restatement of the BC1-BC5 block encoders in src/encode/bc1_to_5.rs and the
// per-format block assembly of src/encode.rs:260-539.
#include "bc_oracle.hpp"

namespace bco {
namespace {

struct Bc15 {
  F block[64];

  // bc1_to_5.rs:14-40
  void load_rgba(const uint8_t* rgba, size_t stride) {
    for (int y = 0; y < 4; y++)
      for (int x = 0; x < 4; x++) {
        const uint8_t* px = rgba + y * stride + x * 4;
        int o = y * 4 + x;
        block[o] = i2f(px[0]);
        block[16 + o] = i2f(px[1]);
        block[32 + o] = i2f(px[2]);
        block[48 + o] = i2f(px[3]);
      }
  }

  // bc1_to_5.rs:42-82 (R or G channel into block[48..64])
  void load_channel(const uint8_t* rgba, size_t stride, int channel) {
    for (int y = 0; y < 4; y++)
      for (int x = 0; x < 4; x++) block[48 + y * 4 + x] = i2f(rgba[y * stride + x * 4 + channel]);
  }

  // bc1_to_5.rs:157-192
  void compute_covar_dc(F covar[6], F dc[3]) const {
    for (int p = 0; p < 3; p++) {
      F acc = 0.0f;
      for (int k = 0; k < 16; k++) acc += block[k + p * 16];
      dc[p] = acc / F(16.0f);
    }
    F c0 = 0.0f, c1 = 0.0f, c2 = 0.0f, c3 = 0.0f, c4 = 0.0f, c5 = 0.0f;
    for (int k = 0; k < 16; k++) {
      F rgb0 = block[k] - dc[0];
      F rgb1 = block[k + 16] - dc[1];
      F rgb2 = block[k + 32] - dc[2];
      c0 += rgb0 * rgb0;
      c1 += rgb0 * rgb1;
      c2 += rgb0 * rgb2;
      c3 += rgb1 * rgb1;
      c4 += rgb1 * rgb2;
      c5 += rgb2 * rgb2;
    }
    covar[0] = c0;
    covar[1] = c1;
    covar[2] = c2;
    covar[3] = c3;
    covar[4] = c4;
    covar[5] = c5;
  }

  // bc1_to_5.rs:194-198
  static void ssymv(F r[3], const F c[6], const F a[3]) {
    r[0] = c[0] * a[0] + c[1] * a[1] + c[2] * a[2];
    r[1] = c[1] * a[0] + c[3] * a[1] + c[4] * a[2];
    r[2] = c[2] * a[0] + c[4] * a[1] + c[5] * a[2];
  }

  // bc1_to_5.rs:200-223
  static void compute_axis3(F axis[3], const F covar[6], int power_iterations) {
    F a[3] = {1.0f, 1.0f, 1.0f};
    for (int i = 0; i < power_iterations; i++) {
      ssymv(axis, covar, a);
      for (int p = 0; p < 3; p++) a[p] = axis[p];
      if (i % 2 == 1) {
        F norm_sq = 0.0f;
        for (int p = 0; p < 3; p++) norm_sq += axis[p] * axis[p];
        F rnorm = F(1.0f) / fsqrt(norm_sq);
        for (int p = 0; p < 3; p++) a[p] *= rnorm;
      }
    }
    for (int p = 0; p < 3; p++) axis[p] = a[p];
  }

  // bc1_to_5.rs:225-256
  void pick_endpoints(F c0[3], F c1[3], const F axis[3], const F dc[3]) const {
    F min_dot = F(256.0f * 256.0f);
    F max_dot = 0.0f;
    for (int y = 0; y < 4; y++)
      for (int x = 0; x < 4; x++) {
        F dot = 0.0f;
        for (int p = 0; p < 3; p++) dot += (block[p * 16 + y * 4 + x] - dc[p]) * axis[p];
        min_dot = fmin_(min_dot, dot);
        max_dot = fmax_(max_dot, dot);
      }
    if (max_dot - min_dot < F(1.0f)) {
      min_dot -= F(0.5f);
      max_dot += F(0.5f);
    }
    F norm_sq = 0.0f;
    for (int p = 0; p < 3; p++) norm_sq += axis[p] * axis[p];
    F rnorm_sq = F(1.0f) / norm_sq;  // f32::recip
    for (int p = 0; p < 3; p++) {
      c0[p] = fclamp(dc[p] + min_dot * rnorm_sq * axis[p], F(0.0f), F(255.0f));
      c1[p] = fclamp(dc[p] + max_dot * rnorm_sq * axis[p], F(0.0f), F(255.0f));
    }
  }

  // bc1_to_5.rs:258-266
  static void dec_rgb565(F c[3], int32_t p) {
    int32_t b5 = p & 31;
    int32_t g6 = (p >> 5) & 63;
    int32_t r5 = (p >> 11) & 31;
    c[0] = i2f((r5 << 3) + (r5 >> 2));
    c[1] = i2f((g6 << 2) + (g6 >> 4));
    c[2] = i2f((b5 << 3) + (b5 >> 2));
  }

  // bc1_to_5.rs:268-278
  static int32_t enc_rgb565(const F c[3]) {
    int32_t r = f2i(c[0]);
    int32_t g = f2i(c[1]);
    int32_t b = f2i(c[2]);
    int32_t r5 = (r * 31 + 128 + ((r * 31) >> 8)) >> 8;
    int32_t g6 = (g * 63 + 128 + ((g * 63) >> 8)) >> 8;
    int32_t b5 = (b * 31 + 128 + ((b * 31) >> 8)) >> 8;
    return (r5 << 11) + (g6 << 5) + b5;
  }

  // bc1_to_5.rs:280-321
  uint32_t fast_quant(int32_t p0, int32_t p1) const {
    F c0[3], c1[3];
    dec_rgb565(c0, p0);
    dec_rgb565(c1, p1);
    F dir[3];
    for (int p = 0; p < 3; p++) dir[p] = c1[p] - c0[p];
    F sq_norm = 0.0f;
    for (int p = 0; p < 3; p++) sq_norm += dir[p] * dir[p];  // powi(2)
    F rsq_norm = F(1.0f) / sq_norm;                            // recip
    for (int p = 0; p < 3; p++) dir[p] *= rsq_norm * F(3.0f);
    F bias = 0.5f;
    for (int p = 0; p < 3; p++) bias -= c0[p] * dir[p];
    uint32_t bits = 0;
    uint32_t scaler = 1;
    for (int k = 0; k < 16; k++) {
      F dot = 0.0f;
      for (int p = 0; p < 3; p++) dot += block[k + p * 16] * dir[p];
      int32_t q = iclamp(f2i(dot + bias), 0, 3);
      bits += (uint32_t)q * scaler;
      scaler *= 4u;  // wrapping_mul
    }
    return bits;
  }

  // bc1_to_5.rs:323-374
  // `colours` (may be null) receives c0 and c1 before the 565 rounding.
  void bc1_refine(int32_t pe[2], uint32_t bits, const F dc[3], F* colours = nullptr) const {
    F c0[3], c1[3];
    if ((bits ^ (bits * 4u)) < 4u) {
      for (int p = 0; p < 3; p++) {
        c0[p] = dc[p];
        c1[p] = dc[p];
      }
    } else {
      F atb1[3] = {0.0f, 0.0f, 0.0f};
      F sum_q = 0.0f;
      F sum_qq = 0.0f;
      uint32_t shifted_bits = bits;
      for (int k = 0; k < 16; k++) {
        F q = u2f(shifted_bits & 3);
        shifted_bits >>= 2;
        F x = F(3.0f) - q;
        sum_q += q;
        sum_qq += q * q;
        for (int p = 0; p < 3; p++) atb1[p] += x * block[k + p * 16];
      }
      F sum[3], atb2[3];
      for (int p = 0; p < 3; p++) {
        sum[p] = dc[p] * F(16.0f);
        atb2[p] = F(3.0f) * sum[p] - atb1[p];
      }
      F cxx = F(16.0f * 9.0f) - F(2.0f * 3.0f) * sum_q + sum_qq;
      F cyy = sum_qq;
      F cxy = F(3.0f) * sum_q - sum_qq;
      F scale = F(3.0f) * (F(1.0f) / (cxx * cyy - cxy * cxy));
      for (int p = 0; p < 3; p++) {
        c0[p] = (atb1[p] * cyy - atb2[p] * cxy) * scale;
        c1[p] = (atb2[p] * cxx - atb1[p] * cxy) * scale;
        c0[p] = fclamp(c0[p], F(0.0f), F(255.0f));
        c1[p] = fclamp(c1[p], F(0.0f), F(255.0f));
      }
    }
    if (colours)
      for (int p = 0; p < 3; p++) {
        colours[p] = c0[p];
        colours[3 + p] = c1[p];
      }
    pe[0] = enc_rgb565(c0);
    pe[1] = enc_rgb565(c1);
  }

  // bc1_to_5.rs:376-384
  static uint32_t fix_qbits(uint32_t qbits) {
    const uint32_t MASK_01B = 0x55555555u;
    const uint32_t MASK_10B = 0xAAAAAAAAu;
    uint32_t qbits0 = qbits & MASK_01B;
    uint32_t qbits1 = qbits & MASK_10B;
    return (qbits1 >> 1) + (qbits1 ^ (qbits0 << 1));
  }

  // bc1_to_5.rs:386-429
  void bc1_core(uint32_t data[2]) const {
    const int power_iterations = 4;
    const int refine_iterations = 1;
    F covar[6], dc[3];
    compute_covar_dc(covar, dc);
    const F EPS = F(1.1920928955078125e-07f);  // f32::EPSILON
    covar[0] += EPS;
    covar[3] += EPS;
    covar[5] += EPS;
    F axis[3] = {0.0f, 0.0f, 0.0f};
    compute_axis3(axis, covar, power_iterations);
    F c0[3], c1[3];
    pick_endpoints(c0, c1, axis, dc);
    int32_t p[2];
    p[0] = enc_rgb565(c0);
    p[1] = enc_rgb565(c1);
    if (p[0] < p[1]) {
      int32_t t = p[0];
      p[0] = p[1];
      p[1] = t;
    }
    data[0] = ((uint32_t)p[1] << 16) | (uint32_t)p[0];
    data[1] = fast_quant(p[0], p[1]);
    for (int it = 0; it < refine_iterations; it++) {
      bc1_refine(p, data[1], dc);
      if (p[0] < p[1]) {
        int32_t t = p[0];
        p[0] = p[1];
        p[1] = t;
      }
      data[0] = ((uint32_t)p[1] << 16) | (uint32_t)p[0];
      data[1] = fast_quant(p[0], p[1]);
    }
    data[1] = fix_qbits(data[1]);
  }

  // bc1_to_5.rs:431-472 (works on block[48..64])
  void bc3_alpha(uint32_t data[2]) const {
    F ep[2] = {F(255.0f), F(0.0f)};
    for (int k = 0; k < 16; k++) {
      ep[0] = fmin_(ep[0], block[48 + k]);
      ep[1] = fmax_(ep[1], block[48 + k]);
    }
    if (ep[0] == ep[1]) ep[1] = ep[0] + F(0.1f);
    uint32_t qblock[2] = {0, 0};
    F scale = F(7.0f) / (ep[1] - ep[0]);
    for (int k = 0; k < 16; k++) {
      F v = block[48 + k];
      F proj = (v - ep[0]) * scale + F(0.5f);
      int32_t q = iclamp(f2i(proj), 0, 7);
      q = 7 - q;
      if (q > 0) q += 1;
      if (q == 8) q = 1;
      qblock[k / 8] |= ((uint32_t)q) << ((k % 8) * 3);
    }
    data[0] = (uclamp(f2u(ep[0]), 0, 255) << 8) | uclamp(f2u(ep[1]), 0, 255);
    data[0] |= qblock[0] << 16;
    data[1] = qblock[0] >> 16;
    data[1] |= qblock[1] << 8;
  }
};

void store_words(uint8_t* out, const uint32_t* w, int n) {
  for (int i = 0; i < n; i++) {
    out[i * 4 + 0] = (uint8_t)w[i];
    out[i * 4 + 1] = (uint8_t)(w[i] >> 8);
    out[i * 4 + 2] = (uint8_t)(w[i] >> 16);
    out[i * 4 + 3] = (uint8_t)(w[i] >> 24);
  }
}

// bc1_to_5.rs:84-115
void load_block_alpha_4bit(const uint8_t* rgba, size_t stride, uint32_t alpha_bits[2]) {
  alpha_bits[0] = 0;
  alpha_bits[1] = 0;
  for (int y = 0; y < 4; y++)
    for (int x = 0; x < 4; x++) {
      F alpha = i2f(rgba[y * stride + x * 4 + 3]) / F(255.0f);
      uint32_t alpha4 = f2u(alpha * F(15.0f));
      int bit_position = y * 16 + x * 4;
      if (bit_position < 32) alpha_bits[0] |= alpha4 << bit_position;
      else alpha_bits[1] |= alpha4 << (bit_position - 32);
    }
}

}  // namespace

// encode.rs:281-285
void bc1_encode_block(const uint8_t* rgba, size_t stride, uint8_t out[8]) {
  Bc15 c;
  for (int i = 0; i < 64; i++) c.block[i] = 0.0f;
  c.load_rgba(rgba, stride);
  uint32_t d[2];
  c.bc1_core(d);
  store_words(out, d, 2);
}

// encode.rs:322-337
void bc2_encode_block(const uint8_t* rgba, size_t stride, uint8_t out[16]) {
  Bc15 c;
  for (int i = 0; i < 64; i++) c.block[i] = 0.0f;
  uint32_t d[4];
  uint32_t a[2];
  load_block_alpha_4bit(rgba, stride, a);
  d[0] = a[0];
  d[1] = a[1];
  c.load_rgba(rgba, stride);
  uint32_t col[2];
  c.bc1_core(col);
  d[2] = col[0];
  d[3] = col[1];
  store_words(out, d, 4);
}

// encode.rs:384-398
void bc3_encode_block(const uint8_t* rgba, size_t stride, uint8_t out[16]) {
  Bc15 c;
  for (int i = 0; i < 64; i++) c.block[i] = 0.0f;
  c.load_rgba(rgba, stride);
  uint32_t d[4];
  uint32_t a[2], col[2];
  c.bc3_alpha(a);
  c.bc1_core(col);
  d[0] = a[0];
  d[1] = a[1];
  d[2] = col[0];
  d[3] = col[1];
  store_words(out, d, 4);
}

// encode.rs:445-455
void bc4_encode_block(const uint8_t* rgba, size_t stride, uint8_t out[8]) {
  Bc15 c;
  for (int i = 0; i < 64; i++) c.block[i] = 0.0f;
  c.load_channel(rgba, stride, 0);
  uint32_t d[2];
  c.bc3_alpha(d);
  store_words(out, d, 2);
}

// encode.rs:498-513
void bc5_encode_block(const uint8_t* rgba, size_t stride, uint8_t out[16]) {
  Bc15 c;
  for (int i = 0; i < 64; i++) c.block[i] = 0.0f;
  uint32_t d[4], r[2], g[2];
  c.load_channel(rgba, stride, 0);
  c.bc3_alpha(r);
  c.load_channel(rgba, stride, 1);
  c.bc3_alpha(g);
  d[0] = r[0];
  d[1] = r[1];
  d[2] = g[0];
  d[3] = g[1];
  store_words(out, d, 4);
}

}  // namespace bco

using namespace bco;

extern "C" {

// Single primitives of bc1_to_5.rs for the oracle's pin tests.  `block` is
// the 64-float channel-major block (R[16], G[16], B[16], A[16]).

// enc_rgb565 (bc1_to_5.rs:268-278) of one colour.
int32_t oracle_bc1_enc_rgb565(float r, float g, float b) {
  F c[3] = {F(r), F(g), F(b)};
  return Bc15::enc_rgb565(c);
}

// dec_rgb565 (bc1_to_5.rs:258-266).
void oracle_bc1_dec_rgb565(int32_t p, float* rgb3) {
  F c[3];
  Bc15::dec_rgb565(c, p);
  for (int i = 0; i < 3; i++) rgb3[i] = raw(c[i]);
}

// fast_quant (bc1_to_5.rs:280-321): 2-bit indices of the block's texels
// against the 565 endpoints p0, p1.
uint32_t oracle_bc1_fast_quant(const float* block64, int32_t p0, int32_t p1) {
  Bc15 c;
  for (int i = 0; i < 64; i++) c.block[i] = F(block64[i]);
  return c.fast_quant(p0, p1);
}

// bc1_refine (bc1_to_5.rs:323-374): the least-squares endpoint colours c0,
// c1 before the 565 rounding (clamped to 0..255), and their 565 codes.
void oracle_bc1_refine(const float* block64, uint32_t bits, float* c0c1_6, int32_t* pe2) {
  Bc15 c;
  for (int i = 0; i < 64; i++) c.block[i] = F(block64[i]);
  F covar[6], dc[3];
  c.compute_covar_dc(covar, dc);
  int32_t pe[2];
  F col[6];
  c.bc1_refine(pe, bits, dc, col);
  pe2[0] = pe[0];
  pe2[1] = pe[1];
  if (c0c1_6)
    for (int i = 0; i < 6; i++) c0c1_6[i] = raw(col[i]);
}

// The BC1 principal axis: compute_covar_dc + eps + compute_axis3 with 4
// iterations (bc1_to_5.rs:386-397); also returns the mean colour.
void oracle_bc1_axis(const float* block64, float* axis3, float* dc3) {
  Bc15 c;
  for (int i = 0; i < 64; i++) c.block[i] = F(block64[i]);
  F covar[6], dc[3], axis[3] = {0.0f, 0.0f, 0.0f};
  c.compute_covar_dc(covar, dc);
  const F EPS = F(1.1920928955078125e-07f);
  covar[0] += EPS;
  covar[3] += EPS;
  covar[5] += EPS;
  Bc15::compute_axis3(axis, covar, 4);
  for (int p = 0; p < 3; p++) {
    axis3[p] = raw(axis[p]);
    dc3[p] = raw(dc[p]);
  }
}

}  // extern "C"
