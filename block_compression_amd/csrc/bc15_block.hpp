// BC1-BC5 block encoders: one 4x4 block per lane.
//
// Follow src/encode/bc1_to_5.rs operation for operation (f32, no FMA, same
// evaluation order), so the blocks equal the crate's CPU encoder output:
//   BC1 colour: PCA axis (4 power iterations), endpoints from the projection
//   extents, 565 rounding, one least-squares refinement (bc1_to_5.rs:386-429);
//   BC3/BC4/BC5 channel block: min/max endpoints and 8-level indices
//   (bc1_to_5.rs:431-472); BC2 alpha: truncating 4-bit alpha (:84-115).
#pragma once
#include "bcn_common.hpp"

namespace bcn {

// bc1_to_5.rs:268-278.  c in [0, 255] or NaN (-> 0, Rust `as i32`).
BCN_HD int enc_rgb565(float c0, float c1, float c2) {
  int r = qclamp(c0, 0.0f, 255.0f), g = qclamp(c1, 0.0f, 255.0f), b = qclamp(c2, 0.0f, 255.0f);
  int r5 = (r * 31 + 128 + ((r * 31) >> 8)) >> 8;
  int g6 = (g * 63 + 128 + ((g * 63) >> 8)) >> 8;
  int b5 = (b * 31 + 128 + ((b * 31) >> 8)) >> 8;
  return (r5 << 11) + (g6 << 5) + b5;
}

// fast_quant (bc1_to_5.rs:280-321): 2-bit index of every texel along p0->p1.
BCN_HD uint32_t bc1_fast_quant(const uint32_t px[16], int p0, int p1) {
  // dec_rgb565 (bc1_to_5.rs:258-266)
  float c0[3], c1[3];
  {
    int b5 = p0 & 31, g6 = (p0 >> 5) & 63, r5 = (p0 >> 11) & 31;
    c0[0] = (float)((r5 << 3) + (r5 >> 2));
    c0[1] = (float)((g6 << 2) + (g6 >> 4));
    c0[2] = (float)((b5 << 3) + (b5 >> 2));
    b5 = p1 & 31;
    g6 = (p1 >> 5) & 63;
    r5 = (p1 >> 11) & 31;
    c1[0] = (float)((r5 << 3) + (r5 >> 2));
    c1[1] = (float)((g6 << 2) + (g6 >> 4));
    c1[2] = (float)((b5 << 3) + (b5 >> 2));
  }
  float dir[3];
#pragma unroll
  for (int p = 0; p < 3; p++) dir[p] = c1[p] - c0[p];
  // Integer-valued, exact in any order.
  float sq_norm = dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2];
  float rsq_norm = rcp_cr(sq_norm);
  float t = rsq_norm * 3.0f;
#pragma unroll
  for (int p = 0; p < 3; p++) dir[p] = dir[p] * t;
  float bias = 0.5f;
#pragma unroll
  for (int p = 0; p < 3; p++) bias = bias - c0[p] * dir[p];
  uint32_t bits = 0;
#pragma unroll
  for (int k = 0; k < 16; k++) {
    uint32_t tx = fresh(px[k]);
    // The reference's sum starts at 0.0; a -0.0 first term only differs in
    // the sign of a zero, which `as i32` maps to 0 either way.
    float dot = chf(tx, 0) * dir[0];
    dot = dot + chf(tx, 1) * dir[1];
    dot = dot + chf(tx, 2) * dir[2];
    uint32_t q = (uint32_t)qclamp(dot + bias, 0.0f, 3.0f);
    bits |= q << (2 * k);
  }
  return bits;
}

// compress_block_bc1_core (bc1_to_5.rs:386-429) -> the two BC1 words.
BCN_HD void bc1_core(const uint32_t px[16], uint32_t out[2]) {
  // compute_covar_dc (bc1_to_5.rs:157-192).  The channel sums are exact
  // integers (<= 16 * 255), one v_dot4_u32_u8 per group of four texels; the
  // covariance accumulates in f32 in texel order.
  int s[3];
  {
    uint32_t cm[4][4], us[3] = {0, 0, 0};
    channel_major(px, cm);
#pragma unroll
    for (int g = 0; g < 4; g++)
#pragma unroll
      for (int p = 0; p < 3; p++) us[p] = udot4(cm[g][p], 0x01010101u, us[p]);
#pragma unroll
    for (int p = 0; p < 3; p++) s[p] = (int)us[p];
  }
  float dc[3] = {(float)s[0] / 16.0f, (float)s[1] / 16.0f, (float)s[2] / 16.0f};
  float c0 = 0.0f, c1 = 0.0f, c2 = 0.0f, c3 = 0.0f, c4 = 0.0f, c5 = 0.0f;
#pragma unroll
  for (int k = 0; k < 16; k++) {
    uint32_t t = fresh(px[k]);
    float r = chf(t, 0) - dc[0], g = chf(t, 1) - dc[1], b = chf(t, 2) - dc[2];
    c0 = c0 + r * r;
    c1 = c1 + r * g;
    c2 = c2 + r * b;
    c3 = c3 + g * g;
    c4 = c4 + g * b;
    c5 = c5 + b * b;
  }
  const float eps = 1.1920928955078125e-07f;  // f32::EPSILON
  c0 += eps;
  c3 += eps;
  c5 += eps;
  // compute_axis3 (bc1_to_5.rs:200-223), 4 power iterations.
  float a0 = 1.0f, a1 = 1.0f, a2 = 1.0f;
#pragma unroll
  for (int i = 0; i < 4; i++) {
    float x0 = c0 * a0 + c1 * a1 + c2 * a2;
    float x1 = c1 * a0 + c3 * a1 + c4 * a2;
    float x2 = c2 * a0 + c4 * a1 + c5 * a2;
    a0 = x0;
    a1 = x1;
    a2 = x2;
    if (i % 2 == 1) {
      float norm_sq = x0 * x0 + x1 * x1 + x2 * x2;
      float rnorm = rcp_cr(sqrt_cr(norm_sq));
      a0 *= rnorm;
      a1 *= rnorm;
      a2 *= rnorm;
    }
  }
  // pick_endpoints (bc1_to_5.rs:225-256).
  float min_dot = 256.0f * 256.0f, max_dot = 0.0f;
#pragma unroll
  for (int k = 0; k < 16; k++) {
    uint32_t t = fresh(px[k]);
    float dot = (chf(t, 0) - dc[0]) * a0;
    dot = dot + (chf(t, 1) - dc[1]) * a1;
    dot = dot + (chf(t, 2) - dc[2]) * a2;
    min_dot = fminf(min_dot, dot);
    max_dot = fmaxf(max_dot, dot);
  }
  if (max_dot - min_dot < 1.0f) {
    min_dot -= 0.5f;
    max_dot += 0.5f;
  }
  float norm_sq = a0 * a0 + a1 * a1 + a2 * a2;
  float rnorm_sq = rcp_cr(norm_sq);
  float axis[3] = {a0, a1, a2};
  float e0[3], e1[3];
#pragma unroll
  for (int p = 0; p < 3; p++) {
    e0[p] = rclamp(dc[p] + min_dot * rnorm_sq * axis[p], 0.0f, 255.0f);
    e1[p] = rclamp(dc[p] + max_dot * rnorm_sq * axis[p], 0.0f, 255.0f);
  }
  int p0 = enc_rgb565(e0[0], e0[1], e0[2]);
  int p1 = enc_rgb565(e1[0], e1[1], e1[2]);
  if (p0 < p1) {
    int tt = p0;
    p0 = p1;
    p1 = tt;
  }
  uint32_t bits = bc1_fast_quant(px, p0, p1);
  // One bc1_refine (bc1_to_5.rs:323-374).
  {
    float r0[3], r1[3];
    if ((bits ^ (bits * 4u)) < 4u) {
#pragma unroll
      for (int p = 0; p < 3; p++) r0[p] = r1[p] = dc[p];
    } else {
      // Sums of products of small integers: exact (<= 16*3*255), one
      // v_dot4_u32_u8 per group of four texels over channel-major bytes and
      // the group's indices as bytes (x = 3 - q = q ^ 3 for 2-bit q).
      int atb1[3], sum_q, sum_qq;
      {
        uint32_t cm[4][4], ua[3] = {0, 0, 0}, uq = 0, uqq = 0;
        channel_major(px, cm);
#pragma unroll
        for (int g = 0; g < 4; g++) {
          const uint32_t b = (bits >> (8 * g)) & 0xFFu;
          const uint32_t qb = (b & 3u) | ((b & 0xCu) << 6) | ((b & 0x30u) << 12) | ((b & 0xC0u) << 18);
          uq = udot4(qb, 0x01010101u, uq);
          uqq = udot4(qb, qb, uqq);
#pragma unroll
          for (int p = 0; p < 3; p++) ua[p] = udot4(cm[g][p], qb ^ 0x03030303u, ua[p]);
        }
#pragma unroll
        for (int p = 0; p < 3; p++) atb1[p] = (int)ua[p];
        sum_q = (int)uq;
        sum_qq = (int)uqq;
      }
      float cxx = (float)(144 - 6 * sum_q + sum_qq);
      float cyy = (float)sum_qq;
      float cxy = (float)(3 * sum_q - sum_qq);
      float scale = 3.0f * rcp_cr(cxx * cyy - cxy * cxy);
#pragma unroll
      for (int p = 0; p < 3; p++) {
        // sum[p] = dc[p] * 16.0 == s[p] exactly; atb2 = 3*sum - atb1 exact.
        float fa1 = (float)atb1[p];
        float fa2 = (float)(3 * s[p] - atb1[p]);
        r0[p] = rclamp((fa1 * cyy - fa2 * cxy) * scale, 0.0f, 255.0f);
        r1[p] = rclamp((fa2 * cxx - fa1 * cxy) * scale, 0.0f, 255.0f);
      }
    }
    p0 = enc_rgb565(r0[0], r0[1], r0[2]);
    p1 = enc_rgb565(r1[0], r1[1], r1[2]);
    if (p0 < p1) {
      int tt = p0;
      p0 = p1;
      p1 = tt;
    }
    bits = bc1_fast_quant(px, p0, p1);
  }
  out[0] = ((uint32_t)p1 << 16) | (uint32_t)p0;
  // fix_qbits (bc1_to_5.rs:376-384)
  uint32_t q0 = bits & 0x55555555u, q1 = bits & 0xAAAAAAAAu;
  out[1] = (q1 >> 1) + (q1 ^ (q0 << 1));
}

// compress_block_bc3_alpha (bc1_to_5.rs:431-472) on channel byte `ch` of the
// texels -> the two words of a BC3 alpha / BC4 block.
BCN_HD void bc3_alpha_block(const uint32_t px[16], int ch, uint32_t out[2]) {
  int mn = 255, mx = 0;
#pragma unroll
  for (int k = 0; k < 16; k++) {
    int v = chi(px[k], ch);
    mn = v < mn ? v : mn;
    mx = v > mx ? v : mx;
  }
  float ep0 = (float)mn, ep1 = (float)mx;
  if (ep0 == ep1) ep1 = ep0 + 0.1f;
  float scale = 7.0f / (ep1 - ep0);
  uint32_t qb0 = 0, qb1 = 0;
#pragma unroll
  for (int k = 0; k < 16; k++) {
    float proj = (chf(px[k], ch) - ep0) * scale + 0.5f;
    int q = qclamp(proj, 0.0f, 7.0f);
    q = 7 - q;
    if (q > 0) q += 1;
    if (q == 8) q = 1;
    if (k < 8) qb0 |= (uint32_t)q << (3 * k);
    else qb1 |= (uint32_t)q << (3 * (k - 8));
  }
  // ep0 / ep1 are integers (or mn + 0.1, which truncates to mn) in [0, 255].
  uint32_t e0 = (uint32_t)mn, e1 = (uint32_t)ep1;
  out[0] = (e0 << 8) | e1 | (qb0 << 16);
  out[1] = (qb0 >> 16) | (qb1 << 8);
}

// load_block_alpha_4bit (bc1_to_5.rs:84-115) -> the two BC2 alpha words.
BCN_HD void bc2_alpha(const uint32_t px[16], uint32_t out[2]) {
  out[0] = out[1] = 0;
#pragma unroll
  for (int k = 0; k < 16; k++) {
    float a = chf(px[k], 3) / 255.0f;
    uint32_t a4 = (uint32_t)qclamp(a * 15.0f, 0.0f, 15.0f);
    if (k < 8) out[0] |= a4 << (4 * k);
    else out[1] |= a4 << (4 * (k - 8));
  }
}

}  // namespace bcn
