// Differentiable views of BC7 and BC1 blocks: the decode as a function of
// continuous endpoints with the block's structure held fixed, its adjoint,
// and a gradient (Adam) endpoint refinement for BC7.  No counterpart in the
// reference crate; the decodes differentiated are decode_block_bc7
// (src/decode/block.rs:801-1252) and decode_color_block (:124-197), restated
// in decode_block.hpp.
//
// A block's *form* is everything but the endpoint values: mode, partition,
// rotation, index selector, p-bits and both index sets.  Its parameters are
// the endpoint values x[e][c] (endpoint e = 2*subset + {0,1}, channel c) in
// the mode's quantised units, 0 .. 2^bits - 1, without the p-bit.
//
// Forward, straight-through: q = clamp(rint(x)) and the exact integer decode.
// Forward, smooth: the decode's linear relaxation
//     byte(x)   = (pbit ? 2x + p : x) * 255 / (2^n - 1),  n = stored bits (+1 with a p-bit)
//     texel     = ((64 - w) * byte_lo + w * byte_hi) / 64
// (bit replication left-aligns and repeats the MSBs, which is v * 255/(2^n-1)
// up to rounding; the interpolation drops only its +32 >> 6 rounding).
// Backward: the transpose of the smooth map's Jacobian, which does not depend
// on x; with the straight-through forward it is the usual straight-through
// gradient.  Channels without stored alpha (modes 0-3) decode alpha as 255
// and have no parameter.
#pragma once
#include "bcn_common.hpp"
#include "decode_block.hpp"

namespace bcn {

struct Bc7Form {
  uint32_t mode;          // 0..7; 8 = a block with no valid mode (decodes to zero)
  uint32_t part, rot, isb;
  int q[6][4];            // quantised endpoint values without p-bit
  int p[6];               // p-bit of each endpoint (0 when the mode has none)
  uint32_t idx[2];        // primary index of texel k at bits 4*(k&7) of word k>>3
  uint32_t idx2[2];       // secondary (modes 4, 5)
};

// kBc7Modes[m] as a compile-time value, for code specialised per mode.
BCN_HD constexpr Bc7ModeInfo bc7_mode_c(int m) {
  return m == 0   ? Bc7ModeInfo{3, 4, 0, 0, 4, 0, 1, 3, 0}
         : m == 1 ? Bc7ModeInfo{2, 6, 0, 0, 6, 0, 2, 3, 0}
         : m == 2 ? Bc7ModeInfo{3, 6, 0, 0, 5, 0, 0, 2, 0}
         : m == 3 ? Bc7ModeInfo{2, 6, 0, 0, 7, 0, 1, 2, 0}
         : m == 4 ? Bc7ModeInfo{1, 0, 2, 1, 5, 6, 0, 2, 3}
         : m == 5 ? Bc7ModeInfo{1, 0, 2, 0, 7, 8, 0, 2, 2}
         : m == 6 ? Bc7ModeInfo{1, 0, 0, 0, 7, 7, 1, 4, 0}
                  : Bc7ModeInfo{2, 6, 0, 0, 5, 5, 1, 2, 0};
}

// Stored bits of channel c (alpha = 3) of the mode; 0 = not stored.
BCN_HD int bc7_stored_bits(const Bc7ModeInfo& mi, int c) { return c < 3 ? mi.cb : mi.ab; }

BCN_HD int bc7_qmax(const Bc7ModeInfo& mi, int c) { return (1 << bc7_stored_bits(mi, c)) - 1; }

// Anchor texels of the form's partition (k == 0 is always one).
BCN_HD void bc7_anchors(const Bc7ModeInfo& mi, uint32_t part, int& a1, int& a2) {
  a1 = mi.ns >= 2 ? kPart.anchor1[mi.ns == 3 ? 64 + part : part] : -1;
  a2 = mi.ns == 3 ? kPart.anchor2[64 + part] : -1;
}

BCN_HD uint32_t bc7_pattern(const Bc7ModeInfo& mi, uint32_t part) {
  return mi.ns == 1 ? 0u : kPart.subset[mi.ns == 3 ? 64 + part : part];
}

// Reads a block's form (the field order of decode_bc7).
BCN_HD void bc7_parse(const uint32_t in[4], Bc7Form& f) {
  uint32_t first = in[0] & 0xFFu;
  f.part = f.rot = f.isb = 0;
  f.idx[0] = f.idx[1] = f.idx2[0] = f.idx2[1] = 0;
#pragma unroll
  for (int e = 0; e < 6; e++) {
    f.p[e] = 0;
#pragma unroll
    for (int c = 0; c < 4; c++) f.q[e][c] = 0;
  }
  if (first == 0) {
    f.mode = 8;
    return;
  }
  f.mode = (uint32_t)__builtin_ctz(first);
  const Bc7ModeInfo mi = kBc7Modes[f.mode];
  BitReader br(in);
  br.skip((uint32_t)f.mode + 1);
  f.part = br.read(mi.pb);
  f.rot = br.read(mi.rb);
  f.isb = br.read(mi.isb);
  const int ne = mi.ns * 2;
#pragma unroll
  for (int c = 0; c < 3; c++)
#pragma unroll
    for (int e = 0; e < 6; e++)
      if (e < ne) f.q[e][c] = (int)br.read(mi.cb);
#pragma unroll
  for (int e = 0; e < 6; e++)
    if (e < ne && mi.ab) f.q[e][3] = (int)br.read(mi.ab);
  if (mi.pbit == 1) {
#pragma unroll
    for (int e = 0; e < 6; e++)
      if (e < ne) f.p[e] = (int)br.read(1);
  } else if (mi.pbit == 2) {
#pragma unroll
    for (int sb = 0; sb < 2; sb++) {
      int b = (int)br.read(1);
      f.p[2 * sb] = f.p[2 * sb + 1] = b;
    }
  }
  int a1, a2;
  bc7_anchors(mi, f.part, a1, a2);
#pragma unroll
  for (int k = 0; k < 16; k++) {
    bool anchor = k == 0 || k == a1 || k == a2;
    f.idx[k >> 3] |= br.read(anchor ? mi.ib - 1 : mi.ib) << (4 * (k & 7));
  }
  if (mi.ib2) {
#pragma unroll
    for (int k = 0; k < 16; k++) f.idx2[k >> 3] |= br.read(k == 0 ? mi.ib2 - 1 : mi.ib2) << (4 * (k & 7));
  }
}

// Writes a form back (inverse of bc7_parse for valid modes).  Anchor indices
// are written without their MSB, which a valid block has clear.
BCN_HD void bc7_pack(const Bc7Form& f, uint32_t out[4]) {
  if (f.mode > 7) {
    out[0] = out[1] = out[2] = out[3] = 0;
    return;
  }
  const Bc7ModeInfo mi = kBc7Modes[f.mode];
  BitWriter w;
  w.put(f.mode + 1, 1u << f.mode);
  w.put(mi.pb, f.part);
  w.put(mi.rb, f.rot);
  w.put(mi.isb, f.isb);
  const int ne = mi.ns * 2;
#pragma unroll
  for (int c = 0; c < 3; c++)
#pragma unroll
    for (int e = 0; e < 6; e++)
      if (e < ne) w.put(mi.cb, (uint32_t)f.q[e][c]);
#pragma unroll
  for (int e = 0; e < 6; e++)
    if (e < ne && mi.ab) w.put(mi.ab, (uint32_t)f.q[e][3]);
  if (mi.pbit == 1) {
#pragma unroll
    for (int e = 0; e < 6; e++)
      if (e < ne) w.put(1, (uint32_t)f.p[e]);
  } else if (mi.pbit == 2) {
    w.put(1, (uint32_t)f.p[0]);
    w.put(1, (uint32_t)f.p[2]);
  }
  int a1, a2;
  bc7_anchors(mi, f.part, a1, a2);
#pragma unroll
  for (int k = 0; k < 16; k++) {
    bool anchor = k == 0 || k == a1 || k == a2;
    w.put(anchor ? mi.ib - 1 : mi.ib, (f.idx[k >> 3] >> (4 * (k & 7))) & 15u);
  }
  if (mi.ib2) {
#pragma unroll
    for (int k = 0; k < 16; k++) w.put(k == 0 ? mi.ib2 - 1 : mi.ib2, (f.idx2[k >> 3] >> (4 * (k & 7))) & 15u);
  }
  out[0] = (uint32_t)w.lo;
  out[1] = (uint32_t)(w.lo >> 32);
  out[2] = (uint32_t)w.hi;
  out[3] = (uint32_t)(w.hi >> 32);
}

// Endpoint byte of quantised value q (the decoder's expansion).
BCN_HD int bc7_ep_byte(const Bc7ModeInfo& mi, int c, int q, int p) {
  int bits = bc7_stored_bits(mi, c);
  if (bits == 0) return 255;
  if (mi.pbit == 1 || (mi.pbit == 2 && c < 3)) return expand_bits((q << 1) | p, bits + 1);
  return expand_bits(q, bits);
}

// d byte / d x of the smooth relaxation, and its value at x.
BCN_HD float bc7_ep_slope(const Bc7ModeInfo& mi, int c) {
  int bits = bc7_stored_bits(mi, c);
  if (bits == 0) return 0.0f;
  bool pb = mi.pbit == 1 || (mi.pbit == 2 && c < 3);
  int n = pb ? bits + 1 : bits;
  return (pb ? 2.0f : 1.0f) * 255.0f / (float)((1 << n) - 1);
}
BCN_HD float bc7_ep_smooth(const Bc7ModeInfo& mi, int c, float x, int p) {
  int bits = bc7_stored_bits(mi, c);
  if (bits == 0) return 255.0f;
  bool pb = mi.pbit == 1 || (mi.pbit == 2 && c < 3);
  int n = pb ? bits + 1 : bits;
  return (pb ? 2.0f * x + (float)p : x) * (255.0f / (float)((1 << n) - 1));
}

BCN_HD int bc7_texel_subset(uint32_t pattern, int k) { return (int)((pattern >> (2 * k)) & 3u); }

BCN_HD void bc7_texel_weights(const Bc7Form& f, const Bc7ModeInfo& mi, int k, int& wc, int& wa) {
  int i1 = (int)((f.idx[k >> 3] >> (4 * (k & 7))) & 15u);
  int w = bptc_weight_rt(mi.ib, i1);
  if (mi.ib2 == 0) {
    wc = wa = w;
  } else {
    int i2 = (int)((f.idx2[k >> 3] >> (4 * (k & 7))) & 15u);
    int w2 = bptc_weight_rt(mi.ib2, i2);
    wc = f.isb ? w2 : w;
    wa = f.isb ? w : w2;
  }
}

// Output channel o of the decoded texel reads internal channel bc7_rot_src(rot, o).
BCN_HD int bc7_rot_src(uint32_t rot, int o) {
  if (rot == 0) return o;
  if (o == 3) return (int)rot - 1;
  return o == (int)rot - 1 ? 3 : o;
}

// Forward: texels[k][o] (0..255 units) of endpoint parameters x[6][4].
// smooth = false: straight-through (exact decode of clamp(rint(x))).
BCN_HD void bc7_forward(const Bc7Form& f, const float x[6][4], bool smooth, float texels[16][4]) {
  if (f.mode > 7) {
#pragma unroll
    for (int k = 0; k < 16; k++)
#pragma unroll
      for (int o = 0; o < 4; o++) texels[k][o] = 0.0f;
    return;
  }
  const Bc7ModeInfo mi = kBc7Modes[f.mode];
  float eb[6][4];
#pragma unroll
  for (int e = 0; e < 6; e++)
#pragma unroll
    for (int c = 0; c < 4; c++) {
      if (smooth) {
        eb[e][c] = bc7_ep_smooth(mi, c, x[e][c], f.p[e]);
      } else {
        int qm = bc7_qmax(mi, c);
        int q = (int)fminf(fmaxf(rintf(x[e][c]), 0.0f), (float)qm);
        eb[e][c] = (float)bc7_ep_byte(mi, c, q, f.p[e]);
      }
    }
  const uint32_t pattern = bc7_pattern(mi, f.part);
#pragma unroll
  for (int k = 0; k < 16; k++) {
    int s = bc7_texel_subset(pattern, k);
    int wc, wa;
    bc7_texel_weights(f, mi, k, wc, wa);
    float c4[4];
#pragma unroll
    for (int c = 0; c < 4; c++) {
      float lo = s == 0 ? eb[0][c] : (s == 1 ? eb[2][c] : eb[4][c]);
      float hi = s == 0 ? eb[1][c] : (s == 1 ? eb[3][c] : eb[5][c]);
      int w = c < 3 ? wc : wa;
      if (smooth) {
        c4[c] = ((float)(64 - w) * lo + (float)w * hi) / 64.0f;
      } else {
        c4[c] = (float)interp((int)lo, (int)hi, w);
      }
    }
#pragma unroll
    for (int o = 0; o < 4; o++) {
      int src = bc7_rot_src(f.rot, o);
      texels[k][o] = src == 0 ? c4[0] : (src == 1 ? c4[1] : (src == 2 ? c4[2] : c4[3]));
    }
  }
}

// Adjoint: grad_x[6][4] = J^T g for g[k][o] = dL/d texels[k][o].
BCN_HD void bc7_backward(const Bc7Form& f, const float g[16][4], float grad_x[6][4]) {
#pragma unroll
  for (int e = 0; e < 6; e++)
#pragma unroll
    for (int c = 0; c < 4; c++) grad_x[e][c] = 0.0f;
  if (f.mode > 7) return;
  const Bc7ModeInfo mi = kBc7Modes[f.mode];
  const uint32_t pattern = bc7_pattern(mi, f.part);
  float acc[6][4];
#pragma unroll
  for (int e = 0; e < 6; e++)
#pragma unroll
    for (int c = 0; c < 4; c++) acc[e][c] = 0.0f;
#pragma unroll
  for (int k = 0; k < 16; k++) {
    int s = bc7_texel_subset(pattern, k);
    int wc, wa;
    bc7_texel_weights(f, mi, k, wc, wa);
    // Internal channel c of this texel feeds output channel o with src(o) == c.
    float gi[4];
#pragma unroll
    for (int c = 0; c < 4; c++) {
      float v = 0.0f;
#pragma unroll
      for (int o = 0; o < 4; o++) v = bc7_rot_src(f.rot, o) == c ? g[k][o] : v;
      gi[c] = v;
    }
#pragma unroll
    for (int c = 0; c < 4; c++) {
      int w = c < 3 ? wc : wa;
      float glo = gi[c] * ((float)(64 - w) / 64.0f), ghi = gi[c] * ((float)w / 64.0f);
      acc[0][c] += s == 0 ? glo : 0.0f;
      acc[1][c] += s == 0 ? ghi : 0.0f;
      acc[2][c] += s == 1 ? glo : 0.0f;
      acc[3][c] += s == 1 ? ghi : 0.0f;
      acc[4][c] += s == 2 ? glo : 0.0f;
      acc[5][c] += s == 2 ? ghi : 0.0f;
    }
  }
  const int ne = mi.ns * 2;
#pragma unroll
  for (int e = 0; e < 6; e++)
#pragma unroll
    for (int c = 0; c < 4; c++) grad_x[e][c] = e < ne ? acc[e][c] * bc7_ep_slope(mi, c) : 0.0f;
}

// Refinement settings (mirrors bcn_refine_params).
struct RefineParams {
  uint32_t steps;
  float lr, beta1, beta2, eps;
  uint32_t channels;  // 3: RGB error, 4: RGBA error
  uint32_t reassign;  // re-pick every index after each `reassign` steps (0: indices fixed)
  uint32_t pbits;     // try flipping each p-bit after each `pbits` steps, keeping flips that help (0: never)
};

// One pass of the refinement at parameters x: the exact (straight-through)
// decode, its squared error against the target over the first `channels`
// output channels, and grad[6][4], the adjoint of that error (d/dx of
// sum r^2 through the relaxation).  Target value (texel k, channel o) is
// target[(4k + o) * tstride].  Returns the squared error.
template <bool GRAD>
BCN_HD float bc7_refine_pass(const Bc7Form& f, const Bc7ModeInfo& mi, uint32_t pattern,
                             const float x[6][4], const float* target, int tstride, int channels,
                             float grad[6][4]) {
  int eb[6][4];
#pragma unroll
  for (int e = 0; e < 6; e++)
#pragma unroll
    for (int c = 0; c < 4; c++) {
      int q = (int)fminf(fmaxf(rintf(x[e][c]), 0.0f), (float)bc7_qmax(mi, c));
      eb[e][c] = bc7_ep_byte(mi, c, q, f.p[e]);
      grad[e][c] = 0.0f;
    }
  // Squared error per internal channel, combined as (c0 + c1) + (c2 + c3):
  // the order a four-lane-per-block kernel reduces in.
  float sse_c[4] = {0.0f, 0.0f, 0.0f, 0.0f};
  // Per-texel subsets and weights are re-derived from laundered words in
  // every pass (a few instructions) rather than hoisted out of the step loop
  // by the compiler, where they would hold 48 registers.
  Bc7Form fl = f;
  fl.idx[0] = fresh(f.idx[0]);
  fl.idx[1] = fresh(f.idx[1]);
  fl.idx2[0] = fresh(f.idx2[0]);
  fl.idx2[1] = fresh(f.idx2[1]);
  const uint32_t pat = fresh(pattern);
  const float* tg = fresh_ptr(target);
#pragma unroll
  for (int k = 0; k < 16; k++) {
    int s = bc7_texel_subset(pat, k);
    int wc, wa;
    bc7_texel_weights(fl, mi, k, wc, wa);
#pragma unroll
    for (int c = 0; c < 4; c++) {
      // Output channel o that internal channel c lands in.
      int o = f.rot == 0 ? c : (c == 3 ? (int)f.rot - 1 : (c == (int)f.rot - 1 ? 3 : c));
      int lo = s == 0 ? eb[0][c] : (s == 1 ? eb[2][c] : eb[4][c]);
      int hi = s == 0 ? eb[1][c] : (s == 1 ? eb[3][c] : eb[5][c]);
      int w = c < 3 ? wc : wa;
      float r = o < channels ? (float)interp(lo, hi, w) - tg[(4 * k + o) * tstride] : 0.0f;
      sse_c[c] += r * r;
      if (!GRAD) continue;
      float glo = 2.0f * r * ((float)(64 - w) / 64.0f), ghi = 2.0f * r * ((float)w / 64.0f);
      grad[0][c] += s == 0 ? glo : 0.0f;
      grad[1][c] += s == 0 ? ghi : 0.0f;
      grad[2][c] += s == 1 ? glo : 0.0f;
      grad[3][c] += s == 1 ? ghi : 0.0f;
      grad[4][c] += s == 2 ? glo : 0.0f;
      grad[5][c] += s == 2 ? ghi : 0.0f;
    }
  }
  if (GRAD) {
    const int ne = mi.ns * 2;
#pragma unroll
    for (int e = 0; e < 6; e++)
#pragma unroll
      for (int c = 0; c < 4; c++) grad[e][c] = e < ne ? grad[e][c] * bc7_ep_slope(mi, c) : 0.0f;
  }
  return (sse_c[0] + sse_c[1]) + (sse_c[2] + sse_c[3]);
}

// Index of texel k in a 4-bit-per-texel word pair.  Written with a select
// between the two words rather than w[k >> 3], so that a loop over k that is
// not unrolled does not force the array (and the form holding it) out of
// registers.
BCN_HD uint32_t nib(const uint32_t w[2], int k) {
  uint32_t word = k < 8 ? w[0] : w[1];
  return (word >> (4 * (k & 7))) & 15u;
}
BCN_HD void set_nib(uint32_t w[2], int k, uint32_t v) {
  const uint32_t sh = 4u * (uint32_t)(k & 7), keep = ~(15u << sh);
  const uint32_t w0 = (w[0] & keep) | (v << sh), w1 = (w[1] & keep) | (v << sh);
  w[0] = k < 8 ? w0 : w[0];
  w[1] = k < 8 ? w[1] : w1;
}

// For the endpoints q = clamp(rint(x)), gives every texel the index with the
// smallest exact error (per index set: the colour channels and, in modes 4/5,
// alpha on its own).  Anchor indices may come out with their MSB set;
// bc7_normalize_anchors fixes that before packing.
BCN_HD void bc7_reassign_indices(Bc7Form& f, const Bc7ModeInfo& mi, uint32_t pattern,
                                 const float x[6][4], const float* target, int tstride, int channels) {
  int eb[6][4];
#pragma unroll
  for (int e = 0; e < 6; e++)
#pragma unroll
    for (int c = 0; c < 4; c++) {
      int q = (int)fminf(fmaxf(rintf(x[e][c]), 0.0f), (float)bc7_qmax(mi, c));
      eb[e][c] = bc7_ep_byte(mi, c, q, f.p[e]);
    }
  const float* tg = fresh_ptr(target);
  // Index set A (bits ib, weights from idx) covers colour (and alpha when
  // ib2 == 0, or alpha alone when isb selects it); set B (idx2) the rest.
  const int bits_a = mi.ib, bits_b = mi.ib2;
  for (int k = 0; k < 16; k++) {
    int s = bc7_texel_subset(pattern, k);
    int lo[4], hi[4];
    float t[4];
    bool scored[4];
#pragma unroll
    for (int c = 0; c < 4; c++) {
      lo[c] = s == 0 ? eb[0][c] : (s == 1 ? eb[2][c] : eb[4][c]);
      hi[c] = s == 0 ? eb[1][c] : (s == 1 ? eb[3][c] : eb[5][c]);
      int o = f.rot == 0 ? c : (c == 3 ? (int)f.rot - 1 : (c == (int)f.rot - 1 ? 3 : c));
      scored[c] = o < channels;
      t[c] = scored[c] ? tg[(4 * k + o) * tstride] : 0.0f;
    }
    // Channels scored by set A: c in [0, 4) if no second set; else colour
    // (c < 3) when isb == 0, alpha (c == 3) when isb == 1.
    float best_a = 3.4e38f;
    uint32_t ia = 0;
    for (int i = 0; i < (1 << bits_a); i++) {
      int w = bptc_weight_rt(bits_a, i);
      float ec[4];
#pragma unroll
      for (int c = 0; c < 4; c++) {
        bool in_a = bits_b == 0 || (f.isb ? c == 3 : c < 3);
        float d = (float)interp(lo[c], hi[c], w) - t[c];
        ec[c] = (in_a && scored[c]) ? d * d : 0.0f;
      }
      float err = (ec[0] + ec[1]) + (ec[2] + ec[3]);
      if (err < best_a) {
        best_a = err;
        ia = (uint32_t)i;
      }
    }
    set_nib(f.idx, k, ia);
    if (bits_b) {
      float best_b = 3.4e38f;
      uint32_t ib = 0;
      for (int i = 0; i < (1 << bits_b); i++) {
        int w = bptc_weight_rt(bits_b, i);
        float ec[4];
#pragma unroll
        for (int c = 0; c < 4; c++) {
          bool in_b = f.isb ? c < 3 : c == 3;
          float d = (float)interp(lo[c], hi[c], w) - t[c];
          ec[c] = (in_b && scored[c]) ? d * d : 0.0f;
        }
        float err = (ec[0] + ec[1]) + (ec[2] + ec[3]);
        if (err < best_b) {
          best_b = err;
          ib = (uint32_t)i;
        }
      }
      set_nib(f.idx2, k, ib);
    }
  }
}

// Tries each p-bit of the form flipped (per endpoint, or per subset where the
// mode shares one) at endpoints x and keeps a flip when the exact error drops.
BCN_HD void bc7_flip_pbits(Bc7Form& f, const Bc7ModeInfo& mi, uint32_t pattern, const float x[6][4],
                           const float* target, int tstride, int channels) {
  if (mi.pbit == 0) return;
  float g[6][4];
  float cur = bc7_refine_pass<false>(f, mi, pattern, x, target, tstride, channels, g);
  const int slots = mi.pbit == 1 ? 2 * mi.ns : mi.ns;
  // Unrolled with constant indices so that the form stays in registers.
#pragma unroll
  for (int i = 0; i < 6; i++) {
    if (i >= slots) continue;
    const int e0 = mi.pbit == 1 ? i : 2 * i, e1 = mi.pbit == 1 ? i : 2 * i + 1;
    f.p[e0] ^= 1;
    if (e1 != e0) f.p[e1] ^= 1;
    float sse = bc7_refine_pass<false>(f, mi, pattern, x, target, tstride, channels, g);
    if (sse < cur) {
      cur = sse;
    } else {
      f.p[e0] ^= 1;
      if (e1 != e0) f.p[e1] ^= 1;
    }
  }
}

// Makes a form encodable: BC7 stores each anchor index without its MSB, so
// where an anchor's index has the MSB set, that subset's endpoints (and
// per-endpoint p-bits) are exchanged and its indices mirrored, which decodes
// to the same texels.  For modes 4/5 each index set is handled with the
// channels it drives.
BCN_HD void bc7_normalize_anchors(Bc7Form& f) {
  if (f.mode > 7) return;
  const Bc7ModeInfo mi = kBc7Modes[f.mode];
  const uint32_t pattern = bc7_pattern(mi, f.part);
  int a1, a2;
  bc7_anchors(mi, f.part, a1, a2);
  // Loops unrolled with constant indices so that the form stays in registers.
#pragma unroll
  for (int s = 0; s < 3; s++) {
    if (s >= (int)mi.ns) continue;
    int a = s == 0 ? 0 : (s == 1 ? a1 : a2);
    const uint32_t top = (1u << mi.ib) - 1;
    if (nib(f.idx, a) <= top / 2) continue;
#pragma unroll
    for (int c = 0; c < 4; c++) {
      bool driven = mi.ib2 == 0 || (f.isb ? c == 3 : c < 3);
      int lo = f.q[2 * s][c], hi = f.q[2 * s + 1][c];
      f.q[2 * s][c] = driven ? hi : lo;
      f.q[2 * s + 1][c] = driven ? lo : hi;
    }
    int tp = f.p[2 * s];
    f.p[2 * s] = f.p[2 * s + 1];
    f.p[2 * s + 1] = tp;
#pragma unroll
    for (int k = 0; k < 16; k++)
      if (bc7_texel_subset(pattern, k) == s) set_nib(f.idx, k, top - nib(f.idx, k));
  }
  if (mi.ib2) {
    const uint32_t top = (1u << mi.ib2) - 1;
    if (nib(f.idx2, 0) > top / 2) {
#pragma unroll
      for (int c = 0; c < 4; c++) {
        bool driven = f.isb ? c < 3 : c == 3;
        int lo = f.q[0][c], hi = f.q[1][c];
        f.q[0][c] = driven ? hi : lo;
        f.q[1][c] = driven ? lo : hi;
      }
#pragma unroll
      for (int k = 0; k < 16; k++) set_nib(f.idx2, k, top - nib(f.idx2, k));
    }
  }
}

// Gradient endpoint refinement of one block of mode M (form f, parsed from
// `in`) against the target (0..255 units, element (k, o) at
// target[(4k + o) * tstride]): Adam on x from the block's own endpoints,
// straight-through forward, the adjoint above for the gradient of the squared
// error, with optional index / p-bit reassignment.  Every iterate is scored
// by its exact decode, and the best one (the start included) is kept, so the
// result never has a larger error than the input block.  Returns the error
// of the kept block; *sse0 gets the input's.  Specialised per mode: only the
// endpoint values the mode stores are parameters.
template <int M>
BCN_HD float bc7_refine_form(Bc7Form& f, const uint32_t in[4], const float* target, int tstride,
                             const RefineParams& rp, uint32_t out[4], float* sse0) {
  constexpr Bc7ModeInfo mi = bc7_mode_c(M);
  constexpr int ne = 2 * mi.ns;
  out[0] = in[0];
  out[1] = in[1];
  out[2] = in[2];
  out[3] = in[3];
  const uint32_t pattern = bc7_pattern(mi, f.part);
  const int channels = (int)rp.channels;
  float x[6][4], m[6][4], v[6][4], g[6][4];
  uint32_t bq[6];  // best endpoints, channel c in byte c
#pragma unroll
  for (int e = 0; e < 6; e++) {
    bq[e] = 0;
#pragma unroll
    for (int c = 0; c < 4; c++) {
      x[e][c] = (float)f.q[e][c];
      m[e][c] = 0.0f;
      v[e][c] = 0.0f;
      bq[e] |= (uint32_t)f.q[e][c] << (8 * c);
    }
  }
  float best = bc7_refine_pass<true>(f, mi, pattern, x, target, tstride, channels, g);
  *sse0 = best;
  const float best0 = best;
  uint32_t bidx[2] = {f.idx[0], f.idx[1]}, bidx2[2] = {f.idx2[0], f.idx2[1]};
  uint32_t bp = 0;  // best p-bits, bit e
#pragma unroll
  for (int e = 0; e < 6; e++) bp |= (uint32_t)f.p[e] << e;
  float b1t = 1.0f, b2t = 1.0f;
  for (uint32_t it = 0; it < rp.steps; it++) {
    b1t *= rp.beta1;
    b2t *= rp.beta2;
    const float c1 = 1.0f / (1.0f - b1t), c2 = 1.0f / (1.0f - b2t);
#pragma unroll
    for (int e = 0; e < ne; e++)
#pragma unroll
      for (int c = 0; c < 4; c++) {
        if (bc7_stored_bits(mi, c) == 0) continue;
        m[e][c] = rp.beta1 * m[e][c] + (1.0f - rp.beta1) * g[e][c];
        v[e][c] = rp.beta2 * v[e][c] + (1.0f - rp.beta2) * g[e][c] * g[e][c];
        float nx = x[e][c] - rp.lr * (m[e][c] * c1) / (sqrtf(v[e][c] * c2) + rp.eps);
        x[e][c] = fminf(fmaxf(nx, -0.49f), (float)bc7_qmax(mi, c) + 0.49f);
      }
    if (rp.reassign && (it + 1) % rp.reassign == 0)
      bc7_reassign_indices(f, mi, pattern, x, target, tstride, channels);
    if (mi.pbit && rp.pbits && (it + 1) % rp.pbits == 0)
      bc7_flip_pbits(f, mi, pattern, x, target, tstride, channels);
    float sse = bc7_refine_pass<true>(f, mi, pattern, x, target, tstride, channels, g);
    if (sse < best) {
      best = sse;
#pragma unroll
      for (int e = 0; e < ne; e++) {
        bq[e] = 0;
#pragma unroll
        for (int c = 0; c < 4; c++)
          bq[e] |= (uint32_t)fminf(fmaxf(rintf(x[e][c]), 0.0f), (float)bc7_qmax(mi, c)) << (8 * c);
      }
      bidx[0] = f.idx[0];
      bidx[1] = f.idx[1];
      bidx2[0] = f.idx2[0];
      bidx2[1] = f.idx2[1];
      bp = 0;
#pragma unroll
      for (int e = 0; e < 6; e++) bp |= (uint32_t)f.p[e] << e;
    }
  }
  if (best < best0) {
#pragma unroll
    for (int e = 0; e < 6; e++)
#pragma unroll
      for (int c = 0; c < 4; c++) f.q[e][c] = (int)((bq[e] >> (8 * c)) & 0xFFu);
    f.idx[0] = bidx[0];
    f.idx[1] = bidx[1];
    f.idx2[0] = bidx2[0];
    f.idx2[1] = bidx2[1];
#pragma unroll
    for (int e = 0; e < 6; e++) f.p[e] = (int)((bp >> e) & 1u);
    bc7_normalize_anchors(f);
    bc7_pack(f, out);
  }
  return best;
}

// Error of a block with no valid mode (it decodes to zero); the block is kept.
BCN_HD float bc7_refine_invalid(const uint32_t in[4], const float* target, int tstride,
                                const RefineParams& rp, uint32_t out[4], float* sse0) {
  float s = 0.0f;
#pragma unroll
  for (int k = 0; k < 16; k++)
#pragma unroll
    for (int o = 0; o < 4; o++) {
      float t = target[(4 * k + o) * tstride];
      s += o < (int)rp.channels ? t * t : 0.0f;
    }
  out[0] = in[0];
  out[1] = in[1];
  out[2] = in[2];
  out[3] = in[3];
  *sse0 = s;
  return s;
}

// bc7_refine_form for any block (parses and dispatches on the mode).
BCN_HD float bc7_refine_block(const uint32_t in[4], const float* target, int tstride,
                              const RefineParams& rp, uint32_t out[4], float* sse0) {
  Bc7Form f;
  bc7_parse(in, f);
  switch (f.mode) {
    case 0: return bc7_refine_form<0>(f, in, target, tstride, rp, out, sse0);
    case 1: return bc7_refine_form<1>(f, in, target, tstride, rp, out, sse0);
    case 2: return bc7_refine_form<2>(f, in, target, tstride, rp, out, sse0);
    case 3: return bc7_refine_form<3>(f, in, target, tstride, rp, out, sse0);
    case 4: return bc7_refine_form<4>(f, in, target, tstride, rp, out, sse0);
    case 5: return bc7_refine_form<5>(f, in, target, tstride, rp, out, sse0);
    case 6: return bc7_refine_form<6>(f, in, target, tstride, rp, out, sse0);
    case 7: return bc7_refine_form<7>(f, in, target, tstride, rp, out, sse0);
    default: return bc7_refine_invalid(in, target, tstride, rp, out, sse0);
  }
}

// ---- BC1 -------------------------------------------------------------------------
//
// Form: the index word and the colour mode (4 colours when c0 > c1 as 565
// words, else 3 colours and transparent black), held fixed.  Parameters:
// x[e][c], endpoint e in {0, 1}, channel c in {R, G, B}, in 565 units
// (0..31, 0..63, 0..31).  Straight-through forward: q = clamp(rint(x)) and the
// decoder's integer arithmetic for the FIXED mode.  Smooth forward: byte =
// x * 255 / (2^bits - 1); the third and fourth colours (2 lo + hi) / 3 and
// (lo + 2 hi) / 3, or (lo + hi) / 2 and 0 in 3-colour mode.  Alpha carries no
// parameter.

struct Bc1Form {
  uint32_t four;   // 1: four-colour mode
  int q[2][3];
  uint32_t idx;    // 2 bits per texel
};

BCN_HD void bc1_parse(const uint32_t in[2], Bc1Form& f) {
  uint32_t c0 = in[0] & 0xFFFFu, c1 = in[0] >> 16;
  f.four = c0 > c1 ? 1u : 0u;
  f.q[0][0] = (int)((c0 >> 11) & 31);
  f.q[0][1] = (int)((c0 >> 5) & 63);
  f.q[0][2] = (int)(c0 & 31);
  f.q[1][0] = (int)((c1 >> 11) & 31);
  f.q[1][1] = (int)((c1 >> 5) & 63);
  f.q[1][2] = (int)(c1 & 31);
  f.idx = in[1];
}

BCN_HD int bc1_qmax(int c) { return c == 1 ? 63 : 31; }

BCN_HD void bc1_forward(const Bc1Form& f, const float x[2][3], bool smooth, float texels[16][4]) {
  float col[4][4];
  if (smooth) {
    float b[2][3];
#pragma unroll
    for (int e = 0; e < 2; e++)
#pragma unroll
      for (int c = 0; c < 3; c++) b[e][c] = x[e][c] * (255.0f / (float)bc1_qmax(c));
#pragma unroll
    for (int c = 0; c < 3; c++) {
      col[0][c] = b[0][c];
      col[1][c] = b[1][c];
      col[2][c] = f.four ? (2.0f * b[0][c] + b[1][c]) / 3.0f : (b[0][c] + b[1][c]) / 2.0f;
      col[3][c] = f.four ? (b[0][c] + 2.0f * b[1][c]) / 3.0f : 0.0f;
    }
  } else {
    int q[2][3];
#pragma unroll
    for (int e = 0; e < 2; e++)
#pragma unroll
      for (int c = 0; c < 3; c++) q[e][c] = (int)fminf(fmaxf(rintf(x[e][c]), 0.0f), (float)bc1_qmax(c));
    // decode_color_block (block.rs:124-197) with the mode held fixed.
#pragma unroll
    for (int e = 0; e < 2; e++) {
      col[e][0] = (float)((q[e][0] * 527 + 23) >> 6);
      col[e][1] = (float)((q[e][1] * 259 + 33) >> 6);
      col[e][2] = (float)((q[e][2] * 527 + 23) >> 6);
    }
    if (f.four) {
      col[2][0] = (float)(((2 * q[0][0] + q[1][0]) * 351 + 61) >> 7);
      col[2][1] = (float)(((2 * q[0][1] + q[1][1]) * 2763 + 1039) >> 11);
      col[2][2] = (float)(((2 * q[0][2] + q[1][2]) * 351 + 61) >> 7);
      col[3][0] = (float)(((q[0][0] + 2 * q[1][0]) * 351 + 61) >> 7);
      col[3][1] = (float)(((q[0][1] + 2 * q[1][1]) * 2763 + 1039) >> 11);
      col[3][2] = (float)(((q[0][2] + 2 * q[1][2]) * 351 + 61) >> 7);
    } else {
      col[2][0] = (float)(((q[0][0] + q[1][0]) * 1053 + 125) >> 8);
      col[2][1] = (float)(((q[0][1] + q[1][1]) * 4145 + 1019) >> 11);
      col[2][2] = (float)(((q[0][2] + q[1][2]) * 1053 + 125) >> 8);
      col[3][0] = col[3][1] = col[3][2] = 0.0f;
    }
  }
  col[0][3] = col[1][3] = col[2][3] = 255.0f;
  col[3][3] = f.four ? 255.0f : 0.0f;
#pragma unroll
  for (int k = 0; k < 16; k++) {
    uint32_t i = (f.idx >> (2 * k)) & 3u;
#pragma unroll
    for (int c = 0; c < 4; c++)
      texels[k][c] = i == 0 ? col[0][c] : (i == 1 ? col[1][c] : (i == 2 ? col[2][c] : col[3][c]));
  }
}

BCN_HD void bc1_backward(const Bc1Form& f, const float g[16][4], float grad_x[2][3]) {
  // d colour_i / d byte_{lo,hi} per palette entry.
  const float w2lo = f.four ? 2.0f / 3.0f : 0.5f, w2hi = f.four ? 1.0f / 3.0f : 0.5f;
  const float w3lo = f.four ? 1.0f / 3.0f : 0.0f, w3hi = f.four ? 2.0f / 3.0f : 0.0f;
  float acc[2][3] = {{0.0f, 0.0f, 0.0f}, {0.0f, 0.0f, 0.0f}};
#pragma unroll
  for (int k = 0; k < 16; k++) {
    uint32_t i = (f.idx >> (2 * k)) & 3u;
    float dlo = i == 0 ? 1.0f : (i == 1 ? 0.0f : (i == 2 ? w2lo : w3lo));
    float dhi = i == 0 ? 0.0f : (i == 1 ? 1.0f : (i == 2 ? w2hi : w3hi));
#pragma unroll
    for (int c = 0; c < 3; c++) {
      acc[0][c] += g[k][c] * dlo;
      acc[1][c] += g[k][c] * dhi;
    }
  }
#pragma unroll
  for (int e = 0; e < 2; e++)
#pragma unroll
    for (int c = 0; c < 3; c++) grad_x[e][c] = acc[e][c] * (255.0f / (float)bc1_qmax(c));
}

// ---- texel-side adjoint -------------------------------------------------------
//
// For texture training: the decoded texels as a function of the *source*
// texels, with the block's form held fixed and its endpoints re-fitted to the
// source by least squares -- the fit the reference's refinement step makes
// (opt_endpoints, src/encode/common.rs:142-206), here without quantisation.
// For one subset and one index set the fitted values are the orthogonal
// projection of the subset's source values onto span{1, w} (w_k: texel k's
// interpolation weight), so the map is linear, its Jacobian is that
// projection, and the adjoint is the same projection (it is symmetric):
//     (H g)_k = mean(g) + (w_k - mean(w)) * S_wg / S_ww,
//     S_wg = sum_j (w_j - mean(w)) g_j,  S_ww = sum_j (w_j - mean(w))^2,
// over the subset's texels, and mean(g) when every w is the same (the fit is
// then the mean).  span{1, w} does not change under w -> a w + b, so the
// weights are used in their integer units.  Decoded channels that do not
// depend on the source (alpha of BC7 modes 0-3 and of BC1, BC1's 3-colour
// black) and blocks with no valid mode get zero.

// r[k] = (H g)_k for the texels k of `member` (bit k); other r[k] untouched.
// Sums in texel order.
BCN_HD void project_span_1w(uint32_t member, const float w[16], const float g[16], float r[16]) {
  float n = 0.0f, sw = 0.0f, sg = 0.0f;
#pragma unroll
  for (int k = 0; k < 16; k++) {
    if ((member >> k) & 1u) {
      n += 1.0f;
      sw += w[k];
      sg += g[k];
    }
  }
  if (n == 0.0f) return;
  const float mw = sw / n, mg = sg / n;
  float sww = 0.0f, swg = 0.0f;
#pragma unroll
  for (int k = 0; k < 16; k++) {
    if ((member >> k) & 1u) {
      const float d = w[k] - mw;
      sww += d * d;
      swg += d * g[k];
    }
  }
  const float beta = sww > 0.0f ? swg / sww : 0.0f;
#pragma unroll
  for (int k = 0; k < 16; k++)
    if ((member >> k) & 1u) r[k] = mg + (w[k] - mw) * beta;
}

// gt[k][o] = dL/d source[k][o] for g[k][o] = dL/d decoded[k][o] (BC7).  The
// rotation swaps a colour channel with alpha on the way in and back on the
// way out, so output channel o is fitted from source channel o with the
// weights of internal channel bc7_rot_src(rot, o).
BCN_HD void bc7_texel_backward(const Bc7Form& f, const float g[16][4], float gt[16][4]) {
#pragma unroll
  for (int k = 0; k < 16; k++)
#pragma unroll
    for (int o = 0; o < 4; o++) gt[k][o] = 0.0f;
  if (f.mode > 7) return;
  const Bc7ModeInfo mi = kBc7Modes[f.mode];
  const uint32_t pattern = bc7_pattern(mi, f.part);
  float wc[16], wa[16];
  uint32_t member[3] = {0u, 0u, 0u};
#pragma unroll
  for (int k = 0; k < 16; k++) {
    int a, b;
    bc7_texel_weights(f, mi, k, a, b);
    wc[k] = (float)a;
    wa[k] = (float)b;
    const int s = bc7_texel_subset(pattern, k);
    member[0] |= (s == 0 ? 1u : 0u) << k;
    member[1] |= (s == 1 ? 1u : 0u) << k;
    member[2] |= (s == 2 ? 1u : 0u) << k;
  }
#pragma unroll
  for (int o = 0; o < 4; o++) {
    const int c = bc7_rot_src(f.rot, o);
    if (bc7_stored_bits(mi, c) == 0) continue;  // alpha 255 whatever the source
    float w[16], go[16], r[16];
#pragma unroll
    for (int k = 0; k < 16; k++) {
      w[k] = c < 3 ? wc[k] : wa[k];
      go[k] = g[k][o];
      r[k] = 0.0f;
    }
#pragma unroll
    for (int s = 0; s < 3; s++)
      if (s < (int)mi.ns) project_span_1w(member[s], w, go, r);
#pragma unroll
    for (int k = 0; k < 16; k++) gt[k][o] = r[k];
  }
}

// The same for BC1: one subset, palette weights in thirds (four-colour mode:
// indices 0, 1, 2, 3 at 0, 3, 1, 2) or halves (three-colour: 0, 2, 1; index 3
// is black whatever the source and is left out of the fit).
BCN_HD void bc1_texel_backward(const Bc1Form& f, const float g[16][4], float gt[16][4]) {
  float w[16];
  uint32_t member = 0u;
#pragma unroll
  for (int k = 0; k < 16; k++) {
    const uint32_t i = (f.idx >> (2 * k)) & 3u;
    w[k] = i == 0 ? 0.0f : (i == 1 ? (f.four ? 3.0f : 2.0f) : (i == 2 ? 1.0f : 2.0f));
    member |= (f.four || i != 3u ? 1u : 0u) << k;
  }
#pragma unroll
  for (int k = 0; k < 16; k++) gt[k][3] = 0.0f;
#pragma unroll
  for (int c = 0; c < 3; c++) {
    float go[16], r[16];
#pragma unroll
    for (int k = 0; k < 16; k++) {
      go[k] = g[k][c];
      r[k] = 0.0f;
    }
    project_span_1w(member, w, go, r);
#pragma unroll
    for (int k = 0; k < 16; k++) gt[k][c] = r[k];
  }
}

}  // namespace bcn
