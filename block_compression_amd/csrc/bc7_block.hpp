// BC7 block encoder: one 4x4 block per lane.
//
// Follows the reference crate's BC7 search (src/encode/bc7.rs, with the shared
// primitives of src/encode/common.rs) decision for decision, so the emitted
// 128-bit block is byte-identical to `compress_rgba8` on the CPU:
//   mode 0 (+2) over partitions -> mode 1/3 (and 7) over the partitions with
//   the smallest PCA split bound -> mode 4/5 over rotations and index
//   precision -> mode 6, a later mode winning only with a strictly smaller
//   error (bc7.rs:853-867).
//
// GPU-side choices, none of which changes a result:
//   * texels stay packed (one RGBA8 dword per texel, 16 registers);
//   * every error the search compares is a sum of squared differences of
//     integers with total below 2^24 (at most 16 texels x 4 channels x 255^2),
//     so it is exact in f32 and is computed in int32; the same holds for the
//     masked moments and the least-squares sums (bounds stated where used);
//   * the winner's parameters are kept and the bitstream is written once at
//     the end; anchor index MSBs (zero after the endpoint swap) are simply not
//     emitted, which yields the same bits as the reference's write-then-shift
//     (common.rs:549-562);
//   * the 64 mode-1/3 partition keys live in a caller-provided scratch row
//     (LDS on the GPU) and the top-k are selected by ascending key; keys are
//     unique (partition id in the low 6 bits), so this equals the reference's
//     partial selection sort (common.rs:126-140).
#pragma once
#include "bcn_common.hpp"

// Occupancy target of kernels running the BC7 search, in waves per SIMD.
// Unconstrained, the block search took about 240 VGPRs (2 waves); at 3 waves
// it fits the 168-VGPR cap (167 used, no spill in the per-format kernel).  On
// 4096x4096 opaque_basic, 2/3/4 waves ran 7.04/6.91/6.95 ms per launch with
// an earlier version of the search.  The 8 KiB of LDS per 64-lane workgroup
// allows 5.
#ifndef BCN_BC7_WAVES
#define BCN_BC7_WAVES 3
#endif

namespace bcn {

// BC7Settings, 19 x u32 in the reference's repr(C) order (src/settings.rs:80-90).
struct Bc7Params {
  uint32_t refine_iterations[8];
  uint32_t mode_selection[4];
  uint32_t skip_mode2;
  uint32_t fast_skip_threshold_mode1;
  uint32_t fast_skip_threshold_mode3;
  uint32_t fast_skip_threshold_mode7;
  uint32_t mode45_channel0;
  uint32_t refine_iterations_channel;
  uint32_t channels;
};

// Current winner.  qep holds endpoint e (0..5) channel p (0..3) at byte e*4+p
// of word e; for modes 4/5 the scalar channel's two endpoints sit in the
// channel-3 bytes of endpoints 0 and 1.  A candidate replaces the winner if
// its error is smaller, or equal with an earlier position in the reference's
// search order (bc7.rs:853-867, :740-753): that is the reference's strict-<,
// first-wins rule, independent of the order the kernel evaluates candidates.
struct Bc7Best {
  int err;
  int order;  // position of the candidate in the reference's search order
  uint32_t mode, part, rotation, swap;
  uint32_t qep[6];
  uint32_t qblock[2];
  uint32_t aqblock[2];
};

// ---- PCA fit (common.rs:294-394, 483-518) -------------------------------------

// Power iteration with renormalisation every second step (common.rs:260-292).
template <int CH>
BCN_HD void compute_axis(float axis[4], const float c[10], int iters) {
  float a0 = 1.0f, a1 = 1.0f, a2 = 1.0f, a3 = 1.0f;
  for (int i = 0; i < iters; i++) {
    float x0, x1, x2, x3 = 0.0f;
    if (CH == 3) {
      x0 = c[0] * a0 + c[1] * a1 + c[2] * a2;
      x1 = c[1] * a0 + c[4] * a1 + c[5] * a2;
      x2 = c[2] * a0 + c[5] * a1 + c[7] * a2;
    } else {
      x0 = c[0] * a0 + c[1] * a1 + c[2] * a2 + c[3] * a3;
      x1 = c[1] * a0 + c[4] * a1 + c[5] * a2 + c[6] * a3;
      x2 = c[2] * a0 + c[5] * a1 + c[7] * a2 + c[8] * a3;
      x3 = c[3] * a0 + c[6] * a1 + c[8] * a2 + c[9] * a3;
    }
    a0 = x0;
    a1 = x1;
    a2 = x2;
    a3 = x3;
    if (i % 2 == 1) {
      // The reference starts the sum at 0.0; 0.0 + x == x for x = sq(..) >= 0.
      float norm_sq = sq(x0) + sq(x1) + sq(x2);
      if (CH == 4) norm_sq = norm_sq + sq(x3);
      float rnorm = rsqrt_cr(norm_sq);
      a0 *= rnorm;
      a1 *= rnorm;
      a2 *= rnorm;
      a3 *= rnorm;
    }
  }
  axis[0] = a0;
  axis[1] = a1;
  axis[2] = a2;
  axis[3] = a3;
}

// Integer masked moments; see compute_stats_masked (common.rs:294-332).
// Bounds: sums <= 16*255 = 4080, sums of products <= 16*255^2 = 1,040,400,
// both < 2^24, so the reference's f32 accumulation is exact and equals these.
struct Moments {
  int n, s[4], q[10];
};

// Four texels at a time: the masked texel words are transposed to one word
// per channel (byte i = texel i of the group, two v_perm_b32 per channel),
// and every sum of products over the group is one v_dot4_u32_u8.
template <int CH, int ROT = 0>
BCN_HD Moments moments(const uint32_t px[16], uint32_t mask, uint32_t sh = 0) {
  launder_texels(px);
  Moments m{};
  m.n = __builtin_popcount(mask & 0xFFFFu);
  uint32_t s[4] = {0, 0, 0, 0}, q[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
  for (int g = 0; g < 4; g++) {
    uint32_t t[4];
#pragma unroll
    for (int i = 0; i < 4; i++) t[i] = texel<ROT>(px, 4 * g + i, sh);
    // (t0.R, t1.R, t0.B, t1.B), (t0.G, t1.G, t0.A, t1.A), then with t2, t3;
    // the texels outside the mask are cleared in the channel words by one
    // AND with the group's byte mask (no per-texel selects, whose lane masks
    // the compiler would otherwise keep live across loops).
    const uint32_t mb = ((((mask >> (4 * g)) & 0xFu) * 0x00204081u) & 0x01010101u) * 0xFFu;
    const uint32_t a = perm_b32(t[1], t[0], 0x06020400u), b = perm_b32(t[1], t[0], 0x07030501u);
    const uint32_t c = perm_b32(t[3], t[2], 0x06020400u), d = perm_b32(t[3], t[2], 0x07030501u);
    const uint32_t r = perm_b32(c, a, 0x05040100u) & mb, bl = perm_b32(c, a, 0x07060302u) & mb;
    const uint32_t gr = perm_b32(d, b, 0x05040100u) & mb;
    s[0] = udot4(r, 0x01010101u, s[0]);
    s[1] = udot4(gr, 0x01010101u, s[1]);
    s[2] = udot4(bl, 0x01010101u, s[2]);
    q[0] = udot4(r, r, q[0]);
    q[1] = udot4(r, gr, q[1]);
    q[2] = udot4(r, bl, q[2]);
    q[4] = udot4(gr, gr, q[4]);
    q[5] = udot4(gr, bl, q[5]);
    q[7] = udot4(bl, bl, q[7]);
    if (CH == 4) {
      const uint32_t al = perm_b32(d, b, 0x07060302u) & mb;
      s[3] = udot4(al, 0x01010101u, s[3]);
      q[3] = udot4(r, al, q[3]);
      q[6] = udot4(gr, al, q[6]);
      q[8] = udot4(bl, al, q[8]);
      q[9] = udot4(al, al, q[9]);
    }
  }
#pragma unroll
  for (int p = 0; p < 4; p++) m.s[p] = (int)s[p];
#pragma unroll
  for (int i = 0; i < 10; i++) m.q[i] = (int)q[i];
  return m;
}

// moments() from channel-major bytes: the masked channel bytes are one AND
// with the group's byte mask, then one v_dot4_u32_u8 per sum.  UNIFORM: the
// mask is the same in every lane (the shape loops of modes 0/2 and the
// partition bounds), so a group of four texels outside the mask is skipped
// and one inside it needs no AND, by scalar branches.
template <int CH, bool UNIFORM = false>
BCN_HD Moments moments_cm(const uint32_t cm[4][4], uint32_t mask) {
  Moments m{};
  m.n = __builtin_popcount(mask & 0xFFFFu);
  uint32_t s[4] = {0, 0, 0, 0}, q[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
  for (int g = 0; g < 4; g++) {
    const uint32_t nib = (mask >> (4 * g)) & 0xFu;
    if (UNIFORM && nib == 0u) continue;
    const uint32_t mb = ((nib * 0x00204081u) & 0x01010101u) * 0xFFu;
    uint32_t x[4];
#pragma unroll
    for (int p = 0; p < CH; p++) x[p] = (UNIFORM && nib == 0xFu) ? cm[g][p] : cm[g][p] & mb;
    s[0] = udot4(x[0], 0x01010101u, s[0]);
    s[1] = udot4(x[1], 0x01010101u, s[1]);
    s[2] = udot4(x[2], 0x01010101u, s[2]);
    q[0] = udot4(x[0], cm[g][0], q[0]);
    q[1] = udot4(x[0], cm[g][1], q[1]);
    q[2] = udot4(x[0], cm[g][2], q[2]);
    q[4] = udot4(x[1], cm[g][1], q[4]);
    q[5] = udot4(x[1], cm[g][2], q[5]);
    q[7] = udot4(x[2], cm[g][2], q[7]);
    if (CH == 4) {
      s[3] = udot4(x[3], 0x01010101u, s[3]);
      q[3] = udot4(x[0], cm[g][3], q[3]);
      q[6] = udot4(x[1], cm[g][3], q[6]);
      q[8] = udot4(x[2], cm[g][3], q[8]);
      q[9] = udot4(x[3], cm[g][3], q[9]);
    }
  }
#pragma unroll
  for (int p = 0; p < 4; p++) m.s[p] = (int)s[p];
#pragma unroll
  for (int i = 0; i < 10; i++) m.q[i] = (int)q[i];
  return m;
}

// covar_from_stats (common.rs:334-350), f32 with the reference's order.  The
// products of sums are integers <= 4080^2 < 2^24 and n is in [1, 16], so
// x / n is div_rcp with rn = 1.0f / n.  POW2: n is a power of two, rn exact,
// and x / n is the exact product x * rn.
template <int CH, bool POW2 = false>
BCN_HD void covar_from_fmoments(float c[10], const float s[4], const float q[10], float s14, float rn) {
  const float s10 = s[0], s11 = s[1], s12 = s[2];
  auto quo = [&](float x) { return POW2 ? x * rn : div_rcp(x, s14, rn); };
  c[0] = q[0] - quo(s10 * s10);
  c[1] = q[1] - quo(s10 * s11);
  c[2] = q[2] - quo(s10 * s12);
  c[4] = q[4] - quo(s11 * s11);
  c[5] = q[5] - quo(s11 * s12);
  c[7] = q[7] - quo(s12 * s12);
  if (CH == 4) {
    float s13 = s[3];
    c[3] = q[3] - quo(s10 * s13);
    c[6] = q[6] - quo(s11 * s13);
    c[8] = q[8] - quo(s12 * s13);
    c[9] = q[9] - quo(s13 * s13);
  } else {
    c[3] = c[6] = c[8] = c[9] = 0.0f;
  }
}
template <int CH, bool POW2 = false>
BCN_HD void covar_from_moments(float c[10], const int s[4], const int q[10], float s14, float rn) {
  float fs[4], fq[10];
#pragma unroll
  for (int i = 0; i < 4; i++) fs[i] = (float)s[i];
#pragma unroll
  for (int i = 0; i < 10; i++) fq[i] = (float)q[i];
  covar_from_fmoments<CH, POW2>(c, fs, fq, s14, rn);
}

// Texel counts of the two subsets of each 2-subset shape as floats, with
// their reciprocals 1.0f / n (correctly rounded, as rcp_normal gives them):
// one scalar load per shape, indexed by the loop counter, instead of a
// conversion and a reciprocal per lane.
struct PartCounts {
  float n[64][2], rcp[64][2];
};
constexpr PartCounts make_part_counts() {
  PartCounts t{};
  for (int i = 0; i < 64; i++) {
    const int n0 = __builtin_popcount(shape_mask(i, 0));
    const int c[2] = {n0, 16 - n0};
    for (int j = 0; j < 2; j++) {
      t.n[i][j] = (float)c[j];
      t.rcp[i][j] = 1.0f / (float)c[j];
    }
  }
  return t;
}
BCN_TABLE PartCounts kPartCounts = make_part_counts();

// block_segment_core (common.rs:483-518), first part: the masked texels'
// mean dc and principal axis.
template <int CH>
BCN_HD void pca_axis_moments(const Moments& m, float axis[4], float dc[4]) {
  float s14 = (float)m.n, rn = rcp_normal(s14);  // n in [1, 16]
#pragma unroll
  for (int p = 0; p < 4; p++) dc[p] = p < CH ? div_rcp((float)m.s[p], s14, rn) : 0.0f;
  float c[10];
  covar_from_moments<CH>(c, m.s, m.q, s14, rn);
  const float inv_var = 1.0f / (256.0f * 256.0f);
  const float eps = 0.001f * 0.001f;
#pragma unroll
  for (int i = 0; i < 10; i++) c[i] *= inv_var;
  c[0] += eps;
  c[4] += eps;
  c[7] += eps;
  c[9] += eps;
  compute_axis<CH>(axis, c, 8);
}

template <int CH, int ROT = 0>
BCN_HD void pca_axis(const uint32_t px[16], uint32_t mask, float axis[4], float dc[4],
                     uint32_t sh = 0) {
  launder_texels(px);
  pca_axis_moments<CH>(moments<CH, ROT>(px, mask, sh), axis, dc);
}

// The rest of block_segment_core: the masked texels' projections onto the
// axis give the extent, and the endpoints are the extent's ends (widened to
// at least 1) on the axis through dc; before the BC7 clamp.
template <int CH, int ROT = 0>
BCN_HD void pca_extent(const uint32_t px[16], uint32_t mask, const float axis[4], const float dc[4],
                       float e0[4], float e1[4], uint32_t sh = 0) {
  launder_texels(px);
  float ext0 = INFINITY, ext1 = -INFINITY;
#pragma unroll
  for (int k = 0; k < 16; k++) {
    // Only texels in the mask are projected.  Where the mask is the same for
    // the whole wave (the partition loops of modes 0 and 2 run partition p on
    // every lane) this is a scalar branch that skips the texel; otherwise it
    // is an execution mask over the same instructions.
    if (!((mask >> k) & 1u)) continue;
    // The reference's running sum starts at 0.0; dropping that leading add
    // can only turn a +0.0 into -0.0, and the sign of a zero projection never
    // reaches an output (ext*axis + dc rounds it away).
    uint32_t t = texel<ROT>(px, k, sh);
    float dot = axis[0] * (chf(t, 0) - dc[0]);
    dot = dot + axis[1] * (chf(t, 1) - dc[1]);
    dot = dot + axis[2] * (chf(t, 2) - dc[2]);
    if (CH == 4) dot = dot + axis[3] * (chf(t, 3) - dc[3]);
    ext0 = fminf(ext0, dot);
    ext1 = fmaxf(ext1, dot);
  }
  if (ext1 - ext0 < 1.0f) {
    ext0 -= 0.5f;
    ext1 += 0.5f;
  }
#pragma unroll
  for (int p = 0; p < 4; p++) {
    e0[p] = p < CH ? ext0 * axis[p] + dc[p] : 0.0f;
    e1[p] = p < CH ? ext1 * axis[p] + dc[p] : 0.0f;
  }
}

template <int CH, int ROT = 0>
BCN_HD void pca_segment(const uint32_t px[16], uint32_t mask, float e0[4], float e1[4],
                        uint32_t sh = 0) {
  float axis[4], dc[4];
  pca_axis<CH, ROT>(px, mask, axis, dc, sh);
  pca_extent<CH, ROT>(px, mask, axis, dc, e0, e1, sh);
}

// block_segment (bc7.rs:371-379): PCA endpoints clamped to [0, 255].
template <int CH, int ROT = 0>
BCN_HD void bc7_segment(const uint32_t px[16], uint32_t mask, float e0[4], float e1[4],
                        uint32_t sh = 0) {
  pca_segment<CH, ROT>(px, mask, e0, e1, sh);
#pragma unroll
  for (int p = 0; p < CH; p++) {
    e0[p] = rclamp(e0[p], 0.0f, 255.0f);
    e1[p] = rclamp(e1[p], 0.0f, 255.0f);
  }
}

// ---- endpoint quantisation (bc7.rs:97-237) ---------------------------------------

// unpack_to_byte (bc7.rs:97-100): vv + (vv >> bits) with vv = v << (8 - bits).
// For 0 <= v < 2^bits and 4 <= bits <= 8 the two terms share no bits and
// vv >> bits == v >> (2 * bits - 8), so it is one shift and one shift-or
// (checked at compile time below).
BCN_HD constexpr int unpack_to_byte(int v, int bits) {
  return (v << (8 - bits)) | (v >> (2 * bits - 8));
}

constexpr bool unpack_to_byte_holds() {
  for (int bits = 4; bits <= 8; bits++)
    for (int v = 0; v < (1 << bits); v++) {
      const int vv = v << (8 - bits);
      if (unpack_to_byte(v, bits) != vv + (vv >> bits)) return false;
    }
  return true;
}
static_assert(unpack_to_byte_holds(), "unpack_to_byte as shift-or");

// The two p-bit candidates of one channel code in ep_quant0367 / ep_quant1
// (bc7.rs:117-123, 159-165): qb_b = 2 * clamp((int)((y - b) / 2 + 0.5), 0,
// L - 1) + b for y = ep / 255 * (2L - 1), the int conversion truncating and
// saturating, NaN to 0.
//   b = 0: y / 2 is exact (or |y| < 2^-125 and the sum is 0.5 either way), so
//          one fused y * 0.5 + 0.5 rounds like the reference's two steps;
//   b = 1: for y >= 1/2, y - 1, the halving and the + 0.5 (giving y / 2) are
//          all exact, so the code is 2 * floor(y / 2) + 1 = floor(y) | 1 with
//          y clamped to [0, 2L - 1]; y < 1/2 (and NaN) gives 1 either way.
// Checked over every float with 2^-10 <= |y| in [-4, 2L + 4] (and a sample
// of smaller ones) against the direct expression in the tests.
template <int L>
BCN_HD void pbit_codes(float y, int& qb0, int& qb1) {
  qb0 = qclamp(__builtin_fmaf(y, 0.5f, 0.5f), 0.0f, (float)(L - 1)) * 2;
  qb1 = qclamp(y, 0.0f, (float)(2 * L - 1)) | 1;
}

// ep_quant0367 for one endpoint pair (bc7.rs:102-148) followed by ep_dequant.
// QCH channels are quantised (4 when the alpha endpoint is emitted), ECH
// channels enter the p-bit error.  Outputs quantised q and dequantised d.
template <int MODE, int QCH, int ECH>
BCN_HD void quant_pair_0367(const float e0[4], const float e1[4], int q0[4], int q1[4],
                            int d0[4], int d1[4]) {
  constexpr int BITS = MODE == 0 ? 4 : (MODE == 7 ? 5 : 7);
  constexpr int L = 1 << BITS;
  constexpr int L2 = L * 2 - 1;
#pragma unroll
  for (int i = 0; i < 2; i++) {
    const float* ep = i == 0 ? e0 : e1;
    int qb[2][4];
    float eb[2][4];
#pragma unroll
    for (int p = 0; p < QCH; p++) {
      pbit_codes<L>(div255(ep[p]) * (float)L2, qb[0][p], qb[1][p]);
#pragma unroll
      for (int b = 0; b < 2; b++)
        eb[b][p] = MODE == 0 ? (float)unpack_to_byte(qb[b][p], 5) : (float)qb[b][p];
    }
    float err0 = sq(ep[0] - eb[0][0]), err1 = sq(ep[0] - eb[1][0]);
#pragma unroll
    for (int p = 1; p < ECH; p++) {
      err0 = err0 + sq(ep[p] - eb[0][p]);
      err1 = err1 + sq(ep[p] - eb[1][p]);
    }
    bool t0 = err0 < err1;
    int* qo = i == 0 ? q0 : q1;
    int* dq = i == 0 ? d0 : d1;
#pragma unroll
    for (int p = 0; p < QCH; p++) {
      qo[p] = t0 ? qb[0][p] : qb[1][p];
      // ep_dequant (bc7.rs:210-232): modes 3 and 6 are 8-bit already.
      dq[p] = MODE == 0 ? unpack_to_byte(qo[p], 5) : (MODE == 7 ? unpack_to_byte(qo[p], 6) : qo[p]);
    }
  }
}

// ep_quant1 (bc7.rs:150-178) for one 2-endpoint pair: one p-bit shared by the
// pair, chosen by the RGB error of both endpoints; then ep_dequant (7 bits).
BCN_HD void quant_pair_1(const float e0[4], const float e1[4], int q0[4], int q1[4], int d0[4],
                         int d1[4]) {
  int qb[2][2][3];
  float eb[2][2][3];
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int p = 0; p < 3; p++) {
      pbit_codes<64>(div255(i == 0 ? e0[p] : e1[p]) * 127.0f, qb[0][i][p], qb[1][i][p]);
#pragma unroll
      for (int b = 0; b < 2; b++) eb[b][i][p] = (float)unpack_to_byte(qb[b][i][p], 7);
    }
  float err0 = sq(e0[0] - eb[0][0][0]), err1 = sq(e0[0] - eb[1][0][0]);
#pragma unroll
  for (int j = 0; j < 2; j++)
#pragma unroll
    for (int p = 0; p < 3; p++) {
      if (j == 0 && p == 0) continue;
      float v = j == 0 ? e0[p] : e1[p];
      err0 = err0 + sq(v - eb[0][j][p]);
      err1 = err1 + sq(v - eb[1][j][p]);
    }
  bool t0 = err0 < err1;
#pragma unroll
  for (int p = 0; p < 3; p++) {
    q0[p] = t0 ? qb[0][0][p] : qb[1][0][p];
    q1[p] = t0 ? qb[0][1][p] : qb[1][1][p];
    d0[p] = unpack_to_byte(q0[p], 7);
    d1[p] = unpack_to_byte(q1[p], 7);
  }
}

// ep_quant245 (bc7.rs:180-189) + ep_dequant: plain rounding to BITS bits.
template <int BITS>
BCN_HD void quant_pair_245(const float e0[4], const float e1[4], int q0[4], int q1[4], int d0[4],
                           int d1[4]) {
  constexpr int L = 1 << BITS;
#pragma unroll
  for (int p = 0; p < 3; p++) {
    q0[p] = qclamp(div255(e0[p]) * (float)(L - 1) + 0.5f, 0.0f, (float)(L - 1));
    q1[p] = qclamp(div255(e1[p]) * (float)(L - 1) + 0.5f, 0.0f, (float)(L - 1));
    d0[p] = unpack_to_byte(q0[p], BITS);
    d1[p] = unpack_to_byte(q1[p], BITS);
  }
}

template <int MODE, int CH>
BCN_HD void quant_pair(const float e0[4], const float e1[4], int q0[4], int q1[4], int d0[4],
                       int d1[4]) {
  if (MODE == 0 || MODE == 3) quant_pair_0367<MODE, 3, 3>(e0, e1, q0, q1, d0, d1);
  else if (MODE == 7) quant_pair_0367<7, 4, 4>(e0, e1, q0, q1, d0, d1);
  else if (MODE == 6) quant_pair_0367<6, 4, CH>(e0, e1, q0, q1, d0, d1);
  else if (MODE == 1) quant_pair_1(e0, e1, q0, q1, d0, d1);
  else if (MODE == 5) quant_pair_245<7>(e0, e1, q0, q1, d0, d1);
  else quant_pair_245<5>(e0, e1, q0, q1, d0, d1);  // modes 2, 4
}

// ---- index selection (common.rs:422-481) ----------------------------------------

// A lane's row of 18 words in LDS (word w at p[w * stride]) holding the
// subsets' index-search data of a 2- or 3-subset block_quant: a texel then
// takes its subset's words with one address computation and three or four
// LDS reads instead of six to fourteen selects.
struct SubsetRow {
  uint32_t* p = nullptr;
  int stride = 0;
};

// The first index candidate of a 2-bit search, clamp((int)(fl(proj / div) *
// 4 + 0.5), 1, 3) for integer proj and div = |e1 - e0|^2 in [0, 260100], is
// 1 + [proj >= t2] + [proj >= t3].  With r = 4 proj / div, the two roundings
// of the float form move r + 0.5 by less than 2^-24 (2|r| + 1) <= 1 / (2 div)
// (|proj| <= 510 sqrt(div)), and an r + 0.5 that is not an integer is at least
// 1 / (2 div) from one, while an integer r + 0.5 is computed exactly (proj /
// div is then a dyadic fraction); so the float form's floor is that of
// r + 0.5, and r + 0.5 >= m exactly when proj >= ceil((2m - 1) div / 8).
// div == 0 (proj == 0, a NaN quotient, index 1): thresholds never reached.
// tests/kernel_check checks both thresholds of every div against the float
// form.
BCN_HD void quant4_thresholds(int div, int& t2, int& t3) {
  t2 = div == 0 ? 0x7FFFFFFF : (3 * div + 7) >> 3;
  t3 = div == 0 ? 0x7FFFFFFF : (5 * div + 7) >> 3;
}

// The first index candidate of a 3-bit search, clamp((int)(fl(proj / div) *
// 8 + 0.5), 1, 7), is floor(r + 0.5) clamped, r = 8 proj / div (the argument
// of quant4_thresholds, with L = 8: the float form's error stays below
// 1 / (2 div) for div <= 260100).  floor(r + 0.5) = floor(N / (4 div)) with
// N = 32 proj + 2 div + 1, and N / (4 div) (odd over even) is at least
// 1 / (4 div) from every integer, while fl(N * fl(1 / (4 div))) is within
// 8.5 * 2^-23 < 1 / (4 * 260100) of it below 8.5 (N is exact: |N| < 2^23).
// So q1 = clamp(fl(N * c), 1, 7) truncated, c = 1.0f / (4 div); div == 0
// (proj == 0, index 1) takes c = 0.  tests/kernel_check checks all six
// thresholds of every div against the float form.
BCN_HD void quant8_consts(int div, int& c1, float& rq) {
  c1 = 2 * div + 1;
  rq = div == 0 ? 0.0f : rcp_normal(4.0f * (float)div);  // 4 div <= 1040400
}
BCN_HD int quant8_index(int proj, int c1, float rq) {
  return qclamp((float)(proj * 32 + c1) * rq, 1.0f, 7.0f);
}

// Index per texel against dequantised (integer) endpoints d[j][0|1][p] of
// subset j, pattern giving each texel's subset.  Returns the exact SSE.
// Bounds: |proj|, div <= 4*255^2 and the error sum <= 16*4*255^2 < 2^24.
// 2 or 3 subsets: `row` must be set (SubsetRow).
template <int CH, int BITS, int NSUB, int ROT = 0>
BCN_HD int block_quant(const uint32_t px[16], const int d[3][2][4], uint32_t pattern,
                       uint32_t qblock[2], uint32_t sh = 0, SubsetRow row = SubsetRow()) {
  launder_texels(px);
  constexpr int L = 1 << BITS;
  // One and two subsets work from the texel itself: proj = t.de - a.de and
  // the rounding term (32 + 64 a) - 64 t, with K = a.de and 32 + 64 a kept
  // per subset (seven words).  Three subsets keep six words per subset in
  // the 18-word row and work from b = t - a.
  constexpr bool KFORM = NSUB != 3;
  constexpr int W = KFORM ? 7 : 6;
  // 2-bit indices: q1 from two integer thresholds on proj (quant4_thresholds);
  // 3-bit: from one reciprocal (quant8_index).
  constexpr bool ITHR = L == 4, IRCP = L == 8;
  // Per subset: endpoint 0 and the endpoint difference as (R, B) and (G, A)
  // pairs of 16-bit lanes (A lanes 0 for 3 channels).  Lane values: endpoints
  // 0..255, differences and endpoint-minus-texel in [-255, 255],
  // (w*diff + 32) in [-16288, 16352], interpolated minus texel in
  // [-510, 510]: all exact in 16 bits, squares and products in the 32-bit
  // dot products.  32 + 64 a is at most 16352.
  uint32_t ea_rb[3] = {0, 0, 0}, ea_ga[3] = {0, 0, 0}, de_rb[3] = {0, 0, 0}, de_ga[3] = {0, 0, 0};
  int negk[3] = {0, 0, 0};
  float fdiv[3] = {1.0f, 1.0f, 1.0f}, frcp[3] = {1.0f, 1.0f, 1.0f};
#pragma unroll
  for (int j = 0; j < NSUB; j++) {
    ea_rb[j] = pk16(d[j][0][0], d[j][0][2]);
    ea_ga[j] = pk16(d[j][0][1], CH == 4 ? d[j][0][3] : 0);
    de_rb[j] = pk16(d[j][1][0] - d[j][0][0], d[j][1][2] - d[j][0][2]);
    de_ga[j] = pk16(d[j][1][1] - d[j][0][1], CH == 4 ? d[j][1][3] - d[j][0][3] : 0);
    int div = dot2_16(de_ga[j], de_ga[j], dot2_16(de_rb[j], de_rb[j], 0));
    fdiv[j] = (float)div;
    // div == 0 gives +inf here and NaN below (proj is then 0), as 0 / 0 does.
    // div in [1, 260100], or 0: then rcp_normal gives NaN instead of +inf,
    // and proj is 0, so the quotient below is NaN either way.
    frcp[j] = rcp_normal(fdiv[j]);
    if (ITHR) {
      int t2, t3;
      quant4_thresholds(div, t2, t3);
      fdiv[j] = __builtin_bit_cast(float, t2);
      frcp[j] = __builtin_bit_cast(float, t3);
    }
    if (IRCP) {
      int c1;
      quant8_consts(div, c1, frcp[j]);
      fdiv[j] = __builtin_bit_cast(float, c1);
    }
    if (KFORM) {
      // |a.de| <= 4 * 255 * 255.
      negk[j] = -dot2_16(ea_ga[j], de_ga[j], dot2_16(ea_rb[j], de_rb[j], 0));
      ea_rb[j] = pk16(32 + 64 * d[j][0][0], 32 + 64 * d[j][0][2]);
      ea_ga[j] = pk16(32 + 64 * d[j][0][1], 32 + (CH == 4 ? 64 * d[j][0][3] : 0));
    }
  }
  if (NSUB >= 2) {
#pragma unroll
    for (int j = 0; j < NSUB; j++) {
      uint32_t* r = row.p + W * j * row.stride;
      r[0] = ea_rb[j];
      r[row.stride] = ea_ga[j];
      r[2 * row.stride] = de_rb[j];
      r[3 * row.stride] = de_ga[j];
      r[4 * row.stride] = __builtin_bit_cast(uint32_t, fdiv[j]);
      r[5 * row.stride] = __builtin_bit_cast(uint32_t, frcp[j]);
      if (KFORM) r[6 * row.stride] = (uint32_t)negk[j];
    }
  }
  const uint32_t ga_mask = CH == 4 ? 0x00FF00FFu : 0x000000FFu;
  const uint32_t m64 = fresh(0xFFC0FFC0u);
  int total = 0;
  qblock[0] = 0;
  qblock[1] = 0;
#pragma unroll
  for (int k = 0; k < 16; k++) {
    int j = NSUB == 1 ? 0 : (int)((pattern >> (2 * k)) & 3u);
    // ea: endpoint 0 (six-word form) or 32 + 64 * endpoint 0 (seven-word).
    uint32_t erb, ega, drb, dga;
    float div, rdiv;
    int nk;
    if (NSUB >= 2) {
      const uint32_t* r = row.p + W * j * row.stride;
      erb = r[0];
      ega = r[row.stride];
      drb = r[2 * row.stride];
      dga = r[3 * row.stride];
      div = __builtin_bit_cast(float, r[4 * row.stride]);
      rdiv = __builtin_bit_cast(float, r[5 * row.stride]);
      nk = KFORM ? (int)r[6 * row.stride] : 0;
    } else {
      erb = ea_rb[0];
      ega = ea_ga[0];
      drb = de_rb[0];
      dga = de_ga[0];
      div = fdiv[0];
      rdiv = frcp[0];
      nk = negk[0];
    }
    uint32_t t = texel<ROT>(px, k, sh);
    const uint32_t trb = t & 0x00FF00FFu, tga = (t >> 8) & ga_mask;
    // proj = sum (texel - e0) * (e1 - e0), an exact integer either way.
    int proj;
    uint32_t c_rb, c_ga;
    // The reference's ((64 - w) * a + w * e1 + 32) >> 6 with e1 = a + diff:
    // the 64 * a term is a multiple of 64 and >> is floor division, so it
    // equals a + ((w * diff + 32) >> 6); minus the texel t that is
    // (w * diff + 32 - 64 * (t - a)) >> 6, with |w * diff|, |64 * (t - a)|
    // <= 16320; c = 32 - 64 * (t - a) per lane.
    if (KFORM) {
      proj = dot2_16(tga, dga, dot2_16(trb, drb, nk));
      c_rb = pk_mad16(trb, m64, erb);
      c_ga = pk_mad16(tga, m64, ega);
    } else {
      const uint32_t brb = pk_sub16(trb, erb), bga = pk_sub16(tga, ega);
      proj = dot2_16(bga, dga, dot2_16(brb, drb, 0));
      c_rb = pk_round_off16(brb, m64);
      c_ga = pk_round_off16(bga, m64);
    }
    int q1;
    if (ITHR) {
      q1 = 1 + (int)(proj >= __builtin_bit_cast(int, div)) + (int)(proj >= __builtin_bit_cast(int, rdiv));
    } else if (IRCP) {
      q1 = quant8_index(proj, __builtin_bit_cast(int, div), rdiv);
    } else {
      // |proj| <= |texel - e0| |e1 - e0| <= 510 sqrt(div) (Cauchy-Schwarz).
      float fp = div_rcp((float)proj, div, rdiv);
      // fp * L is exact (a power-of-two multiple of a quotient in [-510, 510]),
      // so the fused form rounds once, as fp * L + 0.5 does.
      q1 = qclamp(__builtin_fmaf(fp, (float)L, 0.5f), 1.0f, (float)(L - 1));
    }
    const uint32_t wp = weight_pair<BITS>(q1);
    const uint32_t t0rb = pk_wmad16<0>(wp, drb, c_rb), t0ga = pk_wmad16<0>(wp, dga, c_ga);
    const uint32_t t1rb = pk_wmad16<1>(wp, drb, c_rb), t1ga = pk_wmad16<1>(wp, dga, c_ga);
    // The error sums continue the running total (every partial sum < 2^23),
    // so the smaller of the two is the new total.
    const int err0 = dot2_16(t0ga, t0ga, dot2_16(t0rb, t0rb, total));
    const int err1 = dot2_16(t1ga, t1ga, dot2_16(t1rb, t1rb, total));
    const bool lo = err0 < err1;
    total = lo ? err0 : err1;
    qblock[k >> 3] |= (uint32_t)(q1 - (int)lo) << (4 * (k & 7));
  }
  return total;
}

// block_quant of a mode 4/5 candidate's colour part for a 3-channel block
// when the rotated channel's dequantised endpoints are both 255.  The rotated
// block holds 255 in that channel for every texel (bc7.rs:690-695), so the
// channel adds (255 - 255) * 0 to every projection and to div, and its
// interpolated value is 255 for every index: its error is 0.  The search is
// then block_quant on the two other channels alone, in one pair of 16-bit
// lanes: a2 / b2 hold their endpoints 0 / 1 (low lane, high lane), and `sel`
// is the v_perm_b32 selector that takes their bytes from a texel word.  Same
// projections, thresholds, candidate errors and ties as block_quant, so the
// same indices and SSE.
template <int BITS>
BCN_HD int block_quant_pair(const uint32_t px[16], uint32_t a2, uint32_t b2, uint32_t sel,
                            uint32_t qblock[2]) {
  launder_texels(px);
  constexpr int L = 1 << BITS;
  static_assert(L == 4 || L == 8, "2- and 3-bit indices (modes 4 and 5)");
  const uint32_t de = pk_sub16(b2, a2);
  const int div = dot2_16(de, de, 0);
  int t2 = 0, t3 = 0, c1 = 0;
  float rq = 0.0f;
  if (L == 4) quant4_thresholds(div, t2, t3);
  else quant8_consts(div, c1, rq);
  // proj = t.de - a.de and the rounding term (32 + 64 a) - 64 t, as the
  // one-subset block_quant forms them.
  const int nk = -dot2_16(a2, de, 0);
  const uint32_t base = pk_mad16(a2, 0x00400040u, 0x00200020u);
  const uint32_t m64 = fresh(0xFFC0FFC0u);
  int total = 0;
  qblock[0] = 0;
  qblock[1] = 0;
#pragma unroll
  for (int k = 0; k < 16; k++) {
    const uint32_t tp = perm_b32(0u, px[k], sel);
    const int proj = dot2_16(tp, de, nk);
    const uint32_t c = pk_mad16(tp, m64, base);
    const int q1 = L == 4 ? 1 + (int)(proj >= t2) + (int)(proj >= t3) : quant8_index(proj, c1, rq);
    const uint32_t wp = weight_pair<BITS>(q1);
    const uint32_t t0 = pk_wmad16<0>(wp, de, c), t1 = pk_wmad16<1>(wp, de, c);
    const int err0 = dot2_16(t0, t0, total);
    const int err1 = dot2_16(t1, t1, total);
    const bool lo = err0 < err1;
    total = lo ? err0 : err1;
    qblock[k >> 3] |= (uint32_t)(q1 - (int)lo) << (4 * (k & 7));
  }
  return total;
}

// The colour-part block_quant of a mode 4/5 candidate: block_quant_pair when
// every lane of the wave has the rotated channel at 255/255 (3-channel
// rotation, ROT 1), else the full 3-channel search of the rotated block.
template <int BITS, int ROT>
BCN_HD int block_quant45(const uint32_t px[16], const int d[3][2][4], uint32_t rot, uint32_t sh,
                         uint32_t qblock[2]) {
  if (ROT == 1) {
    // The endpoints as plain registers first (fresh): a select between
    // elements of d would otherwise become a run-time index into d, which
    // moves the array to private memory.
    uint32_t a[3], b[3];
#pragma unroll
    for (int p = 0; p < 3; p++) {
      a[p] = fresh((uint32_t)d[0][0][p]);
      b[p] = fresh((uint32_t)d[0][1][p]);
    }
    const uint32_t ar = rot == 0 ? a[0] : (rot == 1 ? a[1] : a[2]);
    const uint32_t br = rot == 0 ? b[0] : (rot == 1 ? b[1] : b[2]);
    if (wave_all(ar == 255u && br == 255u)) {
      // The other two channels, in channel order: (G, B), (R, B) or (R, G).
      const uint32_t a2 = rot == 0 ? (a[1] | (a[2] << 16)) : (rot == 1 ? (a[0] | (a[2] << 16)) : (a[0] | (a[1] << 16)));
      const uint32_t b2 = rot == 0 ? (b[1] | (b[2] << 16)) : (rot == 1 ? (b[0] | (b[2] << 16)) : (b[0] | (b[1] << 16)));
      const uint32_t sel = rot == 0 ? 0x0C020C01u : (rot == 1 ? 0x0C020C00u : 0x0C010C00u);
      return block_quant_pair<BITS>(px, a2, b2, sel, qblock);
    }
  }
  return block_quant<3, BITS, 1, ROT>(px, d, 0u, qblock, sh);
}

// opt_endpoints (common.rs:142-205): least-squares endpoints of the masked
// texels for fixed indices.  Sums are integers (<= 16*15*255 = 61,200);
// cxx, cyy, cxy and the determinant are exact small integers (det <=
// 3600^2 < 2^24); the solution keeps the reference's f32 rounding order.
template <int CH, int BITS, int ROT = 0>
BCN_HD void opt_endpoints(const uint32_t px[16], const uint32_t qblock[2], uint32_t mask,
                          float e0[4], float e1[4], uint32_t sh = 0) {
  launder_texels(px);
  constexpr int L1 = (1 << BITS) - 1;
  // Four texels at a time, as in moments(): channel-major bytes of the masked
  // texels, index bytes, and one v_dot4_u32_u8 per sum.  x = L1 - q is
  // q ^ L1 (q <= L1 = 2^BITS - 1).
  const int cnt = __builtin_popcount(mask & 0xFFFFu);
  uint32_t uq = 0, uqq = 0, us[4] = {0, 0, 0, 0}, ua[4] = {0, 0, 0, 0};
#pragma unroll
  for (int g = 0; g < 4; g++) {
    uint32_t t[4];
#pragma unroll
    for (int i = 0; i < 4; i++) t[i] = texel<ROT>(px, 4 * g + i, sh);
    // Texels outside the mask cleared in the channel words (see moments).
    const uint32_t mb = ((((mask >> (4 * g)) & 0xFu) * 0x00204081u) & 0x01010101u) * 0xFFu;
    const uint32_t a = perm_b32(t[1], t[0], 0x06020400u), b = perm_b32(t[1], t[0], 0x07030501u);
    const uint32_t c = perm_b32(t[3], t[2], 0x06020400u), d = perm_b32(t[3], t[2], 0x07030501u);
    uint32_t ch[4];
    ch[0] = perm_b32(c, a, 0x05040100u) & mb;
    ch[1] = perm_b32(d, b, 0x05040100u) & mb;
    ch[2] = perm_b32(c, a, 0x07060302u) & mb;
    ch[3] = CH == 4 ? perm_b32(d, b, 0x07060302u) & mb : 0u;
    // Index nibbles of the group's texels to bytes, and their mask.
    const uint32_t nib = (qblock[g >> 1] >> (16 * (g & 1))) & 0xFFFFu;
    const uint32_t qb = (nib & 0xFu) | ((nib & 0xF0u) << 4) | ((nib & 0xF00u) << 8) | ((nib & 0xF000u) << 12);
    const uint32_t qm = qb & mb, xb = qb ^ (0x01010101u * (uint32_t)L1);
    uq = udot4(qm, 0x01010101u, uq);
    uqq = udot4(qm, qm, uqq);
#pragma unroll
    for (int p = 0; p < CH; p++) {
      us[p] = udot4(ch[p], 0x01010101u, us[p]);
      ua[p] = udot4(ch[p], xb, ua[p]);
    }
  }
  // Bounds: sum_q <= 240, sum_qq <= 3600, sum <= 4080, atb1 <= 61,200.
  const int sum_q = (int)uq, sum_qq = (int)uqq;
  int sum[4], atb1[4];
#pragma unroll
  for (int p = 0; p < 4; p++) {
    sum[p] = (int)us[p];
    atb1[p] = (int)ua[p];
  }
  int cxx = mul24i(cnt, L1 * L1) - mul24i(2 * L1, sum_q) + sum_qq;
  int cyy = sum_qq;
  int cxy = mul24i(L1, sum_q) - sum_qq;
  int det = mul24i(cxx, cyy) - mul24i(cxy, cxy);
  float scale = (float)L1 / (float)det;
  float fcxx = (float)cxx, fcyy = (float)cyy, fcxy = (float)cxy;
  float fcnt = (float)cnt, rcnt = rcp_normal(fcnt);  // cnt in [1, 16]
#pragma unroll
  for (int p = 0; p < CH; p++) {
    float a1 = (float)atb1[p];
    float a2 = (float)(mul24(L1, sum[p]) - atb1[p]);
    float x0 = (a1 * fcyy - a2 * fcxy) * scale;
    float x1 = (a2 * fcxx - a1 * fcxy) * scale;
    // |det| < 0.001 with an integer det: flatten to the mean.
    float mean = div_rcp((float)sum[p], fcnt, rcnt);  // sum <= 4080, cnt in [1, 16]
    e0[p] = det == 0 ? mean : x0;
    e1[p] = det == 0 ? mean : x1;
  }
}

// ---- scalar channel of modes 4/5 (bc7.rs:239-369) ----------------------------------

struct ChannelFit {
  int err;
  int q0, q1;
  uint32_t qblock[2];
};

// Channel byte at `sh` of texel k: the scalar channel of a mode-4/5 candidate.
BCN_HD int chan(const uint32_t px[16], int k, uint32_t sh) {
  return (int)((px[k] >> sh) & 0xFFu);  // px laundered by the caller
}

// The first index candidate of channel_opt_quant's 2-bit search,
// q1 = clamp((int)(fl(fl(u / den) * 4) + 0.5), 1, 3) for u = texel - e0 and
// den = fl(d + 0.001f), d = e1 - e0 (u, d integers in [-255, 255]), is
// 1 + [y >= 2] + [y >= 3] with y the rounded sum.  With r = u / den the two
// roundings move y from 4 r + 0.5 by less than 2^-21 (4 |r| + 1), so
// y >= m exactly when u / den >= t = (2m - 1) / 8 up to a relative 2^-20;
// and den t = d (2m - 1) / 8 + (0.001f + rho) t, rho the rounding of den
// (|rho| <= 2^-24 * 255.001), i.e. a multiple of 1/8 plus an offset in
// [0.00036, 0.00064], which keeps den t at least 0.0003 from every integer
// while the rounding slack is below 255 * 2^-20 * 0.625 < 0.0002.  So for
// d >= 0 (den > 0) y >= m exactly when u > T_m = floor(d (2m - 1) / 8), and
// for d < 0 exactly when u <= T_m.  Both are sign tests on a 16-bit sum in
// the high half of a word: with the texel in both halves (v2 = v | v << 16),
// (v2 ^ s) + k_m has bit 31 set exactly then, where s = 0xFFFF0000 and
// k_m = (T_m + e0 + 1) << 16 for d >= 0 (the high half is then
// 65535 - v + T_m + e0 + 1, below 2^16 exactly when v > T_m + e0), s = 0 and
// k_m = -(T_m + e0 + 1) << 16 for d < 0; the low half (v) never carries, and
// every sum stays within 16-bit range.  tests/kernel_check checks every
// (e0, e1, texel) against the float form.
BCN_HD void chan_q4_consts(int e0, int e1, uint32_t& s, uint32_t& k2, uint32_t& k3) {
  const int d = e1 - e0;
  const int t2 = ((3 * d) >> 3) + e0 + 1, t3 = ((5 * d) >> 3) + e0 + 1;
  s = d >= 0 ? 0xFFFF0000u : 0u;
  k2 = (uint32_t)(d >= 0 ? t2 : -t2) << 16;
  k3 = (uint32_t)(d >= 0 ? t3 : -t3) << 16;
}
BCN_HD int chan_q4_index(uint32_t v2, uint32_t s, uint32_t k2, uint32_t k3) {
  const uint32_t x = v2 ^ s;
  return 1 + (int)((x + k2) >> 31) + (int)((x + k3) >> 31);
}

// The first index candidate of channel_opt_quant's 3-bit search,
// clamp((int)(fl(fl(u / den) * 8) + 0.5), 1, 7) with u = texel - e0 and
// den = fl(e1 - e0 + 0.001f), from one multiply-add and one product:
// clamp(fl(fl(8 texel + c) * r)) with c = fl(den / 2 - 8 e0) and r the
// correctly rounded 1 / den.  The exact value (8 u + den / 2) / den stays at
// least 0.0015 / |den| from every integer (8 u is a multiple of 8 and
// den (m - 1/2) a multiple of 1/2 plus (m - 1/2)(0.001f + rho)), while the
// three roundings move it by less than 2^-22 * 2200 / |den|.
// tests/kernel_check checks every (e0, e1, texel) against the float form.
BCN_HD void chan_q8_consts(int e0, int e1, float& c, float& r) {
  const float den = (float)(e1 - e0) + 0.001f;  // |den| in [0.001, 255.001]
  c = 0.5f * den - 8.0f * (float)e0;
  r = rcp_normal(den);
}
BCN_HD int chan_q8_index(float v, float c, float r) {
  return qclamp(__builtin_fmaf(v, 8.0f, c) * r, 1.0f, 7.0f);
}

// channel_opt_quant (bc7.rs:278-320).
template <int BITS>
BCN_HD int channel_quant(const uint32_t px[16], uint32_t sh, int e0, int e1, uint32_t qblock[2]) {
  launder_texels(px);
  constexpr int L = 1 << BITS;
  // |den| in [0.001, 255.001]
  float den = (float)(e1 - e0) + 0.001f, rden = rcp_normal(den);
  const float fe0 = (float)e0;
  // 2-bit searches: the first candidate from two sign tests (chan_q4_consts).
  uint32_t q4s = 0, q4k2 = 0, q4k3 = 0;
  if (BITS == 2) chan_q4_consts(e0, e1, q4s, q4k2, q4k3);
  // 3-bit searches: one multiply-add and one product (chan_q8_consts).
  float q8c = 0.0f, q8r = 0.0f;
  if (BITS == 3) chan_q8_consts(e0, e1, q8c, q8r);
  qblock[0] = 0;
  qblock[1] = 0;
  // Both candidates' interpolated-minus-texel values in the two 16-bit lanes:
  // as in block_quant, a + ((w * (b - a) + 32) >> 6) - texel =
  // (w * (b - a) + (32 + 64 * a) - 64 * texel) >> 6, all terms within 2^15.
  const uint32_t de2 = pk16(e1 - e0, e1 - e0), base2 = pk16(32 + 64 * e0, 32 + 64 * e0);
  const uint32_t m64 = fresh(0xFFC0FFC0u);  // -64 per lane, see pk_round_off16
  // Byte sh / 8 of a texel in both lanes.
  const uint32_t sel = ((sh >> 3) * 0x00010001u) | 0x0C000C00u;
  int total = 0;
#pragma unroll
  for (int k = 0; k < 16; k++) {
    const uint32_t cv2 = perm_b32(0u, px[k], sel);
    int q1;
    if (BITS == 2) {
      q1 = chan_q4_index(cv2, q4s, q4k2, q4k3);
    } else if (BITS == 3) {
      q1 = chan_q8_index(chf(cv2, 0), q8c, q8r);
    } else {
      // (float)(texel - e0) as a difference of exact small integers.
      float proj = div_rcp(chf(cv2, 0) - fe0, den, rden);
      // proj * L is exact (|proj| <= 255000, a power-of-two factor), so the
      // fused form rounds once, as proj * L + 0.5 does.
      q1 = qclamp(__builtin_fmaf(proj, (float)L, 0.5f), 1.0f, (float)(L - 1));
    }
    const uint32_t t = pk_mad_shr6(weight_pair<BITS>(q1), de2, pk_mad16(cv2, m64, base2));
    // |t| <= 255, so the squares fit unsigned 16-bit lanes.  Each error on
    // top of the running total (sums <= 16 * 255^2); the smaller is the new
    // total.
    const uint32_t sq = pk_sq16(t);
    const int err0 = (int)(sq & 0xFFFFu) + total, err1 = (int)(sq >> 16) + total;
    const bool lo = err0 < err1;
    total = lo ? err0 : err1;
    qblock[k >> 3] |= (uint32_t)(q1 - (int)lo) << (4 * (k & 7));
  }
  return total;
}

// channel_quant_dequant (bc7.rs:268-276) of one endpoint.
template <int EPBITS>
BCN_HD int channel_qd(float e, int& q) {
  constexpr int E = 1 << EPBITS;
  q = qclamp(div255(e) * (float)(E - 1) + 0.5f, 0.0f, (float)(E - 1));
  return unpack_to_byte(q, EPBITS);
}

// opt_channel (bc7.rs:239-266) with channel_opt_endpoints (bc7.rs:322-369).
template <int BITS, int EPBITS>
BCN_HD ChannelFit opt_channel(const uint32_t px[16], uint32_t sh, uint32_t refine) {
  launder_texels(px);
  constexpr int L1 = (1 << BITS) - 1;
  int mn = 255, mx = 0;
#pragma unroll
  for (int k = 0; k < 16; k++) {
    int v = chan(px, k, sh);
    mn = v < mn ? v : mn;
    mx = v > mx ? v : mx;
  }
  ChannelFit r;
  int e0 = channel_qd<EPBITS>((float)mn, r.q0);
  int e1 = channel_qd<EPBITS>((float)mx, r.q1);
  r.err = channel_quant<BITS>(px, sh, e0, e1, r.qblock);
  if (refine == 0) return r;
  // The channel bytes of texels 4g .. 4g + 3 in one word per group, and their
  // sum; the refinement's other sums are then one v_dot4_u32_u8 per group,
  // as in opt_endpoints.
  const uint32_t b = sh >> 3, sel = b | ((b + 4) << 8) | 0x0C0C0000u;
  uint32_t cw[4], usum = 0;
#pragma unroll
  for (int g = 0; g < 4; g++) {
    const uint32_t lo = perm_b32(px[4 * g + 1], px[4 * g], sel);
    const uint32_t hi = perm_b32(px[4 * g + 3], px[4 * g + 2], sel);
    cw[g] = perm_b32(hi, lo, 0x05040100u);
    usum = udot4(cw[g], 0x01010101u, usum);
  }
  const int sum = (int)usum;  // <= 16 * 255
#pragma unroll 1
  for (uint32_t it = 0; it < refine; it++) {
    // Sums <= 16*15*255; cxx, cyy, cxy, det exact as in opt_endpoints.
    uint32_t uq = 0, uqq = 0, ua = 0;
#pragma unroll
    for (int g = 0; g < 4; g++) {
      const uint32_t nib = (r.qblock[g >> 1] >> (16 * (g & 1))) & 0xFFFFu;
      const uint32_t qb = (nib & 0xFu) | ((nib & 0xF0u) << 4) | ((nib & 0xF00u) << 8) | ((nib & 0xF000u) << 12);
      uq = udot4(qb, 0x01010101u, uq);
      uqq = udot4(qb, qb, uqq);
      // L1 - q == q ^ L1 for q <= L1 = 2^BITS - 1.
      ua = udot4(cw[g], qb ^ (0x01010101u * (uint32_t)L1), ua);
    }
    const int sum_q = (int)uq, sum_qq = (int)uqq, atb1 = (int)ua;
    int cxx = 16 * L1 * L1 - mul24i(2 * L1, sum_q) + sum_qq;
    int cyy = sum_qq;
    int cxy = mul24i(L1, sum_q) - sum_qq;
    int det = mul24i(cxx, cyy) - mul24i(cxy, cxy);
    float scale = (float)L1 / (float)det;
    float a1 = (float)atb1, a2 = (float)(mul24(L1, sum) - atb1);
    float x0 = (a1 * (float)cyy - a2 * (float)cxy) * scale;
    float x1 = (a2 * (float)cxx - a1 * (float)cxy) * scale;
    x0 = rclamp(x0, 0.0f, 255.0f);
    x1 = rclamp(x1, 0.0f, 255.0f);
    float mean = (float)sum / 16.0f;
    x0 = det == 0 ? mean : x0;
    x1 = det == 0 ? mean : x1;
    e0 = channel_qd<EPBITS>(x0, r.q0);
    e1 = channel_qd<EPBITS>(x1, r.q1);
    r.err = channel_quant<BITS>(px, sh, e0, e1, r.qblock);
  }
  return r;
}

// ---- candidate evaluation ---------------------------------------------------------

// A fitted candidate of modes 0-3 / 7: quantised endpoints and indices.
struct PartFit {
  int err;
  uint32_t q[3][2];  // endpoint i of subset j: channel p in byte p
  uint32_t qblock[2];
};

BCN_HD bool better(const Bc7Best& b, int err, int order) {
  return err < b.err || (err == b.err && order < b.order);
}

BCN_HD uint32_t pack4(const int q[4], int ch) {
  uint32_t w = 0;
#pragma unroll
  for (int p = 0; p < 4; p++)
    if (p < ch) w |= (uint32_t)q[p] << (8 * p);
  return w;
}

// Quantise + index a partition from its subsets' PCA endpoints e[j][0|1]
// (the tail of bc7_enc_mode01237_part_fast, bc7.rs:535-537).
template <int MODE>
BCN_HD void fit_from_segments(const uint32_t px[16], int part_id, const float e[3][2][4],
                              PartFit& f, SubsetRow row = SubsetRow()) {
  constexpr int PAIRS = (MODE == 0 || MODE == 2) ? 3 : 2;
  constexpr int BITS = (MODE == 0 || MODE == 1) ? 3 : 2;
  constexpr int CH = MODE == 7 ? 4 : 3;
  int d[3][2][4];
#pragma unroll
  for (int j = 0; j < PAIRS; j++) {
    int q0[4], q1[4];
    quant_pair<MODE, CH>(e[j][0], e[j][1], q0, q1, d[j][0], d[j][1]);
    f.q[j][0] = pack4(q0, CH);
    f.q[j][1] = pack4(q1, CH);
  }
  f.err = block_quant<CH, BITS, PAIRS>(px, d, kPart.subset[part_id], f.qblock, 0u, row);
}

// bc7_enc_mode01237_part_fast (bc7.rs:517-538) for partition `part_id`
// (already offset by 64 for 3-subset modes).
template <int MODE>
BCN_HD void fit_partition(const uint32_t px[16], int part_id, PartFit& f, SubsetRow row = SubsetRow()) {
  constexpr int PAIRS = (MODE == 0 || MODE == 2) ? 3 : 2;
  constexpr int CH = MODE == 7 ? 4 : 3;
  float e[3][2][4];
  if (PAIRS == 2) {
    // Both subsets' moments from one channel-major copy: subset 0's masked,
    // subset 1's as the block's minus subset 0's (exact integers).
    const uint32_t mask0 = part_mask(part_id, 0);
    Moments ms[2];
    {
      uint32_t cm[4][4];
      channel_major(px, cm);
      ms[0] = moments_cm<CH>(cm, mask0);
      const Moments full = moments_cm<CH>(cm, 0xFFFFu);
      ms[1].n = full.n - ms[0].n;
#pragma unroll
      for (int i = 0; i < 4; i++) ms[1].s[i] = full.s[i] - ms[0].s[i];
#pragma unroll
      for (int i = 0; i < 10; i++) ms[1].q[i] = full.q[i] - ms[0].q[i];
    }
#pragma unroll
    for (int j = 0; j < 2; j++) {
      float axis[4], dc[4];
      pca_axis_moments<CH>(ms[j], axis, dc);
      pca_extent<CH>(px, j == 0 ? mask0 : (~mask0 & 0xFFFFu), axis, dc, e[j][0], e[j][1]);
#pragma unroll
      for (int p = 0; p < CH; p++) {
        e[j][0][p] = rclamp(e[j][0][p], 0.0f, 255.0f);
        e[j][1][p] = rclamp(e[j][1][p], 0.0f, 255.0f);
      }
    }
  } else {
#pragma unroll
    for (int j = 0; j < PAIRS; j++) bc7_segment<CH>(px, part_mask(part_id, j), e[j][0], e[j][1]);
  }
  fit_from_segments<MODE>(px, part_id, e, f, row);
}

// The refinement step of bc7_enc_mode01237 (bc7.rs:571-600).
template <int MODE>
BCN_HD void refine_partition(const uint32_t px[16], int part_id, const uint32_t qblock_in[2],
                             PartFit& f, SubsetRow row = SubsetRow()) {
  constexpr int PAIRS = (MODE == 0 || MODE == 2) ? 3 : 2;
  constexpr int BITS = (MODE == 0 || MODE == 1) ? 3 : 2;
  constexpr int CH = MODE == 7 ? 4 : 3;
  int d[3][2][4];
#pragma unroll
  for (int j = 0; j < PAIRS; j++) {
    float e0[4], e1[4];
    int q0[4], q1[4];
    opt_endpoints<CH, BITS>(px, qblock_in, part_mask(part_id, j), e0, e1);
    quant_pair<MODE, CH>(e0, e1, q0, q1, d[j][0], d[j][1]);
    f.q[j][0] = pack4(q0, CH);
    f.q[j][1] = pack4(q1, CH);
  }
  f.err = block_quant<CH, BITS, PAIRS>(px, d, kPart.subset[part_id], f.qblock, 0u, row);
}

template <int MODE>
BCN_HD void store_partfit(Bc7Best& b, const PartFit& f, int err, int part_id, int order) {
  constexpr int PAIRS = (MODE == 0 || MODE == 2) ? 3 : 2;
  b.err = err;
  b.order = order;
  b.mode = MODE;
  b.part = (uint32_t)part_id;
#pragma unroll
  for (int j = 0; j < 3; j++)
#pragma unroll
    for (int i = 0; i < 2; i++) b.qep[j * 2 + i] = j < PAIRS ? f.q[j][i] : 0u;
  b.qblock[0] = f.qblock[0];
  b.qblock[1] = f.qblock[1];
}

// Refinement and final comparison of bc7_enc_mode01237 (bc7.rs:571-609) for
// the best candidate bf of partition part_id.
template <int MODE>
BCN_HD void finish_mode01237(const uint32_t px[16], const Bc7Params& s, PartFit bf, int part_id,
                             int opaque_err, int order, Bc7Best& best, SubsetRow row = SubsetRow()) {
#pragma unroll 1
  for (uint32_t it = 0; it < s.refine_iterations[MODE]; it++) {
    PartFit f;
    refine_partition<MODE>(px, part_id, bf.qblock, f, row);
    if (f.err < bf.err) bf = f;
  }
  int err = MODE != 7 ? bf.err + opaque_err : bf.err;
  if (better(best, err, order)) store_partfit<MODE>(best, bf, err, part_id, order);
}

// ---- subset fits shared between the 3-subset shapes of modes 0 and 2 -------------
//
// A subset's PCA endpoints and their quantisation depend only on its texel
// mask.  Of the 48 subset masks of the 16 shapes mode 0 tries, 12 repeat an
// earlier one (52 of the 192 of mode 2's 64 shapes).  Every lane runs the same
// shape in these loops, so a compile-time plan says per shape and subset
// whether to fit and keep the result for a later repeat, only fit, or take a
// kept result.  A kept result is its two endpoints' 5-bit channel codes, 15
// bits each, in two entries of the lane's partition-key row, which modes
// 1/3/7 use only afterwards.
enum : uint8_t { kFitOnly = 0, kFitKeep = 1, kReuse = 2 };

struct SubsetPlan {
  uint8_t op[64][3];    // kFitOnly, kFitKeep or kReuse
  uint8_t slot[64][3];  // kept result in row entries 2*slot and 2*slot + 1
  int slots;            // slots held at once, at most
};

// Shapes 64 .. 64 + count - 1 in order.  A kept result stays until the last
// repeat of its mask has read it; that slot may then be reused by a later
// subset of the same shape, whose write follows the read.
constexpr SubsetPlan make_subset_plan(int count) {
  SubsetPlan pl{};
  uint32_t m[64][3] = {};
  for (int p = 0; p < count; p++)
    for (int j = 0; j < 3; j++) m[p][j] = shape_mask(64 + p, j);
  bool busy[32] = {};
  int held = 0;
  for (int p = 0; p < count; p++)
    for (int j = 0; j < 3; j++) {
      int prev = -1;
      bool later = false;
      for (int q = 0; q < count; q++)
        for (int i = 0; i < 3; i++)
          if (m[q][i] == m[p][j]) {
            if (q < p && pl.op[q][i] != kFitOnly) prev = pl.slot[q][i];
            if (q > p) later = true;
          }
      if (prev >= 0) {
        pl.op[p][j] = kReuse;
        pl.slot[p][j] = (uint8_t)prev;
        if (!later) {
          busy[prev] = false;
          held--;
        }
      } else if (later) {
        int f = 0;
        while (busy[f]) f++;
        busy[f] = true;
        held++;
        pl.slots = held > pl.slots ? held : pl.slots;
        pl.op[p][j] = kFitKeep;
        pl.slot[p][j] = (uint8_t)f;
      }
    }
  return pl;
}

constexpr SubsetPlan kPlanMode0Host = make_subset_plan(16);
constexpr SubsetPlan kPlanMode2Host = make_subset_plan(64);
// 6 and 21 slots: 12 and 42 of the row's 64 entries.
static_assert(kPlanMode0Host.slots <= 32 && kPlanMode2Host.slots <= 32,
              "kept subset fits must fit the 32-word partition-key row");
BCN_TABLE SubsetPlan kPlanMode0 = make_subset_plan(16);
BCN_TABLE SubsetPlan kPlanMode2 = make_subset_plan(64);

// Three 5-bit channel codes in 15 bits, and back to one byte per channel.
BCN_HD uint32_t code15(const int q[4]) {
  return (uint32_t)q[0] | ((uint32_t)q[1] << 5) | ((uint32_t)q[2] << 10);
}
BCN_HD uint32_t code15_bytes(uint32_t c) {
  return (c & 0x1Fu) | ((c << 3) & 0x1F00u) | ((c << 6) & 0x1F0000u);
}

// fit_partition for shape 64 + part of mode 0 or 2 following the plan; both
// modes dequantise their 5-bit codes with unpack_to_byte(q, 5).
template <int MODE>
BCN_HD void fit_partition_planned(const uint32_t px[16], int part, uint32_t* row, int stride,
                                  SubsetRow srow, PartFit& f) {
  constexpr int BITS = MODE == 0 ? 3 : 2;
  const int part_id = part + 64;
  int d[3][2][4];
  uint32_t cm[4][4];
  channel_major(px, cm);
#pragma unroll
  for (int j = 0; j < 3; j++) {
    const int op = MODE == 0 ? kPlanMode0.op[part][j] : kPlanMode2.op[part][j];
    const int slot = MODE == 0 ? kPlanMode0.slot[part][j] : kPlanMode2.slot[part][j];
    if (op == kReuse) {
      // A kept fit: both endpoints' 15-bit codes in one word of the row.
      const uint32_t c01 = row[slot * stride], c0 = c01 & 0xFFFFu, c1 = c01 >> 16;
      f.q[j][0] = code15_bytes(c0);
      f.q[j][1] = code15_bytes(c1);
#pragma unroll
      for (int p = 0; p < 3; p++) {
        d[j][0][p] = unpack_to_byte((int)((c0 >> (5 * p)) & 0x1Fu), 5);
        d[j][1][p] = unpack_to_byte((int)((c1 >> (5 * p)) & 0x1Fu), 5);
      }
    } else {
      // bc7_segment<3>, its moments from the channel-major copy.
      const uint32_t mask = part_mask(part_id, j);
      float e0[4], e1[4], axis[4], dc[4];
      pca_axis_moments<3>(moments_cm<3, true>(cm, mask), axis, dc);
      pca_extent<3>(px, mask, axis, dc, e0, e1);
#pragma unroll
      for (int p = 0; p < 3; p++) {
        e0[p] = rclamp(e0[p], 0.0f, 255.0f);
        e1[p] = rclamp(e1[p], 0.0f, 255.0f);
      }
      int q0[4], q1[4];
      quant_pair<MODE, 3>(e0, e1, q0, q1, d[j][0], d[j][1]);
      f.q[j][0] = pack4(q0, 3);
      f.q[j][1] = pack4(q1, 3);
      if (op == kFitKeep) {
        row[slot * stride] = code15(q0) | (code15(q1) << 16);
      }
    }
  }
  f.err = block_quant<3, BITS, 3>(px, d, kPart.subset[part_id], f.qblock, 0u, srow);
}

// bc7_enc_mode02's calls (bc7.rs:612-620): every partition 0..count-1, in
// order (count is 16 for mode 0 and 64 for mode 2).
template <int MODE>
BCN_HD void enc_mode02(const uint32_t px[16], const Bc7Params& s, int count, int opaque_err,
                       int order, uint32_t* row, int stride, SubsetRow srow, Bc7Best& best) {
  PartFit bf;
  bf.err = 0x7FFFFFFF;
  int best_part = 0;
#pragma unroll 1
  for (int part = 0; part < count; part++) {
    PartFit f;
    fit_partition_planned<MODE>(px, part, row, stride, srow, f);
    if (f.err < bf.err) {
      bf = f;
      best_part = part + 64;
    }
  }
  finish_mode01237<MODE>(px, s, bf, best_part, opaque_err, order, best, srow);
}

// The lane's 64 partition keys in 32 words (word w at p[w * stride]): the
// key of partition `part` in the low (even part) or high (odd part) 16 bits
// of word part / 2, as bound * 32 + part / 2.  For 8-bit texels a bound is at
// most sqrt(2 * 4 * 255^2 * 16/4 / 65536) * 256 < 1450, so a key is below
// 1450 * 32 < 2^16 and 0xFFFF (a taken entry) sorts after all of them.
// Within one half the keys order like the reference's part + bound * 64,
// i.e. by (bound, part), since all its parts have the same parity.
struct KeyRow {
  uint32_t* p;
  int stride;
  BCN_HD void put(int part, uint32_t v16) const {
    reinterpret_cast<uint16_t*>(p + (part >> 1) * stride)[part & 1] = (uint16_t)v16;
  }
};

// PCA split bounds of the 64 two-subset partitions (bc7.rs:629-638,
// 824-843) into the key row.
template <int CH>
BCN_HD void partition_bounds(const uint32_t px[16], KeyRow keys) {
  uint32_t cm[4][4];
  channel_major(px, cm);
  // The block's moments as floats, converted once: every moment is an
  // integer below 2^24, so float differences of them are exact.
  float ffs[4], ffq[10];
  {
    const Moments full = moments_cm<CH>(cm, 0xFFFFu);
#pragma unroll
    for (int i = 0; i < 4; i++) ffs[i] = (float)full.s[i];
#pragma unroll
    for (int i = 0; i < 10; i++) ffq[i] = (float)full.q[i];
  }
#pragma unroll 1
  for (int part = 0; part < 64; part++) {
    uint32_t mask = kPart.mask01[part] & 0xFFFFu;
    Moments m = moments_cm<CH, true>(cm, mask);
    float fs0[4], fq0[10];
#pragma unroll
    for (int i = 0; i < 4; i++) fs0[i] = (float)m.s[i];
#pragma unroll
    for (int i = 0; i < 10; i++) fq0[i] = (float)m.q[i];
    float bound = 0.0f;
#pragma unroll
    for (int side = 0; side < 2; side++) {
      float fs[4], fq[10];
      int n;
      if (side == 0) {
        n = m.n;
#pragma unroll
        for (int i = 0; i < 4; i++) fs[i] = fs0[i];
#pragma unroll
        for (int i = 0; i < 10; i++) fq[i] = fq0[i];
      } else {
        // full_stats - stats: exact differences of integers below 2^24, taken
        // in float, which saves converting a second set of moments.
        n = 16 - m.n;
#pragma unroll
        for (int i = 0; i < 4; i++) fs[i] = ffs[i] - fs0[i];
#pragma unroll
        for (int i = 0; i < 10; i++) fq[i] = ffq[i] - fq0[i];
      }
      float c[10];
      // n in [1, 15] and the same in every lane (the shape is the loop's).
      const float fn = in_vgpr(kPartCounts.n[part][side]), rn = in_vgpr(kPartCounts.rcp[part][side]);
      if ((n & (n - 1)) == 0) covar_from_fmoments<CH, true>(c, fs, fq, fn, rn);
      else covar_from_fmoments<CH>(c, fs, fq, fn, rn);
      // get_pca_bound (common.rs:208-245).
      const float inv_var = 1.0f / (256.0f * 256.0f);
      const float eps = 0.001f * 0.001f;
#pragma unroll
      for (int i = 0; i < 10; i++) c[i] *= inv_var;
      c[0] += eps;
      c[4] += eps;
      c[7] += eps;
      float axis[4];
      compute_axis<CH>(axis, c, 4);
      float a0, a1, a2, a3 = 0.0f;
      if (CH == 3) {
        a0 = c[0] * axis[0] + c[1] * axis[1] + c[2] * axis[2];
        a1 = c[1] * axis[0] + c[4] * axis[1] + c[5] * axis[2];
        a2 = c[2] * axis[0] + c[5] * axis[1] + c[7] * axis[2];
      } else {
        a0 = c[0] * axis[0] + c[1] * axis[1] + c[2] * axis[2] + c[3] * axis[3];
        a1 = c[1] * axis[0] + c[4] * axis[1] + c[5] * axis[2] + c[6] * axis[3];
        a2 = c[2] * axis[0] + c[5] * axis[1] + c[7] * axis[2] + c[8] * axis[3];
        a3 = c[3] * axis[0] + c[6] * axis[1] + c[8] * axis[2] + c[9] * axis[3];
      }
      float sq_sum = sq(a0) + sq(a1) + sq(a2);
      if (CH == 4) sq_sum = sq_sum + sq(a3);
      float lambda = sqrt_cr(sq_sum);
      float bb = c[0] + c[4] + c[7];
      if (CH == 4) bb = bb + c[9];
      bb = bb - lambda;
      bb = fmaxf(bb, 0.0f);
      // `bound += b`: the first add is 0.0 + b == b for b >= +0.0.
      bound = side == 0 ? bb : bound + bb;
    }
    float bound12 = sqrt_cr(bound) * 256.0f;
    // finite, 0 <= (int)bound12 < 1450
    keys.put(part, (uint32_t)(int)bound12 * 32u + (uint32_t)(part >> 1));
  }
}

// Next partition in ascending (bound, part) order -- the reference's
// partial_sort_list of part + bound * 64 (common.rs:126-140), consumed one
// entry at a time; marks it taken with 0xFFFF.  One packed minimum per word
// gives the smallest even-part key (low half) and odd-part key (high half);
// keys are unique within a half.  Between the two, lo = bound_a * 32 + w_a
// (part 2 w_a) and hi = bound_b * 32 + w_b (part 2 w_b + 1) compare like
// (bound, part): equal bounds and w_a <= w_b mean 2 w_a < 2 w_b + 1, and
// w_a > w_b means 2 w_b + 1 < 2 w_a.  So lo <= hi picks the even part.
BCN_HD int select_next(KeyRow keys) {
  uint32_t best = 0xFFFFFFFFu;
#pragma unroll 8
  for (int w = 0; w < 32; w++) best = pk_min_u16(best, keys.p[w * keys.stride]);
  const uint32_t lo = best & 0xFFFFu, hi = best >> 16;
  const bool even = lo <= hi;
  const int part = (int)(((even ? lo : hi) & 31u) * 2u + (even ? 0u : 1u));
  keys.put(part, 0xFFFFu);
  return part;
}

// fminf / fmaxf into a float in LDS: one non-returning LDS atomic (only the
// owning lane touches the word, so it is an ordinary read-modify-write).
BCN_HD void lds_fmin(float* p, float v) {
#if defined(__HIP_DEVICE_COMPILE__)
  __hip_atomic_fetch_min(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
#else
  *p = fminf(*p, v);
#endif
}
BCN_HD void lds_fmax(float* p, float v) {
#if defined(__HIP_DEVICE_COMPILE__)
  __hip_atomic_fetch_max(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
#else
  *p = fmaxf(*p, v);
#endif
}

// bc7_segment<3> of both subsets of a 2-subset shape whose masks differ
// between lanes (mask0: subset 0; subset 1 is the rest).  Each texel is
// projected once, onto its own subset's axis, instead of all 16 texels once
// per subset under an execution mask: the subsets' axes and means go to the
// lane's LDS row (words 8j + 0..2 and 8j + 3..5), and each projection,
// computed in pca_segment's order, updates its subset's extent in words
// 8j + 6 and 8j + 7 by an LDS min / max.  A minimum or maximum does not
// depend on the order of its operands (the sign of a zero extent never
// reaches an output, as in pca_segment), so the endpoints are bc7_segment's.
BCN_HD void bc7_segment_pair3(const uint32_t px[16], uint32_t mask0, float e[2][2][4], SubsetRow row) {
  float* r = reinterpret_cast<float*>(row.p);
  const int st = row.stride;
  // Both subsets' moments from one channel-major copy of the block: subset
  // 0's masked, subset 1's as the block's minus subset 0's (exact integers).
  Moments ms[2];
  {
    uint32_t cm[4][4];
    channel_major(px, cm);
    ms[0] = moments_cm<3>(cm, mask0);
    const Moments full = moments_cm<3>(cm, 0xFFFFu);
    ms[1].n = full.n - ms[0].n;
#pragma unroll
    for (int i = 0; i < 4; i++) ms[1].s[i] = full.s[i] - ms[0].s[i];
#pragma unroll
    for (int i = 0; i < 10; i++) ms[1].q[i] = full.q[i] - ms[0].q[i];
  }
#pragma unroll
  for (int j = 0; j < 2; j++) {
    float axis[4], dc[4];
    pca_axis_moments<3>(ms[j], axis, dc);
#pragma unroll
    for (int p = 0; p < 3; p++) {
      r[(8 * j + p) * st] = axis[p];
      r[(8 * j + 3 + p) * st] = dc[p];
    }
    r[(8 * j + 6) * st] = INFINITY;
    r[(8 * j + 7) * st] = -INFINITY;
  }
  launder_texels(px);
#pragma unroll
  for (int k = 0; k < 16; k++) {
    float* q = r + (((mask0 >> k) & 1u) ? 0 : 8 * st);
    const uint32_t t = texel<0>(px, k, 0u);
    float dot = q[0] * (chf(t, 0) - q[3 * st]);
    dot = dot + q[st] * (chf(t, 1) - q[4 * st]);
    dot = dot + q[2 * st] * (chf(t, 2) - q[5 * st]);
    lds_fmin(q + 6 * st, dot);
    lds_fmax(q + 7 * st, dot);
  }
#pragma unroll
  for (int j = 0; j < 2; j++) {
    const float* q = r + 8 * j * st;
    float ext0 = q[6 * st], ext1 = q[7 * st];
    if (ext1 - ext0 < 1.0f) {
      ext0 -= 0.5f;
      ext1 += 0.5f;
    }
#pragma unroll
    for (int p = 0; p < 4; p++) {
      const float a = p < 3 ? q[p * st] : 0.0f, d = p < 3 ? q[(3 + p) * st] : 0.0f;
      e[j][0][p] = p < 3 ? rclamp(ext0 * a + d, 0.0f, 255.0f) : 0.0f;
      e[j][1][p] = p < 3 ? rclamp(ext1 * a + d, 0.0f, 255.0f) : 0.0f;
    }
  }
}

// bc7_enc_mode13 (bc7.rs:622-655).  Mode 1 takes the t1 smallest keys and
// mode 3 the t3 smallest, from the same ascending list; a partition's two
// PCA fits (3 channels, its two subset masks) are identical for both modes,
// so each listed partition is fitted once and quantised for both.
BCN_HD void enc_mode13(const uint32_t px[16], const Bc7Params& s, KeyRow keys, SubsetRow row,
                       int opaque_err, Bc7Best& best) {
  const int t1 = (int)s.fast_skip_threshold_mode1, t3 = (int)s.fast_skip_threshold_mode3;
  if (t1 == 0 && t3 == 0) return;
  partition_bounds<3>(px, keys);
  const int n = t1 > t3 ? t1 : t3;
  PartFit b1, b3;
  b1.err = b3.err = 0x7FFFFFFF;
  int p1 = 0, p3 = 0;
#pragma unroll 1
  for (int r = 0; r < n; r++) {
    int part = select_next(keys);
    float e[3][2][4];
    bc7_segment_pair3(px, part_mask(part, 0), e, row);
    if (r < t1) {
      PartFit f;
      fit_from_segments<1>(px, part, e, f, row);
      if (f.err < b1.err) {
        b1 = f;
        p1 = part;
      }
    }
    if (r < t3) {
      // Keep the mode-3 fit after the mode-1 fit (see `after`).
#pragma unroll
      for (int j = 0; j < 2; j++)
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
          for (int p = 0; p < 3; p++) after(e[j][i][p], b1.err);
      PartFit f;
      fit_from_segments<3>(px, part, e, f, row);
      if (f.err < b3.err) {
        b3 = f;
        p3 = part;
      }
    }
  }
  if (t1 != 0) finish_mode01237<1>(px, s, b1, p1, opaque_err, 2, best, row);
  if (t3 != 0) finish_mode01237<3>(px, s, b3, p3, opaque_err, 3, best, row);
}

// bc7_enc_mode7 (bc7.rs:819-851).
template <int CH>
BCN_HD void enc_mode7(const uint32_t px[16], const Bc7Params& s, KeyRow keys, SubsetRow row,
                      Bc7Best& best) {
  const int t7 = (int)s.fast_skip_threshold_mode7;
  if (t7 == 0) return;
  partition_bounds<CH>(px, keys);
  PartFit bf;
  bf.err = 0x7FFFFFFF;
  int p7 = 0;
#pragma unroll 1
  for (int r = 0; r < t7; r++) {
    int part = select_next(keys);
    PartFit f;
    fit_partition<7>(px, part, f, row);
    if (f.err < bf.err) {
      bf = f;
      p7 = part;
    }
  }
  finish_mode01237<7>(px, s, bf, p7, 0, 4, best, row);
}

// bc7_enc_mode45_candidate (bc7.rs:657-733) for rotation `rotation`: the
// rotated block is read through texel<ROT> (channel byte sh = 8*rotation),
// the scalar channel is byte sh of the original texels, and se0/se1 are the
// rotated block's clamped PCA endpoints (block_segment of the rotated block is
// the same for mode 4 swap 0/1 and mode 5).
template <int MODE, int SWAP, int ROT>
BCN_HD void mode45_candidate(const uint32_t px[16], uint32_t sh, const float se0[4],
                             const float se1[4], const Bc7Params& s, uint32_t rotation, int order,
                             Bc7Best& best) {
  constexpr int BITS = SWAP ? 3 : 2;
  constexpr int ABITS = SWAP ? 2 : (MODE == 4 ? 3 : 2);
  constexpr int AEPBITS = MODE == 4 ? 6 : 8;
  float e0[4] = {se0[0], se0[1], se0[2], se0[3]}, e1[4] = {se1[0], se1[1], se1[2], se1[3]};
  int q[2][4], d[3][2][4];
  quant_pair<MODE, 3>(e0, e1, q[0], q[1], d[0][0], d[0][1]);
  uint32_t qblock[2];
  int err = block_quant45<BITS, ROT>(px, d, rotation, sh, qblock);
#pragma unroll 1
  for (uint32_t it = 0; it < s.refine_iterations[MODE]; it++) {
    opt_endpoints<3, BITS, ROT>(px, qblock, 0xFFFFu, e0, e1, sh);
    quant_pair<MODE, 3>(e0, e1, q[0], q[1], d[0][0], d[0][1]);
    err = block_quant45<BITS, ROT>(px, d, rotation, sh, qblock);
  }
  ChannelFit a = opt_channel<ABITS, AEPBITS>(px, sh, s.refine_iterations_channel);
  err += a.err;
  if (better(best, err, order)) {
    best.err = err;
    best.order = order;
    best.mode = MODE;
    best.rotation = rotation;
    best.swap = SWAP;
    best.qep[0] = (uint32_t)q[0][0] | ((uint32_t)q[0][1] << 8) | ((uint32_t)q[0][2] << 16) |
                  ((uint32_t)a.q0 << 24);
    best.qep[1] = (uint32_t)q[1][0] | ((uint32_t)q[1][1] << 8) | ((uint32_t)q[1][2] << 16) |
                  ((uint32_t)a.q1 << 24);
    best.qblock[0] = qblock[0];
    best.qblock[1] = qblock[1];
    best.aqblock[0] = a.qblock[0];
    best.aqblock[1] = a.qblock[1];
  }
}

template <int ROT>
BCN_HD void mode45_rotation(const uint32_t px[16], const Bc7Params& s, uint32_t rot, Bc7Best& best) {
  const uint32_t sh = 8u * rot;
  float e0[4], e1[4];
  bc7_segment<3, ROT>(px, 0xFFFFu, e0, e1, sh);
  const int r = (int)(rot - s.mode45_channel0);
  mode45_candidate<4, 0, ROT>(px, sh, e0, e1, s, rot, 10 + 2 * r, best);
  mode45_candidate<4, 1, ROT>(px, sh, e0, e1, s, rot, 11 + 2 * r, best);
  mode45_candidate<5, 0, ROT>(px, sh, e0, e1, s, rot, 20 + r, best);
}

// bc7_enc_mode45 (bc7.rs:735-760).  Per rotation: the rotated block's PCA
// fit once, then mode 4 swap 0, mode 4 swap 1 and mode 5.  The reference
// visits all mode-4 candidates before the mode-5 ones; the `order` values
// keep its tie-breaking (mode 4: 10 + 2*r + swap, mode 5: 20 + r).
template <int CH>
BCN_HD void enc_mode45(const uint32_t px[16], const Bc7Params& s, Bc7Best& best) {
#pragma unroll 1
  for (uint32_t rot = s.mode45_channel0; rot < s.channels; rot++) {
    if (rot == 3) mode45_rotation<0>(px, s, rot, best);  // no rotation (channels == 4)
    else if (CH == 4) mode45_rotation<2>(px, s, rot, best);
    else mode45_rotation<1>(px, s, rot, best);
  }
}

// bc7_enc_mode6 (bc7.rs:762-817).
template <int CH>
BCN_HD void enc_mode6(const uint32_t px[16], const Bc7Params& s, Bc7Best& best) {
  float e0[4], e1[4];
  bc7_segment<CH>(px, 0xFFFFu, e0, e1);
  if (CH == 3) {
    e0[3] = 255.0f;
    e1[3] = 255.0f;
  }
  int q[2][4], d[3][2][4];
  quant_pair<6, CH>(e0, e1, q[0], q[1], d[0][0], d[0][1]);
  uint32_t qblock[2];
  int err = block_quant<CH, 4, 1>(px, d, 0u, qblock);
#pragma unroll 1
  for (uint32_t it = 0; it < s.refine_iterations[6]; it++) {
    opt_endpoints<CH, 4>(px, qblock, 0xFFFFu, e0, e1);
    if (CH == 3) {
      // opt_endpoints leaves the alpha endpoints alone: they keep the
      // previous dequantised values.
      e0[3] = (float)d[0][0][3];
      e1[3] = (float)d[0][1][3];
    }
    quant_pair<6, CH>(e0, e1, q[0], q[1], d[0][0], d[0][1]);
    err = block_quant<CH, 4, 1>(px, d, 0u, qblock);
  }
  if (better(best, err, 30)) {
    best.err = err;
    best.order = 30;
    best.mode = 6;
    best.qep[0] = (uint32_t)q[0][0] | ((uint32_t)q[0][1] << 8) | ((uint32_t)q[0][2] << 16) |
                  ((uint32_t)q[0][3] << 24);
    best.qep[1] = (uint32_t)q[1][0] | ((uint32_t)q[1][1] << 8) | ((uint32_t)q[1][2] << 16) |
                  ((uint32_t)q[1][3] << 24);
    best.qblock[0] = qblock[0];
    best.qblock[1] = qblock[1];
  }
}

// ---- bitstream (bc7.rs:381-515, common.rs:520-613) ------------------------------

BCN_HD uint32_t qbyte(const uint32_t qep[6], int e, int p) { return (qep[e] >> (8 * p)) & 0xFFu; }

// Index of texel k.  k is a run-time value for the anchor texels: a select
// rather than qb[k >> 3], so that the winner's record stays in registers (a
// run-time array index puts the whole record in private memory).
BCN_HD uint32_t qidx(const uint32_t qb[2], int k) { return ((k < 8 ? qb[0] : qb[1]) >> (4 * (k & 7))) & 15u; }

BCN_HD void write_out(const BitWriter& w, uint32_t out[4]) {
  out[0] = (uint32_t)w.lo;
  out[1] = (uint32_t)(w.lo >> 32);
  out[2] = (uint32_t)w.hi;
  out[3] = (uint32_t)(w.hi >> 32);
}

template <int MODE>
BCN_HD void code_mode01237(const Bc7Best& b, uint32_t out[4]) {
  constexpr int bits = (MODE == 0 || MODE == 1) ? 3 : 2;
  constexpr int pairs = (MODE == 0 || MODE == 2) ? 3 : 2;
  constexpr int channels = MODE == 7 ? 4 : 3;
  constexpr uint32_t levels = 1u << bits;
  const int part = (int)b.part;
  uint32_t qep[6];
#pragma unroll
  for (int e = 0; e < 6; e++) qep[e] = b.qep[e];
  // Endpoint swap so that each subset's anchor index has MSB 0
  // (bc7_code_apply_swap_mode01237, common.rs:583-613).
  const int a1 = kPart.anchor1[part];
  const int a2 = pairs > 2 ? kPart.anchor2[part] : -1;
  uint32_t flips = 0;
#pragma unroll
  for (int j = 0; j < pairs; j++) {
    int k0 = j == 0 ? 0 : (j == 1 ? a1 : a2);
    if (qidx(b.qblock, k0) >= levels / 2) {
      uint32_t t = qep[2 * j];
      qep[2 * j] = qep[2 * j + 1];
      qep[2 * j + 1] = t;
      flips |= part_mask(part, j);
    }
  }
  BitWriter w;
  w.put((uint32_t)MODE + 1, 1u << MODE);
  if (MODE == 0) w.put(4, (uint32_t)(part & 15));
  else w.put(6, (uint32_t)(part & 63));
#pragma unroll
  for (int p = 0; p < channels; p++)
#pragma unroll
    for (int e = 0; e < pairs * 2; e++) {
      uint32_t v = qbyte(qep, e, p);
      if (MODE == 0) w.put(4, v >> 1);
      else if (MODE == 1) w.put(6, v >> 1);
      else if (MODE == 2) w.put(5, v);
      else if (MODE == 3) w.put(7, v >> 1);
      else w.put(5, v >> 1);
    }
  if (MODE == 1)
#pragma unroll
    for (int j = 0; j < 2; j++) w.put(1, qbyte(qep, 2 * j, 0) & 1u);
  if (MODE == 0 || MODE == 3 || MODE == 7)
#pragma unroll
    for (int e = 0; e < pairs * 2; e++) w.put(1, qbyte(qep, e, 0) & 1u);
#pragma unroll
  for (int k = 0; k < 16; k++) {
    uint32_t q = qidx(b.qblock, k);
    if ((flips >> k) & 1u) q = levels - 1 - q;
    bool anchor = k == 0 || k == a1 || k == a2;
    w.put(anchor ? (uint32_t)bits - 1 : (uint32_t)bits, q);
  }
  write_out(w, out);
}

BCN_HD void code_mode45(const Bc7Best& b, uint32_t out[4]) {
  const int mode = (int)b.mode;
  uint32_t qep[2] = {b.qep[0] & 0xFFFFFFu, b.qep[1] & 0xFFFFFFu};
  uint32_t aqep[2] = {b.qep[0] >> 24, b.qep[1] >> 24};
  uint32_t qblock[2] = {b.qblock[0], b.qblock[1]};
  uint32_t aqblock[2] = {b.aqblock[0], b.aqblock[1]};
  const uint32_t bits = 2, abits = mode == 4 ? 3 : 2;
  const uint32_t epbits = mode == 4 ? 5 : 7, aepbits = mode == 4 ? 6 : 8;
  // bc7_code_apply_swap_mode456 (common.rs:564-581); with swap set the vector
  // and scalar index sets trade places first (bc7.rs:455-463).
  if (b.swap) {
    uint32_t t0 = qblock[0], t1 = qblock[1];
    qblock[0] = aqblock[0];
    qblock[1] = aqblock[1];
    aqblock[0] = t0;
    aqblock[1] = t1;
  }
  uint32_t* vq = b.swap ? aqep : qep;        // endpoints tied to qblock's swap test
  uint32_t* sq2 = b.swap ? qep : aqep;       // endpoints tied to aqblock's swap test
  uint32_t lv = 1u << bits, la = 1u << abits;
  if ((qblock[0] & 15u) >= lv / 2) {
    uint32_t t = vq[0];
    vq[0] = vq[1];
    vq[1] = t;
    for (int i = 0; i < 2; i++) qblock[i] = 0x11111111u * (lv - 1) - qblock[i];
  }
  if ((aqblock[0] & 15u) >= la / 2) {
    uint32_t t = sq2[0];
    sq2[0] = sq2[1];
    sq2[1] = t;
    for (int i = 0; i < 2; i++) aqblock[i] = 0x11111111u * (la - 1) - aqblock[i];
  }
  BitWriter w;
  w.put((uint32_t)mode + 1, 1u << mode);
  w.put(2, (b.rotation + 1) & 3u);
  if (mode == 4) w.put(1, b.swap);
  for (int p = 0; p < 3; p++) {
    w.put(epbits, (qep[0] >> (8 * p)) & 0xFFu);
    w.put(epbits, (qep[1] >> (8 * p)) & 0xFFu);
  }
  w.put(aepbits, aqep[0]);
  w.put(aepbits, aqep[1]);
  for (int k = 0; k < 16; k++) w.put(k == 0 ? bits - 1 : bits, qidx(qblock, k));
  for (int k = 0; k < 16; k++) w.put(k == 0 ? abits - 1 : abits, qidx(aqblock, k));
  write_out(w, out);
}

BCN_HD void code_mode6(const Bc7Best& b, uint32_t out[4]) {
  uint32_t qep[2] = {b.qep[0], b.qep[1]};
  uint32_t qblock[2] = {b.qblock[0], b.qblock[1]};
  if ((qblock[0] & 15u) >= 8u) {
    uint32_t t = qep[0];
    qep[0] = qep[1];
    qep[1] = t;
    for (int i = 0; i < 2; i++) qblock[i] = 0x11111111u * 15u - qblock[i];
  }
  BitWriter w;
  w.put(7, 64);
  for (int p = 0; p < 4; p++) {
    w.put(7, ((qep[0] >> (8 * p)) & 0xFFu) >> 1);
    w.put(7, ((qep[1] >> (8 * p)) & 0xFFu) >> 1);
  }
  w.put(1, qep[0] & 1u);
  w.put(1, qep[1] & 1u);
  for (int k = 0; k < 16; k++) w.put(k == 0 ? 3 : 4, qidx(qblock, k));
  write_out(w, out);
}

// compress_block_bc7_core (bc7.rs:853-867) for one block.  `keys` is a
// 32-word scratch row (the partition keys of modes 1/3/7 as KeyRow, the kept
// subset fits of modes 0/2) and `sdata` an 18-word one (SubsetRow), both read
// and written with `stride`.
template <int CH>
BCN_HD void encode_bc7_block(const uint32_t px[16], const Bc7Params& s, uint32_t* keys, int stride,
                             uint32_t* sdata, uint32_t out[4]) {
  const SubsetRow row{sdata, stride};
  Bc7Best best;
  best.err = 0x7FFFFFFF;  // stands for the reference's f32::INFINITY
  best.order = 0x7FFFFFFF;
  best.mode = 8;
  best.part = best.rotation = best.swap = 0;
  for (int e = 0; e < 6; e++) best.qep[e] = 0;
  best.qblock[0] = best.qblock[1] = best.aqblock[0] = best.aqblock[1] = 0;

  // compute_opaque_err (bc7.rs:869-879): exact integer (<= 16*255^2).
  int opaque_err = 0;
  if (CH == 4)
#pragma unroll
    for (int k = 0; k < 16; k++) {
      int t = chi(px[k], 3) - 255;
      opaque_err += t * t;
    }

  // Each mode group starts from laundered texels (see `fresh`), so the
  // compiler cannot carry values derived from them from one group into the
  // next and keep them live across the whole search.
  uint32_t t[16];
#pragma unroll
  for (int k = 0; k < 16; k++) t[k] = fresh(px[k]);
  if (s.mode_selection[0] != 0) {
    enc_mode02<0>(t, s, 16, opaque_err, 0, keys, stride, row, best);
    if (s.skip_mode2 == 0) enc_mode02<2>(t, s, 64, opaque_err, 1, keys, stride, row, best);
  }
#pragma unroll
  for (int k = 0; k < 16; k++) t[k] = fresh(t[k]);
  if (s.mode_selection[1] != 0) {
    enc_mode13(t, s, KeyRow{keys, stride}, row, opaque_err, best);
#pragma unroll
    for (int k = 0; k < 16; k++) t[k] = fresh(t[k]);
    enc_mode7<CH>(t, s, KeyRow{keys, stride}, row, best);
  }
#pragma unroll
  for (int k = 0; k < 16; k++) t[k] = fresh(t[k]);
  if (s.mode_selection[2] != 0) enc_mode45<CH>(t, s, best);
#pragma unroll
  for (int k = 0; k < 16; k++) t[k] = fresh(t[k]);
  if (s.mode_selection[3] != 0) enc_mode6<CH>(t, s, best);

  switch (best.mode) {
    case 0: code_mode01237<0>(best, out); break;
    case 1: code_mode01237<1>(best, out); break;
    case 2: code_mode01237<2>(best, out); break;
    case 3: code_mode01237<3>(best, out); break;
    case 7: code_mode01237<7>(best, out); break;
    case 4:
    case 5: code_mode45(best, out); break;
    case 6: code_mode6(best, out); break;
    default:
      // No candidate beat the initial error: the reference leaves its
      // all-zero data words in place.
      out[0] = out[1] = out[2] = out[3] = 0;
  }
}

}  // namespace bcn
