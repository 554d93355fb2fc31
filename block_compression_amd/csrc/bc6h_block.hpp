// BC6H (unsigned half float) block encoder: one 4x4 block per lane.
//
// Follows src/encode/bc6h.rs decision for decision.  Unlike BC7, the block
// values here are (f16 bits / 31) * 64 (bc6h.rs:802), not integers, so every
// moment, projection and error is accumulated in f32 in the reference's
// order; only the quantised-endpoint arithmetic is integer.
//
// Work shared bit-exactly across modes: the 32 partition keys and their
// top-k order depend only on the block (bc6h.rs:288-302 recomputes them for
// each of the six two-subset modes), so they are computed once per block.
//
// The endpoint packing writes each mode's fields in the order the BC6H
// decoder reads them (the format's bit layout), from a field table: instead
// of one hand-written routine per mode (bc6h.rs:465-791), field f of mode m
// is (channel, endpoint, first bit, bit count, reversed).
#pragma once
#include "bcn_common.hpp"

namespace bcn {

// BC6HSettings, 5 x u32 repr(C) (src/settings.rs:9-15).
struct Bc6hParams {
  uint32_t slow_mode;
  uint32_t fast_mode;
  uint32_t refine_iterations_1p;
  uint32_t refine_iterations_2p;
  uint32_t fast_skip_threshold;
};

// Endpoint field layout of the 14 modes, in bitstream order after the mode
// bits: channel | endpoint << 2 | first_bit << 4 | bits << 9 | reversed << 15.
// One constexpr statement of the layout: the encoder's packer reads the
// device copy kBc6hFields, the decoder a table derived from it.
struct Bc6hFieldTable {
  uint16_t f[14][23];
};
constexpr Bc6hFieldTable kBc6hFieldSrc = {{
    {0x0249, 0x024A, 0x024E, 0x1400, 0x1401, 0x1402, 0x0A04, 0x024D, 0x0809, 0x0A05, 0x020E, 0x080D, 0x0A06, 0x021E, 0x080A, 0x0A08, 0x022E, 0x0A0C, 0x023E},
    {0x0259, 0x024D, 0x025D, 0x0E00, 0x020E, 0x021E, 0x024A, 0x0E01, 0x025A, 0x022E, 0x0249, 0x0E02, 0x023E, 0x025E, 0x024E, 0x0C04, 0x0809, 0x0C05, 0x080D, 0x0C06, 0x080A, 0x0C08, 0x0C0C},
    {0x1400, 0x1401, 0x1402, 0x0A04, 0x02A0, 0x0809, 0x0805, 0x02A1, 0x020E, 0x080D, 0x0806, 0x02A2, 0x021E, 0x080A, 0x0A08, 0x022E, 0x0A0C, 0x023E},
    {0x1400, 0x1401, 0x1402, 0x0804, 0x02A0, 0x024D, 0x0809, 0x0A05, 0x02A1, 0x080D, 0x0806, 0x02A2, 0x021E, 0x080A, 0x0808, 0x020E, 0x022E, 0x080C, 0x0249, 0x023E},
    {0x1400, 0x1401, 0x1402, 0x0804, 0x02A0, 0x024A, 0x0809, 0x0805, 0x02A1, 0x020E, 0x080D, 0x0A06, 0x02A2, 0x080A, 0x0808, 0x021E, 0x022E, 0x080C, 0x024E, 0x023E},
    {0x1200, 0x024A, 0x1201, 0x0249, 0x1202, 0x024E, 0x0A04, 0x024D, 0x0809, 0x0A05, 0x020E, 0x080D, 0x0A06, 0x021E, 0x080A, 0x0A08, 0x022E, 0x0A0C, 0x023E},
    {0x1000, 0x024D, 0x024A, 0x1001, 0x022E, 0x0249, 0x1002, 0x023E, 0x024E, 0x0C04, 0x0809, 0x0A05, 0x020E, 0x080D, 0x0A06, 0x021E, 0x080A, 0x0C08, 0x0C0C},
    {0x1000, 0x020E, 0x024A, 0x1001, 0x0259, 0x0249, 0x1002, 0x025D, 0x024E, 0x0A04, 0x024D, 0x0809, 0x0C05, 0x080D, 0x0A06, 0x021E, 0x080A, 0x0A08, 0x022E, 0x0A0C, 0x023E},
    {0x1000, 0x021E, 0x024A, 0x1001, 0x025A, 0x0249, 0x1002, 0x025E, 0x024E, 0x0A04, 0x024D, 0x0809, 0x0A05, 0x020E, 0x080D, 0x0C06, 0x080A, 0x0A08, 0x022E, 0x0A0C, 0x023E},
    {0x0C00, 0x024D, 0x020E, 0x021E, 0x024A, 0x0C01, 0x0259, 0x025A, 0x022E, 0x0249, 0x0C02, 0x025D, 0x023E, 0x025E, 0x024E, 0x0C04, 0x0809, 0x0C05, 0x080D, 0x0C06, 0x080A, 0x0C08, 0x0C0C},
    {0x1400, 0x1401, 0x1402, 0x1404, 0x1405, 0x1406},
    {0x1400, 0x1401, 0x1402, 0x1204, 0x02A0, 0x1205, 0x02A1, 0x1206, 0x02A2},
    {0x1400, 0x1401, 0x1402, 0x1004, 0x84A0, 0x1005, 0x84A1, 0x1006, 0x84A2},
    {0x1400, 0x1401, 0x1402, 0x0804, 0x8CA0, 0x0805, 0x8CA1, 0x0806, 0x8CA2},
}};
BCN_TABLE Bc6hFieldTable kBc6hFields = kBc6hFieldSrc;
BCN_TABLE uint8_t kBc6hFieldCount[14] = {19, 23, 18, 20, 20, 19, 19, 21, 21, 23, 6, 9, 9, 9};
// Mode bits (bc6h.rs:131-135; 2 bits for modes 0 and 1, else 5).
BCN_TABLE uint8_t kBc6hPrefix[14] = {0, 1, 2, 6, 10, 14, 18, 22, 26, 30, 3, 7, 11, 15};
// Base-endpoint bits and delta bits of R, G, B per mode (the format's
// per-mode precision table), one constexpr statement for encoder and decoder.
struct Bc6hPrecTable {
  uint8_t base[14];
  uint8_t delta[3][14];
};
constexpr Bc6hPrecTable kBc6hPrecSrc = {{10, 7, 11, 11, 11, 9, 8, 8, 8, 6, 10, 11, 12, 16},
                                        {{5, 6, 5, 4, 4, 5, 6, 5, 5, 6, 10, 9, 8, 4},
                                         {5, 6, 4, 5, 4, 5, 5, 6, 5, 6, 10, 9, 8, 4},
                                         {5, 6, 4, 4, 5, 5, 5, 5, 6, 6, 10, 9, 8, 4}}};
BCN_TABLE Bc6hPrecTable kBc6hPrec = kBc6hPrecSrc;

// bc6h.rs:137-156 (constants as Rust evaluates them, in f32)
BCN_HD float bc6h_span(int m) {
  switch (m) {
    case 0: return 0.9f * 65535.0f / 64.0f;
    case 1: return 0.9f * 65535.0f / 4.0f;
    case 2: return 0.8f * 65535.0f / 256.0f;
    case 5: return 0.9f * 65535.0f / 32.0f;
    case 6: return 0.9f * 65535.0f / 16.0f;
    case 9: return 65535.0f;
    case 10: return 65535.0f;
    case 11: return 0.95f * 65535.0f / 8.0f;
    case 12: return 0.95f * 65535.0f / 32.0f;
    case 13: return 6.0f;
    default: return -1.0f;
  }
}

// bc6h.rs:158-164 (endpoint bits of the mode family)
BCN_HD int bc6h_mode_bits(int m) {
  switch (m) {
    case 0: return 10;
    case 1: return 7;
    case 2: return 11;
    case 5: return 9;
    case 6: return 8;
    case 9: return 6;
    case 10: return 10;
    case 11: return 11;
    case 12: return 12;
    default: return 16;  // 13
  }
}

// bc6h.rs:199-211
BCN_HD int unpack_to_uf16(int v, int bits) {
  if (bits >= 15) return v;
  if (v == 0) return 0;
  if (v == (1 << bits) - 1) return 0xFFFF;
  return (v * 2 + 1) << (15 - bits);
}

struct Bc6hBlock {
  float blk[48];  // channel p of texel k at blk[p*16 + k]
};

// compute_stats_masked (common.rs:294-332) in f32, texel order.  Adding the
// flag-multiplied zero of a masked-out texel is an exact no-op, so only
// masked texels are accumulated.
struct FStats {
  float s[15];
};

BCN_HD FStats fstats(const Bc6hBlock& b, uint32_t mask) {
  FStats st;
#pragma unroll
  for (int i = 0; i < 15; i++) st.s[i] = 0.0f;
#pragma unroll
  for (int k = 0; k < 16; k++) {
    // A branch, not a select: where the mask is the same on every lane (the
    // partition-key loop) the scalar unit skips the texel, elsewhere the
    // accumulation runs under the execution mask.
    if (!((mask >> k) & 1u)) continue;
    float r = b.blk[k], g = b.blk[16 + k], bb = b.blk[32 + k];
    st.s[14] = st.s[14] + 1.0f;
    st.s[10] = st.s[10] + r;
    st.s[11] = st.s[11] + g;
    st.s[12] = st.s[12] + bb;
    st.s[0] = st.s[0] + r * r;
    st.s[1] = st.s[1] + r * g;
    st.s[2] = st.s[2] + r * bb;
    st.s[4] = st.s[4] + g * g;
    st.s[5] = st.s[5] + g * bb;
    st.s[7] = st.s[7] + bb * bb;
  }
  return st;
}

// covar_from_stats (common.rs:334-350), 3 channels.
BCN_HD void fcovar(float c[10], const float* s) {
  c[0] = s[0] - s[10] * s[10] / s[14];
  c[1] = s[1] - s[10] * s[11] / s[14];
  c[2] = s[2] - s[10] * s[12] / s[14];
  c[4] = s[4] - s[11] * s[11] / s[14];
  c[5] = s[5] - s[11] * s[12] / s[14];
  c[7] = s[7] - s[12] * s[12] / s[14];
  c[3] = c[6] = c[8] = c[9] = 0.0f;
}

// Power iteration for 3 channels (common.rs:260-292).
BCN_HD void axis3(float axis[3], const float c[10], int iters) {
  float a0 = 1.0f, a1 = 1.0f, a2 = 1.0f;
  for (int i = 0; i < iters; i++) {
    float x0 = c[0] * a0 + c[1] * a1 + c[2] * a2;
    float x1 = c[1] * a0 + c[4] * a1 + c[5] * a2;
    float x2 = c[2] * a0 + c[5] * a1 + c[7] * a2;
    a0 = x0;
    a1 = x1;
    a2 = x2;
    if (i % 2 == 1) {
      float rnorm = 1.0f / sqrtf(sq(x0) + sq(x1) + sq(x2));
      a0 *= rnorm;
      a1 *= rnorm;
      a2 *= rnorm;
    }
  }
  axis[0] = a0;
  axis[1] = a1;
  axis[2] = a2;
}

// get_pca_bound (common.rs:208-245), 3 channels.
BCN_HD float pca_bound3(const float covar[10]) {
  const float inv_var = 1.0f / (256.0f * 256.0f);
  const float eps = 0.001f * 0.001f;
  float c[10];
#pragma unroll
  for (int i = 0; i < 10; i++) c[i] = covar[i] * inv_var;
  c[0] += eps;
  c[4] += eps;
  c[7] += eps;
  float ax[3];
  axis3(ax, c, 4);
  float a0 = c[0] * ax[0] + c[1] * ax[1] + c[2] * ax[2];
  float a1 = c[1] * ax[0] + c[4] * ax[1] + c[5] * ax[2];
  float a2 = c[2] * ax[0] + c[5] * ax[1] + c[7] * ax[2];
  float lambda = sqrtf(sq(a0) + sq(a1) + sq(a2));
  float b = c[0] + c[4] + c[7];
  b = b - lambda;
  return fmaxf(b, 0.0f);
}

// block_segment_core (common.rs:483-518), no clamp (BC6H).
BCN_HD void fsegment(const Bc6hBlock& b, uint32_t mask, float e0[3], float e1[3]) {
  FStats st = fstats(b, mask);
  float dc[3];
#pragma unroll
  for (int p = 0; p < 3; p++) dc[p] = st.s[10 + p] / st.s[14];
  float c[10];
  fcovar(c, st.s);
  const float inv_var = 1.0f / (256.0f * 256.0f);
  const float eps = 0.001f * 0.001f;
#pragma unroll
  for (int i = 0; i < 10; i++) c[i] *= inv_var;
  c[0] += eps;
  c[4] += eps;
  c[7] += eps;
  float ax[3];
  axis3(ax, c, 8);
  float ext0 = INFINITY, ext1 = -INFINITY;
#pragma unroll
  for (int k = 0; k < 16; k++) {
    // Leading 0.0 of the reference's sum dropped: only a zero's sign can
    // differ, and ext*axis + dc absorbs it.
    if (!((mask >> k) & 1u)) continue;  // a branch, as in fstats
    float dot = ax[0] * (b.blk[k] - dc[0]);
    dot = dot + ax[1] * (b.blk[16 + k] - dc[1]);
    dot = dot + ax[2] * (b.blk[32 + k] - dc[2]);
    ext0 = fminf(ext0, dot);
    ext1 = fmaxf(ext1, dot);
  }
  if (ext1 - ext0 < 1.0f) {
    ext0 -= 0.5f;
    ext1 += 0.5f;
  }
#pragma unroll
  for (int p = 0; p < 3; p++) {
    e0[p] = ext0 * ax[p] + dc[p];
    e1[p] = ext1 * ax[p] + dc[p];
  }
}

// opt_endpoints (common.rs:142-205), 3 channels, f32 sums in texel order.
template <int BITS>
BCN_HD void fopt_endpoints(const Bc6hBlock& b, const uint32_t qblock[2], uint32_t mask,
                           float e0[3], float e1[3]) {
  constexpr int L1 = (1 << BITS) - 1;
  float atb1[3] = {0.0f, 0.0f, 0.0f}, sum[3] = {0.0f, 0.0f, 0.0f};
  float sum_q = 0.0f, sum_qq = 0.0f, sum4 = 0.0f;
#pragma unroll
  for (int k = 0; k < 16; k++) {
    bool on = (mask >> k) & 1u;
    int qi = (int)((qblock[k >> 3] >> (4 * (k & 7))) & 15u);
    float q = (float)qi;
    float x = (float)(L1 - qi);  // (levels - 1) as f32 - q, exact
    sum_q = on ? sum_q + q : sum_q;
    sum_qq = on ? sum_qq + q * q : sum_qq;
    sum4 = on ? sum4 + 1.0f : sum4;
#pragma unroll
    for (int p = 0; p < 3; p++) {
      float v = b.blk[p * 16 + k];
      sum[p] = on ? sum[p] + v : sum[p];
      atb1[p] = on ? atb1[p] + x * v : atb1[p];
    }
  }
  float fl1 = (float)L1;
  float cxx = sum4 * sq(fl1) - 2.0f * fl1 * sum_q + sum_qq;
  float cyy = sum_qq;
  float cxy = fl1 * sum_q - sum_qq;
  float det = cxx * cyy - cxy * cxy;
  float scale = fl1 / det;
  bool flat = fabsf(det) < 0.001f;
#pragma unroll
  for (int p = 0; p < 3; p++) {
    float atb2 = fl1 * sum[p] - atb1[p];
    float x0 = (atb1[p] * cyy - atb2 * cxy) * scale;
    float x1 = (atb2 * cxx - atb1[p] * cxy) * scale;
    float mean = sum[p] / sum4;
    e0[p] = flat ? mean : x0;
    e1[p] = flat ? mean : x1;
  }
}

// block_quant (common.rs:422-481) for 3 channels and NSUB subsets whose
// dequantised endpoints are d[j][0|1][p] (integers as f32 < 2^16).
template <int BITS, int NSUB>
BCN_HD float fblock_quant(const Bc6hBlock& b, const int d[2][2][3], uint32_t pattern,
                          uint32_t qblock[2]) {
  constexpr int L = 1 << BITS;
  float total = 0.0f;
  qblock[0] = qblock[1] = 0;
  // Per subset: endpoint 0 and the endpoint difference as f32 (the
  // projection), and as integers the difference and 32 + 64 * endpoint 0:
  // a + ((w * (b - a) + 32) >> 6) == (w * (b - a) + 32 + 64 * a) >> 6, since
  // 64 * a is a multiple of 64 and >> is floor division.  Endpoints are
  // 16-bit (signed: within +-32767), so |w * (b - a)| and |64 * a| are below
  // 2^23 and the 24-bit multiply-add is exact.
  float fdiv[2], fe0[2][3], fde[2][3];
  int ide[2][3], ic0[2][3];
#pragma unroll
  for (int j = 0; j < NSUB; j++) {
#pragma unroll
    for (int p = 0; p < 3; p++) {
      fe0[j][p] = (float)d[j][0][p];
      fde[j][p] = (float)d[j][1][p] - fe0[j][p];
      ide[j][p] = d[j][1][p] - d[j][0][p];
      ic0[j][p] = 32 + 64 * d[j][0][p];
    }
    float dv = sq(fde[j][0]);
    dv = dv + sq(fde[j][1]);
    dv = dv + sq(fde[j][2]);
    fdiv[j] = dv;
  }
#pragma unroll
  for (int k = 0; k < 16; k++) {
    int j = NSUB == 1 ? 0 : (int)((pattern >> (2 * k)) & 1u);
    float ea[3], de[3];
    int id[3], ic[3];
#pragma unroll
    for (int p = 0; p < 3; p++) {
      ea[p] = j == 0 ? fe0[0][p] : fe0[NSUB - 1][p];
      de[p] = j == 0 ? fde[0][p] : fde[NSUB - 1][p];
      id[p] = j == 0 ? ide[0][p] : ide[NSUB - 1][p];
      ic[p] = j == 0 ? ic0[0][p] : ic0[NSUB - 1][p];
    }
    float div = j == 0 ? fdiv[0] : fdiv[NSUB - 1];
    // Leading 0.0 dropped: a -0.0 first term only flips the sign of a zero
    // projection, which rounds to the same index.  (eb - ea) is the exact
    // difference of two integers below 2^17.
    float proj = (b.blk[k] - ea[0]) * de[0];
    proj = proj + (b.blk[16 + k] - ea[1]) * de[1];
    proj = proj + (b.blk[32 + k] - ea[2]) * de[2];
    proj = proj / div;
    // proj * L is exact (a power-of-two scaling), so the fused form rounds
    // once, as proj * L + 0.5 does.
    int q1 = qclamp(__builtin_fmaf(proj, (float)L, 0.5f), 1.0f, (float)(L - 1));
    int w0 = weight_mad<BITS>(q1 - 1), w1 = weight_mad<BITS>(q1);
    float err0 = 0.0f, err1 = 0.0f;
#pragma unroll
    for (int p = 0; p < 3; p++) {
      float v0 = (float)((mul24(w0, id[p]) + ic[p]) >> 6);
      float v1 = (float)((mul24(w1, id[p]) + ic[p]) >> 6);
      err0 = err0 + sq(v0 - b.blk[p * 16 + k]);
      err1 = err1 + sq(v1 - b.blk[p * 16 + k]);
    }
    bool lo = err0 < err1;
    total = total + (lo ? err0 : err1);
    qblock[k >> 3] |= (uint32_t)(lo ? q1 - 1 : q1) << (4 * (k & 7));
  }
  return total;
}

struct Bc6hState {
  float rgb_bounds[6];
  float max_span;
  int max_span_idx;
  int epb;      // endpoint bits of the current mode
  int mode;     // current mode (0..13)
  int qlo[3], qhi[3];  // qbounds
};

// Signed (SF16) encoding -- not in the reference, whose encoder is unsigned
// only (its decoder reads signed blocks, decode/block.rs:266, 697-719).  The
// same search over the signed value domain: a texel of half bits h is
// +-(|h| / 31) * 32 (the inverse of the signed finish_unquantize,
// block.rs:1391-1397); an endpoint's magnitude quantises to 0..2^(b-1)-1 over
// 0..32767 as the unsigned one does over 0..65535, and dequantises as the
// decoder's signed unquantize (block.rs:1369-1389); the interpolation is the
// decoder's floor (fblock_quant's >> 6).  Endpoints are stored as 16-bit two's
// complement, so every packed field (within the low 16 bits) is exact.
BCN_HD int quant_signed(float e, int bits) {
  const int top = (1 << (bits - 1)) - 1;
  int m = qclamp(fabsf(e) / 32767.0f * (float)top + 0.5f, 0.0f, (float)top);
  return e < 0.0f ? -m : m;
}

BCN_HD int unquant_signed(int v, int bits) {
  if (bits >= 16) return v;
  int m = v < 0 ? -v : v, u;
  if (m == 0)
    u = 0;
  else if (m >= (1 << (bits - 1)) - 1)
    u = 0x7FFF;
  else
    u = ((m << 15) + 0x4000) >> (bits - 1);
  return v < 0 ? -u : u;
}

// ep_quant_dequant_bc6h (bc6h.rs:213-239) of one endpoint.
template <bool SGN>
BCN_HD void bc6h_qd(const Bc6hState& st, const float e[3], int q[3], int d[3]) {
  const int levels = 1 << st.epb;
#pragma unroll
  for (int p = 0; p < 3; p++) {
    int v = SGN ? quant_signed(e[p], st.epb)
                : qclamp(div65535(e[p]) * (float)(levels - 1) + 0.5f, 0.0f,
                         (float)(levels - 1));
    v = iclamp(v, st.qlo[p], st.qhi[p]);
    q[p] = v;
    d[p] = SGN ? unquant_signed(v, st.epb) : unpack_to_uf16(v, st.epb);
  }
}

// compute_qbounds_core (bc6h.rs:175-185) with per-channel spans.
template <bool SGN>
BCN_HD void bc6h_qbounds(Bc6hState& st, const float span[3]) {
  const int levels = 1 << st.epb;
#pragma unroll
  for (int p = 0; p < 3; p++) {
    float middle = (st.rgb_bounds[p] + st.rgb_bounds[3 + p]) / 2.0f;
    float lo = middle - span[p] / 2.0f;
    float hi = middle + span[p] / 2.0f;
    // ep_quant_bc6h_8 (bc6h.rs:166-173)
    if (SGN) {
      st.qlo[p] = quant_signed(lo, st.epb);
      st.qhi[p] = quant_signed(hi, st.epb);
    } else {
      st.qlo[p] = qclamp(div65535(lo) * (float)(levels - 1) + 0.5f, 0.0f,
                         (float)(levels - 1));
      st.qhi[p] = qclamp(div65535(hi) * (float)(levels - 1) + 0.5f, 0.0f,
                         (float)(levels - 1));
    }
  }
}

// bc6h_test_mode (bc6h.rs:414-446).  Returns false if the span test skips it.
template <bool SGN>
BCN_HD bool bc6h_set_mode(Bc6hState& st, int m, float margin) {
  float span = bc6h_span(m);
  if (st.max_span * margin > span) return false;
  st.epb = bc6h_mode_bits(m);
  float sp[3] = {span, span, span};
  if (m >= 10 || m <= 1 || m == 5 || m == 9) {
    st.mode = m;
  } else {
    st.mode = m + st.max_span_idx;
    if (st.max_span_idx < 3) sp[st.max_span_idx] *= 2.0f;
  }
  bc6h_qbounds<SGN>(st, sp);
  return true;
}

// The current winner: mode, partition, quantised endpoints (16 bits each,
// endpoint e channel p at half-word e*3+p), indices.
struct Bc6hBest {
  float err;
  int mode, part;
  uint32_t qep[6];
  uint32_t qblock[2];
};

BCN_HD void store_qep(uint32_t w[6], const int q[4][3]) {
#pragma unroll
  for (int i = 0; i < 6; i++) {
    int a = (2 * i) / 3, pa = (2 * i) % 3, b2 = (2 * i + 1) / 3, pb = (2 * i + 1) % 3;
    w[i] = ((uint32_t)q[a][pa] & 0xFFFFu) | ((uint32_t)q[b2][pb] << 16);
  }
}

BCN_HD int load_qep(const uint32_t w[6], int e, int p) {
  int h = e * 3 + p;
  return (int)((w[h >> 1] >> (16 * (h & 1))) & 0xFFFFu);
}

// bc6h_enc_2p_part_fast (bc6h.rs:304-330) after the two segments: quantise
// the subsets' PCA endpoints e[j][0|1] for the current mode and index the
// block.  Returns the error; q and qb are the candidate.
template <bool SGN>
BCN_HD float bc6h_fit_2p(const Bc6hBlock& b, const Bc6hState& st, int part, const float e[2][2][3],
                         int q[4][3], uint32_t qb[2]) {
  int d[2][2][3];
#pragma unroll
  for (int j = 0; j < 2; j++) {
    bc6h_qd<SGN>(st, e[j][0], q[2 * j], d[j][0]);
    bc6h_qd<SGN>(st, e[j][1], q[2 * j + 1], d[j][1]);
  }
  return fblock_quant<3, 2>(b, d, kPart.subset[part], qb);
}

// The refinement and final comparison of bc6h_enc_2p_list (bc6h.rs:356-388)
// for the best candidate of its list loop (best_err, best_part, bq, bqb).
template <bool SGN>
BCN_HD void bc6h_finish_2p(const Bc6hBlock& b, const Bc6hParams& s, const Bc6hState& st,
                           float best_err, int best_part, int bq[4][3], uint32_t bqb[2],
                           Bc6hBest& best) {
#pragma unroll 1
  for (uint32_t it = 0; it < s.refine_iterations_2p; it++) {
    float e[2][2][3];
#pragma unroll
    for (int j = 0; j < 2; j++) fopt_endpoints<3>(b, bqb, part_mask(best_part, j), e[j][0], e[j][1]);
    int q[4][3];
    uint32_t qb[2];
    float err = bc6h_fit_2p<SGN>(b, st, best_part, e, q, qb);
    if (err < best_err) {
      best_err = err;
#pragma unroll
      for (int e2 = 0; e2 < 4; e2++)
#pragma unroll
        for (int p = 0; p < 3; p++) bq[e2][p] = q[e2][p];
      bqb[0] = qb[0];
      bqb[1] = qb[1];
    }
  }
  if (best_err < best.err) {
    best.err = best_err;
    best.mode = st.mode;
    best.part = best_part;
    store_qep(best.qep, bq);
    best.qblock[0] = bqb[0];
    best.qblock[1] = bqb[1];
  }
}

// bc6h_enc_2p_list (bc6h.rs:326-388) for the current mode over the candidate
// partitions in keys[0..count), the best candidate kept in registers.
template <bool SGN>
BCN_HD void bc6h_enc_2p_list(const Bc6hBlock& b, const Bc6hParams& s, const Bc6hState& st,
                             const int* keys, int stride, int count, Bc6hBest& best) {
  if (count == 0) return;
  float best_err = INFINITY;
  int best_part = keys[0] & 31;
  int bq[4][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  uint32_t bqb[2] = {0, 0};
#pragma unroll 1
  for (int i = 0; i < count; i++) {
    const int part = keys[i * stride] & 31;
    float e[2][2][3];
#pragma unroll
    for (int j = 0; j < 2; j++) fsegment(b, part_mask(part, j), e[j][0], e[j][1]);
    int q[4][3];
    uint32_t qb[2];
    const float err = bc6h_fit_2p<SGN>(b, st, part, e, q, qb);
    if (err < best_err) {
      best_err = err;
      best_part = part;
#pragma unroll
      for (int e2 = 0; e2 < 4; e2++)
#pragma unroll
        for (int p = 0; p < 3; p++) bq[e2][p] = q[e2][p];
      bqb[0] = qb[0];
      bqb[1] = qb[1];
    }
  }
  bc6h_finish_2p<SGN>(b, s, st, best_err, best_part, bq, bqb, best);
}

// Entries of the per-lane row the block encoder uses: the 32 partition keys,
// then an error and a partition for each of up to 6 two-subset modes.
constexpr int kBc6hRow = 32 + 2 * 6;

// Successive bc6h_enc_2p_list calls (bc6h.rs:326-388) for the two-subset
// modes modes[0..n) (mode numbers as bc6h_set_mode takes them, each passing
// its span test) over the same candidate list.  A partition's two PCA segments depend only on the
// block and the subset masks, so the loops are swapped: each candidate's
// segments are fitted once and quantised for every mode, and each mode's
// best error and partition (strict <, first in list order, as in the list
// loop of bc6h_enc_2p_list, bc6h.rs:332-355) are kept in the lane's row
// after the 32 keys.  Then, in the given order, each mode's winner is fitted
// again (the same computation, so the same candidate), refined and compared
// with the block's best (bc6h_finish_2p).
template <bool SGN>
BCN_HD void bc6h_enc_2p_modes(const Bc6hBlock& b, const Bc6hParams& s, Bc6hState& st,
                              const int* modes, int n, int* keys, int stride, int count,
                              Bc6hBest& best) {
  if (count == 0) return;
  int* row = keys + 32 * stride;  // mode i: error bits at row[2i], partition at row[2i + 1]
#pragma unroll 1
  for (int m = 0; m < n; m++) {
    row[(2 * m) * stride] = __builtin_bit_cast(int, INFINITY);
    row[(2 * m + 1) * stride] = keys[0] & 31;
  }
#pragma unroll 1
  for (int i = 0; i < count; i++) {
    const int part = keys[i * stride] & 31;
    float e[2][2][3];
#pragma unroll
    for (int j = 0; j < 2; j++) fsegment(b, part_mask(part, j), e[j][0], e[j][1]);
#pragma unroll 1
    for (int m = 0; m < n; m++) {
      bc6h_set_mode<SGN>(st, modes[m], 0.0f);
      int q[4][3];
      uint32_t qb[2];
      const float err = bc6h_fit_2p<SGN>(b, st, part, e, q, qb);
      if (err < __builtin_bit_cast(float, row[(2 * m) * stride])) {
        row[(2 * m) * stride] = __builtin_bit_cast(int, err);
        row[(2 * m + 1) * stride] = part;
      }
    }
  }
#pragma unroll 1
  for (int m = 0; m < n; m++) {
    bc6h_set_mode<SGN>(st, modes[m], 0.0f);
    const int part = row[(2 * m + 1) * stride];
    float e[2][2][3];
#pragma unroll
    for (int j = 0; j < 2; j++) fsegment(b, part_mask(part, j), e[j][0], e[j][1]);
    int bq[4][3];
    uint32_t bqb[2];
    const float err = bc6h_fit_2p<SGN>(b, st, part, e, bq, bqb);
    bc6h_finish_2p<SGN>(b, s, st, err, part, bq, bqb, best);
  }
}

// bc6h_enc_1p (bc6h.rs:390-412).
// seg0/seg1: the full block's segment (fsegment(b, 0xFFFF)), which does not
// depend on the mode, so the caller fits it once for modes 10-13.
template <bool SGN>
BCN_HD void bc6h_enc_1p(const Bc6hBlock& b, const Bc6hParams& s, const Bc6hState& st,
                        const float seg0[3], const float seg1[3], Bc6hBest& best) {
  float e0[3] = {seg0[0], seg0[1], seg0[2]}, e1[3] = {seg1[0], seg1[1], seg1[2]};
  int q[4][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  int d[2][2][3];
  bc6h_qd<SGN>(st, e0, q[0], d[0][0]);
  bc6h_qd<SGN>(st, e1, q[1], d[0][1]);
  uint32_t qb[2];
  float err = fblock_quant<4, 1>(b, d, 0u, qb);
#pragma unroll 1
  for (uint32_t it = 0; it < s.refine_iterations_1p; it++) {
    fopt_endpoints<4>(b, qb, 0xFFFFu, e0, e1);
    bc6h_qd<SGN>(st, e0, q[0], d[0][0]);
    bc6h_qd<SGN>(st, e1, q[1], d[0][1]);
    err = fblock_quant<4, 1>(b, d, 0u, qb);
  }
  if (err < best.err) {
    best.err = err;
    best.mode = st.mode;
    best.part = 0;
    store_qep(best.qep, q);
    best.qblock[0] = qb[0];
    best.qblock[1] = qb[1];
  }
}

// The 32 two-subset partition keys part + bound*64 (bc6h.rs:289-298), then
// the `count` smallest in ascending order (keys are unique).
BCN_HD void bc6h_keys(const Bc6hBlock& b, int* keys, int stride, int count) {
  FStats full = fstats(b, 0xFFFFu);
#pragma unroll 1
  for (int part = 0; part < 32; part++) {
    FStats st = fstats(b, kPart.mask01[part] & 0xFFFFu);
    float c1[10], c2[10], rest[15];
    fcovar(c1, st.s);
#pragma unroll
    for (int i = 0; i < 15; i++) rest[i] = full.s[i] - st.s[i];
    fcovar(c2, rest);
    float bound = pca_bound3(c1);  // 0.0 + b1 == b1 (b1 >= +0.0)
    bound = bound + pca_bound3(c2);
    keys[part * stride] = part + (int)(sqrtf(bound) * 256.0f) * 64;
  }
  int last = -2147483647 - 1;
#pragma unroll 1
  for (int t = 0; t < count; t++) {
    int bv = 0x7FFFFFFF, bi = 0;
    for (int i = t; i < 32; i++) {
      int v = keys[i * stride];
      bool take = v < bv && v > last;
      bv = take ? v : bv;
      bi = take ? i : bi;
    }
    int tmp = keys[t * stride];
    keys[t * stride] = bv;
    keys[bi * stride] = tmp;
    last = bv;
  }
}

// Bitstream of the winner (bc6h.rs:241-286 with the field-table packer).
BCN_HD void bc6h_code(const Bc6hBest& best, uint32_t out[4]) {
  const int m = best.mode;
  const bool two = m < 10;
  int q[4][3];
#pragma unroll
  for (int e = 0; e < 4; e++)
#pragma unroll
    for (int p = 0; p < 3; p++) q[e][p] = load_qep(best.qep, e, p);
  uint32_t qb[2] = {best.qblock[0], best.qblock[1]};
  uint32_t flips = 0;
  const int bits = two ? 3 : 4;
  if (two) {
    // bc7_code_apply_swap_mode01237 for two subsets (common.rs:583-613).
#pragma unroll
    for (int j = 0; j < 2; j++) {
      int k0 = j == 0 ? 0 : kPart.anchor1[best.part];
      uint32_t qk = (qb[k0 >> 3] >> (4 * (k0 & 7))) & 15u;
      if (qk >= 4u) {
        for (int p = 0; p < 3; p++) {
          int t = q[2 * j][p];
          q[2 * j][p] = q[2 * j + 1][p];
          q[2 * j + 1][p] = t;
        }
        flips |= part_mask(best.part, j);
      }
    }
  } else if ((qb[0] & 15u) >= 8u) {
    // bc7_code_apply_swap_mode456 (common.rs:564-581).
    for (int p = 0; p < 3; p++) {
      int t = q[0][p];
      q[0][p] = q[1][p];
      q[1][p] = t;
    }
    qb[0] = 0x11111111u * 15u - qb[0];
    qb[1] = 0x11111111u * 15u - qb[1];
  }
  // Field values: endpoint 0 absolute; the others as deltas from endpoint 0
  // truncated to the mode's delta bits, except modes 9 and 10 (absolute).
  uint32_t f[12];  // endpoint e, channel p at f[e*3 + p]
#pragma unroll
  for (int e = 0; e < 4; e++)
#pragma unroll
    for (int p = 0; p < 3; p++) {
      int v = q[e][p];
      if (e > 0 && m != 9 && m != 10) v = (v - q[0][p]) & ((1 << kBc6hPrec.delta[p][m]) - 1);
      f[e * 3 + p] = (uint32_t)v;
    }
  BitWriter w;
  w.put(m <= 1 ? 2 : 5, kBc6hPrefix[m]);
  for (int i = 0; i < kBc6hFieldCount[m]; i++) {
    uint32_t t = kBc6hFields.f[m][i];
    uint32_t c = t & 3u, e = (t >> 2) & 3u, first = (t >> 4) & 31u, n = (t >> 9) & 63u;
    uint32_t slot = e * 3 + c, v = 0;
#pragma unroll
    for (uint32_t j = 0; j < 12; j++) v = j == slot ? f[j] : v;
    if (t >> 15) {
      // Reversed field: the bits from first+n-1 down to first.
      for (uint32_t k = 0; k < n; k++) w.put(1, (v >> (first + n - 1 - k)) & 1u);
    } else {
      w.put(n, (v >> first) & ((1u << n) - 1u));
    }
  }
  if (two) w.put(5, (uint32_t)best.part);
  int a1 = two ? kPart.anchor1[best.part] : -1;
  for (int k = 0; k < 16; k++) {
    uint32_t qk = (qb[k >> 3] >> (4 * (k & 7))) & 15u;
    if ((flips >> k) & 1u) qk = (uint32_t)((1 << bits) - 1) - qk;
    bool anchor = k == 0 || k == a1;
    w.put(anchor ? (uint32_t)bits - 1 : (uint32_t)bits, qk);
  }
  out[0] = (uint32_t)w.lo;
  out[1] = (uint32_t)(w.lo >> 32);
  out[2] = (uint32_t)w.hi;
  out[3] = (uint32_t)(w.hi >> 32);
}

// compress_bc6h_core (bc6h.rs:821-860) for one block of f16 bit patterns
// (channel p of texel k at bits[p*16+k]); SGN: the signed (SF16) encoding.
// SLOW: the settings' slow_mode is set (the kernels are instantiated per
// path, each with the registers its own search needs).
template <bool SGN, bool SLOW>
BCN_HD void encode_bc6h_block(const uint16_t bits[48], const Bc6hParams& s, int* keys, int stride,
                              uint32_t out[4]) {
  Bc6hBlock b;
  Bc6hState st;
  // bc6h_setup (bc6h.rs:793-819)
  for (int p = 0; p < 3; p++) {
    st.rgb_bounds[p] = 65535.0f;
    st.rgb_bounds[3 + p] = SGN ? -65535.0f : 0.0f;
  }
#pragma unroll
  for (int p = 0; p < 3; p++)
#pragma unroll
    for (int k = 0; k < 16; k++) {
      const uint32_t h = bits[p * 16 + k];
      float v;
      if (SGN) {
        v = ((float)(h & 0x7FFFu) / 31.0f) * 32.0f;
        if (h & 0x8000u) v = -v;
      } else {
        v = ((float)h / 31.0f) * 64.0f;
      }
      b.blk[p * 16 + k] = v;
      st.rgb_bounds[p] = fminf(st.rgb_bounds[p], v);
      st.rgb_bounds[3 + p] = fmaxf(st.rgb_bounds[3 + p], v);
    }
  st.max_span = 0.0f;
  st.max_span_idx = 0;
  for (int p = 0; p < 3; p++) {
    float span = st.rgb_bounds[3 + p] - st.rgb_bounds[p];
    if (span > st.max_span) {
      st.max_span_idx = p;
      st.max_span = span;
    }
  }
  st.epb = 0;
  st.mode = 0;
  Bc6hBest best;
  best.err = INFINITY;
  best.mode = -1;
  best.part = 0;
  for (int i = 0; i < 6; i++) best.qep[i] = 0;
  best.qblock[0] = best.qblock[1] = 0;

  const int count = (int)s.fast_skip_threshold;
  float seg0[3], seg1[3];
  if (SLOW) {
    // Modes 0, 1, 2, 5, 6, 9 in turn (every span test passes with margin 0).
    const int modes[6] = {0, 1, 2, 5, 6, 9};
    bc6h_keys(b, keys, stride, count);
    bc6h_enc_2p_modes<SGN>(b, s, st, modes, 6, keys, stride, count, best);
    fsegment(b, 0xFFFFu, seg0, seg1);
#pragma unroll 1
    for (int m = 10; m < 14; m++)
      if (bc6h_set_mode<SGN>(st, m, 0.0f)) bc6h_enc_1p<SGN>(b, s, st, seg0, seg1, best);
  } else {
    if (s.fast_skip_threshold > 0) {
      // The span tests only move the current mode; the last one that passes
      // is the mode bc6h_enc_2p_list encodes with (bc6h.rs:836-851).
      bc6h_set_mode<SGN>(st, 9, 0.0f);
      if (s.fast_mode != 0) bc6h_set_mode<SGN>(st, 1, 1.0f);
      bc6h_set_mode<SGN>(st, 6, 1.0f / 1.2f);
      bc6h_set_mode<SGN>(st, 5, 1.0f / 1.2f);
      bc6h_set_mode<SGN>(st, 0, 1.0f / 1.2f);
      bc6h_set_mode<SGN>(st, 2, 1.0f);
      bc6h_keys(b, keys, stride, count);
      bc6h_enc_2p_list<SGN>(b, s, st, keys, stride, count, best);
      if (s.fast_mode == 0 && bc6h_set_mode<SGN>(st, 1, 0.0f))
        bc6h_enc_2p_list<SGN>(b, s, st, keys, stride, count, best);
    }
    bc6h_set_mode<SGN>(st, 10, 0.0f);
    bc6h_set_mode<SGN>(st, 11, 1.0f);
    bc6h_set_mode<SGN>(st, 12, 1.0f);
    bc6h_set_mode<SGN>(st, 13, 1.0f);
    fsegment(b, 0xFFFFu, seg0, seg1);
    bc6h_enc_1p<SGN>(b, s, st, seg0, seg1, best);
  }
  if (best.mode < 0) {
    out[0] = out[1] = out[2] = out[3] = 0;
  } else {
    bc6h_code(best, out);
  }
}

// The two paths chosen at run time (host checks).
template <bool SGN>
BCN_HD void encode_bc6h_block(const uint16_t bits[48], const Bc6hParams& s, int* keys, int stride,
                              uint32_t out[4]) {
  if (s.slow_mode != 0) encode_bc6h_block<SGN, true>(bits, s, keys, stride, out);
  else encode_bc6h_block<SGN, false>(bits, s, keys, stride, out);
}

}  // namespace bcn
