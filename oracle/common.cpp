// TEST INFRASTRUCTURE ONLY -- the CPU oracle (see bc_oracle.hpp).
//
// Restatement of the shared encoder primitives in src/encode/common.rs.
// Each function names the lines it follows.
#include "bc_oracle.hpp"

namespace bco {

#ifdef BC_OPCOUNT
thread_local uint64_t g_ops = 0;
#endif

// common.rs:6-21
int32_t get_unquant_value(uint32_t bits, int32_t index) {
  static const int32_t T2[16] = {0, 21, 43, 64, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  static const int32_t T3[16] = {0, 9, 18, 27, 37, 46, 55, 64, 0, 0, 0, 0, 0, 0, 0, 0};
  static const int32_t T4[16] = {0, 4, 9, 13, 17, 21, 26, 30, 34, 38, 43, 47, 51, 55, 60, 64};
  if (bits == 2) return T2[index];
  if (bits == 3) return T3[index];
  return T4[index];
}

// Packed BC7/BC6H partition tables, in the reference's packing
// (common.rs:24-44, 50-70, 86-96): 2 bits of subset per texel; the texel
// masks of subsets 0 and 1 in the low / high 16 bits; anchor texels of
// subsets 1 and 2 in the high / low nibble.
static const uint32_t PATTERN_TABLE[128] = {
    0x50505050, 0x40404040, 0x54545454, 0x54505040, 0x50404000, 0x55545450, 0x55545040, 0x54504000,
    0x50400000, 0x55555450, 0x55544000, 0x54400000, 0x55555440, 0x55550000, 0x55555500, 0x55000000,
    0x55150100, 0x00004054, 0x15010000, 0x00405054, 0x00004050, 0x15050100, 0x05010000, 0x40505054,
    0x00404050, 0x05010100, 0x14141414, 0x05141450, 0x01155440, 0x00555500, 0x15014054, 0x05414150,
    0x44444444, 0x55005500, 0x11441144, 0x05055050, 0x05500550, 0x11114444, 0x41144114, 0x44111144,
    0x15055054, 0x01055040, 0x05041050, 0x05455150, 0x14414114, 0x50050550, 0x41411414, 0x00141400,
    0x00041504, 0x00105410, 0x10541000, 0x04150400, 0x50410514, 0x41051450, 0x05415014, 0x14054150,
    0x41050514, 0x41505014, 0x40011554, 0x54150140, 0x50505500, 0x00555050, 0x15151010, 0x54540404,
    0xAA685050, 0x6A5A5040, 0x5A5A4200, 0x5450A0A8, 0xA5A50000, 0xA0A05050, 0x5555A0A0, 0x5A5A5050,
    0xAA550000, 0xAA555500, 0xAAAA5500, 0x90909090, 0x94949494, 0xA4A4A4A4, 0xA9A59450, 0x2A0A4250,
    0xA5945040, 0x0A425054, 0xA5A5A500, 0x55A0A0A0, 0xA8A85454, 0x6A6A4040, 0xA4A45000, 0x1A1A0500,
    0x0050A4A4, 0xAAA59090, 0x14696914, 0x69691400, 0xA08585A0, 0xAA821414, 0x50A4A450, 0x6A5A0200,
    0xA9A58000, 0x5090A0A8, 0xA8A09050, 0x24242424, 0x00AA5500, 0x24924924, 0x24499224, 0x50A50A50,
    0x500AA550, 0xAAAA4444, 0x66660000, 0xA5A0A5A0, 0x50A050A0, 0x69286928, 0x44AAAA44, 0x66666600,
    0xAA444444, 0x54A854A8, 0x95809580, 0x96969600, 0xA85454A8, 0x80959580, 0xAA141414, 0x96960000,
    0xAAAA1414, 0xA05050A0, 0xA0A5A5A0, 0x96000000, 0x40804080, 0xA9A8A9A8, 0xAAAAAA44, 0x2A4A5254,
};

static const uint32_t PATTERN_MASK_TABLE[128] = {
    0xCCCC3333, 0x88887777, 0xEEEE1111, 0xECC81337, 0xC880377F, 0xFEEC0113, 0xFEC80137, 0xEC80137F,
    0xC80037FF, 0xFFEC0013, 0xFE80017F, 0xE80017FF, 0xFFE80017, 0xFF0000FF, 0xFFF0000F, 0xF0000FFF,
    0xF71008EF, 0x008EFF71, 0x71008EFF, 0x08CEF731, 0x008CFF73, 0x73108CEF, 0x3100CEFF, 0x8CCE7331,
    0x088CF773, 0x3110CEEF, 0x66669999, 0x366CC993, 0x17E8E817, 0x0FF0F00F, 0x718E8E71, 0x399CC663,
    0xAAAA5555, 0xF0F00F0F, 0x5A5AA5A5, 0x33CCCC33, 0x3C3CC3C3, 0x55AAAA55, 0x96966969, 0xA55A5AA5,
    0x73CE8C31, 0x13C8EC37, 0x324CCDB3, 0x3BDCC423, 0x69969669, 0xC33C3CC3, 0x99666699, 0x0660F99F,
    0x0272FD8D, 0x04E4FB1B, 0x4E40B1BF, 0x2720D8DF, 0xC93636C9, 0x936C6C93, 0x39C6C639, 0x639C9C63,
    0x93366CC9, 0x9CC66339, 0x817E7E81, 0xE71818E7, 0xCCF0330F, 0x0FCCF033, 0x774488BB, 0xEE2211DD,
    0x08CC0133, 0x8CC80037, 0xCC80006F, 0xEC001331, 0x330000FF, 0x00CC3333, 0xFF000033, 0xCCCC0033,
    0x0F0000FF, 0x0FF0000F, 0x00F0000F, 0x44443333, 0x66661111, 0x22221111, 0x136C0013, 0x008C8C63,
    0x36C80137, 0x08CEC631, 0x3330000F, 0xF0000333, 0x00EE1111, 0x88880077, 0x22C0113F, 0x443088CF,
    0x0C22F311, 0x03440033, 0x69969009, 0x9960009F, 0x03303443, 0x00660699, 0xC22C3113, 0x8C0000EF,
    0x1300007F, 0xC4003331, 0x004C1333, 0x22229999, 0x00F0F00F, 0x24929249, 0x29429429, 0xC30C30C3,
    0xC03C3C03, 0x00AA0055, 0xAA0000FF, 0x30300303, 0xC0C03333, 0x90900909, 0xA00A5005, 0xAAA0000F,
    0x0AAA0555, 0xE0E01111, 0x70700707, 0x6660000F, 0x0EE01111, 0x07707007, 0x06660999, 0x660000FF,
    0x00660099, 0x0CC03333, 0x03303003, 0x60000FFF, 0x80807777, 0x10100101, 0x000A0005, 0x08CE8421,
};

static const uint32_t SKIP_TABLE[128] = {
    0xF0, 0xF0, 0xF0, 0xF0, 0xF0, 0xF0, 0xF0, 0xF0, 0xF0, 0xF0, 0xF0, 0xF0, 0xF0, 0xF0, 0xF0, 0xF0,
    0xF0, 0x20, 0x80, 0x20, 0x20, 0x80, 0x80, 0xF0, 0x20, 0x80, 0x20, 0x20, 0x80, 0x80, 0x20, 0x20,
    0xF0, 0xF0, 0x60, 0x80, 0x20, 0x80, 0xF0, 0xF0, 0x20, 0x80, 0x20, 0x20, 0x20, 0xF0, 0xF0, 0x60,
    0x60, 0x20, 0x60, 0x80, 0xF0, 0xF0, 0x20, 0x20, 0xF0, 0xF0, 0xF0, 0xF0, 0xF0, 0x20, 0x20, 0xF0,
    0x3F, 0x38, 0xF8, 0xF3, 0x8F, 0x3F, 0xF3, 0xF8, 0x8F, 0x8F, 0x6F, 0x6F, 0x6F, 0x5F, 0x3F, 0x38,
    0x3F, 0x38, 0x8F, 0xF3, 0x3F, 0x38, 0x6F, 0xA8, 0x53, 0x8F, 0x86, 0x6A, 0x8F, 0x5F, 0xFA, 0xF8,
    0x8F, 0xF3, 0x3F, 0x5A, 0x6A, 0xA8, 0x89, 0xFA, 0xF6, 0x3F, 0xF8, 0x5F, 0xF3, 0xF6, 0xF6, 0xF8,
    0x3F, 0xF3, 0x5F, 0x5F, 0x5F, 0x8F, 0x5F, 0xAF, 0x5F, 0xAF, 0x8F, 0xDF, 0xF3, 0xCF, 0x3F, 0x38,
};

// common.rs:23-47
uint32_t get_pattern(int32_t part_id) { return PATTERN_TABLE[part_id]; }

// common.rs:49-83
uint32_t get_pattern_mask(int32_t part_id, uint32_t j) {
  uint32_t mask_packed = PATTERN_MASK_TABLE[part_id];
  uint32_t mask0 = mask_packed & 0xFFFF;
  uint32_t mask1 = mask_packed >> 16;
  if (j == 2) return ~mask0 & ~mask1;
  if (j == 0) return mask0;
  return mask1;
}

// common.rs:85-101
void get_skips(int32_t part_id, uint32_t skips[3]) {
  uint32_t skip_packed = SKIP_TABLE[part_id];
  skips[0] = 0;
  skips[1] = skip_packed >> 4;
  skips[2] = skip_packed & 15;
}

// common.rs:103-109
void put_bits(uint32_t data[5], uint32_t* pos, uint32_t bits, uint32_t v) {
  data[*pos / 32] |= v << (*pos % 32);
  if (*pos % 32 + bits > 32) {
    // Rust's `v >> 32` would overflow; the branch is only reached with
    // *pos % 32 > 0, so the shift is 1..31.
    data[*pos / 32 + 1] |= v >> (32 - *pos % 32);
  }
  *pos += bits;
}

// common.rs:111-124
void data_shl_1bit_from(uint32_t data[5], size_t from_bits) {
  if (from_bits < 96) {
    uint32_t shifted = (data[2] >> 1) | (data[3] << 31);
    uint32_t mask = ((1u << (from_bits - 64)) - 1) >> 1;
    data[2] = (mask & data[2]) | (~mask & shifted);
    data[3] = (data[3] >> 1) | (data[4] << 31);
    data[4] >>= 1;
  } else if (from_bits < 128) {
    uint32_t shifted = (data[3] >> 1) | (data[4] << 31);
    uint32_t mask = ((1u << (from_bits - 96)) - 1) >> 1;
    data[3] = (mask & data[3]) | (~mask & shifted);
    data[4] >>= 1;
  }
}

// common.rs:126-140 (selection sort of the first `partial_count` entries)
void partial_sort_list(int32_t* list, size_t length, uint32_t partial_count) {
  for (size_t k = 0; k < (size_t)partial_count; k++) {
    size_t best_idx = k;
    int32_t best_value = list[k];
    for (size_t i = k + 1; i < length; i++) {
      if (best_value > list[i]) {
        best_value = list[i];
        best_idx = i;
      }
    }
    int32_t t = list[k];
    list[k] = list[best_idx];
    list[best_idx] = t;
  }
}

// common.rs:142-205
void opt_endpoints(F* ep, const F block[64], uint32_t bits, const uint32_t qblock[2],
                   uint32_t mask, size_t channels) {
  int32_t levels = 1 << bits;
  F atb1[4] = {0.0f, 0.0f, 0.0f, 0.0f};
  F sum_q = 0.0f;
  F sum_qq = 0.0f;
  F sum[5] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f};

  uint32_t mask_shifted = mask << 1;
  for (int k1 = 0; k1 < 2; k1++) {
    uint32_t qbits_shifted = qblock[k1];
    for (int k2 = 0; k2 < 8; k2++) {
      int k = k1 * 8 + k2;
      F q = u2f(qbits_shifted & 15);
      qbits_shifted >>= 4;
      mask_shifted >>= 1;
      if ((mask_shifted & 1) == 0) continue;
      F x = i2f(levels - 1) - q;
      sum_q += q;
      sum_qq += q * q;
      sum[4] += F(1.0f);
      for (size_t p = 0; p < channels; p++) {
        sum[p] += block[k + p * 16];
        atb1[p] += x * block[k + p * 16];
      }
    }
  }

  F atb2[4] = {0.0f, 0.0f, 0.0f, 0.0f};
  for (size_t p = 0; p < channels; p++) atb2[p] = i2f(levels - 1) * sum[p] - atb1[p];

  F cxx = sum[4] * sq(i2f(levels - 1)) - F(2.0f) * i2f(levels - 1) * sum_q + sum_qq;
  F cyy = sum_qq;
  F cxy = i2f(levels - 1) * sum_q - sum_qq;
  F scale = i2f(levels - 1) / (cxx * cyy - cxy * cxy);

  for (size_t p = 0; p < channels; p++) {
    ep[p] = (atb1[p] * cyy - atb2[p] * cxy) * scale;
    ep[4 + p] = (atb2[p] * cxx - atb1[p] * cxy) * scale;
  }

  if (fabs_(cxx * cyy - cxy * cxy) < F(0.001f)) {
    for (size_t p = 0; p < channels; p++) {
      ep[p] = sum[p] / sum[4];
      ep[4 + p] = ep[p];
    }
  }
}

// common.rs:247-251
static void ssymv3(F a[4], const F covar[10], const F b[4]) {
  a[0] = covar[0] * b[0] + covar[1] * b[1] + covar[2] * b[2];
  a[1] = covar[1] * b[0] + covar[4] * b[1] + covar[5] * b[2];
  a[2] = covar[2] * b[0] + covar[5] * b[1] + covar[7] * b[2];
}

// common.rs:253-258
static void ssymv4(F a[4], const F covar[10], const F b[4]) {
  a[0] = covar[0] * b[0] + covar[1] * b[1] + covar[2] * b[2] + covar[3] * b[3];
  a[1] = covar[1] * b[0] + covar[4] * b[1] + covar[5] * b[2] + covar[6] * b[3];
  a[2] = covar[2] * b[0] + covar[5] * b[1] + covar[7] * b[2] + covar[8] * b[3];
  a[3] = covar[3] * b[0] + covar[6] * b[1] + covar[8] * b[2] + covar[9] * b[3];
}

// common.rs:260-292 (power iteration, renormalised every second step)
void compute_axis(F axis[4], const F covar[10], uint32_t power_iterations, size_t channels) {
  F a_vec[4] = {1.0f, 1.0f, 1.0f, 1.0f};
  for (uint32_t i = 0; i < power_iterations; i++) {
    if (channels == 3) ssymv3(axis, covar, a_vec);
    else if (channels == 4) ssymv4(axis, covar, a_vec);
    for (size_t p = 0; p < channels; p++) a_vec[p] = axis[p];
    if (i % 2 == 1) {
      F norm_sq = 0.0f;
      for (size_t p = 0; p < channels; p++) norm_sq += sq(axis[p]);
      F rnorm = F(1.0f) / fsqrt(norm_sq);
      for (size_t p = 0; p < channels; p++) a_vec[p] *= rnorm;
    }
  }
  for (size_t p = 0; p < channels; p++) axis[p] = a_vec[p];
}

// common.rs:208-245
F get_pca_bound(const F covar[10], size_t channels) {
  const uint32_t POWER_ITERATIONS = 4;
  F covar_scaled[10];
  for (int i = 0; i < 10; i++) covar_scaled[i] = covar[i];
  F inv_var = F(1.0f / (256.0f * 256.0f));
  for (int i = 0; i < 10; i++) covar_scaled[i] *= inv_var;

  const F EPS = F(0.001f * 0.001f);  // sq(0.001), a const fn evaluated in f32
  covar_scaled[0] += EPS;
  covar_scaled[4] += EPS;
  covar_scaled[7] += EPS;

  F axis[4] = {0.0f, 0.0f, 0.0f, 0.0f};
  compute_axis(axis, covar_scaled, POWER_ITERATIONS, channels);

  F a_vec[4] = {0.0f, 0.0f, 0.0f, 0.0f};
  if (channels == 3) ssymv3(a_vec, covar_scaled, axis);
  else if (channels == 4) ssymv4(a_vec, covar_scaled, axis);

  F sq_sum = 0.0f;
  for (size_t p = 0; p < channels; p++) sq_sum += sq(a_vec[p]);
  F lambda = fsqrt(sq_sum);

  F bound = covar_scaled[0] + covar_scaled[4] + covar_scaled[7];
  if (channels == 4) bound += covar_scaled[9];
  bound -= lambda;
  return fmax_(bound, F(0.0f));
}

// common.rs:294-332
void compute_stats_masked(F stats[15], const F block[64], uint32_t mask, size_t channels) {
  uint32_t mask_shifted = mask << 1;
  for (int k = 0; k < 16; k++) {
    mask_shifted >>= 1;
    F flag = u2f(mask_shifted & 1);
    F rgba[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    for (size_t p = 0; p < channels; p++) rgba[p] = block[k + p * 16] * flag;
    stats[14] += flag;

    stats[10] += rgba[0];
    stats[11] += rgba[1];
    stats[12] += rgba[2];

    stats[0] += rgba[0] * rgba[0];
    stats[1] += rgba[0] * rgba[1];
    stats[2] += rgba[0] * rgba[2];

    stats[4] += rgba[1] * rgba[1];
    stats[5] += rgba[1] * rgba[2];

    stats[7] += rgba[2] * rgba[2];

    if (channels == 4) {
      stats[13] += rgba[3];
      stats[3] += rgba[0] * rgba[3];
      stats[6] += rgba[1] * rgba[3];
      stats[8] += rgba[2] * rgba[3];
      stats[9] += rgba[3] * rgba[3];
    }
  }
}

// common.rs:334-350
void covar_from_stats(F covar[10], const F stats[15], size_t channels) {
  covar[0] = stats[0] - stats[10] * stats[10] / stats[14];
  covar[1] = stats[1] - stats[10] * stats[11] / stats[14];
  covar[2] = stats[2] - stats[10] * stats[12] / stats[14];

  covar[4] = stats[4] - stats[11] * stats[11] / stats[14];
  covar[5] = stats[5] - stats[11] * stats[12] / stats[14];

  covar[7] = stats[7] - stats[12] * stats[12] / stats[14];

  if (channels == 4) {
    covar[3] = stats[3] - stats[10] * stats[13] / stats[14];
    covar[6] = stats[6] - stats[11] * stats[13] / stats[14];
    covar[8] = stats[8] - stats[12] * stats[13] / stats[14];
    covar[9] = stats[9] - stats[13] * stats[13] / stats[14];
  }
}

// common.rs:352-368
static void compute_covar_dc_masked(F covar[10], F dc[4], const F block[64], uint32_t mask,
                                    size_t channels) {
  F stats[15];
  for (int i = 0; i < 15; i++) stats[i] = 0.0f;
  compute_stats_masked(stats, block, mask, channels);
  for (size_t p = 0; p < channels; p++) dc[p] = stats[10 + p] / stats[14];
  covar_from_stats(covar, stats, channels);
}

// common.rs:370-394
void block_pca_axis(F axis[4], F dc[4], const F block[64], uint32_t mask, size_t channels) {
  const uint32_t POWER_ITERATIONS = 8;
  F covar[10];
  for (int i = 0; i < 10; i++) covar[i] = 0.0f;
  compute_covar_dc_masked(covar, dc, block, mask, channels);

  const F INV_VAR = F(1.0f / (256.0f * 256.0f));
  for (int i = 0; i < 10; i++) covar[i] *= INV_VAR;

  const F EPS = F(0.001f * 0.001f);
  covar[0] += EPS;
  covar[4] += EPS;
  covar[7] += EPS;
  covar[9] += EPS;

  compute_axis(axis, covar, POWER_ITERATIONS, channels);
}

// common.rs:396-420
F block_pca_bound_split(const F block[64], uint32_t mask, const F full_stats[15],
                        size_t channels) {
  F stats[15];
  for (int i = 0; i < 15; i++) stats[i] = 0.0f;
  compute_stats_masked(stats, block, mask, channels);

  F covar1[10];
  for (int i = 0; i < 10; i++) covar1[i] = 0.0f;
  covar_from_stats(covar1, stats, channels);

  for (int i = 0; i < 15; i++) stats[i] = full_stats[i] - stats[i];

  F covar2[10];
  for (int i = 0; i < 10; i++) covar2[i] = 0.0f;
  covar_from_stats(covar2, stats, channels);

  F bound = 0.0f;
  bound += get_pca_bound(covar1, channels);
  bound += get_pca_bound(covar2, channels);
  return fsqrt(bound) * F(256.0f);
}

// common.rs:422-481.  `floor_interp` rounds the interpolation down (>> 6, as
// the decoder does) instead of Rust's truncating `/ 64`; the two differ only
// for negative sums, which occur in the signed BC6H extension alone.
F block_quant(uint32_t qblock[2], const F block[64], uint32_t bits, const F* ep,
              uint32_t pattern, size_t channels, bool floor_interp) {
  F total_err = 0.0f;
  int32_t levels = 1 << bits;
  qblock[0] = 0;
  qblock[1] = 0;

  uint32_t pattern_shifted = pattern;
  for (int k = 0; k < 16; k++) {
    size_t j = pattern_shifted & 3;
    pattern_shifted >>= 2;

    F proj = 0.0f;
    F div = 0.0f;
    for (size_t p = 0; p < channels; p++) {
      F ep_a = ep[8 * j + p];
      F ep_b = ep[8 * j + 4 + p];
      proj += (block[k + p * 16] - ep_a) * (ep_b - ep_a);
      div += sq(ep_b - ep_a);
    }
    proj /= div;

    int32_t q1 = f2i(proj * i2f(levels) + F(0.5f));
    int32_t q1_clamped = iclamp(q1, 1, levels - 1);

    F err0 = 0.0f;
    F err1 = 0.0f;
    int32_t w0 = get_unquant_value(bits, q1_clamped - 1);
    int32_t w1 = get_unquant_value(bits, q1_clamped);

    for (size_t p = 0; p < channels; p++) {
      F ep_a = ep[8 * j + p];
      F ep_b = ep[8 * j + 4 + p];
      int32_t s0 = (64 - w0) * f2i(ep_a) + w0 * f2i(ep_b) + 32;
      int32_t s1 = (64 - w1) * f2i(ep_a) + w1 * f2i(ep_b) + 32;
      F dec_v0 = i2f(floor_interp ? s0 >> 6 : s0 / 64);
      F dec_v1 = i2f(floor_interp ? s1 >> 6 : s1 / 64);
      err0 += sq(dec_v0 - block[k + p * 16]);
      err1 += sq(dec_v1 - block[k + p * 16]);
    }

    F best_err = err1;
    int32_t best_q = q1_clamped;
    if (err0 < err1) {
      best_err = err0;
      best_q = q1_clamped - 1;
    }

    qblock[k / 8] |= ((uint32_t)best_q) << (4 * (k % 8));
    total_err += best_err;
  }
  return total_err;
}

// common.rs:483-518
void block_segment_core(F* ep, const F block[64], uint32_t mask, size_t channels) {
  F axis[4] = {0.0f, 0.0f, 0.0f, 0.0f};
  F dc[4] = {0.0f, 0.0f, 0.0f, 0.0f};
  block_pca_axis(axis, dc, block, mask, channels);

  F ext[2] = {F(INFINITY), F(-INFINITY)};

  uint32_t mask_shifted = mask << 1;
  for (int k = 0; k < 16; k++) {
    mask_shifted >>= 1;
    if ((mask_shifted & 1) == 0) continue;
    F dot = 0.0f;
    for (size_t p = 0; p < channels; p++) dot += axis[p] * (block[16 * p + k] - dc[p]);
    ext[0] = fmin_(ext[0], dot);
    ext[1] = fmax_(ext[1], dot);
  }

  if (ext[1] - ext[0] < F(1.0f)) {
    ext[0] -= F(0.5f);
    ext[1] += F(0.5f);
  }

  for (int i = 0; i < 2; i++)
    for (size_t p = 0; p < channels; p++) ep[4 * i + p] = ext[i] * axis[p] + dc[p];
}

// common.rs:520-547
void bc7_code_qblock(uint32_t data[5], uint32_t* qpos, const uint32_t qblock[2],
                     uint32_t bits, uint32_t flips) {
  uint32_t levels = 1u << bits;
  uint32_t flips_shifted = flips;
  for (int k1 = 0; k1 < 2; k1++) {
    uint32_t qbits_shifted = qblock[k1];
    for (int k2 = 0; k2 < 8; k2++) {
      uint32_t q = qbits_shifted & 15;
      if ((flips_shifted & 1) > 0) q = (levels - 1) - q;
      if (k1 == 0 && k2 == 0) put_bits(data, qpos, bits - 1, q);
      else put_bits(data, qpos, bits, q);
      qbits_shifted >>= 4;
      flips_shifted >>= 1;
    }
  }
}

// common.rs:549-562
void bc7_code_adjust_skip_mode01237(uint32_t data[5], size_t mode, int32_t part_id) {
  size_t pairs = (mode == 0 || mode == 2) ? 3 : 2;
  size_t bits = (mode == 0 || mode == 1) ? 3 : 2;
  uint32_t skips[3];
  get_skips(part_id, skips);
  if (pairs > 2 && skips[1] < skips[2]) {
    uint32_t t = skips[1];
    skips[1] = skips[2];
    skips[2] = t;
  }
  for (size_t i = 1; i < pairs; i++) {
    size_t k = skips[i];
    data_shl_1bit_from(data, 128 + (pairs - 1) - (15 - k) * bits);
  }
}

// common.rs:564-581
void bc7_code_apply_swap_mode456(int32_t* qep, size_t channels, uint32_t qblock[2],
                                 uint32_t bits) {
  uint32_t levels = 1u << bits;
  if ((qblock[0] & 15) >= levels / 2) {
    for (size_t p = 0; p < channels; p++) {
      int32_t t = qep[p];
      qep[p] = qep[channels + p];
      qep[channels + p] = t;
    }
    for (int i = 0; i < 2; i++) qblock[i] = (0x11111111u * (levels - 1)) - qblock[i];
  }
}

// common.rs:583-613
uint32_t bc7_code_apply_swap_mode01237(int32_t qep[24], const uint32_t qblock[2],
                                       size_t mode, int32_t part_id) {
  uint32_t bits = (mode == 0 || mode == 1) ? 3 : 2;
  size_t pairs = (mode == 0 || mode == 2) ? 3 : 2;
  uint32_t flips = 0;
  uint32_t levels = 1u << bits;
  uint32_t skips[3];
  get_skips(part_id, skips);
  for (size_t j = 0; j < pairs; j++) {
    size_t k0 = skips[j];
    uint32_t q = (qblock[k0 >> 3] << (28 - (k0 & 7) * 4)) >> 28;
    if (q >= levels / 2) {
      for (int p = 0; p < 4; p++) {
        int32_t t = qep[8 * j + p];
        qep[8 * j + p] = qep[8 * j + 4 + p];
        qep[8 * j + 4 + p] = t;
      }
      flips |= get_pattern_mask(part_id, (uint32_t)j);
    }
  }
  return flips;
}

}  // namespace bco
