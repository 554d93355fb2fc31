// TEST INFRASTRUCTURE ONLY -- the CPU oracle (see bc_oracle.hpp).
//
// Restatement of the BC7 block encoder in src/encode/bc7.rs.  Each function
// names the lines it follows.
#include "bc_oracle.hpp"

namespace bco {
namespace {

// bc7.rs:4-12
struct Mode45Parameters {
  int32_t qep[8] = {0};
  uint32_t qblock[2] = {0, 0};
  int32_t aqep[2] = {0, 0};
  uint32_t aqblock[2] = {0, 0};
  uint32_t rotation = 0;
  uint32_t swap = 0;
};

// bc7.rs:97-100
int32_t unpack_to_byte(int32_t v, uint32_t bits) {
  int32_t vv = v << (8 - bits);
  return vv + (vv >> bits);
}

// bc7.rs:102-148
void ep_quant0367(int32_t* qep, const F* ep, size_t mode, size_t channels) {
  int32_t bits = mode == 0 ? 4 : (mode == 7 ? 5 : 7);
  int32_t levels = 1 << bits;
  int32_t levels2 = levels * 2 - 1;
  for (int i = 0; i < 2; i++) {
    int32_t qep_b[8] = {0};
    for (int b = 0; b < 2; b++) {
      for (int p = 0; p < 4; p++) {
        int32_t v = f2i((ep[i * 4 + p] / F(255.0f) * i2f(levels2) - i2f(b)) / F(2.0f) +
                        F(0.5f)) * 2 + b;
        qep_b[b * 4 + p] = iclamp(v, b, levels2 - 1 + b);
      }
    }
    F ep_b[8];
    for (int j = 0; j < 8; j++) ep_b[j] = i2f(qep_b[j]);
    if (mode == 0)
      for (int j = 0; j < 8; j++) ep_b[j] = i2f(unpack_to_byte(qep_b[j], 5));

    F err0 = 0.0f;
    F err1 = 0.0f;
    for (size_t p = 0; p < channels; p++) {
      err0 += sq(ep[i * 4 + p] - ep_b[p]);
      err1 += sq(ep[i * 4 + p] - ep_b[4 + p]);
    }
    bool take0 = err0 < err1;
    for (int p = 0; p < 4; p++) qep[i * 4 + p] = take0 ? qep_b[p] : qep_b[4 + p];
  }
}

// bc7.rs:150-178
void ep_quant1(int32_t* qep, const F* ep) {
  int32_t qep_b[16] = {0};
  for (int b = 0; b < 2; b++) {
    for (int i = 0; i < 8; i++) {
      int32_t v = f2i((ep[i] / F(255.0f) * F(127.0f) - i2f(b)) / F(2.0f) + F(0.5f)) * 2 + b;
      qep_b[b * 8 + i] = iclamp(v, b, 126 + b);
    }
  }
  F ep_b[16];
  for (int k = 0; k < 16; k++) ep_b[k] = i2f(unpack_to_byte(qep_b[k], 7));

  F err0 = 0.0f;
  F err1 = 0.0f;
  for (int j = 0; j < 2; j++) {
    for (int p = 0; p < 3; p++) {
      err0 += sq(ep[j * 4 + p] - ep_b[j * 4 + p]);
      err1 += sq(ep[j * 4 + p] - ep_b[8 + j * 4 + p]);
    }
  }
  bool take0 = err0 < err1;
  for (int i = 0; i < 8; i++) qep[i] = take0 ? qep_b[i] : qep_b[8 + i];
}

// bc7.rs:180-189
void ep_quant245(int32_t* qep, const F* ep, size_t mode) {
  int32_t bits = mode == 5 ? 7 : 5;
  int32_t levels = 1 << bits;
  for (int i = 0; i < 8; i++) {
    int32_t v = f2i(ep[i] / F(255.0f) * i2f(levels - 1) + F(0.5f));
    qep[i] = iclamp(v, 0, levels - 1);
  }
}

const size_t PAIRS_TABLE[8] = {3, 2, 3, 2, 1, 1, 1, 2};

// bc7.rs:191-208
void ep_quant(int32_t* qep, F* ep, size_t mode, size_t channels) {
  size_t pairs = PAIRS_TABLE[mode];
  if (mode == 0 || mode == 3 || mode == 6 || mode == 7) {
    for (size_t i = 0; i < pairs; i++) ep_quant0367(qep + i * 8, ep + i * 8, mode, channels);
  } else if (mode == 1) {
    for (size_t i = 0; i < pairs; i++) ep_quant1(qep + i * 8, ep + i * 8);
  } else if (mode == 2 || mode == 4 || mode == 5) {
    for (size_t i = 0; i < pairs; i++) ep_quant245(qep + i * 8, ep + i * 8, mode);
  }
}

// bc7.rs:210-232
void ep_dequant(F* ep, const int32_t* qep, size_t mode) {
  size_t pairs = PAIRS_TABLE[mode];
  if (mode == 3 || mode == 6) {
    for (size_t i = 0; i < 8 * pairs; i++) ep[i] = i2f(qep[i]);
  } else if (mode == 1 || mode == 5) {
    for (size_t i = 0; i < 8 * pairs; i++) ep[i] = i2f(unpack_to_byte(qep[i], 7));
  } else if (mode == 0 || mode == 2 || mode == 4) {
    for (size_t i = 0; i < 8 * pairs; i++) ep[i] = i2f(unpack_to_byte(qep[i], 5));
  } else if (mode == 7) {
    for (size_t i = 0; i < 8 * pairs; i++) ep[i] = i2f(unpack_to_byte(qep[i], 6));
  }
}

// bc7.rs:234-237
void ep_quant_dequant(int32_t* qep, F* ep, size_t mode, size_t channels) {
  ep_quant(qep, ep, mode, channels);
  ep_dequant(ep, qep, mode);
}

// bc7.rs:268-276
void channel_quant_dequant(int32_t qep[2], F ep[2], uint32_t epbits) {
  int32_t elevels = 1 << epbits;
  for (int i = 0; i < 2; i++) {
    int32_t v = f2i(ep[i] / F(255.0f) * i2f(elevels - 1) + F(0.5f));
    qep[i] = iclamp(v, 0, elevels - 1);
    ep[i] = i2f(unpack_to_byte(qep[i], epbits));
  }
}

// bc7.rs:278-320
F channel_opt_quant(uint32_t qblock[2], const F channel_block[16], uint32_t bits,
                    const F ep[2]) {
  int32_t levels = 1 << bits;
  qblock[0] = 0;
  qblock[1] = 0;
  F total_err = 0.0f;
  for (int k = 0; k < 16; k++) {
    F proj = (channel_block[k] - ep[0]) / (ep[1] - ep[0] + F(0.001f));
    int32_t q1 = f2i(proj * i2f(levels) + F(0.5f));
    int32_t q1_clamped = iclamp(q1, 1, levels - 1);

    F err0 = 0.0f;
    F err1 = 0.0f;
    int32_t w0 = get_unquant_value(bits, q1_clamped - 1);
    int32_t w1 = get_unquant_value(bits, q1_clamped);
    F dec_v0 = i2f(((64 - w0) * f2i(ep[0]) + w0 * f2i(ep[1]) + 32) / 64);
    F dec_v1 = i2f(((64 - w1) * f2i(ep[0]) + w1 * f2i(ep[1]) + 32) / 64);
    err0 += sq(dec_v0 - channel_block[k]);
    err1 += sq(dec_v1 - channel_block[k]);

    F best_err = err0 < err1 ? err0 : err1;
    int32_t best_q = err0 < err1 ? q1_clamped - 1 : q1_clamped;

    qblock[k / 8] |= ((uint32_t)best_q) << (4 * (k % 8));
    total_err += best_err;
  }
  return total_err;
}

// bc7.rs:322-369
void channel_opt_endpoints(F ep[2], const F channel_block[16], uint32_t bits,
                           const uint32_t qblock[2]) {
  int32_t levels = 1 << bits;
  F atb1 = 0.0f;
  F sum_q = 0.0f;
  F sum_qq = 0.0f;
  F sum = 0.0f;
  for (int k1 = 0; k1 < 2; k1++) {
    uint32_t qbits_shifted = qblock[k1];
    for (int k2 = 0; k2 < 8; k2++) {
      int k = k1 * 8 + k2;
      F q = u2f(qbits_shifted & 15);
      qbits_shifted >>= 4;
      F x = i2f(levels - 1) - q;
      sum_q += q;
      sum_qq += q * q;
      sum += channel_block[k];
      atb1 += x * channel_block[k];
    }
  }
  F atb2 = i2f(levels - 1) * sum - atb1;
  F cxx = F(16.0f) * sq(i2f(levels - 1)) - F(2.0f) * i2f(levels - 1) * sum_q + sum_qq;
  F cyy = sum_qq;
  F cxy = i2f(levels - 1) * sum_q - sum_qq;
  F scale = i2f(levels - 1) / (cxx * cyy - cxy * cxy);

  ep[0] = (atb1 * cyy - atb2 * cxy) * scale;
  ep[1] = (atb2 * cxx - atb1 * cxy) * scale;
  ep[0] = fclamp(ep[0], F(0.0f), F(255.0f));
  ep[1] = fclamp(ep[1], F(0.0f), F(255.0f));

  if (fabs_(cxx * cyy - cxy * cxy) < F(0.001f)) {
    ep[0] = sum / F(16.0f);
    ep[1] = ep[0];
  }
}

// bc7.rs:371-379
void block_segment(F* ep, const F block[64], uint32_t mask, size_t channels) {
  block_segment_core(ep, block, mask, channels);
  for (int i = 0; i < 2; i++)
    for (size_t p = 0; p < channels; p++) ep[4 * i + p] = fclamp(ep[4 * i + p], F(0.0f), F(255.0f));
}

struct Bc7 {
  F block[64];
  uint32_t data[5];
  F best_err;
  F opaque_err;
  const BC7Settings* settings;

  // bc7.rs:239-266
  F opt_channel(uint32_t qblock[2], int32_t qep[2], const F channel_block[16], uint32_t bits,
                uint32_t epbits) const {
    F ep[2] = {F(255.0f), F(0.0f)};
    for (int k = 0; k < 16; k++) {
      ep[0] = fmin_(ep[0], channel_block[k]);
      ep[1] = fmax_(ep[1], channel_block[k]);
    }
    channel_quant_dequant(qep, ep, epbits);
    F err = channel_opt_quant(qblock, channel_block, bits, ep);
    for (uint32_t it = 0; it < settings->refine_iterations_channel; it++) {
      channel_opt_endpoints(ep, channel_block, bits, qblock);
      channel_quant_dequant(qep, ep, epbits);
      err = channel_opt_quant(qblock, channel_block, bits, ep);
    }
    return err;
  }

  // bc7.rs:381-440
  void code_mode01237(int32_t qep[24], const uint32_t qblock[2], int32_t part_id, size_t mode) {
    uint32_t bits = (mode == 0 || mode == 1) ? 3 : 2;
    size_t pairs = (mode == 0 || mode == 2) ? 3 : 2;
    size_t channels = mode == 7 ? 4 : 3;

    uint32_t flips = bc7_code_apply_swap_mode01237(qep, qblock, mode, part_id);

    std::memset(data, 0, sizeof(data));
    uint32_t pos = 0;
    put_bits(data, &pos, (uint32_t)mode + 1, 1u << mode);
    if (mode == 0) put_bits(data, &pos, 4, (uint32_t)(part_id & 15));
    else put_bits(data, &pos, 6, (uint32_t)(part_id & 63));

    for (size_t p = 0; p < channels; p++) {
      for (size_t j = 0; j < pairs * 2; j++) {
        uint32_t q = (uint32_t)qep[j * 4 + p];
        if (mode == 0) put_bits(data, &pos, 4, q >> 1);
        else if (mode == 1) put_bits(data, &pos, 6, q >> 1);
        else if (mode == 2) put_bits(data, &pos, 5, q);
        else if (mode == 3) put_bits(data, &pos, 7, q >> 1);
        else if (mode == 7) put_bits(data, &pos, 5, q >> 1);
      }
    }
    if (mode == 1)
      for (int j = 0; j < 2; j++) put_bits(data, &pos, 1, (uint32_t)qep[j * 8] & 1);
    if (mode == 0 || mode == 3 || mode == 7)
      for (size_t j = 0; j < pairs * 2; j++) put_bits(data, &pos, 1, (uint32_t)qep[j * 4] & 1);

    bc7_code_qblock(data, &pos, qblock, bits, flips);
    bc7_code_adjust_skip_mode01237(data, mode, part_id);
  }

  // bc7.rs:442-492
  void code_mode45(const Mode45Parameters& params, size_t mode) {
    int32_t qep[8];
    std::memcpy(qep, params.qep, sizeof(qep));
    uint32_t qblock[2] = {params.qblock[0], params.qblock[1]};
    int32_t aqep[2] = {params.aqep[0], params.aqep[1]};
    uint32_t aqblock[2] = {params.aqblock[0], params.aqblock[1]};
    uint32_t rotation = params.rotation;
    uint32_t swap = params.swap;

    uint32_t bits = 2;
    uint32_t abits = mode == 4 ? 3 : 2;
    uint32_t epbits = mode == 4 ? 5 : 7;
    uint32_t aepbits = mode == 4 ? 6 : 8;

    if (swap == 0) {
      bc7_code_apply_swap_mode456(qep, 4, qblock, bits);
      bc7_code_apply_swap_mode456(aqep, 1, aqblock, abits);
    } else {
      for (int i = 0; i < 2; i++) {
        uint32_t t = qblock[i];
        qblock[i] = aqblock[i];
        aqblock[i] = t;
      }
      bc7_code_apply_swap_mode456(aqep, 1, qblock, bits);
      bc7_code_apply_swap_mode456(qep, 4, aqblock, abits);
    }

    std::memset(data, 0, sizeof(data));
    uint32_t pos = 0;
    put_bits(data, &pos, (uint32_t)mode + 1, 1u << mode);
    put_bits(data, &pos, 2, (rotation + 1) & 3);
    if (mode == 4) put_bits(data, &pos, 1, swap);
    for (int p = 0; p < 3; p++) {
      put_bits(data, &pos, epbits, (uint32_t)qep[p]);
      put_bits(data, &pos, epbits, (uint32_t)qep[4 + p]);
    }
    put_bits(data, &pos, aepbits, (uint32_t)aqep[0]);
    put_bits(data, &pos, aepbits, (uint32_t)aqep[1]);
    bc7_code_qblock(data, &pos, qblock, bits, 0);
    bc7_code_qblock(data, &pos, aqblock, abits, 0);
  }

  // bc7.rs:494-515
  void code_mode6(int32_t* qep, uint32_t qblock[2]) {
    bc7_code_apply_swap_mode456(qep, 4, qblock, 4);
    std::memset(data, 0, sizeof(data));
    uint32_t pos = 0;
    put_bits(data, &pos, 7, 64);
    for (int p = 0; p < 4; p++) {
      put_bits(data, &pos, 7, (uint32_t)qep[p] >> 1);
      put_bits(data, &pos, 7, (uint32_t)qep[4 + p] >> 1);
    }
    put_bits(data, &pos, 1, (uint32_t)qep[0] & 1);
    put_bits(data, &pos, 1, (uint32_t)qep[4] & 1);
    bc7_code_qblock(data, &pos, qblock, 4, 0);
  }

  // bc7.rs:517-538
  F enc_mode01237_part_fast(int32_t qep[24], uint32_t qblock[2], int32_t part_id,
                            size_t mode) const {
    uint32_t pattern = get_pattern(part_id);
    uint32_t bits = (mode == 0 || mode == 1) ? 3 : 2;
    size_t pairs = (mode == 0 || mode == 2) ? 3 : 2;
    size_t channels = mode == 7 ? 4 : 3;

    F ep[24];
    for (int i = 0; i < 24; i++) ep[i] = 0.0f;
    for (size_t j = 0; j < pairs; j++) {
      uint32_t mask = get_pattern_mask(part_id, (uint32_t)j);
      block_segment(ep + j * 8, block, mask, channels);
    }
    ep_quant_dequant(qep, ep, mode, channels);
    return block_quant(qblock, block, bits, ep, pattern, channels);
  }

  // bc7.rs:540-610
  void enc_mode01237(size_t mode, const int32_t part_list[64], size_t part_count) {
    if (part_count == 0) return;
    uint32_t bits = (mode == 0 || mode == 1) ? 3 : 2;
    size_t pairs = (mode == 0 || mode == 2) ? 3 : 2;
    size_t channels = mode == 7 ? 4 : 3;

    int32_t best_qep[24] = {0};
    uint32_t best_qblock[2] = {0, 0};
    int32_t best_part_id = -1;
    F best = F(INFINITY);

    for (size_t i = 0; i < part_count; i++) {
      int32_t part_id = part_list[i] & 63;
      part_id = pairs == 3 ? part_id + 64 : part_id;
      int32_t qep[24] = {0};
      uint32_t qblock[2] = {0, 0};
      F err = enc_mode01237_part_fast(qep, qblock, part_id, mode);
      if (err < best) {
        std::memcpy(best_qep, qep, sizeof(int32_t) * 8 * pairs);
        best_qblock[0] = qblock[0];
        best_qblock[1] = qblock[1];
        best_part_id = part_id;
        best = err;
      }
    }

    uint32_t refine_iterations = settings->refine_iterations[mode];
    for (uint32_t it = 0; it < refine_iterations; it++) {
      F ep[24];
      for (int i = 0; i < 24; i++) ep[i] = 0.0f;
      for (size_t j = 0; j < pairs; j++) {
        uint32_t mask = get_pattern_mask(best_part_id, (uint32_t)j);
        opt_endpoints(ep + j * 8, block, bits, best_qblock, mask, channels);
      }
      int32_t qep[24] = {0};
      uint32_t qblock[2] = {0, 0};
      ep_quant_dequant(qep, ep, mode, channels);
      uint32_t pattern = get_pattern(best_part_id);
      F err = block_quant(qblock, block, bits, ep, pattern, channels);
      if (err < best) {
        std::memcpy(best_qep, qep, sizeof(int32_t) * 8 * pairs);
        best_qblock[0] = qblock[0];
        best_qblock[1] = qblock[1];
        best = err;
      }
    }

    if (mode != 7) best += opaque_err;
    if (best < best_err) {
      best_err = best;
      code_mode01237(best_qep, best_qblock, best_part_id, mode);
    }
  }

  // bc7.rs:612-620
  void enc_mode02() {
    int32_t part_list[64];
    for (int part = 0; part < 64; part++) part_list[part] = part;
    enc_mode01237(0, part_list, 16);
    if (settings->skip_mode2 == 0) enc_mode01237(2, part_list, 64);
  }

  // bc7.rs:622-655
  void enc_mode13() {
    if (settings->fast_skip_threshold_mode1 == 0 && settings->fast_skip_threshold_mode3 == 0)
      return;
    F full_stats[15];
    for (int i = 0; i < 15; i++) full_stats[i] = 0.0f;
    compute_stats_masked(full_stats, block, 0xFFFFFFFFu, 3);

    int32_t part_list[64];
    for (int32_t part = 0; part < 64; part++) {
      uint32_t mask = get_pattern_mask(part, 0);
      F bound12 = block_pca_bound_split(block, mask, full_stats, 3);
      int32_t bound = f2i(bound12);
      part_list[part] = part + bound * 64;
    }
    uint32_t partial_count = settings->fast_skip_threshold_mode1 > settings->fast_skip_threshold_mode3
                                 ? settings->fast_skip_threshold_mode1
                                 : settings->fast_skip_threshold_mode3;
    partial_sort_list(part_list, 64, partial_count);
    enc_mode01237(1, part_list, settings->fast_skip_threshold_mode1);
    enc_mode01237(3, part_list, settings->fast_skip_threshold_mode3);
  }

  // bc7.rs:657-733
  void enc_mode45_candidate(Mode45Parameters& best_candidate, F& best, size_t mode,
                            uint32_t rotation, uint32_t swap) const {
    uint32_t bits = 2;
    uint32_t abits = 2;
    uint32_t aepbits = 8;
    if (mode == 4) {
      abits = 3;
      aepbits = 6;
    }
    if (swap == 1) {
      bits = 3;
      abits = 2;
    }

    F candidate_block[64];
    for (int i = 0; i < 64; i++) candidate_block[i] = 0.0f;
    for (int k = 0; k < 16; k++) {
      for (int p = 0; p < 3; p++) candidate_block[k + p * 16] = block[k + p * 16];
      if (rotation < 3) {
        if (settings->channels == 4) candidate_block[k + rotation * 16] = block[k + 3 * 16];
        if (settings->channels == 3) candidate_block[k + rotation * 16] = F(255.0f);
      }
    }

    F ep[8];
    for (int i = 0; i < 8; i++) ep[i] = 0.0f;
    block_segment(ep, candidate_block, 0xFFFFFFFFu, 3);

    int32_t qep[8] = {0};
    ep_quant_dequant(qep, ep, mode, 3);

    uint32_t qblock[2] = {0, 0};
    F err = block_quant(qblock, candidate_block, bits, ep, 0, 3);

    for (uint32_t it = 0; it < settings->refine_iterations[mode]; it++) {
      opt_endpoints(ep, candidate_block, bits, qblock, 0xFFFFFFFFu, 3);
      ep_quant_dequant(qep, ep, mode, 3);
      err = block_quant(qblock, candidate_block, bits, ep, 0, 3);
    }

    F channel_data[16];
    for (int k = 0; k < 16; k++) channel_data[k] = block[k + rotation * 16];

    int32_t aqep[2] = {0, 0};
    uint32_t aqblock[2] = {0, 0};
    err += opt_channel(aqblock, aqep, channel_data, abits, aepbits);

    if (err < best) {
      std::memcpy(best_candidate.qep, qep, sizeof(qep));
      best_candidate.qblock[0] = qblock[0];
      best_candidate.qblock[1] = qblock[1];
      best_candidate.aqblock[0] = aqblock[0];
      best_candidate.aqblock[1] = aqblock[1];
      best_candidate.aqep[0] = aqep[0];
      best_candidate.aqep[1] = aqep[1];
      best_candidate.rotation = rotation;
      best_candidate.swap = swap;
      best = err;
    }
  }

  // bc7.rs:735-760
  void enc_mode45() {
    Mode45Parameters best_candidate;
    F best = best_err;
    uint32_t channel0 = settings->mode45_channel0;
    for (uint32_t p = channel0; p < settings->channels; p++) {
      enc_mode45_candidate(best_candidate, best, 4, p, 0);
      enc_mode45_candidate(best_candidate, best, 4, p, 1);
    }
    if (best < best_err) {
      best_err = best;
      code_mode45(best_candidate, 4);
    }
    for (uint32_t p = channel0; p < settings->channels; p++)
      enc_mode45_candidate(best_candidate, best, 5, p, 0);
    if (best < best_err) {
      best_err = best;
      code_mode45(best_candidate, 5);
    }
  }

  // bc7.rs:762-817
  void enc_mode6() {
    const size_t MODE = 6;
    const uint32_t BITS = 4;
    size_t channels = settings->channels;
    F ep[8];
    for (int i = 0; i < 8; i++) ep[i] = 0.0f;
    block_segment(ep, block, 0xFFFFFFFFu, channels);
    if (settings->channels == 3) {
      ep[3] = F(255.0f);
      ep[7] = F(255.0f);
    }
    int32_t qep[8] = {0};
    ep_quant_dequant(qep, ep, MODE, channels);
    uint32_t qblock[2] = {0, 0};
    F err = block_quant(qblock, block, BITS, ep, 0, channels);
    for (uint32_t it = 0; it < settings->refine_iterations[MODE]; it++) {
      opt_endpoints(ep, block, BITS, qblock, 0xFFFFFFFFu, channels);
      ep_quant_dequant(qep, ep, MODE, channels);
      err = block_quant(qblock, block, BITS, ep, 0, channels);
    }
    if (err < best_err) {
      best_err = err;
      code_mode6(qep, qblock);
    }
  }

  // bc7.rs:819-851
  void enc_mode7() {
    if (settings->fast_skip_threshold_mode7 == 0) return;
    size_t channels = settings->channels;
    F full_stats[15];
    for (int i = 0; i < 15; i++) full_stats[i] = 0.0f;
    compute_stats_masked(full_stats, block, 0xFFFFFFFFu, channels);
    int32_t part_list[64];
    for (int32_t part = 0; part < 64; part++) {
      uint32_t mask = get_pattern_mask(part, 0);
      F bound12 = block_pca_bound_split(block, mask, full_stats, channels);
      int32_t bound = f2i(bound12);
      part_list[part] = part + bound * 64;
    }
    partial_sort_list(part_list, 64, settings->fast_skip_threshold_mode7);
    enc_mode01237(7, part_list, settings->fast_skip_threshold_mode7);
  }

  // bc7.rs:853-867
  void compress_core() {
    if (settings->mode_selection[0] != 0) enc_mode02();
    if (settings->mode_selection[1] != 0) {
      enc_mode13();
      enc_mode7();
    }
    if (settings->mode_selection[2] != 0) enc_mode45();
    if (settings->mode_selection[3] != 0) enc_mode6();
  }

  // bc7.rs:869-879
  void compute_opaque_err() {
    if (settings->channels == 3) {
      opaque_err = 0.0f;
    } else {
      F err = 0.0f;
      for (int k = 0; k < 16; k++) err += sq(block[48 + k] - F(255.0f));
      opaque_err = err;
    }
  }
};

}  // namespace

// bc7.rs:28-63 (new + load_block_interleaved_rgba), then encode.rs:640-644.
float bc7_encode_block(const uint8_t* rgba, size_t stride, const BC7Settings& s,
                       uint8_t out[16]) {
  Bc7 st;
  st.settings = &s;
  std::memset(st.data, 0, sizeof(st.data));
  st.best_err = F(INFINITY);
  st.opaque_err = 0.0f;
  for (int y = 0; y < 4; y++) {
    for (int x = 0; x < 4; x++) {
      const uint8_t* px = rgba + y * stride + x * 4;
      st.block[y * 4 + x] = i2f(px[0]);
      st.block[16 + y * 4 + x] = i2f(px[1]);
      st.block[32 + y * 4 + x] = i2f(px[2]);
      st.block[48 + y * 4 + x] = i2f(px[3]);
    }
  }
  st.compute_opaque_err();
  st.compress_core();
  // bc7.rs:87-95 store_data1: the first four words, little-endian.
  for (int i = 0; i < 4; i++) {
    uint32_t v = st.data[i];
    out[i * 4 + 0] = (uint8_t)v;
    out[i * 4 + 1] = (uint8_t)(v >> 8);
    out[i * 4 + 2] = (uint8_t)(v >> 16);
    out[i * 4 + 3] = (uint8_t)(v >> 24);
  }
  return raw(st.best_err);
}

}  // namespace bco

using namespace bco;

extern "C" {

// Single primitives of bc7.rs for the oracle's pin tests, in plain floats.
// One endpoint pair (8 floats: endpoint A RGBA, endpoint B RGBA) through the
// mode's quantiser: ep_quant0367 (modes 0, 3, 6, 7), ep_quant1 (mode 1) or
// ep_quant245 (modes 2, 4, 5) (bc7.rs:102-189).
void oracle_bc7_ep_quant_pair(int mode, const float* ep8, int channels, int32_t* qep8) {
  F e[8];
  for (int i = 0; i < 8; i++) e[i] = F(ep8[i]);
  for (int i = 0; i < 8; i++) qep8[i] = 0;
  if (mode == 0 || mode == 3 || mode == 6 || mode == 7) ep_quant0367(qep8, e, (size_t)mode, (size_t)channels);
  else if (mode == 1) ep_quant1(qep8, e);
  else ep_quant245(qep8, e, (size_t)mode);
}

// ep_dequant of one pair (bc7.rs:210-232).
void oracle_bc7_ep_dequant_pair(int mode, const int32_t* qep8, float* ep8) {
  F e[8];
  ep_dequant(e, qep8, (size_t)mode);
  for (int i = 0; i < 8; i++) ep8[i] = raw(e[i]);
}

// channel_opt_endpoints (bc7.rs:322-369): scalar-channel least squares.
void oracle_bc7_channel_opt_endpoints(const float* channel16, uint32_t bits, const uint32_t* qblock2,
                                      float* ep2) {
  F c[16], e[2] = {0.0f, 0.0f};
  for (int i = 0; i < 16; i++) c[i] = F(channel16[i]);
  channel_opt_endpoints(e, c, bits, qblock2);
  ep2[0] = raw(e[0]);
  ep2[1] = raw(e[1]);
}

// channel_opt_quant (bc7.rs:278-320).
float oracle_bc7_channel_opt_quant(const float* channel16, uint32_t bits, const float* ep2,
                                   uint32_t* qblock2) {
  F c[16], e[2] = {F(ep2[0]), F(ep2[1])};
  for (int i = 0; i < 16; i++) c[i] = F(channel16[i]);
  return raw(channel_opt_quant(qblock2, c, bits, e));
}

}  // extern "C"
