// CPU build of the product's per-block kernel code (the same headers the HIP
// kernels include), so the tests can compare the kernel math with the oracle
// without a GPU.  Built by tests/kernel_check/build.sh into
// tests/kernel_check/_build/libkernel_check.so.
#include <cmath>
#include <cstring>

#include "../../block_compression_amd/csrc/bc7_block.hpp"

extern "C" int kcheck_bc7_encode(const void* settings, const uint8_t* rgba, uint32_t width,
                                 uint32_t height, uint32_t stride, uint8_t* out) {
  bcn::Bc7Params s;
  std::memcpy(&s, settings, sizeof(s));
  uint32_t bw = width / 4, bh = height / 4;
  for (uint32_t idx = 0; idx < bw * bh; idx++) {
    uint32_t yy = idx / bw, xx = idx % bw;
    uint32_t px[16];
    for (int y = 0; y < 4; y++)
      for (int x = 0; x < 4; x++) {
        const uint8_t* p = rgba + (size_t)(yy * 4 + y) * stride + (xx * 4 + x) * 4;
        px[y * 4 + x] = (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) |
                        ((uint32_t)p[3] << 24);
      }
    uint32_t keys[32];
    uint32_t sdata[18];
    uint32_t o[4];
    if (s.channels == 4) bcn::encode_bc7_block<4>(px, s, keys, 1, sdata, o);
    else bcn::encode_bc7_block<3>(px, s, keys, 1, sdata, o);
    std::memcpy(out + (size_t)idx * 16, o, 16);
  }
  return 0;
}

// pbit_codes<L> against the reference's expression, 2 * clamp((int)((y - b)
// / 2 + 0.5), 0, L - 1) + b with Rust's saturating, NaN-to-0 conversion: every
// float y with 2^-10 <= |y| in [-4, 2L + 4], every 1024th bit pattern of the
// smaller magnitudes (all of which give codes 0 and 1), zeros, NaN and the
// infinities.  Returns the number of mismatches.
template <int L>
static long pbit_mismatches() {
  auto direct = [](float y, int b) {
    const float x = (y - (float)b) / 2.0f + 0.5f;
    long v = x != x ? 0 : (x >= 2147483647.0f ? 2147483647L : (x <= -2147483648.0f ? -2147483648L : (long)x));
    v = v < 0 ? 0 : (v > L - 1 ? L - 1 : v);
    return (int)(2 * v + b);
  };
  long bad = 0;
  auto check = [&](float y) {
    int q0, q1;
    bcn::pbit_codes<L>(y, q0, q1);
    bad += (q0 != direct(y, 0)) + (q1 != direct(y, 1));
  };
  auto bits = [](float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; };
  auto flt = [](uint32_t u) { float f; std::memcpy(&f, &u, 4); return f; };
  const uint32_t tiny = bits(0x1p-10f);
  for (uint32_t u = tiny; u <= bits((float)(2 * L + 4)); u++) check(flt(u));
  for (uint32_t u = tiny; u <= bits(4.0f); u++) check(-flt(u));
  for (uint32_t u = 0; u < tiny; u += 1024) {
    check(flt(u));
    check(-flt(u));
  }
  check(NAN);
  check(INFINITY);
  check(-INFINITY);
  return bad;
}

extern "C" long kcheck_pbit_mismatches(int levels) {
  switch (levels) {
    case 16: return pbit_mismatches<16>();
    case 32: return pbit_mismatches<32>();
    case 64: return pbit_mismatches<64>();
    case 128: return pbit_mismatches<128>();
    default: return -1;
  }
}

#include "../../block_compression_amd/csrc/bc15_block.hpp"

// BC1-BC5 through the same per-block functions as the HIP kernel.
extern "C" int kcheck_bc15_encode(int fmt, const uint8_t* rgba, uint32_t width, uint32_t height,
                                  uint32_t stride, uint8_t* out) {
  uint32_t bw = width / 4, bh = height / 4;
  int words = (fmt == 0 || fmt == 3) ? 2 : 4;
  for (uint32_t idx = 0; idx < bw * bh; idx++) {
    uint32_t yy = idx / bw, xx = idx % bw;
    uint32_t px[16];
    for (int y = 0; y < 4; y++)
      for (int x = 0; x < 4; x++) {
        const uint8_t* p = rgba + (size_t)(yy * 4 + y) * stride + (xx * 4 + x) * 4;
        px[y * 4 + x] = (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) |
                        ((uint32_t)p[3] << 24);
      }
    uint32_t o[4];
    switch (fmt) {
      case 0: bcn::bc1_core(px, o); break;
      case 1: bcn::bc2_alpha(px, o); bcn::bc1_core(px, o + 2); break;
      case 2: bcn::bc3_alpha_block(px, 3, o); bcn::bc1_core(px, o + 2); break;
      case 3: bcn::bc3_alpha_block(px, 0, o); break;
      case 4: bcn::bc3_alpha_block(px, 0, o); bcn::bc3_alpha_block(px, 1, o + 2); break;
      default: return -1;
    }
    std::memcpy(out + (size_t)idx * words * 4, o, words * 4);
  }
  return 0;
}

#include "../../block_compression_amd/csrc/bc6h_block.hpp"

// BC6H through the same per-block function as the HIP kernel.  `rgba16` are
// f16 bit patterns, stride in f16 elements; is_signed selects BC6H_SF16.
extern "C" int kcheck_bc6h_encode(const void* settings, const uint16_t* rgba16, uint32_t width,
                                  uint32_t height, uint32_t stride, uint8_t* out, int is_signed) {
  bcn::Bc6hParams s;
  std::memcpy(&s, settings, sizeof(s));
  uint32_t bw = width / 4, bh = height / 4;
  for (uint32_t idx = 0; idx < bw * bh; idx++) {
    uint32_t yy = idx / bw, xx = idx % bw;
    uint16_t bits[48];
    for (int y = 0; y < 4; y++)
      for (int x = 0; x < 4; x++) {
        const uint16_t* p = rgba16 + (size_t)(yy * 4 + y) * stride + (xx * 4 + x) * 4;
        for (int c = 0; c < 3; c++) bits[c * 16 + y * 4 + x] = p[c];
      }
    int keys[bcn::kBc6hRow];
    uint32_t o[4];
    if (is_signed)
      bcn::encode_bc6h_block<true>(bits, s, keys, 1, o);
    else
      bcn::encode_bc6h_block<false>(bits, s, keys, 1, o);
    std::memcpy(out + (size_t)idx * 16, o, 16);
  }
  return 0;
}

#include "../../block_compression_amd/csrc/decode_block.hpp"

// Decoders: `count` blocks of `bb` bytes -> 16 texels each.  fmt 0..4, 6
// write RGBA8 texels (BC4 -> R,0,0,0; BC5 -> R,G,0,0); fmt 5 writes 48 f16
// bit patterns (RGB per texel), signed if bit 0 of is_signed is set, with
// one-subset blocks through the one-subset path if bit 1 is.
extern "C" int kcheck_decode(int fmt, int is_signed, const uint8_t* blocks, size_t count, void* out) {
  size_t bb = (fmt == 0 || fmt == 3) ? 8 : 16;
  for (size_t i = 0; i < count; i++) {
    uint32_t in[4] = {0, 0, 0, 0};
    std::memcpy(in, blocks + i * bb, bb);
    if (fmt == 5) {
      uint16_t h[48];
      bcn::decode_bc6h(in, (is_signed & 1) != 0, h, (is_signed & 2) != 0);
      std::memcpy((uint16_t*)out + i * 48, h, sizeof(h));
      continue;
    }
    uint32_t px[16];
    // The BC7 decoder's record array, as lane 0 of a workgroup.
    static bcn::Bc7SubsetData sub[3 * bcn::kBc7SubRows];
    switch (fmt) {
      case 0: bcn::decode_rgba8_block<0>(in, px, sub, 0); break;
      case 1: bcn::decode_rgba8_block<1>(in, px, sub, 0); break;
      case 2: bcn::decode_rgba8_block<2>(in, px, sub, 0); break;
      case 3: bcn::decode_rgba8_block<3>(in, px, sub, 0); break;
      case 4: bcn::decode_rgba8_block<4>(in, px, sub, 0); break;
      default: bcn::decode_rgba8_block<6>(in, px, sub, 0); break;
    }
    std::memcpy((uint32_t*)out + i * 16, px, sizeof(px));
  }
  return 0;
}

// ---- decode adjoints and refinement (decode_adjoint.hpp) ----------------------
#include "../../block_compression_amd/csrc/decode_adjoint.hpp"

extern "C" void kcheck_bc7_repack(const uint32_t* blocks, size_t n, uint32_t* out) {
  for (size_t b = 0; b < n; b++) {
    bcn::Bc7Form f;
    bcn::bc7_parse(blocks + 4 * b, f);
    bcn::bc7_pack(f, out + 4 * b);
  }
}

// fmt 6 = BC7 (x: 24 floats per block), 0 = BC1 (x: 6 floats per block).
extern "C" void kcheck_endpoints(int fmt, const uint32_t* blocks, size_t n, float* x) {
  for (size_t b = 0; b < n; b++) {
    if (fmt == 6) {
      bcn::Bc7Form f;
      bcn::bc7_parse(blocks + 4 * b, f);
      for (int e = 0; e < 6; e++)
        for (int c = 0; c < 4; c++) x[b * 24 + e * 4 + c] = (float)f.q[e][c];
    } else {
      bcn::Bc1Form f;
      bcn::bc1_parse(blocks + 2 * b, f);
      for (int e = 0; e < 2; e++)
        for (int c = 0; c < 3; c++) x[b * 6 + e * 3 + c] = (float)f.q[e][c];
    }
  }
}

extern "C" void kcheck_forward(int fmt, const uint32_t* blocks, const float* x, size_t n, int smooth,
                               float* texels) {
  for (size_t b = 0; b < n; b++) {
    float t[16][4];
    if (fmt == 6) {
      bcn::Bc7Form f;
      bcn::bc7_parse(blocks + 4 * b, f);
      float xv[6][4];
      for (int e = 0; e < 6; e++)
        for (int c = 0; c < 4; c++) xv[e][c] = x[b * 24 + e * 4 + c];
      bcn::bc7_forward(f, xv, smooth != 0, t);
    } else {
      bcn::Bc1Form f;
      bcn::bc1_parse(blocks + 2 * b, f);
      float xv[2][3];
      for (int e = 0; e < 2; e++)
        for (int c = 0; c < 3; c++) xv[e][c] = x[b * 6 + e * 3 + c];
      bcn::bc1_forward(f, xv, smooth != 0, t);
    }
    memcpy(texels + b * 64, t, sizeof(t));
  }
}

extern "C" void kcheck_backward(int fmt, const uint32_t* blocks, const float* g, size_t n, float* gx) {
  for (size_t b = 0; b < n; b++) {
    float gv[16][4];
    memcpy(gv, g + b * 64, sizeof(gv));
    if (fmt == 6) {
      bcn::Bc7Form f;
      bcn::bc7_parse(blocks + 4 * b, f);
      float o[6][4];
      bcn::bc7_backward(f, gv, o);
      memcpy(gx + b * 24, o, sizeof(o));
    } else {
      bcn::Bc1Form f;
      bcn::bc1_parse(blocks + 2 * b, f);
      float o[2][3];
      bcn::bc1_backward(f, gv, o);
      memcpy(gx + b * 6, o, sizeof(o));
    }
  }
}

// target: float[n][16][4] in 0..255 units (block-major).
extern "C" void kcheck_texel_backward(int fmt, const uint32_t* blocks, const float* g, size_t n, float* gs) {
  for (size_t b = 0; b < n; b++) {
    float gv[16][4], o[16][4];
    for (int k = 0; k < 16; k++)
      for (int c = 0; c < 4; c++) gv[k][c] = g[b * 64 + k * 4 + c];
    if (fmt == 6) {
      bcn::Bc7Form f;
      bcn::bc7_parse(blocks + 4 * b, f);
      bcn::bc7_texel_backward(f, gv, o);
    } else {
      bcn::Bc1Form f;
      bcn::bc1_parse(blocks + 2 * b, f);
      bcn::bc1_texel_backward(f, gv, o);
    }
    for (int k = 0; k < 16; k++)
      for (int c = 0; c < 4; c++) gs[b * 64 + k * 4 + c] = o[k][c];
  }
}

extern "C" void kcheck_bc7_refine(const uint32_t* blocks, const float* target, size_t n, uint32_t steps,
                                  float lr, uint32_t channels, uint32_t reassign, uint32_t* out, float* sse0,
                                  float* sse1) {
  bcn::RefineParams rp;
  rp.steps = steps;
  rp.lr = lr;
  rp.beta1 = 0.9f;
  rp.beta2 = 0.999f;
  rp.eps = 1e-8f;
  rp.channels = channels;
  rp.reassign = reassign & 0xFFFFu;
  rp.pbits = reassign >> 16;
  for (size_t b = 0; b < n; b++) sse1[b] = bcn::bc7_refine_block(blocks + 4 * b, target + b * 64, 1, rp, out + 4 * b, sse0 + b);
}

// quant4_thresholds against the float form of the 2-bit index candidate,
// floor(fl(fl(proj / div) * 4 + 0.5)) (fused, as the kernels compute it),
// which is monotone in proj: for every div in [1, 260100] and m = 2, 3 the
// float form must be below m at proj = t_m - 1 and reach m at t_m.  Returns
// the number of failures.
extern "C" long kcheck_quant4_mismatches() {
  long bad = 0;
  for (int div = 1; div <= 260100; div++) {
    int t[2];
    bcn::quant4_thresholds(div, t[0], t[1]);
    for (int i = 0; i < 2; i++) {
      const int m = 2 + i;
      volatile float below = (float)(t[i] - 1) / (float)div, at = (float)t[i] / (float)div;
      if (std::floor(std::fma((float)below, 4.0f, 0.5f)) >= (float)m) bad++;
      if (std::floor(std::fma((float)at, 4.0f, 0.5f)) < (float)m) bad++;
    }
  }
  int t2, t3;
  bcn::quant4_thresholds(0, t2, t3);
  if (0 >= t2 || 0 >= t3) bad++;
  return bad;
}

// quant8_consts / quant8_index against the float form of the 3-bit index
// candidate, clamp(floor(fl(fl(proj / div) * 8 + 0.5)), 1, 7); both are
// monotone in proj, so for every div in [1, 260100] and m = 2..7 both must be
// below m at proj = t - 1 and reach m at t, t = ceil((2m - 1) div / 16).
// Returns the number of failures.
extern "C" long kcheck_quant8_mismatches() {
  long bad = 0;
  auto flt = [](int proj, int div) {
    volatile float q = (float)proj / (float)div;
    const float y = std::floor(std::fma((float)q, 8.0f, 0.5f));
    return y < 1.0f ? 1 : (y > 7.0f ? 7 : (int)y);
  };
  for (int div = 1; div <= 260100; div++) {
    int c1;
    float rq;
    bcn::quant8_consts(div, c1, rq);
    for (int m = 2; m <= 7; m++) {
      const int t = ((2 * m - 1) * div + 15) >> 4;
      if (bcn::quant8_index(t - 1, c1, rq) >= m || bcn::quant8_index(t, c1, rq) < m) bad++;
      if (flt(t - 1, div) >= m || flt(t, div) < m) bad++;
    }
  }
  int c1;
  float rq;
  bcn::quant8_consts(0, c1, rq);
  if (bcn::quant8_index(0, c1, rq) != 1) bad++;
  return bad;
}

// chan_q4_consts / chan_q4_index (bc7_block.hpp) against channel_opt_quant's
// float form of the 2-bit index candidate (bc7.rs:292-295),
// clamp((int)(fl(fl(fl(texel - e0) / fl(fl(e1 - e0) + 0.001f)) * 4) + 0.5),
// 1, 3), for every endpoint pair e0, e1 and texel in 0..255.  Returns the
// number of mismatches.
extern "C" long kcheck_chan_q4_mismatches() {
  long bad = 0;
  for (int e0 = 0; e0 < 256; e0++)
    for (int e1 = 0; e1 < 256; e1++) {
      uint32_t s, k2, k3;
      bcn::chan_q4_consts(e0, e1, s, k2, k3);
      volatile float den = (float)(e1 - e0) + 0.001f;
      for (int v = 0; v < 256; v++) {
        volatile float proj = ((float)v - (float)e0) / den;
        volatile float y = (float)proj * 4.0f;
        y = (float)y + 0.5f;
        int q = (int)(float)y;
        q = q < 1 ? 1 : (q > 3 ? 3 : q);
        if (bcn::chan_q4_index((uint32_t)v * 0x10001u, s, k2, k3) != q) bad++;
      }
    }
  return bad;
}

// chan_q8_consts / chan_q8_index (bc7_block.hpp) against channel_opt_quant's
// float form of the 3-bit index candidate (bc7.rs:292-295) for every
// endpoint pair e0, e1 and texel in 0..255.  Returns the number of
// mismatches.
extern "C" long kcheck_chan_q8_mismatches() {
  long bad = 0;
  for (int e0 = 0; e0 < 256; e0++)
    for (int e1 = 0; e1 < 256; e1++) {
      float c, r;
      bcn::chan_q8_consts(e0, e1, c, r);
      volatile float den = (float)(e1 - e0) + 0.001f;
      for (int v = 0; v < 256; v++) {
        volatile float proj = ((float)v - (float)e0) / den;
        volatile float y = (float)proj * 8.0f;
        y = (float)y + 0.5f;
        int q = (int)(float)y;
        q = q < 1 ? 1 : (q > 7 ? 7 : q);
        if (bcn::chan_q8_index((float)v, c, r) != q) bad++;
      }
    }
  return bad;
}

// block_pos (decode_block.hpp) for n (bw, idx) pairs: rows[i], cols[i].
extern "C" void kcheck_block_pos(const uint32_t* bw, const uint32_t* idx, size_t n, uint32_t* rows,
                                 uint32_t* cols) {
  for (size_t i = 0; i < n; i++) bcn::block_pos(bcn::block_rows(bw[i]), idx[i], rows[i], cols[i]);
}
