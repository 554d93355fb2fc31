// TEST INFRASTRUCTURE ONLY -- the CPU oracle (see bc_oracle.hpp).
//
// Image-level drivers (src/encode.rs:83-251, src/decode.rs:200-468) and a C
// ABI for the Python test harness and the bench's CPU-baseline leg.  The
// per-block loop is row-major over blocks, (xx, yy) = (idx % bw, idx / bw)
// (encode.rs:279), optionally spread over OpenMP threads the way the crate's
// rayon feature spreads it (encode.rs:274-286).
#include <omp.h>

#include "bc_oracle.hpp"

using namespace bco;

namespace bco {
extern const uint8_t BC7_PARTITION_SETS[2][64][16];
}

namespace {

// Variant tags shared with include/bcn_mi355x.h (lib.rs:74-103 order).
enum { V_BC1 = 0, V_BC2 = 1, V_BC3 = 2, V_BC4 = 3, V_BC5 = 4, V_BC6H = 5, V_BC7 = 6 };

size_t block_bytes(int v) { return (v == V_BC1 || v == V_BC4) ? 8 : 16; }

// decode.rs:151-159, in f32 (Rust's f32::powf is the C library's powf).
uint8_t linear_to_srgb(float linear) {
  float v = linear <= 0.0031308f ? linear * 12.92f : 1.055f * std::pow(linear, 1.0f / 2.4f) - 0.055f;
  if (v < 0.0f) v = 0.0f;
  if (v > 1.0f) v = 1.0f;
  return (uint8_t)std::round(v * 255.0f);
}

}  // namespace

extern "C" {

int oracle_block_bytes(int variant) { return (int)block_bytes(variant); }

// encode.rs:83-151.  Returns 0, or a negative code on a failed precondition
// (the crate panics there): -1 width/height not a multiple of 4, -2 output
// too small, -3 bad variant.
int oracle_compress_rgba8(int variant, const void* settings, const uint8_t* rgba, uint8_t* out,
                          size_t out_len, uint32_t width, uint32_t height, uint32_t stride,
                          int threads) {
  if (height % 4 != 0 || width % 4 != 0) return -1;
  size_t bw = (width + 3) / 4, bh = (height + 3) / 4;
  if (variant < 0 || variant > 6) return -3;
  size_t bb = block_bytes(variant);
  if (out_len < bw * bh * bb) return -2;
  long nblocks = (long)(bw * bh);
  BC7Settings s7;
  BC6HSettings s6;
  if (variant == V_BC7) std::memcpy(&s7, settings, sizeof(s7));
  if (variant == V_BC6H) std::memcpy(&s6, settings, sizeof(s6));
  if (threads <= 0) threads = 1;
#pragma omp parallel for schedule(static) num_threads(threads)
  for (long idx = 0; idx < nblocks; idx++) {
    size_t yy = (size_t)idx / bw, xx = (size_t)idx % bw;
    const uint8_t* src = rgba + yy * 4 * stride + xx * 16;
    uint8_t* dst = out + (size_t)idx * bb;
    switch (variant) {
      case V_BC1: bc1_encode_block(src, stride, dst); break;
      case V_BC2: bc2_encode_block(src, stride, dst); break;
      case V_BC3: bc3_encode_block(src, stride, dst); break;
      case V_BC4: bc4_encode_block(src, stride, dst); break;
      case V_BC5: bc5_encode_block(src, stride, dst); break;
      case V_BC6H: bc6h_encode_block_8(src, stride, s6, dst); break;
      case V_BC7: bc7_encode_block(src, stride, s7, dst); break;
    }
  }
  return 0;
}

// encode.rs:211-251.  `stride` in f16 elements.
int oracle_compress_rgba16(const void* settings, const uint16_t* rgba16, uint8_t* out,
                           size_t out_len, uint32_t width, uint32_t height, uint32_t stride,
                           int threads) {
  if (height % 4 != 0 || width % 4 != 0) return -1;
  size_t bw = (width + 3) / 4, bh = (height + 3) / 4;
  if (out_len < bw * bh * 16) return -2;
  BC6HSettings s6;
  std::memcpy(&s6, settings, sizeof(s6));
  long nblocks = (long)(bw * bh);
  if (threads <= 0) threads = 1;
#pragma omp parallel for schedule(static) num_threads(threads)
  for (long idx = 0; idx < nblocks; idx++) {
    size_t yy = (size_t)idx / bw, xx = (size_t)idx % bw;
    bc6h_encode_block_16(rgba16 + yy * 4 * stride + xx * 16, stride, s6, out + (size_t)idx * 16);
  }
  return 0;
}

// Signed (SF16) BC6H encode of an RGBA16F image; not in the reference (its
// encoder is unsigned only), see Bc6h::sgn in bc6h.cpp.  Also returns each
// block's search error in `errs` when non-null.
int oracle_compress_rgba16_sf16(const void* settings, const uint16_t* rgba16, uint8_t* out,
                                size_t out_len, uint32_t width, uint32_t height, uint32_t stride,
                                float* errs, int threads) {
  if (height % 4 != 0 || width % 4 != 0) return -1;
  size_t bw = width / 4, bh = height / 4;
  if (out_len < bw * bh * 16) return -2;
  BC6HSettings s6;
  std::memcpy(&s6, settings, sizeof(s6));
  long nblocks = (long)(bw * bh);
  if (threads <= 0) threads = 1;
#pragma omp parallel for schedule(static) num_threads(threads)
  for (long idx = 0; idx < nblocks; idx++) {
    size_t yy = (size_t)idx / bw, xx = (size_t)idx % bw;
    float e = bc6h_encode_block_16(rgba16 + yy * 4 * stride + xx * 16, stride, s6,
                                   out + (size_t)idx * 16, true);
    if (errs) errs[idx] = e;
  }
  return 0;
}

// A listed subset of blocks (by block index) of an RGBA8 image: the sampled
// CPU check of large workloads.  Writes block i's bytes at out + i * bb.
int oracle_compress_rgba8_sample(int variant, const void* settings, const uint8_t* rgba,
                                 uint32_t width, uint32_t stride, const int64_t* block_ids,
                                 size_t count, uint8_t* out, int threads) {
  size_t bw = (width + 3) / 4;
  size_t bb = block_bytes(variant);
  BC7Settings s7;
  BC6HSettings s6;
  if (variant == V_BC7) std::memcpy(&s7, settings, sizeof(s7));
  if (variant == V_BC6H) std::memcpy(&s6, settings, sizeof(s6));
  if (threads <= 0) threads = 1;
#pragma omp parallel for schedule(static) num_threads(threads)
  for (long i = 0; i < (long)count; i++) {
    size_t idx = (size_t)block_ids[i];
    size_t yy = idx / bw, xx = idx % bw;
    const uint8_t* src = rgba + yy * 4 * stride + xx * 16;
    uint8_t* dst = out + (size_t)i * bb;
    switch (variant) {
      case V_BC1: bc1_encode_block(src, stride, dst); break;
      case V_BC2: bc2_encode_block(src, stride, dst); break;
      case V_BC3: bc3_encode_block(src, stride, dst); break;
      case V_BC4: bc4_encode_block(src, stride, dst); break;
      case V_BC5: bc5_encode_block(src, stride, dst); break;
      case V_BC6H: bc6h_encode_block_8(src, stride, s6, dst); break;
      case V_BC7: bc7_encode_block(src, stride, s7, dst); break;
    }
  }
  return 0;
}

// Per-block encoder error (best_err), for the invariant tests.
int oracle_bc7_block_errors(const void* settings, const uint8_t* rgba, uint32_t width,
                            uint32_t height, uint32_t stride, uint8_t* out, float* errs) {
  BC7Settings s7;
  std::memcpy(&s7, settings, sizeof(s7));
  size_t bw = width / 4, bh = height / 4;
  for (size_t idx = 0; idx < bw * bh; idx++) {
    size_t yy = idx / bw, xx = idx % bw;
    errs[idx] = bc7_encode_block(rgba + yy * 4 * stride + xx * 16, stride, s7, out + idx * 16);
  }
  return 0;
}

int oracle_bc6h_block_errors(const void* settings, const uint16_t* rgba16, uint32_t width,
                             uint32_t height, uint32_t stride, uint8_t* out, float* errs) {
  BC6HSettings s6;
  std::memcpy(&s6, settings, sizeof(s6));
  size_t bw = width / 4, bh = height / 4;
  for (size_t idx = 0; idx < bw * bh; idx++) {
    size_t yy = idx / bw, xx = idx % bw;
    errs[idx] = bc6h_encode_block_16(rgba16 + yy * 4 * stride + xx * 16, stride, s6, out + idx * 16);
  }
  return 0;
}

// Block decoders with the reference's signatures (block.rs:50-119, 262, 774, 801).
void oracle_decode_block_bc1(const uint8_t* in, uint8_t* out, size_t pitch) { decode_block_bc1(in, out, pitch); }
void oracle_decode_block_bc2(const uint8_t* in, uint8_t* out, size_t pitch) { decode_block_bc2(in, out, pitch); }
void oracle_decode_block_bc3(const uint8_t* in, uint8_t* out, size_t pitch) { decode_block_bc3(in, out, pitch); }
void oracle_decode_block_bc4(const uint8_t* in, uint8_t* out, size_t pitch) { decode_block_bc4(in, out, pitch); }
void oracle_decode_block_bc5(const uint8_t* in, uint8_t* out, size_t pitch) { decode_block_bc5(in, out, pitch); }
void oracle_decode_block_bc7(const uint8_t* in, uint8_t* out, size_t pitch) { decode_block_bc7(in, out, pitch); }
void oracle_decode_block_bc6h(const uint8_t* in, uint16_t* out, size_t pitch, int is_signed) {
  decode_block_bc6h(in, out, pitch, is_signed != 0);
}
void oracle_decode_block_bc6h_comp(const uint8_t* in, int32_t* comp, int is_signed) {
  uint16_t tmp[48];
  decode_block_bc6h(in, tmp, 12, is_signed != 0, comp);
}
void oracle_decode_block_bc6h_float(const uint8_t* in, float* out, size_t pitch, int is_signed) {
  decode_block_bc6h_float(in, out, pitch, is_signed != 0);
}

// decode.rs:200-232 + 336-387: RGBA8 image, tightly packed.  BC4 -> (R,0,0,0),
// BC5 -> (R,G,0,0) (decode.rs:97-148).  BC6H -> sRGB8 with A = 0
// (decode.rs:151-181; its powf is the platform libm's).
int oracle_decompress_rgba8(int variant, uint32_t width, uint32_t height, const uint8_t* blocks,
                            size_t blocks_len, uint8_t* rgba, size_t rgba_len) {
  size_t bw = (width + 3) / 4, bh = (height + 3) / 4;
  size_t bb = block_bytes(variant);
  if (blocks_len != bw * bh * bb) return -2;
  if (rgba_len != (size_t)width * height * 4) return -4;
  size_t pitch = (size_t)width * 4;
  for (size_t by = 0; by < bh; by++)
    for (size_t bx = 0; bx < bw; bx++) {
      const uint8_t* blk = blocks + (by * bw + bx) * bb;
      uint8_t* o = rgba + by * 4 * pitch + bx * 16;
      switch (variant) {
        case V_BC1: decode_block_bc1(blk, o, pitch); break;
        case V_BC2: decode_block_bc2(blk, o, pitch); break;
        case V_BC3: decode_block_bc3(blk, o, pitch); break;
        case V_BC7: decode_block_bc7(blk, o, pitch); break;
        case V_BC4: {
          uint8_t buf[16];
          decode_block_bc4(blk, buf, 4);
          for (int y = 0; y < 4; y++)
            for (int x = 0; x < 4; x++) {
              uint8_t* p = o + y * pitch + x * 4;
              p[0] = buf[y * 4 + x];
              p[1] = p[2] = p[3] = 0;
            }
          break;
        }
        case V_BC5: {
          uint8_t buf[32];
          decode_block_bc5(blk, buf, 8);
          for (int y = 0; y < 4; y++)
            for (int x = 0; x < 4; x++) {
              uint8_t* p = o + y * pitch + x * 4;
              p[0] = buf[y * 8 + x * 2];
              p[1] = buf[y * 8 + x * 2 + 1];
              p[2] = p[3] = 0;
            }
          break;
        }
        case V_BC6H: {
          float buf[48];
          decode_block_bc6h_float(blk, buf, 12, false);
          for (int y = 0; y < 4; y++)
            for (int x = 0; x < 4; x++) {
              uint8_t* p = o + y * pitch + x * 4;
              for (int c = 0; c < 3; c++) p[c] = linear_to_srgb(buf[y * 12 + x * 3 + c]);
              p[3] = 0;
            }
          break;
        }
        default: return -3;
      }
    }
  return 0;
}

// decompress_rgba16f / decompress_rgba32f + decompress_blocks_as_rgba16f /
// _rgba32f (decode.rs:247-334, :397-468).  BC6H only (-3 otherwise, where
// the reference panics).  The output image has width*4 elements per row and
// each block goes to its RGBA offset, but decode_block_bc6h(_float) writes
// texel j at j*3 (block.rs:753, :787-792): per block row, elements 0..11
// receive RGB and 12..15 keep what the caller's buffer held.
int oracle_decompress_rgba16f(int variant, uint32_t width, uint32_t height, const uint8_t* blocks,
                              size_t blocks_len, uint16_t* rgba, size_t rgba_len) {
  if (variant != V_BC6H) return -3;
  size_t bw = (width + 3) / 4, bh = (height + 3) / 4;
  if (blocks_len != bw * bh * 16) return -2;
  if (rgba_len != (size_t)width * height * 4) return -4;
  size_t pitch = (size_t)width * 4;
  for (size_t by = 0; by < bh; by++)
    for (size_t bx = 0; bx < bw; bx++)
      decode_block_bc6h(blocks + (by * bw + bx) * 16, rgba + by * 4 * pitch + bx * 16, pitch, false);
  return 0;
}

int oracle_decompress_rgba32f(int variant, uint32_t width, uint32_t height, const uint8_t* blocks,
                              size_t blocks_len, float* rgba, size_t rgba_len) {
  if (variant != V_BC6H) return -3;
  size_t bw = (width + 3) / 4, bh = (height + 3) / 4;
  if (blocks_len != bw * bh * 16) return -2;
  if (rgba_len != (size_t)width * height * 4) return -4;
  size_t pitch = (size_t)width * 4;
  for (size_t by = 0; by < bh; by++)
    for (size_t bx = 0; bx < bw; bx++)
      decode_block_bc6h_float(blocks + (by * bw + bx) * 16, rgba + by * 4 * pitch + bx * 16, pitch,
                              false);
  return 0;
}

uint8_t oracle_linear_to_srgb(float linear) { return linear_to_srgb(linear); }

// Primitives exposed for the brute-force and closed-form pin tests.
void oracle_partial_sort_list(int32_t* list, size_t length, uint32_t partial_count) {
  partial_sort_list(list, length, partial_count);
}
void oracle_get_tables(uint32_t* pattern, uint32_t* mask0, uint32_t* mask1, uint32_t* mask2,
                       uint32_t* skips) {
  for (int i = 0; i < 128; i++) {
    pattern[i] = get_pattern(i);
    mask0[i] = get_pattern_mask(i, 0);
    mask1[i] = get_pattern_mask(i, 1);
    mask2[i] = get_pattern_mask(i, 2);
    uint32_t s[3];
    get_skips(i, s);
    skips[i * 3 + 0] = s[0];
    skips[i * 3 + 1] = s[1];
    skips[i * 3 + 2] = s[2];
  }
}
void oracle_get_partition_sets(uint8_t* out /* 2*64*16 */) {
  std::memcpy(out, bco::BC7_PARTITION_SETS, 2 * 64 * 16);
}
uint16_t oracle_f32_to_f16(float x) { return f32_to_f16_bits(x); }
float oracle_f16_to_f32(uint16_t h) { return f16_bits_to_f32(h); }

// Shared encoder primitives of common.rs, one call each, in plain floats.
// `block` is the 64-float channel-major block (R[16], G[16], B[16], A[16]).
void oracle_compute_axis(const float* covar10, uint32_t iterations, uint32_t channels, float* axis4) {
  F c[10], a[4] = {0.0f, 0.0f, 0.0f, 0.0f};
  for (int i = 0; i < 10; i++) c[i] = F(covar10[i]);
  compute_axis(a, c, iterations, channels);
  for (int i = 0; i < 4; i++) axis4[i] = raw(a[i]);
}
float oracle_get_pca_bound(const float* covar10, uint32_t channels) {
  F c[10];
  for (int i = 0; i < 10; i++) c[i] = F(covar10[i]);
  return raw(get_pca_bound(c, channels));
}
void oracle_block_pca_axis(const float* block64, uint32_t mask, uint32_t channels, float* axis4,
                           float* dc4) {
  F b[64], a[4] = {0.0f, 0.0f, 0.0f, 0.0f}, d[4] = {0.0f, 0.0f, 0.0f, 0.0f};
  for (int i = 0; i < 64; i++) b[i] = F(block64[i]);
  block_pca_axis(a, d, b, mask, channels);
  for (int i = 0; i < 4; i++) {
    axis4[i] = raw(a[i]);
    dc4[i] = raw(d[i]);
  }
}
void oracle_block_segment_core(const float* block64, uint32_t mask, uint32_t channels, float* ep8) {
  F b[64], e[8];
  for (int i = 0; i < 64; i++) b[i] = F(block64[i]);
  for (int i = 0; i < 8; i++) e[i] = 0.0f;
  block_segment_core(e, b, mask, channels);
  for (int i = 0; i < 8; i++) ep8[i] = raw(e[i]);
}
void oracle_opt_endpoints(const float* block64, uint32_t bits, const uint32_t* qblock2,
                          uint32_t mask, uint32_t channels, float* ep8) {
  F b[64], e[8];
  for (int i = 0; i < 64; i++) b[i] = F(block64[i]);
  for (int i = 0; i < 8; i++) e[i] = 0.0f;
  opt_endpoints(e, b, bits, qblock2, mask, channels);
  for (int i = 0; i < 8; i++) ep8[i] = raw(e[i]);
}
// block_quant over one subset: ep8 holds endpoints A (0..3) and B (4..7) of
// subset 0; pattern 0 sends every texel to subset 0.
float oracle_block_quant(const float* block64, uint32_t bits, const float* ep8, uint32_t channels,
                         uint32_t* qblock2) {
  F b[64], e[24];
  for (int i = 0; i < 64; i++) b[i] = F(block64[i]);
  for (int i = 0; i < 24; i++) e[i] = i < 8 ? F(ep8[i]) : F(0.0f);
  return raw(block_quant(qblock2, b, bits, e, 0, channels));
}
void oracle_compute_stats_masked(const float* block64, uint32_t mask, uint32_t channels, float* stats15) {
  F b[64], s[15];
  for (int i = 0; i < 64; i++) b[i] = F(block64[i]);
  for (int i = 0; i < 15; i++) s[i] = 0.0f;
  compute_stats_masked(s, b, mask, channels);
  for (int i = 0; i < 15; i++) stats15[i] = raw(s[i]);
}
float oracle_block_pca_bound_split(const float* block64, uint32_t mask, uint32_t channels) {
  F b[64], s[15];
  for (int i = 0; i < 64; i++) b[i] = F(block64[i]);
  for (int i = 0; i < 15; i++) s[i] = 0.0f;
  compute_stats_masked(s, b, 0xFFFFFFFFu, channels);
  return raw(block_pca_bound_split(b, mask, s, channels));
}

#ifdef BC_OPCOUNT
// Counted op total of encoding the listed blocks (single thread).
uint64_t oracle_count_ops_rgba8(int variant, const void* settings, const uint8_t* rgba,
                                uint32_t width, uint32_t stride, const int64_t* block_ids,
                                size_t count) {
  g_ops = 0;
  size_t bb = block_bytes(variant);
  uint8_t* tmp = new uint8_t[bb];
  size_t bw = (width + 3) / 4;
  BC7Settings s7;
  BC6HSettings s6;
  if (variant == V_BC7) std::memcpy(&s7, settings, sizeof(s7));
  if (variant == V_BC6H) std::memcpy(&s6, settings, sizeof(s6));
  for (size_t i = 0; i < count; i++) {
    size_t idx = (size_t)block_ids[i];
    size_t yy = idx / bw, xx = idx % bw;
    const uint8_t* src = rgba + yy * 4 * stride + xx * 16;
    switch (variant) {
      case V_BC1: bc1_encode_block(src, stride, tmp); break;
      case V_BC2: bc2_encode_block(src, stride, tmp); break;
      case V_BC3: bc3_encode_block(src, stride, tmp); break;
      case V_BC4: bc4_encode_block(src, stride, tmp); break;
      case V_BC5: bc5_encode_block(src, stride, tmp); break;
      case V_BC6H: bc6h_encode_block_8(src, stride, s6, tmp); break;
      case V_BC7: bc7_encode_block(src, stride, s7, tmp); break;
    }
  }
  delete[] tmp;
  return g_ops;
}

uint64_t oracle_count_ops_rgba16(const void* settings, const uint16_t* rgba16, uint32_t width,
                                 uint32_t stride, const int64_t* block_ids, size_t count) {
  g_ops = 0;
  uint8_t tmp[16];
  size_t bw = (width + 3) / 4;
  BC6HSettings s6;
  std::memcpy(&s6, settings, sizeof(s6));
  for (size_t i = 0; i < count; i++) {
    size_t idx = (size_t)block_ids[i];
    size_t yy = idx / bw, xx = idx % bw;
    bc6h_encode_block_16(rgba16 + yy * 4 * stride + xx * 16, stride, s6, tmp);
  }
  return g_ops;
}
#endif

}  // extern "C"
