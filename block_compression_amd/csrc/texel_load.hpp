// Texel loaders of the encode kernels: one 4x4 block from a source texture.
//
// The reference's shaders read texels with textureLoad from a texture view
// of any unorm or float format (src/block_compressor.rs:300-313;
// bc7.wgsl:64-78, bc1_to_5.wgsl:38-66, bc6h.wgsl:64-85); a channel the view
// does not have reads as 0 (G, B) or 1 (A).  These loaders give the kernels
// the same texels from the source layouts of bcn_source_format, with the
// format a kernel argument (uniform across the launch, so the branch is
// scalar), straight from HBM with no conversion pass:
//
//   RGBA8  4 bytes per texel          BGRA8  4 bytes, B and R swapped
//   R8     1 byte  -> (r, 0, 0, 255)  RG8    2 bytes -> (r, g, 0, 255)
//
// Every layout reads rows at `pitch` bytes from the block's first texel;
// adjacent lanes read adjacent row segments.  Output: 16 RGBA8 words,
// R in bits 0..7 and A in bits 24..31, texel k = y*4 + x.
#pragma once

#include "bcn_common.hpp"

#if defined(__HIPCC__)
#include "../../include/bcn_mi355x.h"

namespace bcn {

// Bytes per texel of a source format (0 for an unknown format).
BCN_HD uint32_t source_texel_bytes(uint32_t fmt) {
  switch (fmt) {
    case BCN_SRC_RGBA8:
    case BCN_SRC_BGRA8: return 4;
    case BCN_SRC_R8: return 1;
    case BCN_SRC_RG8: return 2;
    case BCN_SRC_RGBA16F: return 8;
    case BCN_SRC_RGBA32F: return 16;
    default: return 0;
  }
}

// `vec`: the row segments are 16-byte aligned (RGBA8 / BGRA8 then load each
// row with one 16-byte load).
__device__ __forceinline__ void load_block_rgba8(const uint8_t* src, uint32_t pitch, uint32_t xx,
                                                 uint32_t yy, uint32_t fmt, bool vec,
                                                 uint32_t px[16]) {
  const uint8_t* base = src + (size_t)yy * 4 * pitch;
  if (fmt == BCN_SRC_R8) {
    const uint8_t* p = base + (size_t)xx * 4;
#pragma unroll
    for (int y = 0; y < 4; y++) {
      uint32_t w = *reinterpret_cast<const uint32_t*>(p + (size_t)y * pitch);
#pragma unroll
      for (int x = 0; x < 4; x++) px[y * 4 + x] = ((w >> (8 * x)) & 0xFFu) | 0xFF000000u;
    }
  } else if (fmt == BCN_SRC_RG8) {
    const uint8_t* p = base + (size_t)xx * 8;
#pragma unroll
    for (int y = 0; y < 4; y++) {
      const uint32_t* r = reinterpret_cast<const uint32_t*>(p + (size_t)y * pitch);
      uint32_t w0 = r[0], w1 = r[1];
      px[y * 4 + 0] = (w0 & 0xFFFFu) | 0xFF000000u;
      px[y * 4 + 1] = (w0 >> 16) | 0xFF000000u;
      px[y * 4 + 2] = (w1 & 0xFFFFu) | 0xFF000000u;
      px[y * 4 + 3] = (w1 >> 16) | 0xFF000000u;
    }
  } else {
    const uint8_t* p = base + (size_t)xx * 16;
#pragma unroll
    for (int y = 0; y < 4; y++) {
      if (vec) {
        uint4 v = *reinterpret_cast<const uint4*>(p + (size_t)y * pitch);
        px[y * 4 + 0] = v.x;
        px[y * 4 + 1] = v.y;
        px[y * 4 + 2] = v.z;
        px[y * 4 + 3] = v.w;
      } else {
        const uint32_t* r = reinterpret_cast<const uint32_t*>(p + (size_t)y * pitch);
#pragma unroll
        for (int x = 0; x < 4; x++) px[y * 4 + x] = r[x];
      }
    }
    if (fmt == BCN_SRC_BGRA8) {
#pragma unroll
      for (int k = 0; k < 16; k++) {
        uint32_t v = px[k];
        px[k] = (v & 0xFF00FF00u) | ((v >> 16) & 0xFFu) | ((v & 0xFFu) << 16);
      }
    }
  }
}

// f32 -> binary16 bit pattern as bc6h.wgsl's f32_to_f16_bits takes it
// (bc6h.wgsl:59-70): every zero reads as +0; other values round to the
// nearest half, ties to even (quantizeToF16), with half subnormals kept.
__device__ __forceinline__ uint16_t f32_to_f16_bits(float f) {
  if (f == 0.0f) return 0;
  return __builtin_bit_cast(uint16_t, (_Float16)f);
}

// A half texel as textureLoad + f32_to_f16_bits hands it to the BC6H search:
// -0 reads as +0 (bc6h.wgsl:61-63), every other pattern as itself.
__device__ __forceinline__ uint16_t half_texel(uint32_t h) {
  return h == 0x8000u ? (uint16_t)0 : (uint16_t)h;
}

// An 8-bit unorm channel as the BC6H search sees it through a unorm view:
// textureLoad's c / 255 in f32 (correctly rounded), then f32_to_f16_bits
// (round to nearest even; every such value is 0 or a normal half).
__device__ __forceinline__ uint16_t unorm8_texel(uint32_t c) {
  return f32_to_f16_bits((float)c / 255.0f);
}

// BC6H source block: the R, G, B half bit patterns, bits[c*16 + k], from an
// RGBA16F, RGBA32F or 8-bit unorm (RGBA8, BGRA8, R8, RG8) layout.
__device__ __forceinline__ void load_block_rgb16f(const uint8_t* src, uint32_t pitch, uint32_t xx,
                                                  uint32_t yy, uint32_t fmt, uint16_t bits[48]) {
  const uint8_t* base = src + (size_t)yy * 4 * pitch;
  if (fmt != BCN_SRC_RGBA32F && fmt != BCN_SRC_RGBA16F) {
    uint32_t px[16];
    load_block_rgba8(src, pitch, xx, yy, fmt, false, px);
#pragma unroll
    for (int k = 0; k < 16; k++) {
      bits[k] = unorm8_texel(px[k] & 0xFFu);
      bits[16 + k] = unorm8_texel((px[k] >> 8) & 0xFFu);
      bits[32 + k] = unorm8_texel((px[k] >> 16) & 0xFFu);
    }
  } else if (fmt == BCN_SRC_RGBA32F) {
    const uint8_t* p = base + (size_t)xx * 64;
#pragma unroll
    for (int y = 0; y < 4; y++) {
      const uint4* row = reinterpret_cast<const uint4*>(p + (size_t)y * pitch);
#pragma unroll
      for (int x = 0; x < 4; x++) {
        uint4 t = row[x];
        int k = y * 4 + x;
        bits[k] = f32_to_f16_bits(__uint_as_float(t.x));
        bits[16 + k] = f32_to_f16_bits(__uint_as_float(t.y));
        bits[32 + k] = f32_to_f16_bits(__uint_as_float(t.z));
      }
    }
  } else {
    const uint8_t* p = base + (size_t)xx * 32;
#pragma unroll
    for (int y = 0; y < 4; y++) {
      const uint2* row = reinterpret_cast<const uint2*>(p + (size_t)y * pitch);
#pragma unroll
      for (int x = 0; x < 4; x++) {
        uint2 t = row[x];  // R | G << 16, B | A << 16
        int k = y * 4 + x;
        bits[k] = half_texel(t.x & 0xFFFFu);
        bits[16 + k] = half_texel(t.x >> 16);
        bits[32 + k] = half_texel(t.y & 0xFFFFu);
      }
    }
  }
}

}  // namespace bcn
#endif
