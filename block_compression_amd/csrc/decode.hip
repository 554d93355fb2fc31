// Block decode kernels: one lane per 4x4 block.
//
// The reference decodes on the CPU only (src/decode.rs); these are the GPU
// counterparts of its image-level drivers decompress_blocks_as_rgba8 /
// _rgba16f / _rgba32f (src/decode.rs:336-468) and of decode_block_bc6h with
// the signed variant.  Each lane reads its 8/16-byte block and writes four
// row segments of 16 bytes (RGBA8), 32 bytes (RGBA16F) or 64 bytes (RGBA32F);
// adjacent lanes write adjacent segments of the same rows.  HBM-bound: 80
// bytes of traffic per BC7 block (16 in, 64 out).
#include <cmath>

#include "decode_block.hpp"
#include "bcn_internal.hpp"

namespace bcn {

constexpr int kDecGroup = 256;

// Output forms of the BC6H kernel.
enum Bc6hOut : int {
  kOutF16Rgba = 0,    // RGBA16F, alpha 1.0
  kOutF32Rgba = 1,    // RGBA32F, alpha 1.0
  kOutF16RefRgb = 2,  // the reference's RGB-at-stride-3 layout in an RGBA16F image
  kOutF32RefRgb = 3   // the same in an RGBA32F image
};

template <int FMT>
__global__ __launch_bounds__(kDecGroup) void decode_rgba8_kernel(const uint8_t* __restrict__ blocks,
                                                                 uint32_t bw, uint32_t nblocks,
                                                                 uint8_t* __restrict__ rgba,
                                                                 uint32_t pitch) {
  uint32_t idx = blockIdx.x * kDecGroup + threadIdx.x;
  if (idx >= nblocks) return;
  constexpr int BB = (FMT == BCN_BC1 || FMT == BCN_BC4) ? 8 : 16;
  uint32_t in[4] = {0, 0, 0, 0};
  const uint32_t* src = reinterpret_cast<const uint32_t*>(blocks + (size_t)idx * BB);
#pragma unroll
  for (int i = 0; i < BB / 4; i++) in[i] = src[i];
  uint32_t px[16];
  if (FMT == BCN_BC1) {
    decode_color(in[0], in[1], false, px);
  } else if (FMT == BCN_BC2 || FMT == BCN_BC3) {
    decode_color(in[2], in[3], true, px);
    if (FMT == BCN_BC2) {
      // decode_sharp_alpha_block (block.rs:200-217): 4-bit alpha x 17
#pragma unroll
      for (int k = 0; k < 16; k++) {
        uint32_t a = ((k < 8 ? in[0] : in[1]) >> (4 * (k & 7))) & 15u;
        px[k] = (px[k] & 0x00FFFFFFu) | ((a * 17u) << 24);
      }
    } else {
      uint8_t a[16];
      decode_channel(in[0], in[1], a);
#pragma unroll
      for (int k = 0; k < 16; k++) px[k] = (px[k] & 0x00FFFFFFu) | ((uint32_t)a[k] << 24);
    }
  } else if (FMT == BCN_BC4) {
    uint8_t r[16];
    decode_channel(in[0], in[1], r);
#pragma unroll
    for (int k = 0; k < 16; k++) px[k] = r[k];  // (R, 0, 0, 0) (decode.rs:97-118)
  } else if (FMT == BCN_BC5) {
    uint8_t r[16], g[16];
    decode_channel(in[0], in[1], r);
    decode_channel(in[2], in[3], g);
#pragma unroll
    for (int k = 0; k < 16; k++) px[k] = r[k] | ((uint32_t)g[k] << 8);  // (R, G, 0, 0)
  } else {
    decode_bc7(in, px);
  }
  uint32_t yy = idx / bw, xx = idx - yy * bw;
  uint8_t* o = rgba + (size_t)yy * 4 * pitch + (size_t)xx * 16;
#pragma unroll
  for (int y = 0; y < 4; y++)
    *reinterpret_cast<uint4*>(o + (size_t)y * pitch) =
        make_uint4(px[y * 4 + 0], px[y * 4 + 1], px[y * 4 + 2], px[y * 4 + 3]);
}

// BC6H -> RGBA8 (decode.rs:161-181): unsigned decode, linear_to_srgb per
// channel, alpha 0.  SGN: signed blocks, whose negative values (and -0) map
// to code 0 as linear_to_srgb clamps them.
template <bool SGN>
__global__ __launch_bounds__(kDecGroup) void decode_bc6h_srgb8_kernel(
    const uint8_t* __restrict__ blocks, uint32_t bw, uint32_t nblocks, uint8_t* __restrict__ rgba,
    uint32_t pitch, Srgb8Table tab) {
  __shared__ uint16_t lt[256];
  lt[threadIdx.x] = tab.t[threadIdx.x];
  __syncthreads();
  uint32_t idx = blockIdx.x * kDecGroup + threadIdx.x;
  if (idx >= nblocks) return;
  const uint32_t* src = reinterpret_cast<const uint32_t*>(blocks + (size_t)idx * 16);
  uint32_t in[4] = {src[0], src[1], src[2], src[3]};
  uint16_t h[48];
  decode_bc6h_t<SGN>(in, h);
  uint32_t yy = idx / bw, xx = idx - yy * bw;
  uint8_t* o = rgba + (size_t)yy * 4 * pitch + (size_t)xx * 16;
  // Negative values (and -0) code as 0, the code of +0: select the input.
  auto code = [&](uint16_t v) -> uint32_t {
    return srgb8_code((SGN && (v & 0x8000u)) ? (uint16_t)0 : v, lt);
  };
#pragma unroll
  for (int y = 0; y < 4; y++) {
    uint32_t px[4];
#pragma unroll
    for (int x = 0; x < 4; x++) {
      int k = y * 4 + x;
      px[x] = code(h[k * 3]) | (code(h[k * 3 + 1]) << 8) | (code(h[k * 3 + 2]) << 16);
    }
    *reinterpret_cast<uint4*>(o + (size_t)y * pitch) = make_uint4(px[0], px[1], px[2], px[3]);
  }
}

template <int OUT, bool SGN>
__global__ __launch_bounds__(kDecGroup) void decode_bc6h_kernel(const uint8_t* __restrict__ blocks,
                                                                uint32_t bw, uint32_t nblocks,
                                                                uint8_t* __restrict__ rgba,
                                                                uint32_t pitch) {
  uint32_t idx = blockIdx.x * kDecGroup + threadIdx.x;
  if (idx >= nblocks) return;
  const uint32_t* src = reinterpret_cast<const uint32_t*>(blocks + (size_t)idx * 16);
  uint32_t in[4] = {src[0], src[1], src[2], src[3]};
  uint16_t h[48];
  decode_bc6h_t<SGN>(in, h);
  constexpr bool F32 = OUT == kOutF32Rgba || OUT == kOutF32RefRgb;
  uint32_t yy = idx / bw, xx = idx - yy * bw;
  uint8_t* o = rgba + (size_t)yy * 4 * pitch + (size_t)xx * (F32 ? 64 : 32);
#pragma unroll
  for (int y = 0; y < 4; y++) {
    uint8_t* r = o + (size_t)y * pitch;
    const uint16_t* hr = h + y * 12;  // the row's 4 texels, RGB interleaved
    if (OUT == kOutF16Rgba) {
      uint4* row = reinterpret_cast<uint4*>(r);
#pragma unroll
      for (int x = 0; x < 4; x += 2)
        row[x / 2] = make_uint4((uint32_t)hr[x * 3] | ((uint32_t)hr[x * 3 + 1] << 16),
                                (uint32_t)hr[x * 3 + 2] | (0x3C00u << 16),
                                (uint32_t)hr[x * 3 + 3] | ((uint32_t)hr[x * 3 + 4] << 16),
                                (uint32_t)hr[x * 3 + 5] | (0x3C00u << 16));
    } else if (OUT == kOutF32Rgba) {
      uint4* row = reinterpret_cast<uint4*>(r);
#pragma unroll
      for (int x = 0; x < 4; x++)
        row[x] = make_uint4(__float_as_uint(half_to_float(hr[x * 3])),
                            __float_as_uint(half_to_float(hr[x * 3 + 1])),
                            __float_as_uint(half_to_float(hr[x * 3 + 2])), 0x3F800000u);
    } else if (OUT == kOutF16RefRgb) {
      // decode_block_bc6h writes texel j at j*3 (block.rs:753) from the
      // block's RGBA16F offset (decode.rs:271-279): 12 halves, the other 4
      // of the block's 16 untouched.
      uint2* row = reinterpret_cast<uint2*>(r);
#pragma unroll
      for (int j = 0; j < 3; j++)
        row[j] = make_uint2((uint32_t)hr[4 * j] | ((uint32_t)hr[4 * j + 1] << 16),
                            (uint32_t)hr[4 * j + 2] | ((uint32_t)hr[4 * j + 3] << 16));
    } else {
      // decode_block_bc6h_float, the same layout in f32 (block.rs:787-792,
      // decode.rs:317-325).
      uint4* row = reinterpret_cast<uint4*>(r);
#pragma unroll
      for (int j = 0; j < 3; j++)
        row[j] = make_uint4(__float_as_uint(half_to_float(hr[4 * j])),
                            __float_as_uint(half_to_float(hr[4 * j + 1])),
                            __float_as_uint(half_to_float(hr[4 * j + 2])),
                            __float_as_uint(half_to_float(hr[4 * j + 3])));
    }
  }
}

// ---- host side --------------------------------------------------------------

// binary16 -> binary32 of a non-negative finite half, exactly.
float half_to_float_host(uint16_t h) {
  int e = h >> 10, m = h & 1023;
  return e == 0 ? std::ldexp((float)m, -24) : std::ldexp((float)(1024 + m), e - 25);
}

// linear_to_srgb (src/decode.rs:151-159) in f32 with the C library's powf,
// the function Rust's f32::powf calls on Linux.
#pragma clang fp contract(off)
uint8_t linear_to_srgb8_host(float lin) {
  float v = lin <= 0.0031308f ? lin * 12.92f : 1.055f * powf(lin, 1.0f / 2.4f) - 0.055f;
  v = v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v);
  return (uint8_t)roundf(v * 255.0f);
}

// Thresholds over every unsigned BC6H output (bit patterns 0..0x7BFF; the
// unsigned finish_unquantize, block.rs:1391-1397, never exceeds 0x7BFF).
const Srgb8Table& srgb8_table() {
  static const Srgb8Table tab = [] {
    Srgb8Table t;
    t.t[0] = 0;
    for (int k = 1; k < 256; k++) t.t[k] = 0xFFFF;
    int k = 1;
    for (uint32_t h = 0; h <= 0x7BFF && k < 256; h++) {
      int code = linear_to_srgb8_host(half_to_float_host((uint16_t)h));
      while (k <= code) t.t[k++] = (uint16_t)h;
    }
    return t;
  }();
  return tab;
}

bool launch_ok() { return hipGetLastError() == hipSuccess; }

// decompress_blocks_as_rgba16f / _rgba32f (decode.rs:397-468): BC6H only
// (BC6H_SF16 blocks decode signed).
int decode_bc6h_image(bcn_variant variant, const void* blocks, size_t blocks_size, uint32_t width,
                      uint32_t height, int is_signed, void* rgba, uint32_t rgba_pitch,
                      uint32_t flags, bool f32, void* hip_stream) {
  if (variant < BCN_BC1 || variant > BCN_BC6H_SF16) return BCN_ERR_INVALID_ARGUMENT;
  if (variant != BCN_BC6H && variant != BCN_BC6H_SF16) return BCN_ERR_UNSUPPORTED;
  if (variant == BCN_BC6H_SF16) is_signed = 1;
  if (flags & ~(uint32_t)BCN_DECODE_REFERENCE_LAYOUT) return BCN_ERR_INVALID_ARGUMENT;
  if (width % 4 || height % 4) return BCN_ERR_NOT_MULTIPLE_OF_4;
  uint32_t bw = width / 4, nblocks = bw * (height / 4);
  if (blocks_size != (size_t)nblocks * 16) return BCN_ERR_BUFFER_TOO_SMALL;
  if ((uint64_t)rgba_pitch < (uint64_t)width * (f32 ? 16 : 8)) return BCN_ERR_INVALID_ARGUMENT;
  if (nblocks == 0) return BCN_OK;
  if (!blocks || !rgba) return BCN_ERR_INVALID_ARGUMENT;
  if ((((uintptr_t)rgba | rgba_pitch) & 15u) || ((uintptr_t)blocks & 3u)) return BCN_ERR_MISALIGNED;
  dim3 grid((nblocks + kDecGroup - 1) / kDecGroup), block(kDecGroup);
  hipStream_t s = static_cast<hipStream_t>(hip_stream);
  const uint8_t* b = static_cast<const uint8_t*>(blocks);
  uint8_t* o = static_cast<uint8_t*>(rgba);
  bool ref = flags & BCN_DECODE_REFERENCE_LAYOUT;
  int out = f32 ? (ref ? kOutF32RefRgb : kOutF32Rgba) : (ref ? kOutF16RefRgb : kOutF16Rgba);
#define BCN_DEC6(O, S) hipLaunchKernelGGL((decode_bc6h_kernel<O, S>), grid, block, 0, s, b, bw, nblocks, o, rgba_pitch)
  if (is_signed) {
    switch (out) {
      case kOutF16Rgba: BCN_DEC6(kOutF16Rgba, true); break;
      case kOutF32Rgba: BCN_DEC6(kOutF32Rgba, true); break;
      case kOutF16RefRgb: BCN_DEC6(kOutF16RefRgb, true); break;
      default: BCN_DEC6(kOutF32RefRgb, true); break;
    }
  } else {
    switch (out) {
      case kOutF16Rgba: BCN_DEC6(kOutF16Rgba, false); break;
      case kOutF32Rgba: BCN_DEC6(kOutF32Rgba, false); break;
      case kOutF16RefRgb: BCN_DEC6(kOutF16RefRgb, false); break;
      default: BCN_DEC6(kOutF32RefRgb, false); break;
    }
  }
#undef BCN_DEC6
  return launch_ok() ? BCN_OK : BCN_ERR_HIP;
}

}  // namespace bcn

using namespace bcn;

extern "C" {

int bcn_decode_rgba8(bcn_variant variant, const void* blocks, size_t blocks_size, uint32_t width,
                     uint32_t height, void* rgba, uint32_t rgba_pitch, void* hip_stream) {
  if (variant < BCN_BC1 || variant > BCN_BC6H_SF16) return BCN_ERR_INVALID_ARGUMENT;
  if (width % 4 || height % 4) return BCN_ERR_NOT_MULTIPLE_OF_4;
  size_t bb = (variant == BCN_BC1 || variant == BCN_BC4) ? 8 : 16;
  uint32_t bw = width / 4, nblocks = bw * (height / 4);
  // decode.rs:343-355: the block slice and the image must have the exact sizes.
  if (blocks_size != (size_t)nblocks * bb) return BCN_ERR_BUFFER_TOO_SMALL;
  if ((uint64_t)rgba_pitch < (uint64_t)width * 4) return BCN_ERR_INVALID_ARGUMENT;
  if (nblocks == 0) return BCN_OK;
  if (!blocks || !rgba) return BCN_ERR_INVALID_ARGUMENT;
  if ((((uintptr_t)rgba | rgba_pitch) & 15u) || ((uintptr_t)blocks & 7u)) return BCN_ERR_MISALIGNED;
  dim3 grid((nblocks + kDecGroup - 1) / kDecGroup), block(kDecGroup);
  hipStream_t s = static_cast<hipStream_t>(hip_stream);
  const uint8_t* b = static_cast<const uint8_t*>(blocks);
  uint8_t* o = static_cast<uint8_t*>(rgba);
  switch (variant) {
    case BCN_BC1: hipLaunchKernelGGL(decode_rgba8_kernel<BCN_BC1>, grid, block, 0, s, b, bw, nblocks, o, rgba_pitch); break;
    case BCN_BC2: hipLaunchKernelGGL(decode_rgba8_kernel<BCN_BC2>, grid, block, 0, s, b, bw, nblocks, o, rgba_pitch); break;
    case BCN_BC3: hipLaunchKernelGGL(decode_rgba8_kernel<BCN_BC3>, grid, block, 0, s, b, bw, nblocks, o, rgba_pitch); break;
    case BCN_BC4: hipLaunchKernelGGL(decode_rgba8_kernel<BCN_BC4>, grid, block, 0, s, b, bw, nblocks, o, rgba_pitch); break;
    case BCN_BC5: hipLaunchKernelGGL(decode_rgba8_kernel<BCN_BC5>, grid, block, 0, s, b, bw, nblocks, o, rgba_pitch); break;
    case BCN_BC6H: hipLaunchKernelGGL(decode_bc6h_srgb8_kernel<false>, grid, block, 0, s, b, bw, nblocks, o, rgba_pitch, srgb8_table()); break;
    case BCN_BC6H_SF16: hipLaunchKernelGGL(decode_bc6h_srgb8_kernel<true>, grid, block, 0, s, b, bw, nblocks, o, rgba_pitch, srgb8_table()); break;
    default: hipLaunchKernelGGL(decode_rgba8_kernel<BCN_BC7>, grid, block, 0, s, b, bw, nblocks, o, rgba_pitch); break;
  }
  return launch_ok() ? BCN_OK : BCN_ERR_HIP;
}

int bcn_decode_rgba16f(bcn_variant variant, const void* blocks, size_t blocks_size, uint32_t width,
                       uint32_t height, void* rgba, uint32_t rgba_pitch, uint32_t flags,
                       void* hip_stream) {
  return decode_bc6h_image(variant, blocks, blocks_size, width, height, 0, rgba, rgba_pitch, flags,
                           false, hip_stream);
}

int bcn_decode_rgba32f(bcn_variant variant, const void* blocks, size_t blocks_size, uint32_t width,
                       uint32_t height, void* rgba, uint32_t rgba_pitch, uint32_t flags,
                       void* hip_stream) {
  return decode_bc6h_image(variant, blocks, blocks_size, width, height, 0, rgba, rgba_pitch, flags,
                           true, hip_stream);
}

int bcn_decode_bc6h_rgba16f(const void* blocks, size_t blocks_size, uint32_t width,
                            uint32_t height, int is_signed, void* rgba, uint32_t rgba_pitch,
                            void* hip_stream) {
  return decode_bc6h_image(BCN_BC6H, blocks, blocks_size, width, height, is_signed, rgba,
                           rgba_pitch, 0, false, hip_stream);
}

void bcn_bc6h_srgb8_thresholds(uint16_t* out) {
  const Srgb8Table& t = srgb8_table();
  for (int k = 0; k < 256; k++) out[k] = t.t[k];
}

}  // extern "C"
