// Block decode kernels: one lane per 4x4 block.
//
// The reference decodes on the CPU only (src/decode.rs); these are the GPU
// counterparts of decompress_blocks_as_rgba8 (src/decode.rs:336-387) and of
// decode_block_bc6h into RGBA16F.  Each lane reads its 8/16-byte block and
// writes four 16-byte (RGBA8) or 32-byte (RGBA16F) row segments; adjacent
// lanes write adjacent segments of the same rows.  HBM-bound: 80 bytes of
// traffic per BC7 block (16 in, 64 out).
#include "decode_block.hpp"
#include "bcn_internal.hpp"

namespace bcn {

constexpr int kDecGroup = 256;

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

__global__ __launch_bounds__(kDecGroup) void decode_bc6h_kernel(const uint8_t* __restrict__ blocks,
                                                                uint32_t bw, uint32_t nblocks,
                                                                int is_signed,
                                                                uint8_t* __restrict__ rgba,
                                                                uint32_t pitch) {
  uint32_t idx = blockIdx.x * kDecGroup + threadIdx.x;
  if (idx >= nblocks) return;
  const uint32_t* src = reinterpret_cast<const uint32_t*>(blocks + (size_t)idx * 16);
  uint32_t in[4] = {src[0], src[1], src[2], src[3]};
  uint16_t h[48];
  decode_bc6h(in, is_signed != 0, h);
  uint32_t yy = idx / bw, xx = idx - yy * bw;
  uint8_t* o = rgba + (size_t)yy * 4 * pitch + (size_t)xx * 32;
#pragma unroll
  for (int y = 0; y < 4; y++) {
    uint4* row = reinterpret_cast<uint4*>(o + (size_t)y * pitch);
#pragma unroll
    for (int x = 0; x < 4; x += 2) {
      int k0 = y * 4 + x, k1 = k0 + 1;
      // RGBA16F texels, alpha 1.0 (0x3C00).
      row[x / 2] = make_uint4((uint32_t)h[k0 * 3] | ((uint32_t)h[k0 * 3 + 1] << 16),
                              (uint32_t)h[k0 * 3 + 2] | (0x3C00u << 16),
                              (uint32_t)h[k1 * 3] | ((uint32_t)h[k1 * 3 + 1] << 16),
                              (uint32_t)h[k1 * 3 + 2] | (0x3C00u << 16));
    }
  }
}

}  // namespace bcn

using namespace bcn;

extern "C" {

int bcn_decode_rgba8(bcn_variant variant, const void* blocks, size_t blocks_size, uint32_t width,
                     uint32_t height, void* rgba, uint32_t rgba_pitch, void* hip_stream) {
  if (variant < BCN_BC1 || variant > BCN_BC7) return BCN_ERR_INVALID_ARGUMENT;
  if (variant == BCN_BC6H) return BCN_ERR_UNSUPPORTED;
  if (width % 4 || height % 4) return BCN_ERR_NOT_MULTIPLE_OF_4;
  size_t bb = (variant == BCN_BC1 || variant == BCN_BC4) ? 8 : 16;
  uint32_t bw = width / 4, nblocks = bw * (height / 4);
  // decode.rs:343-355: the block slice and the image must have the exact sizes.
  if (blocks_size != (size_t)nblocks * bb) return BCN_ERR_BUFFER_TOO_SMALL;
  if (rgba_pitch < width * 4) return BCN_ERR_INVALID_ARGUMENT;
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
    default: hipLaunchKernelGGL(decode_rgba8_kernel<BCN_BC7>, grid, block, 0, s, b, bw, nblocks, o, rgba_pitch); break;
  }
  return hipGetLastError() == hipSuccess ? BCN_OK : BCN_ERR_HIP;
}

int bcn_decode_bc6h_rgba16f(const void* blocks, size_t blocks_size, uint32_t width,
                            uint32_t height, int is_signed, void* rgba, uint32_t rgba_pitch,
                            void* hip_stream) {
  if (width % 4 || height % 4) return BCN_ERR_NOT_MULTIPLE_OF_4;
  uint32_t bw = width / 4, nblocks = bw * (height / 4);
  if (blocks_size != (size_t)nblocks * 16) return BCN_ERR_BUFFER_TOO_SMALL;
  if (rgba_pitch < width * 8) return BCN_ERR_INVALID_ARGUMENT;
  if (nblocks == 0) return BCN_OK;
  if (!blocks || !rgba) return BCN_ERR_INVALID_ARGUMENT;
  if ((((uintptr_t)rgba | rgba_pitch) & 15u) || ((uintptr_t)blocks & 3u)) return BCN_ERR_MISALIGNED;
  dim3 grid((nblocks + kDecGroup - 1) / kDecGroup), block(kDecGroup);
  hipLaunchKernelGGL(decode_bc6h_kernel, grid, block, 0, static_cast<hipStream_t>(hip_stream),
                     static_cast<const uint8_t*>(blocks), bw, nblocks, is_signed,
                     static_cast<uint8_t*>(rgba), rgba_pitch);
  return hipGetLastError() == hipSuccess ? BCN_OK : BCN_ERR_HIP;
}

}  // extern "C"
