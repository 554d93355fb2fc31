// BC6H (unsigned) encode kernel: one lane per 4x4 block, row-major order.
//
// Replaces compress_bc6h (src/shader/bc6h.wgsl:1401-1424).  Each lane reads
// its block of RGBA16F texels as four 32-byte rows (RGBA32F: 64-byte rows,
// rounded to halves; 8-bit unorm layouts: c/255 rounded to halves) and keeps
// the 48 scaled channel values in registers;
// the 32 partition keys live in LDS, one row per lane, interleaved by lane.
#include "bc6h_block.hpp"
#include "bcn_internal.hpp"
#include "texel_load.hpp"

namespace bcn {

constexpr int kBc6hGroup = 64;

// Waves per SIMD.  Unconstrained the block search took 328 registers (256
// VGPRs + 72 AGPRs, one wave); capped at 256 (two waves) it spills about 40
// dwords per lane and ran 20% faster than one wave.  The slow-mode search
// (six two-subset modes) runs 4-6% faster again at three waves (168 VGPRs,
// 140-230 spilled); the other presets are 1-4% slower there and keep two.
#ifndef BCN_BC6H_WAVES
#define BCN_BC6H_WAVES 2
#endif
#ifndef BCN_BC6H_SLOW_WAVES
#define BCN_BC6H_SLOW_WAVES 3
#endif

template <bool SGN, bool SLOW>
__global__ __launch_bounds__(kBc6hGroup) __attribute__((amdgpu_waves_per_eu(SLOW ? BCN_BC6H_SLOW_WAVES : BCN_BC6H_WAVES))) void bc6h_encode_kernel(const uint8_t* __restrict__ src,
                                                                 uint32_t pitch, uint32_t src_format,
                                                                 uint32_t bw, uint32_t nblocks,
                                                                 uint32_t* __restrict__ dst,
                                                                 Bc6hParams s) {
  __shared__ int keys[kBc6hRow * kBc6hGroup];
  uint32_t idx = blockIdx.x * kBc6hGroup + threadIdx.x;
  if (idx >= nblocks) return;
  uint32_t yy = idx / bw, xx = idx - yy * bw;
  uint16_t bits[48];
  load_block_rgb16f(src, pitch, xx, yy, src_format, bits);
  uint32_t out[4];
  encode_bc6h_block<SGN, SLOW>(bits, s, keys + threadIdx.x, kBc6hGroup, out);
  uint32_t* o = dst + (size_t)idx * 4;
#pragma unroll
  for (int i = 0; i < 4; i++) o[i] = out[i];
}

hipError_t launch_bc6h(const Task& t, hipStream_t stream) {
  if (source_texel_bytes(t.format) == 0) return hipErrorNotSupported;
  Bc6hParams s;
  static_assert(sizeof(Bc6hParams) == 20, "BC6HSettings is 5 x u32");
  __builtin_memcpy(&s, t.settings, sizeof(s));
  uint32_t bw = blocks_x(t), nblocks = bw * blocks_y(t);
  if (nblocks == 0) return hipSuccess;
  const uintptr_t align = t.format == BCN_SRC_RGBA32F ? 15u : (t.format == BCN_SRC_RGBA16F ? 7u : 3u);
  if ((((uintptr_t)t.texture | (uintptr_t)t.pitch) & align) != 0) return hipErrorInvalidValue;
  dim3 grid((nblocks + kBc6hGroup - 1) / kBc6hGroup), block(kBc6hGroup);
  uint32_t* dst = reinterpret_cast<uint32_t*>(t.blocks);
  // One instantiation per (signedness, slow_mode != 0).
  auto kernel = t.variant == BCN_BC6H_SF16
                    ? (s.slow_mode != 0 ? bc6h_encode_kernel<true, true> : bc6h_encode_kernel<true, false>)
                    : (s.slow_mode != 0 ? bc6h_encode_kernel<false, true> : bc6h_encode_kernel<false, false>);
  hipLaunchKernelGGL(kernel, grid, block, 0, stream, t.texture, t.pitch, (uint32_t)t.format, bw,
                     nblocks, dst, s);
  return hipGetLastError();
}

}  // namespace bcn
