// BC7 encode kernel: one lane per 4x4 block, row-major block order.
//
// Replaces compress_bc7 (src/shader/bc7.wgsl:1352-1376, one invocation per
// block in 8x8 workgroups) with a 1-D grid of 64-lane workgroups over the
// task's blocks.  Each lane loads its 16 texels as four 16-byte row loads
// (neighbouring lanes read neighbouring 16-byte chunks, so a wave reads four
// contiguous 1 KiB row segments), runs the block search in bc7_block.hpp
// and writes 16 bytes.  The 64 mode-1/3/7 partition keys of each lane live in
// LDS as 16-bit bounds (8 KiB per 64-lane workgroup), interleaved by lane.
#include "bc7_block.hpp"
#include "bcn_internal.hpp"
#include "texel_load.hpp"

namespace bcn {

constexpr int kBc7Wave = 64;


template <int CH, bool VEC>
__global__ __launch_bounds__(kBc7Wave)
__attribute__((amdgpu_waves_per_eu(BCN_BC7_WAVES, BCN_BC7_WAVES))) void bc7_encode_kernel(const uint8_t* __restrict__ src,
                                                              uint32_t pitch, uint32_t src_format,
                                                              uint32_t bw, uint32_t nblocks,
                                                              uint32_t* __restrict__ dst,
                                                              Bc7Params s) {
  __shared__ uint32_t keys[32 * kBc7Wave];
  __shared__ uint32_t sdata[18 * kBc7Wave];
  uint32_t idx = blockIdx.x * kBc7Wave + threadIdx.x;
  if (idx >= nblocks) return;
  uint32_t yy = idx / bw, xx = idx - yy * bw;
  uint32_t px[16];
  load_block_rgba8(src, pitch, xx, yy, src_format, VEC, px);
  uint32_t out[4];
  encode_bc7_block<CH>(px, s, keys + threadIdx.x, kBc7Wave, sdata + threadIdx.x, out);
  uint32_t* o = dst + (size_t)idx * 4;
  if (VEC) {
    *reinterpret_cast<uint4*>(o) = make_uint4(out[0], out[1], out[2], out[3]);
  } else {
#pragma unroll
    for (int i = 0; i < 4; i++) o[i] = out[i];
  }
}

hipError_t launch_bc7(const Task& t, hipStream_t stream) {
  Bc7Params s;
  static_assert(sizeof(Bc7Params) == 76, "BC7Settings is 19 x u32");
  __builtin_memcpy(&s, t.settings, sizeof(s));
  uint32_t bw = blocks_x(t), nblocks = bw * blocks_y(t);
  if (nblocks == 0) return hipSuccess;
  dim3 grid((nblocks + kBc7Wave - 1) / kBc7Wave), block(kBc7Wave);
  bool vec = (((uintptr_t)t.texture | (uintptr_t)t.pitch | (uintptr_t)t.blocks) & 15u) == 0;
  uint32_t* dst = reinterpret_cast<uint32_t*>(t.blocks);
  if (s.channels == 4) {
    if (vec) hipLaunchKernelGGL((bc7_encode_kernel<4, true>), grid, block, 0, stream, t.texture, t.pitch, (uint32_t)t.format, bw, nblocks, dst, s);
    else hipLaunchKernelGGL((bc7_encode_kernel<4, false>), grid, block, 0, stream, t.texture, t.pitch, (uint32_t)t.format, bw, nblocks, dst, s);
  } else {
    if (vec) hipLaunchKernelGGL((bc7_encode_kernel<3, true>), grid, block, 0, stream, t.texture, t.pitch, (uint32_t)t.format, bw, nblocks, dst, s);
    else hipLaunchKernelGGL((bc7_encode_kernel<3, false>), grid, block, 0, stream, t.texture, t.pitch, (uint32_t)t.format, bw, nblocks, dst, s);
  }
  return hipGetLastError();
}

}  // namespace bcn
