// BC1-BC5 encode kernels: one lane per 4x4 block, row-major block order.
//
// Replace compress_bc1..compress_bc5 (src/shader/bc1_to_5.wgsl:454-591).
// Every lane reads its block as four row loads (16 bytes for RGBA8; adjacent
// lanes read adjacent chunks; texel_load.hpp for the other layouts) and
// writes 8 or 16 bytes.  BC4/BC5 touch 64 bytes of RGBA8 input per block (16
// of R8, 32 of RG8) for 8/16 bytes out and are HBM-bound; BC1-BC3 add a short
// PCA fit per block.
#include "bc15_block.hpp"
#include "bcn_internal.hpp"
#include "texel_load.hpp"

namespace bcn {

constexpr int kBc15Group = 256;

template <int FMT, bool VEC>
__global__ __launch_bounds__(kBc15Group) void bc15_encode_kernel(const uint8_t* __restrict__ src,
                                                                 uint32_t pitch, uint32_t src_format,
                                                                 uint32_t bw, uint32_t nblocks,
                                                                 uint32_t* __restrict__ dst) {
  uint32_t idx = blockIdx.x * kBc15Group + threadIdx.x;
  if (idx >= nblocks) return;
  uint32_t yy = idx / bw, xx = idx - yy * bw;
  uint32_t px[16];
  load_block_rgba8(src, pitch, xx, yy, src_format, VEC, px);
  constexpr int WORDS = (FMT == BCN_BC1 || FMT == BCN_BC4) ? 2 : 4;
  uint32_t out[4];
  if (FMT == BCN_BC1) {
    bc1_core(px, out);
  } else if (FMT == BCN_BC2) {
    bc2_alpha(px, out);
    bc1_core(px, out + 2);
  } else if (FMT == BCN_BC3) {
    bc3_alpha_block(px, 3, out);
    bc1_core(px, out + 2);
  } else if (FMT == BCN_BC4) {
    bc3_alpha_block(px, 0, out);
  } else {
    bc3_alpha_block(px, 0, out);
    bc3_alpha_block(px, 1, out + 2);
  }
  uint32_t* o = dst + (size_t)idx * WORDS;
  if (VEC && WORDS == 4) {
    *reinterpret_cast<uint4*>(o) = make_uint4(out[0], out[1], out[2], out[3]);
  } else if (VEC) {
    *reinterpret_cast<uint2*>(o) = make_uint2(out[0], out[1]);
  } else {
#pragma unroll
    for (int i = 0; i < WORDS; i++) o[i] = out[i];
  }
}

template <int FMT>
static hipError_t launch_fmt(const Task& t, uint32_t bw, uint32_t nblocks, bool vec,
                             hipStream_t stream) {
  dim3 grid((nblocks + kBc15Group - 1) / kBc15Group), block(kBc15Group);
  uint32_t* dst = reinterpret_cast<uint32_t*>(t.blocks);
  if (vec)
    hipLaunchKernelGGL((bc15_encode_kernel<FMT, true>), grid, block, 0, stream, t.texture, t.pitch, (uint32_t)t.format, bw, nblocks, dst);
  else
    hipLaunchKernelGGL((bc15_encode_kernel<FMT, false>), grid, block, 0, stream, t.texture, t.pitch, (uint32_t)t.format, bw, nblocks, dst);
  return hipGetLastError();
}

hipError_t launch_bc15(const Task& t, hipStream_t stream) {
  uint32_t bw = blocks_x(t), nblocks = bw * blocks_y(t);
  if (nblocks == 0) return hipSuccess;
  uint32_t align = (t.variant == BCN_BC1 || t.variant == BCN_BC4) ? 7u : 15u;
  bool vec = (((uintptr_t)t.texture | (uintptr_t)t.pitch) & 15u) == 0 &&
             ((uintptr_t)t.blocks & align) == 0;
  switch (t.variant) {
    case BCN_BC1: return launch_fmt<BCN_BC1>(t, bw, nblocks, vec, stream);
    case BCN_BC2: return launch_fmt<BCN_BC2>(t, bw, nblocks, vec, stream);
    case BCN_BC3: return launch_fmt<BCN_BC3>(t, bw, nblocks, vec, stream);
    case BCN_BC4: return launch_fmt<BCN_BC4>(t, bw, nblocks, vec, stream);
    case BCN_BC5: return launch_fmt<BCN_BC5>(t, bw, nblocks, vec, stream);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace bcn
