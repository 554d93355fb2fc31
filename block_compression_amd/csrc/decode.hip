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
#include <mutex>

#include "decode_block.hpp"
#include "bcn_internal.hpp"

namespace bcn {

constexpr int kDecGroup = 256;

// The sRGB8 bucket table (srgb8_word) on each device, uploaded once per
// device before its first BC6H -> RGBA8 decode.
__device__ __align__(16) uint16_t g_srgb8_buckets[kSrgb8Buckets];

// Output forms of the BC6H kernel.
enum Bc6hOut : int {
  kOutF16Rgba = 0,    // RGBA16F, alpha 1.0
  kOutF32Rgba = 1,    // RGBA32F, alpha 1.0
  kOutF16RefRgb = 2,  // the reference's RGB-at-stride-3 layout in an RGBA16F image
  kOutF32RefRgb = 3   // the same in an RGBA32F image
};

// Workgroup size of decode_rgba8_kernel: BC7 runs one wave per group, so each
// wave's LDS records are freed as soon as that wave ends.
constexpr int rgba8_group(int fmt) { return fmt == BCN_BC7 ? 64 : kDecGroup; }

template <int FMT>
__global__ __launch_bounds__(rgba8_group(FMT)) void decode_rgba8_kernel(const uint8_t* __restrict__ blocks,
                                                                        BlockRows rows, uint32_t nblocks,
                                                                        uint8_t* __restrict__ rgba,
                                                                        uint32_t pitch) {
  constexpr int BB = (FMT == BCN_BC1 || FMT == BCN_BC4) ? 8 : 16;
  constexpr int G = rgba8_group(FMT);
  const uint32_t idx = blockIdx.x * G + threadIdx.x;
  if (idx >= nblocks) return;
  const uint32_t* src = reinterpret_cast<const uint32_t*>(blocks + (size_t)idx * BB);
  uint32_t in[4];
#pragma unroll
  for (int i = 0; i < 4; i++) in[i] = i < BB / 4 ? src[i] : 0u;
  uint32_t yy, xx;
  block_pos(rows, idx, yy, xx);
  uint8_t* o = rgba + (size_t)yy * 4 * pitch + (size_t)xx * 16;
  // BC7's three subset records per lane, interleaved by lane (3 KiB).
  static_assert(rgba8_group(BCN_BC7) == (int)kBc7SubRows, "the BC7 record addresses assume 64 lanes");
  __shared__ Bc7SubsetData subs[FMT == BCN_BC7 ? 3 * kBc7SubRows : 1];
  if constexpr (FMT == BCN_BC7) {
    // Each row is stored as soon as it is decoded.
    decode_bc7_rows(in, subs, threadIdx.x, [&](int y, uint32_t a, uint32_t b, uint32_t c, uint32_t e) {
      *reinterpret_cast<uint4*>(o + (size_t)y * pitch) = make_uint4(a, b, c, e);
    });
    return;
  }
  uint32_t px[16];
  decode_rgba8_block<FMT>(in, px, subs, threadIdx.x);
#pragma unroll
  for (int y = 0; y < 4; y++)
    *reinterpret_cast<uint4*>(o + (size_t)y * pitch) =
        make_uint4(px[y * 4 + 0], px[y * 4 + 1], px[y * 4 + 2], px[y * 4 + 3]);
}

// The BC6H decoders' LDS: each lane's 18-word scratch row, interleaved by
// lane.  (The field table is read from constant memory, a row per lane.)
struct Bc6hLds {
  uint32_t scratch[18 * kDecGroup];
};

// One lane's BC6H block to RGBA8 sRGB codes, row by row: each row is stored
// as soon as it is decoded.  ONE: the wave holds only one-subset blocks.
template <bool SGN, bool ONE>
__device__ __forceinline__ void bc6h_srgb8_rows(const uint32_t in[4], int m, Bc6hScratch sc,
                                                const uint16_t* fields, const uint16_t* lb,
                                                uint8_t* o, uint32_t pitch) {
  Bc6hIdx st = decode_bc6h_header<SGN, ONE>(in, m, sc, fields);
  // Negative values (and -0) code as 0, the code of +0.
  auto word = [&](uint16_t v) -> uint32_t {
    const uint32_t x = SGN ? (uint32_t)max((int)(int16_t)v, 0) : (uint32_t)v;
    return srgb8_word(x, lb);
  };
#pragma unroll
  for (int y = 0; y < 4; y++) {
    uint32_t px[4];
#pragma unroll
    for (int x = 0; x < 4; x++) {
      uint16_t h[3];
      decode_bc6h_texel<SGN, ONE>(st, y * 4 + x, sc, h);
      // The high bytes of the three words: R, G, then B; alpha 0.
      const uint32_t rg = perm_b32(word(h[1]), word(h[0]), 0x0C0C0501u);
      px[x] = perm_b32(word(h[2]), rg, 0x0C050100u);
    }
    *reinterpret_cast<uint4*>(o + (size_t)y * pitch) = make_uint4(px[0], px[1], px[2], px[3]);
  }
}

// BC6H -> RGBA8 (decode.rs:161-181): unsigned decode, linear_to_srgb per
// channel, alpha 0.  SGN: signed blocks, whose negative values (and -0) map
// to code 0 as linear_to_srgb clamps them.  Each workgroup copies the sRGB8
// bucket table (7,936 bytes) to LDS; a channel is then one LDS read.
template <bool SGN>
__global__ __launch_bounds__(kDecGroup) void decode_bc6h_srgb8_kernel(
    const uint8_t* __restrict__ blocks, BlockRows rows, uint32_t nblocks, uint8_t* __restrict__ rgba,
    uint32_t pitch) {
  __shared__ __align__(16) uint16_t lb[kSrgb8Buckets];
  __shared__ Bc6hLds lds;
  for (uint32_t i = threadIdx.x; i < kSrgb8Buckets / 8; i += kDecGroup)
    reinterpret_cast<uint4*>(lb)[i] = reinterpret_cast<const uint4*>(g_srgb8_buckets)[i];
  __syncthreads();
  const uint32_t idx = blockIdx.x * kDecGroup + threadIdx.x;
  if (idx >= nblocks) return;
  const uint32_t* src = reinterpret_cast<const uint32_t*>(blocks + (size_t)idx * 16);
  const uint32_t in[4] = {src[0], src[1], src[2], src[3]};
  const Bc6hScratch sc{lds.scratch + threadIdx.x, kDecGroup};
  uint32_t yy, xx;
  block_pos(rows, idx, yy, xx);
  uint8_t* o = rgba + (size_t)yy * 4 * pitch + (size_t)xx * 16;
  const int m = bc6h_mode(in);
  if (wave_all(bc6h_one_subset(m)))
    bc6h_srgb8_rows<SGN, true>(in, m, sc, &kBc6hDec.f[0][0], lb, o, pitch);
  else
    bc6h_srgb8_rows<SGN, false>(in, m, sc, &kBc6hDec.f[0][0], lb, o, pitch);
}

// One block row per lane (N x 16 bytes, lane order) stored through the
// wave's LDS slice, when the wave's 64 block rows are consecutive in memory
// from `dst`: store instruction x then writes bytes x * 1024 .. x * 1024 +
// 1023 of the wave's run (lane L the 16 bytes at x * 1024 + 16 L), instead of
// 16 bytes every N * 16 bytes.  Element e = 64 x + L of the run is chunk
// e % N of lane e / N's row, which the slice holds at index e.
template <int N>
__device__ __forceinline__ void store_row_run(const uint4 (&v)[N], uint4* slice, uint8_t* dst,
                                              uint32_t lane) {
#pragma unroll
  for (int c = 0; c < N; c++) slice[lane * N + c] = v[c];
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
  for (int x = 0; x < N; x++)
    reinterpret_cast<uint4*>(dst)[x * 64 + lane] = slice[x * 64 + lane];
  // The slice is rewritten by the next row only after every lane has read it.
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// One lane's BC6H block to the float outputs, row by row: each row is
// stored as soon as it is decoded, through the wave's row-run slice when
// `run` (the wave's 64 block rows consecutive from run_dst).  ONE: the wave
// holds only one-subset blocks.
template <int OUT, bool SGN, bool ONE, int N>
__device__ __forceinline__ void bc6h_float_rows(const uint32_t in[4], int m, Bc6hScratch sc,
                                                const uint16_t* fields, uint8_t* o, uint32_t pitch,
                                                bool run, uint4* slice, uint8_t* run_dst,
                                                uint32_t lane) {
  constexpr bool kRun = OUT == kOutF16Rgba || OUT == kOutF32Rgba;
  Bc6hIdx st = decode_bc6h_header<SGN, ONE>(in, m, sc, fields);
#pragma unroll
  for (int y = 0; y < 4; y++) {
    uint8_t* r = o + (size_t)y * pitch;
    uint16_t hr[12];  // the row's 4 texels, RGB interleaved
#pragma unroll
    for (int x = 0; x < 4; x++) decode_bc6h_texel<SGN, ONE>(st, y * 4 + x, sc, hr + 3 * x);
    if (kRun) {
      uint4 v[N];
#pragma unroll
      for (int x = 0; x < 4; x++) {
        if (OUT == kOutF16Rgba) {
          const uint2 t = make_uint2((uint32_t)hr[x * 3] | ((uint32_t)hr[x * 3 + 1] << 16),
                                     (uint32_t)hr[x * 3 + 2] | (0x3C00u << 16));
          if (x & 1) v[x / 2].z = t.x, v[x / 2].w = t.y;
          else v[x / 2].x = t.x, v[x / 2].y = t.y;
        } else {
          v[x % N] = make_uint4(__float_as_uint(half_to_float(hr[x * 3])),
                                __float_as_uint(half_to_float(hr[x * 3 + 1])),
                                __float_as_uint(half_to_float(hr[x * 3 + 2])), 0x3F800000u);
        }
      }
      if (run) {
        store_row_run<N>(v, slice, run_dst + (size_t)y * pitch, lane);
      } else {
        uint4* row = reinterpret_cast<uint4*>(r);
#pragma unroll
        for (int c = 0; c < N; c++) row[c] = v[c];
      }
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

template <int OUT, bool SGN>
__global__ __launch_bounds__(kDecGroup) void decode_bc6h_kernel(const uint8_t* __restrict__ blocks,
                                                                BlockRows rows, uint32_t nblocks,
                                                                uint8_t* __restrict__ rgba,
                                                                uint32_t pitch) {
  __shared__ Bc6hLds lds;
  // RGBA outputs: each wave's row-run slice, 64 rows of 32 (16F) or 64 (32F)
  // bytes.
  constexpr bool F32 = OUT == kOutF32Rgba || OUT == kOutF32RefRgb;
  constexpr bool kRun = OUT == kOutF16Rgba || OUT == kOutF32Rgba;
  constexpr int N = F32 ? 4 : 2;  // 16-byte chunks per block row
  __shared__ uint4 runs[kRun ? (kDecGroup / 64) * 64 * N : 1];
  const uint32_t idx = blockIdx.x * kDecGroup + threadIdx.x;
  if (idx >= nblocks) return;
  const uint32_t* src = reinterpret_cast<const uint32_t*>(blocks + (size_t)idx * 16);
  const uint32_t in[4] = {src[0], src[1], src[2], src[3]};
  const Bc6hScratch sc{lds.scratch + threadIdx.x, kDecGroup};
  uint32_t yy, xx;
  block_pos(rows, idx, yy, xx);
  uint8_t* o = rgba + (size_t)yy * 4 * pitch + (size_t)xx * (F32 ? 64 : 32);
  // Whether the wave's 64 blocks all exist and sit in one block row: then its
  // rows are consecutive in memory and go out through the row-run slice.
  const uint32_t lane = threadIdx.x & 63u, first = idx - lane;
  uint32_t fy, fx;
  block_pos(rows, first, fy, fx);
  const bool run = kRun && first + 63 < nblocks && fx + 63 < rows.bw;
  uint4* slice = runs + (threadIdx.x >> 6) * 64 * N;
  uint8_t* run_dst = rgba + (size_t)fy * 4 * pitch + (size_t)fx * (F32 ? 64 : 32);
  const int m = bc6h_mode(in);
  if (wave_all(bc6h_one_subset(m)))
    bc6h_float_rows<OUT, SGN, true, N>(in, m, sc, &kBc6hDec.f[0][0], o, pitch, run, slice, run_dst, lane);
  else
    bc6h_float_rows<OUT, SGN, false, N>(in, m, sc, &kBc6hDec.f[0][0], o, pitch, run, slice, run_dst, lane);
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

// sRGB8 thresholds over every unsigned BC6H output (bit patterns 0..0x7BFF;
// the unsigned finish_unquantize, block.rs:1391-1397, never exceeds 0x7BFF):
// t[k] is the smallest pattern whose code is >= k (t[0] = 0).
struct Srgb8Table {
  uint16_t t[256];
};
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

// The bucket entries of srgb8_word from the thresholds; ok is false if some
// bucket of 8 patterns held two rises (the entries could not express it).
struct Srgb8Buckets {
  uint16_t e[kSrgb8Buckets];
  bool ok;
};
const Srgb8Buckets& srgb8_buckets() {
  static const Srgb8Buckets tab = [] {
    const Srgb8Table& t = srgb8_table();
    auto code = [&](uint32_t h) {
      int c = 0;
      for (int k = 1; k < 256; k++) c += h >= t.t[k] ? 1 : 0;
      return c;
    };
    Srgb8Buckets b;
    b.ok = true;
    for (uint32_t i = 0; i < kSrgb8Buckets; i++) {
      const int base = code(8 * i);
      uint32_t off = 8;
      for (uint32_t j = 1; j < 8; j++) {
        const int c = code(8 * i + j);
        if (c != base && off == 8) off = j;
        if (c != base + (off <= j ? 1 : 0)) b.ok = false;
      }
      b.e[i] = (uint16_t)(256 * base + 32 * (8 - off));
    }
    return b;
  }();
  return tab;
}

// Uploads the bucket table to the current device, once per device (a
// synchronous copy at the first BC6H -> RGBA8 decode there).
int srgb8_buckets_upload() {
  constexpr int kMaxDevices = 64;
  static std::mutex mu;
  static bool done[kMaxDevices] = {};
  const Srgb8Buckets& b = srgb8_buckets();
  if (!b.ok) return BCN_ERR_UNSUPPORTED;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return BCN_ERR_HIP;
  std::lock_guard<std::mutex> lock(mu);
  if (!done[dev]) {
    if (hipMemcpyToSymbol(HIP_SYMBOL(g_srgb8_buckets), b.e, sizeof(b.e), 0, hipMemcpyHostToDevice) != hipSuccess)
      return BCN_ERR_HIP;
    done[dev] = true;
  }
  return BCN_OK;
}

// Block grid of an image: bw x rows blocks, both below 2^24 (block_pos) and
// at most 2^31 blocks in all.
bool block_grid_ok(uint32_t width, uint32_t height) {
  const uint64_t bw = width / 4, bh = height / 4;
  return bw < (1u << 24) && bh < (1u << 24) && bw * bh <= (1ull << 31);
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
  if (!block_grid_ok(width, height)) return BCN_ERR_INVALID_ARGUMENT;
  uint32_t bw = width / 4, nblocks = bw * (height / 4);
  if (blocks_size != (size_t)nblocks * 16) return BCN_ERR_BUFFER_TOO_SMALL;
  if ((uint64_t)rgba_pitch < (uint64_t)width * (f32 ? 16 : 8)) return BCN_ERR_INVALID_ARGUMENT;
  if (nblocks == 0) return BCN_OK;
  if (!blocks || !rgba) return BCN_ERR_INVALID_ARGUMENT;
  if ((((uintptr_t)rgba | rgba_pitch) & 15u) || ((uintptr_t)blocks & 3u)) return BCN_ERR_MISALIGNED;
  dim3 grid((nblocks + kDecGroup - 1) / kDecGroup), block(kDecGroup);
  const BlockRows rows = block_rows(bw);
  hipStream_t s = static_cast<hipStream_t>(hip_stream);
  const uint8_t* b = static_cast<const uint8_t*>(blocks);
  uint8_t* o = static_cast<uint8_t*>(rgba);
  bool ref = flags & BCN_DECODE_REFERENCE_LAYOUT;
  int out = f32 ? (ref ? kOutF32RefRgb : kOutF32Rgba) : (ref ? kOutF16RefRgb : kOutF16Rgba);
#define BCN_DEC6(O, S) hipLaunchKernelGGL((decode_bc6h_kernel<O, S>), grid, block, 0, s, b, rows, nblocks, o, rgba_pitch)
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
  if (!block_grid_ok(width, height)) return BCN_ERR_INVALID_ARGUMENT;
  size_t bb = (variant == BCN_BC1 || variant == BCN_BC4) ? 8 : 16;
  uint32_t bw = width / 4, nblocks = bw * (height / 4);
  // decode.rs:343-355: the block slice and the image must have the exact sizes.
  if (blocks_size != (size_t)nblocks * bb) return BCN_ERR_BUFFER_TOO_SMALL;
  if ((uint64_t)rgba_pitch < (uint64_t)width * 4) return BCN_ERR_INVALID_ARGUMENT;
  if (nblocks == 0) return BCN_OK;
  if (!blocks || !rgba) return BCN_ERR_INVALID_ARGUMENT;
  if ((((uintptr_t)rgba | rgba_pitch) & 15u) || ((uintptr_t)blocks & 7u)) return BCN_ERR_MISALIGNED;
  dim3 grid((nblocks + kDecGroup - 1) / kDecGroup), block(kDecGroup);
  const BlockRows rows = block_rows(bw);
  hipStream_t s = static_cast<hipStream_t>(hip_stream);
  const uint8_t* b = static_cast<const uint8_t*>(blocks);
  uint8_t* o = static_cast<uint8_t*>(rgba);
  if (variant == BCN_BC6H || variant == BCN_BC6H_SF16) {
    const int rc = srgb8_buckets_upload();
    if (rc != BCN_OK) return rc;
  }
  switch (variant) {
    case BCN_BC1: hipLaunchKernelGGL(decode_rgba8_kernel<BCN_BC1>, grid, block, 0, s, b, rows, nblocks, o, rgba_pitch); break;
    case BCN_BC2: hipLaunchKernelGGL(decode_rgba8_kernel<BCN_BC2>, grid, block, 0, s, b, rows, nblocks, o, rgba_pitch); break;
    case BCN_BC3: hipLaunchKernelGGL(decode_rgba8_kernel<BCN_BC3>, grid, block, 0, s, b, rows, nblocks, o, rgba_pitch); break;
    case BCN_BC4: hipLaunchKernelGGL(decode_rgba8_kernel<BCN_BC4>, grid, block, 0, s, b, rows, nblocks, o, rgba_pitch); break;
    case BCN_BC5: hipLaunchKernelGGL(decode_rgba8_kernel<BCN_BC5>, grid, block, 0, s, b, rows, nblocks, o, rgba_pitch); break;
    case BCN_BC6H: hipLaunchKernelGGL(decode_bc6h_srgb8_kernel<false>, grid, block, 0, s, b, rows, nblocks, o, rgba_pitch); break;
    case BCN_BC6H_SF16: hipLaunchKernelGGL(decode_bc6h_srgb8_kernel<true>, grid, block, 0, s, b, rows, nblocks, o, rgba_pitch); break;
    default: {
      constexpr int G = rgba8_group(BCN_BC7);
      hipLaunchKernelGGL(decode_rgba8_kernel<BCN_BC7>, dim3((nblocks + G - 1) / G), dim3(G), 0, s, b, rows,
                         nblocks, o, rgba_pitch);
      break;
    }
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

int bcn_bc6h_srgb8_buckets(uint16_t* out) {
  const Srgb8Buckets& b = srgb8_buckets();
  if (!out) return BCN_ERR_INVALID_ARGUMENT;
  for (uint32_t i = 0; i < kSrgb8Buckets; i++) out[i] = b.e[i];
  return b.ok ? BCN_OK : BCN_ERR_UNSUPPORTED;
}

}  // extern "C"
