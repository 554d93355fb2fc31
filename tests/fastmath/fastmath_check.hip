// Exhaustive GPU check of the kernels' square root and reciprocal
// (sqrt_cr, rcp_cr in block_compression_amd/csrc/bcn_common.hpp) against the
// compiler's correctly rounded sqrtf and 1.0f / x, built with the same flags
// as the library.  Every float bit pattern in [lo, hi) is tested; mismatches
// are counted per exponent field (bits 23..30).  Mode 0 tests the functions as
// the library uses them; mode 1 their unguarded fast paths, to show where the
// guards are needed.  The second check sweeps the BC6H -> RGBA8 decode's sRGB8
// code (srgb8_word in decode_block.hpp, from an LDS copy of the bucket table
// as the kernel reads it) over every non-negative finite half; the third,
// rsqrt_cr against 1.0f / sqrtf(x).
// Test infrastructure only: the library does not link this.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../block_compression_amd/csrc/bcn_common.hpp"
#include "../../block_compression_amd/csrc/decode_block.hpp"

__device__ float fast_sqrt_unguarded(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return bcn::sqrt_markstein(x);
#else
  return x;
#endif
}

__device__ float fast_rcp_unguarded(float x) {
  const float y = __builtin_amdgcn_rcpf(x);
  return __builtin_fmaf(__builtin_fmaf(-x, y, 1.0f), y, y);
}

__global__ void fastmath_kernel(uint32_t lo, uint32_t hi, int mode, unsigned long long* bad_sqrt,
                                unsigned long long* bad_rcp) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t u = (uint64_t)lo + blockIdx.x * blockDim.x + threadIdx.x; u < hi; u += stride) {
    const uint32_t bits = (uint32_t)u;
    const float x = __uint_as_float(bits);
    const uint32_t ex = (bits >> 23) & 0xFFu;
    const float s = mode ? fast_sqrt_unguarded(x) : bcn::sqrt_cr(x);
    const float r = mode ? fast_rcp_unguarded(x) : bcn::rcp_cr(x);
    if (__float_as_uint(s) != __float_as_uint(sqrtf(x))) atomicAdd(&bad_sqrt[ex], 1ull);
    // NaN results compare by class (1/x is never NaN for a non-NaN x here).
    if (__float_as_uint(r) != __float_as_uint(1.0f / x)) atomicAdd(&bad_rcp[ex], 1ull);
  }
}

// Counts for every exponent value (256 each) into host arrays.  Returns 0 or a
// hipError_t.
extern "C" int fastmath_check(uint32_t lo, uint32_t hi, int mode, unsigned long long* bad_sqrt,
                              unsigned long long* bad_rcp) {
  unsigned long long* d = nullptr;
  hipError_t e = hipMalloc(&d, 512 * sizeof(unsigned long long));
  if (e != hipSuccess) return (int)e;
  e = hipMemset(d, 0, 512 * sizeof(unsigned long long));
  if (e == hipSuccess) {
    hipLaunchKernelGGL(fastmath_kernel, dim3(8192), dim3(256), 0, 0, lo, hi, mode, d, d + 256);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e == hipSuccess) e = hipMemcpy(bad_sqrt, d, 256 * sizeof(unsigned long long), hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(bad_rcp, d + 256, 256 * sizeof(unsigned long long), hipMemcpyDeviceToHost);
  (void)hipFree(d);
  return (int)e;
}

// ---- sRGB8 codes of the BC6H -> RGBA8 decode ---------------------------------
// For every half bit pattern h in [0, n): code[h] = high byte of
// srgb8_word(h, table) (block_compression_amd/csrc/decode_block.hpp), with
// `table` the library's bucket entries (bcn_bc6h_srgb8_buckets) in LDS.

__global__ void srgb8_kernel(uint32_t n, const uint16_t* buckets, uint8_t* code) {
  __shared__ uint16_t lb[bcn::kSrgb8Buckets];
  for (uint32_t i = threadIdx.x; i < bcn::kSrgb8Buckets; i += 256) lb[i] = buckets[i];
  __syncthreads();
  const uint32_t h = blockIdx.x * 256 + threadIdx.x;
  if (h >= n) return;
  code[h] = (uint8_t)(bcn::srgb8_word(h, lb) >> 8);
}

extern "C" int srgb8_check(const uint16_t* buckets, uint32_t n, uint8_t* code) {
  if (n > 0x7C00u) return -1;
  uint8_t* d = nullptr;
  hipError_t e = hipMalloc(&d, 0x7C00 + 2 * bcn::kSrgb8Buckets);
  if (e != hipSuccess) return (int)e;
  uint16_t* t = reinterpret_cast<uint16_t*>(d + 0x7C00);
  e = hipMemcpy(t, buckets, 2 * bcn::kSrgb8Buckets, hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(srgb8_kernel, dim3((n + 255) / 256), dim3(256), 0, 0, n, t, d);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e == hipSuccess) e = hipMemcpy(code, d, n, hipMemcpyDeviceToHost);
  (void)hipFree(d);
  return (int)e;
}

// ---- rsqrt_cr (bcn_common.hpp) ------------------------------------------------
// For every float bit pattern in [lo, hi): rsqrt_cr(x) against the compiler's
// correctly rounded 1.0f / sqrtf(x), mismatches counted per exponent field.

__global__ void rsqrt_kernel(uint32_t lo, uint32_t hi, unsigned long long* bad) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t u = (uint64_t)lo + blockIdx.x * blockDim.x + threadIdx.x; u < hi; u += stride) {
    const uint32_t bits = (uint32_t)u;
    const float x = __uint_as_float(bits);
    if (__float_as_uint(bcn::rsqrt_cr(x)) != __float_as_uint(1.0f / sqrtf(x)))
      atomicAdd(&bad[(bits >> 23) & 0xFFu], 1ull);
  }
}

extern "C" int rsqrt_check(uint32_t lo, uint32_t hi, unsigned long long* bad) {
  unsigned long long* d = nullptr;
  hipError_t e = hipMalloc(&d, 256 * sizeof(unsigned long long));
  if (e != hipSuccess) return (int)e;
  e = hipMemset(d, 0, 256 * sizeof(unsigned long long));
  if (e == hipSuccess) {
    hipLaunchKernelGGL(rsqrt_kernel, dim3(8192), dim3(256), 0, 0, lo, hi, d);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e == hipSuccess) e = hipMemcpy(bad, d, 256 * sizeof(unsigned long long), hipMemcpyDeviceToHost);
  (void)hipFree(d);
  return (int)e;
}
