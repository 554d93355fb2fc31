// Shared device helpers and format tables for the BCn kernels.
//
// Numeric contract (what makes the GPU output byte-identical to the reference
// crate's CPU encoder, src/encode/*.rs):
//   * the library is compiled with -ffp-contract=off and correctly rounded f32
//     division and sqrt, so every float expression written here rounds exactly
//     like the Rust source's;
//   * sums of products of integer-valued floats whose every partial sum stays
//     below 2^24 are exact in f32 in any order; those are computed in int32
//     (each site states its bound);
//   * Rust's saturating `as i32` followed by a clamp to [lo, hi] with lo >= 0
//     equals trunc(min(max(x, lo), hi)) including NaN -> lo (qclamp below).
//
// BCN_HD marks functions that are compiled for gfx950 and, in the CPU
// kernel-check build the tests use, for the host.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define BCN_HD __host__ __device__ __forceinline__
#define BCN_TABLE __constant__
#else
#include <cmath>
#define BCN_HD inline
#define BCN_TABLE static const
#endif

namespace bcn {

// ---- numeric helpers --------------------------------------------------------

BCN_HD float sq(float x) { return x * x; }

// Rust `x as i32` then i32::clamp(lo, hi) for integer lo >= 0: NaN -> lo.
BCN_HD int qclamp(float x, float lo, float hi) { return (int)fminf(fmaxf(x, lo), hi); }

// Rust f32::clamp: `if x < lo {lo}; if x > hi {hi}` (NaN passes through).
BCN_HD float rclamp(float x, float lo, float hi) {
  x = x < lo ? lo : x;
  return x > hi ? hi : x;
}

BCN_HD int iclamp(int x, int lo, int hi) { return x < lo ? lo : (x > hi ? hi : x); }

// a * b for operands in [-2^23, 2^23) whose product fits in int32.  gfx950
// multiplies 24-bit operands at full rate (v_mul_i32_i24, v_mad_i32_i24) and
// 32-bit ones at quarter rate (v_mul_lo_u32).  The compiler picks the fast
// form only where instruction selection can prove the range inside one basic
// block, which it mostly cannot here, so the intrinsic is called directly
// (clang has no builtin for it; the asm label names the LLVM intrinsic).
#if defined(__HIP_DEVICE_COMPILE__)
extern "C" __device__ int bcn_llvm_mul_i24(int, int) __asm("llvm.amdgcn.mul.i24");
#endif
BCN_HD int mul24(int a, int b) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __mul24(a, b);
#else
  return a * b;
#endif
}
BCN_HD int mul24i(int a, int b) {
#if defined(__HIP_DEVICE_COMPILE__)
  return bcn_llvm_mul_i24(a, b);
#else
  return a * b;
#endif
}

// ---- two signed 16-bit lanes per word ---------------------------------------
//
// The integer-exact parts of the 8-bit encoders work on channel pairs: an
// RGBA8 texel word gives (R, B) = t & 0x00FF00FF and (G, A) = (t >> 8) &
// 0x00FF00FF.  Callers keep every lane value within [-2^15, 2^15), so the
// packed 16-bit integer instructions (v_pk_*_i16/u16, v_dot2_i32_i16) give the
// same integers as 32-bit arithmetic.  Host builds compute lane by lane.

BCN_HD uint32_t pk16(int lo, int hi) { return ((uint32_t)lo & 0xFFFFu) | ((uint32_t)hi << 16); }

#if defined(__HIP_DEVICE_COMPILE__)
typedef short bcn_s16x2 __attribute__((ext_vector_type(2)));
BCN_HD bcn_s16x2 as_s16x2(uint32_t v) { return __builtin_bit_cast(bcn_s16x2, v); }
BCN_HD uint32_t as_u32(bcn_s16x2 v) { return __builtin_bit_cast(uint32_t, v); }
#else
BCN_HD int lane16(uint32_t v, int h) { return (int)(int16_t)(uint16_t)(v >> (16 * h)); }
#endif

// Lane-wise a - b and a + b.
BCN_HD uint32_t pk_sub16(uint32_t a, uint32_t b) {
#if defined(__HIP_DEVICE_COMPILE__)
  return as_u32(as_s16x2(a) - as_s16x2(b));
#else
  return pk16(lane16(a, 0) - lane16(b, 0), lane16(a, 1) - lane16(b, 1));
#endif
}
BCN_HD uint32_t pk_add16(uint32_t a, uint32_t b) {
#if defined(__HIP_DEVICE_COMPILE__)
  return as_u32(as_s16x2(a) + as_s16x2(b));
#else
  return pk16(lane16(a, 0) + lane16(b, 0), lane16(a, 1) + lane16(b, 1));
#endif
}

// Lane-wise (w * d + 32) >> 6 (arithmetic shift), |w * d + 32| < 2^15.
BCN_HD uint32_t pk_wdelta16(uint32_t w, uint32_t d) {
#if defined(__HIP_DEVICE_COMPILE__)
  const bcn_s16x2 c32 = {32, 32}, c6 = {6, 6};
  return as_u32((as_s16x2(w) * as_s16x2(d) + c32) >> c6);
#else
  return pk16((lane16(w, 0) * lane16(d, 0) + 32) >> 6, (lane16(w, 1) * lane16(d, 1) + 32) >> 6);
#endif
}

// Whether any lane of the wave has `p` (the lane's own value on the host).
BCN_HD bool wave_any(bool p) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __ballot(p) != 0;
#else
  return p;
#endif
}

// Whether every lane of the wave has `p` (the lane's own value on the host).
BCN_HD bool wave_all(bool p) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __ballot(!p) == 0;
#else
  return p;
#endif
}

// Lane-wise minimum of unsigned 16-bit lanes (v_pk_min_u16).
BCN_HD uint32_t pk_min_u16(uint32_t a, uint32_t b) {
#if defined(__HIP_DEVICE_COMPILE__)
  typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));
  return __builtin_bit_cast(uint32_t, __builtin_elementwise_min(__builtin_bit_cast(u16x2, a),
                                                                 __builtin_bit_cast(u16x2, b)));
#else
  const uint32_t lo = (a & 0xFFFFu) < (b & 0xFFFFu) ? (a & 0xFFFFu) : (b & 0xFFFFu);
  const uint32_t hi = (a >> 16) < (b >> 16) ? (a >> 16) : (b >> 16);
  return lo | (hi << 16);
#endif
}

// Lane-wise logical shifts of unsigned 16-bit lanes by per-lane amounts
// (v_pk_lshlrev_b16 / v_pk_lshrrev_b16); amounts below 16.
BCN_HD uint32_t pk_shl16(uint32_t v, uint32_t sh) {
#if defined(__HIP_DEVICE_COMPILE__)
  typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));
  return __builtin_bit_cast(uint32_t, (u16x2)(__builtin_bit_cast(u16x2, v) << __builtin_bit_cast(u16x2, sh)));
#else
  return pk16((int)(((v & 0xFFFFu) << (sh & 0xFu)) & 0xFFFFu), (int)(((v >> 16) << ((sh >> 16) & 0xFu)) & 0xFFFFu));
#endif
}
BCN_HD uint32_t pk_shr16(uint32_t v, uint32_t sh) {
#if defined(__HIP_DEVICE_COMPILE__)
  typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));
  return __builtin_bit_cast(uint32_t, (u16x2)(__builtin_bit_cast(u16x2, v) >> __builtin_bit_cast(u16x2, sh)));
#else
  return pk16((int)((v & 0xFFFFu) >> (sh & 0xFu)), (int)((v >> 16) >> ((sh >> 16) & 0xFu)));
#endif
}

// Byte permute (v_perm_b32): byte i of the result is byte sel.byte_i of
// {src1 (0..3), src0 (4..7)}; selector 12 gives 0.
BCN_HD uint32_t perm_b32(uint32_t src0, uint32_t src1, uint32_t sel) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_perm(src0, src1, sel);
#else
  uint64_t bytes = ((uint64_t)src0 << 32) | src1;
  uint32_t r = 0;
  for (int i = 0; i < 4; i++) {
    uint32_t s = (sel >> (8 * i)) & 0xFFu;
    uint32_t b = s < 8 ? (uint32_t)(bytes >> (8 * s)) & 0xFFu : 0u;
    r |= b << (8 * i);
  }
  return r;
#endif
}

// Bits off .. off + w - 1 of x (v_bfe_u32): the instruction takes off and w
// modulo 32, w = 0 gives 0, and bits past bit 31 read as 0.  The host form
// computes the same for every input.
BCN_HD uint32_t bfe_u32(uint32_t x, uint32_t off, uint32_t w) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_ubfe(x, off, w);
#else
  off &= 31u;
  w &= 31u;
  return w == 0 ? 0u : (x >> off) & ((1u << w) - 1u);
#endif
}

// Sum of the four byte products of a and b, plus c (v_dot4_u32_u8).
BCN_HD uint32_t udot4(uint32_t a, uint32_t b, uint32_t c) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_udot4(a, b, c, false);
#else
  for (int i = 0; i < 4; i++) c += ((a >> (8 * i)) & 0xFFu) * ((b >> (8 * i)) & 0xFFu);
  return c;
#endif
}

// a.lo * b.lo + a.hi * b.hi + c with signed lanes and 32-bit products.
// Every caller's sum stays well inside int32, so the clamp bit (saturation
// of the result) never acts; it makes the compiler select the three-operand
// v_dot2_i32_i16, whose accumulator may be any register or an inline
// constant, instead of the two-operand v_dot2c_i32_i16, which needs a copy
// of the accumulator into the destination whenever the accumulator is a
// constant or is still live (one extra move per dot-product chain).
BCN_HD int dot2_16(uint32_t a, uint32_t b, int c) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_sdot2(as_s16x2(a), as_s16x2(b), c, true);
#else
  return lane16(a, 0) * lane16(b, 0) + lane16(a, 1) * lane16(b, 1) + c;
#endif
}

// a / b from y = 1.0f / b (Markstein's correction step): q = a*y, then the
// exact residual r = a - b*q by FMA, then q + r*y rounded once.  This is not
// correctly rounded for every (a, b), so it is used only at call sites whose
// whole input domain was checked exhaustively against IEEE division
// (tests/divcheck/divcheck.cpp), namely
//   * integer a in [0, 2^24] over integer b in [1, 16];
//   * integer a with |a| <= 510 sqrt(b) over integer b in [1, 260100];
//   * integer a in [-255, 255] over b = k + 0.001f, integer k in [-255, 255];
//   * every finite a over b = 255 or b = 65535, except that a -0 numerator
//     gives +0; each of those sites adds 0.5 (and subtracts a p-bit) after
//     scaling the quotient, which maps -0 and +0 to the same value.
// Three operations instead of the ten of a correctly rounded division.
BCN_HD float div_rcp(float a, float b, float y) {
  float q = a * y;
  float r = __builtin_fmaf(-b, q, a);
  return __builtin_fmaf(r, y, q);
}

#if defined(__HIP_DEVICE_COMPILE__)
// The square root of x >= 2^-100 (finite), correctly rounded, from the
// hardware reciprocal square root y ~ 1/sqrt(x) by Markstein's method (P.
// Markstein, "IA-64 and Elementary Functions", 2000, ch. 8): g = x y and
// h = y / 2 approximate sqrt(x) and 1/(2 sqrt(x)); one coupled correction
// r = 1/2 - g h refines both, and the final FMA from the exact residual
// x - g^2 rounds once.  Equal to sqrtf on every float from 2^-102 up and at 0
// (exhaustive GPU sweep, tests/test_gpu_fastmath.py); the callers pass 0 as
// the smallest normal argument, which gives g = 0 and a result of 0.
// Seven full-rate f32 operations and one transcendental, against the
// compiler's correction of the hardware square root by two candidate
// residuals (compares and selects).
__device__ __forceinline__ float sqrt_markstein(float x) {
  const float y = __builtin_amdgcn_rsqf(x);
  const float g = x * y, h = 0.5f * y;
  const float r = __builtin_fmaf(-g, h, 0.5f);
  const float g1 = __builtin_fmaf(g, r, g), h1 = __builtin_fmaf(h, r, h);
  const float d = __builtin_fmaf(-g1, g1, x);
  return __builtin_fmaf(d, h1, g1);
}
#endif

// sqrtf(x), correctly rounded, for every x: sqrt_markstein on 0 and on
// [2^-100, 2^128) (with the reciprocal square root of 2^-100 for 0, which
// keeps it finite); other inputs (tiny, negative, infinite, NaN) take the compiler's
// sequence, a branch no wave takes in practice.
BCN_HD float sqrt_cr(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  // x in [2^-100, 2^128) as one unsigned range test on the bit pattern
  // (negative x, +-inf and NaN fall outside it), or +0.
  const uint32_t u = __float_as_uint(x);
  if (!(u - 0x0D800000u <= 0x7F7FFFFFu - 0x0D800000u || u == 0u)) return sqrtf(x);
  const float xs = __uint_as_float(u > 0x0D800000u ? u : 0x0D800000u);
  const float y = __builtin_amdgcn_rsqf(xs);
  const float g = x * y, h = 0.5f * y;
  const float r = __builtin_fmaf(-g, h, 0.5f);
  const float g1 = __builtin_fmaf(g, r, g), h1 = __builtin_fmaf(h, r, h);
  const float d = __builtin_fmaf(-g1, g1, x);
  return __builtin_fmaf(d, h1, g1);
#else
  return sqrtf(x);
#endif
}

// 1.0f / x for normal x with |x| < 2^126, where it is correctly rounded: the
// hardware reciprocal and one Newton step with FMA (exhaustive GPU sweep,
// tests/test_gpu_fastmath.py).  Used without a guard where the argument is
// shown to lie in that range (each call site says why); for x = 0 it gives
// NaN instead of +inf.
BCN_HD float rcp_normal(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  const float y = __builtin_amdgcn_rcpf(x);
  return __builtin_fmaf(__builtin_fmaf(-x, y, 1.0f), y, y);
#else
  return 1.0f / x;
#endif
}

// 1.0f / x, correctly rounded, for every x: the hardware reciprocal and one
// Newton step with FMA equal the correctly rounded quotient for every normal
// x with |x| < 2^126 (exhaustive GPU sweep, tests/test_gpu_fastmath.py);
// other inputs take the compiler's division.
BCN_HD float rcp_cr(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  const float ax = fabsf(x);
  if (!(ax >= 0x1p-125f && ax < 0x1p125f)) return 1.0f / x;
  const float y = __builtin_amdgcn_rcpf(x);
  const float e = __builtin_fmaf(-x, y, 1.0f);
  return __builtin_fmaf(e, y, y);
#else
  return 1.0f / x;
#endif
}

// 1.0f / sqrtf(x), both correctly rounded, for every x.  For x in
// [2^-100, 2^128) the square root is sqrt_markstein and lies in
// [2^-50, 2^64], where the hardware reciprocal with one Newton step is the
// correctly rounded 1/s (exhaustive GPU sweep, tests/test_gpu_fastmath.py);
// other inputs take the compiler's sequences, behind one branch.
BCN_HD float rsqrt_cr(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  // x in [2^-100, 2^128): one unsigned range test on the bit pattern.
  if (!(__float_as_uint(x) - 0x0D800000u <= 0x7F7FFFFFu - 0x0D800000u)) return 1.0f / sqrtf(x);
  const float s = sqrt_markstein(x);
  const float y = __builtin_amdgcn_rcpf(s);
  return __builtin_fmaf(__builtin_fmaf(-s, y, 1.0f), y, y);
#else
  return 1.0f / sqrtf(x);
#endif
}

// x / 255.0f and x / 65535.0f for every finite x (see div_rcp).
BCN_HD float div255(float x) { return div_rcp(x, 255.0f, 1.0f / 255.0f); }
BCN_HD float div65535(float x) { return div_rcp(x, 65535.0f, 1.0f / 65535.0f); }

// A copy of x the compiler must treat as unknown: stops it from hoisting
// per-channel texel extractions out of the candidate loops and keeping 48
// unpacked values live; re-extracting a byte costs one instruction.
BCN_HD uint32_t fresh(uint32_t x) {
#if defined(__HIP_DEVICE_COMPILE__)
  asm volatile("" : "+v"(x));
#else
  asm volatile("" : "+r"(x));
#endif
  return x;
}

// A wave-uniform value copied into a vector register.  On gfx950 a VALU
// instruction of the fast kinds (f32 add/mul/fma, integer add, logic) issues
// every ~2.4 cycles per SIMD with vector or constant operands but every ~4.2
// when one operand is a scalar register; a value used by many such
// instructions is cheaper in a vector register (one copy).
BCN_HD float in_vgpr(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  asm volatile("" : "+v"(x));
#endif
  return x;
}
BCN_HD uint32_t in_vgpr(uint32_t x) {
#if defined(__HIP_DEVICE_COMPILE__)
  asm volatile("" : "+v"(x));
#endif
  return x;
}

// A pointer the compiler must treat as unknown (see fresh): loads through it
// cannot be hoisted out of the loop that launders it.
template <typename T>
BCN_HD const T* fresh_ptr(const T* p) {
#if defined(__HIP_DEVICE_COMPILE__)
  asm volatile("" : "+v"(p));
#else
  asm volatile("" : "+r"(p));
#endif
  return p;
}

// Makes x depend on `dep` without changing it: code reading x cannot be
// scheduled before `dep` is computed.  Used to keep two independent fits in
// sequence instead of interleaved (interleaving doubles the live registers).
BCN_HD void after(float& x, int dep) {
#if defined(__HIP_DEVICE_COMPILE__)
  asm volatile("" : "+v"(x) : "v"(dep));
#else
  asm volatile("" : "+r"(x) : "r"(dep));
#endif
}

// Makes the block's 16 texel registers opaque in place.  The old values are
// overwritten, so this costs no instruction (unlike a per-access copy through
// fresh).  Called at the entry of every function that reads texels, it stops
// the compiler from hoisting per-channel extractions out of the callers'
// loops or sharing them between calls, which would keep up to 48 unpacked
// values live.  The array's owner is a non-const register array.
BCN_HD void launder_texels(const uint32_t px[16]) {
  uint32_t* p = const_cast<uint32_t*>(px);
#pragma unroll
  for (int k = 0; k < 16; k++) p[k] = fresh(p[k]);
}

// Texel k of a block as the mode-4/5 candidates see it (bc7.rs:680-696):
// ROT 0 = as stored; ROT 1 = channel byte at `sh` forced to 255 (3-channel
// rotation); ROT 2 = channel byte at `sh` replaced by alpha (4-channel).
// Consumers read only bytes 0..2.
// The caller has laundered px (launder_texels).
template <int ROT>
BCN_HD uint32_t texel(const uint32_t px[16], int k, uint32_t sh) {
  uint32_t t = px[k];
  if (ROT == 1) return t | (0xFFu << sh);
  if (ROT == 2) return (t & ~(0xFFu << sh)) | ((t >> 24) << sh);
  return t;
}

// Byte p of a packed RGBA8 texel (a single v_cvt_f32_ubyteN / v_bfe on gfx950).
BCN_HD float chf(uint32_t px, int p) { return (float)((px >> (8 * p)) & 0xFFu); }
BCN_HD int chi(uint32_t px, int p) { return (int)((px >> (8 * p)) & 0xFFu); }

// The texels as channel-major bytes: cm[g][p] holds channel p of texels
// 4g .. 4g + 3 (byte i = texel 4g + i), two v_perm_b32 per word, so that a
// sum over a group of byte products is one v_dot4_u32_u8.
BCN_HD void channel_major(const uint32_t px[16], uint32_t cm[4][4]) {
#pragma unroll
  for (int g = 0; g < 4; g++) {
    const uint32_t t0 = fresh(px[4 * g]), t1 = fresh(px[4 * g + 1]);
    const uint32_t t2 = fresh(px[4 * g + 2]), t3 = fresh(px[4 * g + 3]);
    const uint32_t a = perm_b32(t1, t0, 0x06020400u), b = perm_b32(t1, t0, 0x07030501u);
    const uint32_t c = perm_b32(t3, t2, 0x06020400u), d = perm_b32(t3, t2, 0x07030501u);
    cm[g][0] = perm_b32(c, a, 0x05040100u);
    cm[g][1] = perm_b32(d, b, 0x05040100u);
    cm[g][2] = perm_b32(c, a, 0x07060302u);
    cm[g][3] = perm_b32(d, b, 0x07060302u);
  }
}

// Index weights of BPTC interpolation (2/3/4-bit: 0,21,43,64 / 0,9,..,64 /
// 0,4,..,64), one byte per index.
BCN_HD int weight2(int i) { return (int)((0x402B1500u >> (8 * i)) & 0xFFu); }
BCN_HD int weight3(int i) { return (int)((0x40372E251B120900ull >> (8 * i)) & 0xFFu); }
BCN_HD int weight4(int i) {
  const uint64_t lo = 0x1E1A15110D090400ull, hi = 0x403C37332F2B2622ull;
  return (int)(((i < 8 ? lo : hi) >> (8 * (i & 7))) & 0xFFu);
}
template <int BITS>
BCN_HD int weight(int i) {
  return BITS == 2 ? weight2(i) : (BITS == 3 ? weight3(i) : weight4(i));
}

// The same weights as one multiply-add and shift: round(64 i / m) with
// m = 2^BITS - 1 equals (i * (4095 / m) + 32) >> 6 for every index of the
// 2-, 3- and 4-bit tables (checked at compile time below).
template <int BITS>
BCN_HD int weight_mad(int i) {
  constexpr int K = BITS == 2 ? 1365 : (BITS == 3 ? 585 : 273);
  return (mul24(i, K) + 32) >> 6;
}

// The 2- and 3-bit weight tables as little-endian bytes: indices 0-3 in the
// low word, 4-7 in the high word.
template <int BITS>
constexpr uint32_t weight_bytes(int word) {
  uint32_t w = 0;
  constexpr int K = BITS == 2 ? 1365 : 585;
  for (int b = 0; b < 4; b++) w |= (uint32_t)(((4 * word + b) * K + 32) >> 6) << (8 * b);
  return w;
}
static_assert(weight_bytes<2>(0) == 0x402B1500u && weight_bytes<3>(0) == 0x1B120900u &&
                  weight_bytes<3>(1) == 0x40372E25u,
              "byte tables of the 2- and 3-bit weights");

// Weights of indices q - 1 and q (1 <= q < 2^BITS) as the low and high 16-bit
// lanes.  2 and 3 bits: one v_perm_b32 gathers bytes q - 1 and q of the byte
// table into the lanes' low bytes (selector 0x0C gives a zero byte), its
// selector (q - 1) | 0x0C << 8 | q << 16 | 0x0C << 24 one 24-bit multiply-add.
// 4 bits: (q - 1) | q << 16 by the multiply-add, then the closed form
// lane-wise (i * K + 32 <= 4127 fits the lanes).
template <int BITS>
BCN_HD uint32_t weight_pair(int q) {
#if defined(__HIP_DEVICE_COMPILE__)
  if (BITS < 4) {
    const uint32_t sel = (uint32_t)mul24(q, 0x10001) + 0x0C000BFFu;
    return perm_b32(weight_bytes<BITS>(BITS == 3), weight_bytes<BITS>(0), sel);
  }
  constexpr int K = 273;
  typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));
  const u16x2 qq = __builtin_bit_cast(u16x2, (uint32_t)(mul24(q, 0x10001) - 1));
  const u16x2 k = {K, K}, c32 = {32, 32}, c6 = {6, 6};
  return __builtin_bit_cast(uint32_t, (u16x2)((qq * k + c32) >> c6));
#else
  return pk16(weight_mad<BITS>(q - 1), weight_mad<BITS>(q));
#endif
}

// Lane-wise (w * d + c) >> 6 (arithmetic shift) with w lane H of wp in both
// lanes.  Callers keep |w * d + c| < 2^15; the 16-bit multiply-add wraps
// modulo 2^16, which is exact for such results.
template <int H>
BCN_HD uint32_t pk_wmad16(uint32_t wp, uint32_t d, uint32_t c) {
#if defined(__HIP_DEVICE_COMPILE__)
  const bcn_s16x2 w = H ? as_s16x2(wp).yy : as_s16x2(wp).xx, c6 = {6, 6};
  return as_u32((w * as_s16x2(d) + as_s16x2(c)) >> c6);
#else
  const int w = lane16(wp, H);
  return pk16((w * lane16(d, 0) + lane16(c, 0)) >> 6, (w * lane16(d, 1) + lane16(c, 1)) >> 6);
#endif
}

// Lane-wise 32 - 64 * b for |b| <= 255: the rounding term of pk_wmad16 with
// b whole units taken off.  m64 is 0xFFC0FFC0 (-64 in both lanes) held in a
// register: with a visible constant the compiler splits the multiply-add into
// a shift and a subtract.
BCN_HD uint32_t pk_round_off16(uint32_t b, uint32_t m64) {
#if defined(__HIP_DEVICE_COMPILE__)
  const bcn_s16x2 c32 = {32, 32};
  return as_u32(as_s16x2(b) * as_s16x2(m64) + c32);
#else
  return pk16(32 + lane16(m64, 0) * lane16(b, 0), 32 + lane16(m64, 1) * lane16(b, 1));
#endif
}

// Lane-wise (a * b + c) >> 6 (arithmetic shift) for |a * b + c| < 2^15.
BCN_HD uint32_t pk_mad_shr6(uint32_t a, uint32_t b, uint32_t c) {
#if defined(__HIP_DEVICE_COMPILE__)
  const bcn_s16x2 c6 = {6, 6};
  return as_u32((as_s16x2(a) * as_s16x2(b) + as_s16x2(c)) >> c6);
#else
  return pk16((lane16(a, 0) * lane16(b, 0) + lane16(c, 0)) >> 6,
              (lane16(a, 1) * lane16(b, 1) + lane16(c, 1)) >> 6);
#endif
}

// Lane-wise a * b + c modulo 2^16 (v_pk_mad_u16).
BCN_HD uint32_t pk_mad16(uint32_t a, uint32_t b, uint32_t c) {
#if defined(__HIP_DEVICE_COMPILE__)
  return as_u32(as_s16x2(a) * as_s16x2(b) + as_s16x2(c));
#else
  return pk16(lane16(a, 0) * lane16(b, 0) + lane16(c, 0), lane16(a, 1) * lane16(b, 1) + lane16(c, 1));
#endif
}

// Lane-wise a * w.lo + c modulo 2^16: both lanes of a times the low lane of
// w (v_pk_mad_u16 with the low half of w read for both lanes).
BCN_HD uint32_t pk_mad16_lo(uint32_t a, uint32_t w, uint32_t c) {
#if defined(__HIP_DEVICE_COMPILE__)
  return as_u32(as_s16x2(a) * as_s16x2(w).xx + as_s16x2(c));
#else
  return pk16(lane16(a, 0) * lane16(w, 0) + lane16(c, 0), lane16(a, 1) * lane16(w, 0) + lane16(c, 1));
#endif
}

// Lane-wise a * a for |a| <= 255 (the square fits an unsigned 16-bit lane).
BCN_HD uint32_t pk_sq16(uint32_t a) {
#if defined(__HIP_DEVICE_COMPILE__)
  typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));
  const u16x2 v = __builtin_bit_cast(u16x2, a);
  return __builtin_bit_cast(uint32_t, (u16x2)(v * v));
#else
  return pk16(lane16(a, 0) * lane16(a, 0), lane16(a, 1) * lane16(a, 1));
#endif
}

constexpr bool weight_mad_holds() {
  constexpr int t2[4] = {0, 21, 43, 64};
  constexpr int t3[8] = {0, 9, 18, 27, 37, 46, 55, 64};
  constexpr int t4[16] = {0, 4, 9, 13, 17, 21, 26, 30, 34, 38, 43, 47, 51, 55, 60, 64};
  for (int i = 0; i < 4; i++)
    if (((i * 1365 + 32) >> 6) != t2[i]) return false;
  for (int i = 0; i < 8; i++)
    if (((i * 585 + 32) >> 6) != t3[i]) return false;
  for (int i = 0; i < 16; i++)
    if (((i * 273 + 32) >> 6) != t4[i]) return false;
  return true;
}
static_assert(weight_mad_holds(), "BPTC weight closed form");

// ---- partition tables -------------------------------------------------------
//
// The BPTC partition shapes (2 subsets: 64 shapes, 3 subsets: 64 shapes) as
// 2 bits of subset index per texel (texel k at bits 2k..2k+1) and the anchor
// (fix-up) texels of subsets 1 and 2.  The encoder's per-subset texel masks
// are derived from the shapes at compile time.
struct PartitionTables {
  uint32_t subset[128];  // [0,64): 2-subset shapes, [64,128): 3-subset shapes
  uint32_t mask01[128];  // texel mask of subset 0 (low 16 bits) and subset 1 (high 16)
  uint8_t anchor1[128];  // anchor texel of subset 1
  uint8_t anchor2[128];  // anchor texel of subset 2 (3-subset shapes; 0 otherwise)
};

constexpr uint32_t kShapes[128] = {
      0x50505050u, 0x40404040u, 0x54545454u, 0x54505040u, 0x50404000u, 0x55545450u, 0x55545040u,
      0x54504000u, 0x50400000u, 0x55555450u, 0x55544000u, 0x54400000u, 0x55555440u, 0x55550000u,
      0x55555500u, 0x55000000u, 0x55150100u, 0x00004054u, 0x15010000u, 0x00405054u, 0x00004050u,
      0x15050100u, 0x05010000u, 0x40505054u, 0x00404050u, 0x05010100u, 0x14141414u, 0x05141450u,
      0x01155440u, 0x00555500u, 0x15014054u, 0x05414150u, 0x44444444u, 0x55005500u, 0x11441144u,
      0x05055050u, 0x05500550u, 0x11114444u, 0x41144114u, 0x44111144u, 0x15055054u, 0x01055040u,
      0x05041050u, 0x05455150u, 0x14414114u, 0x50050550u, 0x41411414u, 0x00141400u, 0x00041504u,
      0x00105410u, 0x10541000u, 0x04150400u, 0x50410514u, 0x41051450u, 0x05415014u, 0x14054150u,
      0x41050514u, 0x41505014u, 0x40011554u, 0x54150140u, 0x50505500u, 0x00555050u, 0x15151010u,
      0x54540404u, 0xAA685050u, 0x6A5A5040u, 0x5A5A4200u, 0x5450A0A8u, 0xA5A50000u, 0xA0A05050u,
      0x5555A0A0u, 0x5A5A5050u, 0xAA550000u, 0xAA555500u, 0xAAAA5500u, 0x90909090u, 0x94949494u,
      0xA4A4A4A4u, 0xA9A59450u, 0x2A0A4250u, 0xA5945040u, 0x0A425054u, 0xA5A5A500u, 0x55A0A0A0u,
      0xA8A85454u, 0x6A6A4040u, 0xA4A45000u, 0x1A1A0500u, 0x0050A4A4u, 0xAAA59090u, 0x14696914u,
      0x69691400u, 0xA08585A0u, 0xAA821414u, 0x50A4A450u, 0x6A5A0200u, 0xA9A58000u, 0x5090A0A8u,
      0xA8A09050u, 0x24242424u, 0x00AA5500u, 0x24924924u, 0x24499224u, 0x50A50A50u, 0x500AA550u,
      0xAAAA4444u, 0x66660000u, 0xA5A0A5A0u, 0x50A050A0u, 0x69286928u, 0x44AAAA44u, 0x66666600u,
      0xAA444444u, 0x54A854A8u, 0x95809580u, 0x96969600u, 0xA85454A8u, 0x80959580u, 0xAA141414u,
      0x96960000u, 0xAAAA1414u, 0xA05050A0u, 0xA0A5A5A0u, 0x96000000u, 0x40804080u, 0xA9A8A9A8u,
      0xAAAAAA44u, 0x2A4A5254u};

// Texel mask (16 bits) of subset j of shape i.
constexpr uint32_t shape_mask(int i, int j) {
  uint32_t m = 0;
  for (int k = 0; k < 16; k++)
    if (((kShapes[i] >> (2 * k)) & 3u) == (uint32_t)j) m |= 1u << k;
  return m;
}

constexpr PartitionTables make_partition_tables() {
  constexpr uint8_t a1[128] = {
      15, 15, 15, 15, 15, 15, 15, 15, 15, 15, 15, 15, 15, 15, 15, 15, 15, 2,  8,  2,  2,  8,
      8,  15, 2,  8,  2,  2,  8,  8,  2,  2,  15, 15, 6,  8,  2,  8,  15, 15, 2,  8,  2,  2,
      2,  15, 15, 6,  6,  2,  6,  8,  15, 15, 2,  2,  15, 15, 15, 15, 15, 2,  2,  15, 3,  3,
      15, 15, 8,  3,  15, 15, 8,  8,  6,  6,  6,  5,  3,  3,  3,  3,  8,  15, 3,  3,  6,  10,
      5,  8,  8,  6,  8,  5,  15, 15, 8,  15, 3,  5,  6,  10, 8,  15, 15, 3,  15, 5,  15, 15,
      15, 15, 3,  15, 5,  5,  5,  8,  5,  10, 5,  10, 8,  13, 15, 12, 3,  3};
  constexpr uint8_t a2[64] = {15, 8,  8,  3,  15, 15, 3,  8,  15, 15, 15, 15, 15, 15, 15, 8,
                              15, 8,  15, 3,  15, 8,  15, 8,  3,  15, 6,  10, 15, 15, 10, 8,
                              15, 3,  15, 10, 10, 8,  9,  10, 6,  15, 8,  15, 3,  6,  6,  8,
                              15, 3,  15, 15, 15, 15, 15, 15, 15, 15, 15, 15, 3,  15, 15, 8};
  PartitionTables t{};
  for (int i = 0; i < 128; i++) {
    t.subset[i] = kShapes[i];
    t.mask01[i] = shape_mask(i, 0) | (shape_mask(i, 1) << 16);
    t.anchor1[i] = a1[i];
    t.anchor2[i] = i < 64 ? 0 : a2[i - 64];
  }
  return t;
}

BCN_TABLE PartitionTables kPart = make_partition_tables();

// Texel mask (16 bits) of subset j of partition `part`.
BCN_HD uint32_t part_mask(int part, int j) {
  uint32_t m = kPart.mask01[part];
  uint32_t m0 = m & 0xFFFFu, m1 = m >> 16;
  return j == 0 ? m0 : (j == 1 ? m1 : (~(m0 | m1) & 0xFFFFu));
}

// ---- 128-bit little-endian bit writer ----------------------------------------

struct BitWriter {
  uint64_t lo = 0, hi = 0;
  uint32_t pos = 0;
  BCN_HD void put(uint32_t nbits, uint32_t v) {
    uint64_t x = (uint64_t)v;
    if (pos < 64) {
      lo |= x << pos;
      if (pos + nbits > 64) hi |= x >> (64 - pos);
    } else {
      hi |= x << (pos - 64);
    }
    pos += nbits;
  }
};

}  // namespace bcn
