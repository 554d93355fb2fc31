// TEST INFRASTRUCTURE ONLY -- the CPU oracle.
//
// An op-for-op C++ restatement of the reference crate's CPU block encoders and
// decoders (block_compression v0.3.0, src/encode/*.rs, src/decode/*.rs).  Only
// tests/, the smoke check and bench.py's cpu_baseline leg may load it, and only
// as the checker or as the timed CPU baseline.  The shipped library never links
// or calls it.
//
// Numeric rules that make it follow the Rust source exactly:
//   * no FMA contraction (built with -ffp-contract=off), no fast-math;
//   * float -> int casts saturate and map NaN to 0 (Rust `as` semantics);
//   * f32::min / f32::max are IEEE minNum / maxNum (fminf / fmaxf);
//   * f32::clamp is `if x < lo {lo}; if x > hi {hi}` (NaN passes through);
//   * every floating-point expression keeps the source's evaluation order.
//
// The float type F is `float` in the normal build.  Built with -DBC_OPCOUNT it
// becomes a counting wrapper: every f32 + - * / sqrt min max abs, compare and
// int<->float conversion adds one to a thread-local counter.  That build
// measures the algorithmic op count per block; integer bit-twiddling is not
// counted.
#pragma once
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstring>

namespace bco {

#ifdef BC_OPCOUNT
extern thread_local uint64_t g_ops;
struct F {
  float v;
  F() : v(0.0f) {}
  F(float x) : v(x) {}
  F(double x) : v((float)x) {}
  explicit F(int x) : v((float)x) {}
  friend F operator+(F a, F b) { ++g_ops; return F(a.v + b.v); }
  friend F operator-(F a, F b) { ++g_ops; return F(a.v - b.v); }
  friend F operator*(F a, F b) { ++g_ops; return F(a.v * b.v); }
  friend F operator/(F a, F b) { ++g_ops; return F(a.v / b.v); }
  friend F operator-(F a) { ++g_ops; return F(-a.v); }
  F& operator+=(F b) { ++g_ops; v += b.v; return *this; }
  F& operator-=(F b) { ++g_ops; v -= b.v; return *this; }
  F& operator*=(F b) { ++g_ops; v *= b.v; return *this; }
  F& operator/=(F b) { ++g_ops; v /= b.v; return *this; }
  friend bool operator<(F a, F b) { ++g_ops; return a.v < b.v; }
  friend bool operator>(F a, F b) { ++g_ops; return a.v > b.v; }
  friend bool operator<=(F a, F b) { ++g_ops; return a.v <= b.v; }
  friend bool operator>=(F a, F b) { ++g_ops; return a.v >= b.v; }
  friend bool operator==(F a, F b) { ++g_ops; return a.v == b.v; }
  friend bool operator!=(F a, F b) { ++g_ops; return a.v != b.v; }
};
static inline float raw(F x) { return x.v; }
static inline void count_op(int n = 1) { g_ops += (uint64_t)n; }
#else
typedef float F;
static inline float raw(F x) { return x; }
static inline void count_op(int = 1) {}
#endif

// ---- Rust numeric semantics ------------------------------------------------

static inline F sq(F x) { return x * x; }

// `x as i32` for f32 x: saturating, NaN -> 0.
static inline int32_t f2i(F x) {
  count_op();
  float v = raw(x);
  if (v != v) return 0;
  if (v >= 2147483648.0f) return INT32_MAX;
  if (v <= -2147483648.0f) return INT32_MIN;
  return (int32_t)v;
}

// `x as u32` for f32 x: saturating, NaN -> 0, negative -> 0.
static inline uint32_t f2u(F x) {
  count_op();
  float v = raw(x);
  if (v != v) return 0;
  if (v >= 4294967296.0f) return UINT32_MAX;
  if (v <= 0.0f) return 0;
  return (uint32_t)v;
}

static inline F i2f(int32_t x) { count_op(); return F((float)x); }
static inline F u2f(uint32_t x) { count_op(); return F((float)x); }

static inline F fmin_(F a, F b) { count_op(); return F(std::fmin(raw(a), raw(b))); }
static inline F fmax_(F a, F b) { count_op(); return F(std::fmax(raw(a), raw(b))); }
static inline F fabs_(F a) { count_op(); return F(std::fabs(raw(a))); }
static inline F fsqrt(F a) { count_op(); return F(std::sqrt(raw(a))); }

// f32::clamp
static inline F fclamp(F x, F lo, F hi) {
  if (x < lo) x = lo;
  if (x > hi) x = hi;
  return x;
}

// i32::clamp / u32::clamp
static inline int32_t iclamp(int32_t x, int32_t lo, int32_t hi) {
  return x < lo ? lo : (x > hi ? hi : x);
}
static inline uint32_t uclamp(uint32_t x, uint32_t lo, uint32_t hi) {
  return x < lo ? lo : (x > hi ? hi : x);
}

// ---- settings (byte layout of src/settings.rs:9-15, 80-90) -----------------

struct BC6HSettings {
  uint32_t slow_mode;
  uint32_t fast_mode;
  uint32_t refine_iterations_1p;
  uint32_t refine_iterations_2p;
  uint32_t fast_skip_threshold;
};

struct BC7Settings {
  uint32_t refine_iterations[8];
  uint32_t mode_selection[4];
  uint32_t skip_mode2;
  uint32_t fast_skip_threshold_mode1;
  uint32_t fast_skip_threshold_mode3;
  uint32_t fast_skip_threshold_mode7;
  uint32_t mode45_channel0;
  uint32_t refine_iterations_channel;
  uint32_t channels;
};

static_assert(sizeof(BC6HSettings) == 20, "BC6HSettings is 5 x u32");
static_assert(sizeof(BC7Settings) == 76, "BC7Settings is 19 x u32");

// ---- shared primitives (src/encode/common.rs) ------------------------------

int32_t get_unquant_value(uint32_t bits, int32_t index);
uint32_t get_pattern(int32_t part_id);
uint32_t get_pattern_mask(int32_t part_id, uint32_t j);
void get_skips(int32_t part_id, uint32_t skips[3]);
void put_bits(uint32_t data[5], uint32_t* pos, uint32_t bits, uint32_t v);
void data_shl_1bit_from(uint32_t data[5], size_t from_bits);
void partial_sort_list(int32_t* list, size_t length, uint32_t partial_count);
void opt_endpoints(F* ep, const F block[64], uint32_t bits, const uint32_t qblock[2],
                   uint32_t mask, size_t channels);
F get_pca_bound(const F covar[10], size_t channels);
void compute_axis(F axis[4], const F covar[10], uint32_t power_iterations, size_t channels);
void compute_stats_masked(F stats[15], const F block[64], uint32_t mask, size_t channels);
void covar_from_stats(F covar[10], const F stats[15], size_t channels);
void block_pca_axis(F axis[4], F dc[4], const F block[64], uint32_t mask, size_t channels);
F block_pca_bound_split(const F block[64], uint32_t mask, const F full_stats[15],
                        size_t channels);
F block_quant(uint32_t qblock[2], const F block[64], uint32_t bits, const F* ep,
              uint32_t pattern, size_t channels,
              bool floor_interp = false);
void block_segment_core(F* ep, const F block[64], uint32_t mask, size_t channels);
void bc7_code_qblock(uint32_t data[5], uint32_t* qpos, const uint32_t qblock[2],
                     uint32_t bits, uint32_t flips);
void bc7_code_adjust_skip_mode01237(uint32_t data[5], size_t mode, int32_t part_id);
void bc7_code_apply_swap_mode456(int32_t* qep, size_t channels, uint32_t qblock[2],
                                 uint32_t bits);
uint32_t bc7_code_apply_swap_mode01237(int32_t qep[24], const uint32_t qblock[2],
                                       size_t mode, int32_t part_id);

// ---- per-block encoders -----------------------------------------------------

// BC7 (src/encode/bc7.rs): `rgba` points at the block's top-left texel.
// Writes 16 bytes.  Returns the final best_err (for invariant tests).
float bc7_encode_block(const uint8_t* rgba, size_t stride, const BC7Settings& s,
                       uint8_t out[16]);

// BC1-BC5 (src/encode/bc1_to_5.rs).
void bc1_encode_block(const uint8_t* rgba, size_t stride, uint8_t out[8]);
void bc2_encode_block(const uint8_t* rgba, size_t stride, uint8_t out[16]);
void bc3_encode_block(const uint8_t* rgba, size_t stride, uint8_t out[16]);
void bc4_encode_block(const uint8_t* rgba, size_t stride, uint8_t out[8]);
void bc5_encode_block(const uint8_t* rgba, size_t stride, uint8_t out[16]);

// BC6H (src/encode/bc6h.rs).  `rgba16` are f16 bit patterns, stride in f16
// elements.
float bc6h_encode_block_16(const uint16_t* rgba16, size_t stride, const BC6HSettings& s,
                           uint8_t out[16], bool is_signed = false);
float bc6h_encode_block_8(const uint8_t* rgba, size_t stride, const BC6HSettings& s,
                          uint8_t out[16]);

// ---- decoders (src/decode/block.rs) ------------------------------------------

void decode_block_bc1(const uint8_t* in, uint8_t* out, size_t pitch);
void decode_block_bc2(const uint8_t* in, uint8_t* out, size_t pitch);
void decode_block_bc3(const uint8_t* in, uint8_t* out, size_t pitch);
void decode_block_bc4(const uint8_t* in, uint8_t* out, size_t pitch);
void decode_block_bc5(const uint8_t* in, uint8_t* out, size_t pitch);
void decode_block_bc6h(const uint8_t* in, uint16_t* out, size_t pitch, bool is_signed,
                       int32_t* comp = nullptr);
void decode_block_bc6h_float(const uint8_t* in, float* out, size_t pitch, bool is_signed);
void decode_block_bc7(const uint8_t* in, uint8_t* out, size_t pitch);

// IEEE binary16 helpers (the `half` crate's from_f32 / to_f32 are RNE / exact).
uint16_t f32_to_f16_bits(float x);
float f16_bits_to_f32(uint16_t h);

}  // namespace bco
