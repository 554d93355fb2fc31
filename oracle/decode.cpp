// TEST INFRASTRUCTURE ONLY -- the CPU oracle (see bc_oracle.hpp).
//
// Restatement of the block decoders in src/decode/block.rs (the crate's port
// of bcdec v0.98).  All integer arithmetic; bit-exact by construction.
#include "bc_oracle.hpp"

namespace bco {
namespace {

// block.rs:1341-1403: 128-bit little-endian bit reader.
struct BitStream {
  uint64_t low, high;
  explicit BitStream(const uint8_t* d) {
    low = 0;
    high = 0;
    for (int i = 7; i >= 0; i--) low = (low << 8) | d[i];
    for (int i = 15; i >= 8; i--) high = (high << 8) | d[i];
  }
  uint32_t read_bits(uint32_t n) {
    if (n == 0) return 0;
    uint64_t mask = (n >= 64) ? ~0ull : ((1ull << n) - 1);
    uint32_t bits = (uint32_t)(low & mask);
    low >>= n;
    low |= (high & mask) << (64 - n);
    high >>= n;
    return bits;
  }
  int32_t read_bit_i32() { return (int32_t)read_bits(1); }
  int32_t read_bits_i32(uint32_t n) { return (int32_t)read_bits(n); }
  int32_t read_bits_reversed(uint32_t n) {
    int32_t bits = read_bits_i32(n);
    int32_t result = 0;
    for (uint32_t i = 0; i < n; i++) {
      result <<= 1;
      result |= bits & 1;
      bits >>= 1;
    }
    return result;
  }
};

// block.rs:122-197
void decode_color_block(const uint8_t* in, uint8_t* out, size_t pitch, bool opaque_mode) {
  uint32_t ref[4];
  uint32_t c0 = (uint32_t)in[0] | ((uint32_t)in[1] << 8);
  uint32_t c1 = (uint32_t)in[2] | ((uint32_t)in[3] << 8);
  uint32_t r0 = (c0 >> 11) & 0x1F, g0 = (c0 >> 5) & 0x3F, b0 = c0 & 0x1F;
  uint32_t r1 = (c1 >> 11) & 0x1F, g1 = (c1 >> 5) & 0x3F, b1 = c1 & 0x1F;

  uint32_t r = (r0 * 527 + 23) >> 6, g = (g0 * 259 + 33) >> 6, b = (b0 * 527 + 23) >> 6;
  ref[0] = 0xFF000000u | (b << 16) | (g << 8) | r;
  r = (r1 * 527 + 23) >> 6;
  g = (g1 * 259 + 33) >> 6;
  b = (b1 * 527 + 23) >> 6;
  ref[1] = 0xFF000000u | (b << 16) | (g << 8) | r;

  if (c0 > c1 || opaque_mode) {
    r = ((2 * r0 + r1) * 351 + 61) >> 7;
    g = ((2 * g0 + g1) * 2763 + 1039) >> 11;
    b = ((2 * b0 + b1) * 351 + 61) >> 7;
    ref[2] = 0xFF000000u | (b << 16) | (g << 8) | r;
    r = ((r0 + r1 * 2) * 351 + 61) >> 7;
    g = ((g0 + g1 * 2) * 2763 + 1039) >> 11;
    b = ((b0 + b1 * 2) * 351 + 61) >> 7;
    ref[3] = 0xFF000000u | (b << 16) | (g << 8) | r;
  } else {
    r = ((r0 + r1) * 1053 + 125) >> 8;
    g = ((g0 + g1) * 4145 + 1019) >> 11;
    b = ((b0 + b1) * 1053 + 125) >> 8;
    ref[2] = 0xFF000000u | (b << 16) | (g << 8) | r;
    ref[3] = 0;
  }

  uint32_t idx = (uint32_t)in[4] | ((uint32_t)in[5] << 8) | ((uint32_t)in[6] << 16) |
                 ((uint32_t)in[7] << 24);
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++) {
      uint32_t c = ref[idx & 3];
      uint8_t* o = out + i * pitch + j * 4;
      o[0] = (uint8_t)c;
      o[1] = (uint8_t)(c >> 8);
      o[2] = (uint8_t)(c >> 16);
      o[3] = (uint8_t)(c >> 24);
      idx >>= 2;
    }
}

// block.rs:200-217
void decode_sharp_alpha_block(const uint8_t* in, uint8_t* out, size_t pitch) {
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++) {
      int byte_index = i * 2 + (j / 2);
      int shift = (j % 2) * 4;
      uint8_t a = (in[byte_index] >> shift) & 0x0F;
      out[i * pitch + j * 4 + 3] = (uint8_t)(a * 17);
    }
}

// block.rs:219-258
void decode_smooth_alpha_block(const uint8_t* in, uint8_t* out, size_t pitch, size_t px_size) {
  uint64_t block = 0;
  for (int i = 7; i >= 0; i--) block = (block << 8) | in[i];
  uint32_t a[8];
  a[0] = (uint32_t)(block & 0xFF);
  a[1] = (uint32_t)((block >> 8) & 0xFF);
  if (a[0] > a[1]) {
    a[2] = (6 * a[0] + a[1]) / 7;
    a[3] = (5 * a[0] + 2 * a[1]) / 7;
    a[4] = (4 * a[0] + 3 * a[1]) / 7;
    a[5] = (3 * a[0] + 4 * a[1]) / 7;
    a[6] = (2 * a[0] + 5 * a[1]) / 7;
    a[7] = (a[0] + 6 * a[1]) / 7;
  } else {
    a[2] = (4 * a[0] + a[1]) / 5;
    a[3] = (3 * a[0] + 2 * a[1]) / 5;
    a[4] = (2 * a[0] + 3 * a[1]) / 5;
    a[5] = (a[0] + 4 * a[1]) / 5;
    a[6] = 0x00;
    a[7] = 0xFF;
  }
  uint64_t indices = block >> 16;
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++) {
      out[i * pitch + j * px_size] = (uint8_t)a[indices & 7];
      indices >>= 3;
    }
}

// block.rs:1256-1258
int32_t interpolate(int32_t a, int32_t b, const int32_t* w, int32_t index) {
  return (a * (64 - w[index]) + b * w[index] + 32) >> 6;
}

// block.rs:1262-1266
int32_t extend_sign(int32_t val, int32_t bits) {
  return (int32_t)((uint32_t)val << (32 - bits)) >> (32 - bits);
}

// block.rs:1270-1280
int32_t transform_inverse(int32_t val, int32_t a0, int32_t bits, bool is_signed) {
  int32_t t = (val + a0) & ((1 << bits) - 1);
  return is_signed ? extend_sign(t, bits) : t;
}

// block.rs:1284-1316
int32_t unquantize(int32_t val, int32_t bits, bool is_signed) {
  if (!is_signed) {
    if (bits >= 15) return val;
    if (val == 0) return 0;
    if (val == ((1 << bits) - 1)) return 0xFFFF;
    return ((val << 16) + 0x8000) >> bits;
  }
  if (bits >= 16) return val;
  bool s = val < 0;
  int32_t v = s ? -val : val;
  int32_t unq;
  if (v == 0) unq = 0;
  else if (v >= ((1 << (bits - 1)) - 1)) unq = 0x7FFF;
  else unq = ((v << 15) + 0x4000) >> (bits - 1);
  return s ? -unq : unq;
}

// block.rs:1320-1339
uint16_t finish_unquantize(int32_t val, bool is_signed) {
  if (!is_signed) return (uint16_t)((val * 31) >> 6);
  int32_t scaled = val < 0 ? -(((-val) * 31) >> 5) : (val * 31) >> 5;
  uint32_t sign_bit = 0;
  int32_t mag = scaled;
  if (scaled < 0) {
    sign_bit = 0x8000;
    mag = -scaled;
  }
  return (uint16_t)(sign_bit | (uint32_t)mag);
}

}  // namespace

// block.rs:806-952, exported for the table cross-checks in tests.  The
// BC6H decoder's 32-entry table (block.rs:270-316) equals the first 32
// two-subset entries, so BC6H indexes this one.
extern const uint8_t BC7_PARTITION_SETS[2][64][16];
const uint8_t BC7_PARTITION_SETS[2][64][16] = {
    {
        {128, 0, 1, 1, 0, 0, 1, 1, 0, 0, 1, 1, 0, 0, 1, 129},
        {128, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0, 129},
        {128, 1, 1, 1, 0, 1, 1, 1, 0, 1, 1, 1, 0, 1, 1, 129},
        {128, 0, 0, 1, 0, 0, 1, 1, 0, 0, 1, 1, 0, 1, 1, 129},
        {128, 0, 0, 0, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 1, 129},
        {128, 0, 1, 1, 0, 1, 1, 1, 0, 1, 1, 1, 1, 1, 1, 129},
        {128, 0, 0, 1, 0, 0, 1, 1, 0, 1, 1, 1, 1, 1, 1, 129},
        {128, 0, 0, 0, 0, 0, 0, 1, 0, 0, 1, 1, 0, 1, 1, 129},
        {128, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 0, 0, 1, 129},
        {128, 0, 1, 1, 0, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 129},
        {128, 0, 0, 0, 0, 0, 0, 1, 0, 1, 1, 1, 1, 1, 1, 129},
        {128, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 0, 1, 1, 129},
        {128, 0, 0, 1, 0, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 129},
        {128, 0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 1, 1, 1, 129},
        {128, 0, 0, 0, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 129},
        {128, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 129},
        {128, 0, 0, 0, 1, 0, 0, 0, 1, 1, 1, 0, 1, 1, 1, 129},
        {128, 1, 129, 1, 0, 0, 0, 1, 0, 0, 0, 0, 0, 0, 0, 0},
        {128, 0, 0, 0, 0, 0, 0, 0, 129, 0, 0, 0, 1, 1, 1, 0},
        {128, 1, 129, 1, 0, 0, 1, 1, 0, 0, 0, 1, 0, 0, 0, 0},
        {128, 0, 129, 1, 0, 0, 0, 1, 0, 0, 0, 0, 0, 0, 0, 0},
        {128, 0, 0, 0, 1, 0, 0, 0, 129, 1, 0, 0, 1, 1, 1, 0},
        {128, 0, 0, 0, 0, 0, 0, 0, 129, 0, 0, 0, 1, 1, 0, 0},
        {128, 1, 1, 1, 0, 0, 1, 1, 0, 0, 1, 1, 0, 0, 0, 129},
        {128, 0, 129, 1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0, 0},
        {128, 0, 0, 0, 1, 0, 0, 0, 129, 0, 0, 0, 1, 1, 0, 0},
        {128, 1, 129, 0, 0, 1, 1, 0, 0, 1, 1, 0, 0, 1, 1, 0},
        {128, 0, 129, 1, 0, 1, 1, 0, 0, 1, 1, 0, 1, 1, 0, 0},
        {128, 0, 0, 1, 0, 1, 1, 1, 129, 1, 1, 0, 1, 0, 0, 0},
        {128, 0, 0, 0, 1, 1, 1, 1, 129, 1, 1, 1, 0, 0, 0, 0},
        {128, 1, 129, 1, 0, 0, 0, 1, 1, 0, 0, 0, 1, 1, 1, 0},
        {128, 0, 129, 1, 1, 0, 0, 1, 1, 0, 0, 1, 1, 1, 0, 0},
        {128, 1, 0, 1, 0, 1, 0, 1, 0, 1, 0, 1, 0, 1, 0, 129},
        {128, 0, 0, 0, 1, 1, 1, 1, 0, 0, 0, 0, 1, 1, 1, 129},
        {128, 1, 0, 1, 1, 0, 129, 0, 0, 1, 0, 1, 1, 0, 1, 0},
        {128, 0, 1, 1, 0, 0, 1, 1, 129, 1, 0, 0, 1, 1, 0, 0},
        {128, 0, 129, 1, 1, 1, 0, 0, 0, 0, 1, 1, 1, 1, 0, 0},
        {128, 1, 0, 1, 0, 1, 0, 1, 129, 0, 1, 0, 1, 0, 1, 0},
        {128, 1, 1, 0, 1, 0, 0, 1, 0, 1, 1, 0, 1, 0, 0, 129},
        {128, 1, 0, 1, 1, 0, 1, 0, 1, 0, 1, 0, 0, 1, 0, 129},
        {128, 1, 129, 1, 0, 0, 1, 1, 1, 1, 0, 0, 1, 1, 1, 0},
        {128, 0, 0, 1, 0, 0, 1, 1, 129, 1, 0, 0, 1, 0, 0, 0},
        {128, 0, 129, 1, 0, 0, 1, 0, 0, 1, 0, 0, 1, 1, 0, 0},
        {128, 0, 129, 1, 1, 0, 1, 1, 1, 1, 0, 1, 1, 1, 0, 0},
        {128, 1, 129, 0, 1, 0, 0, 1, 1, 0, 0, 1, 0, 1, 1, 0},
        {128, 0, 1, 1, 1, 1, 0, 0, 1, 1, 0, 0, 0, 0, 1, 129},
        {128, 1, 1, 0, 0, 1, 1, 0, 1, 0, 0, 1, 1, 0, 0, 129},
        {128, 0, 0, 0, 0, 1, 129, 0, 0, 1, 1, 0, 0, 0, 0, 0},
        {128, 1, 0, 0, 1, 1, 129, 0, 0, 1, 0, 0, 0, 0, 0, 0},
        {128, 0, 129, 0, 0, 1, 1, 1, 0, 0, 1, 0, 0, 0, 0, 0},
        {128, 0, 0, 0, 0, 0, 129, 0, 0, 1, 1, 1, 0, 0, 1, 0},
        {128, 0, 0, 0, 0, 1, 0, 0, 129, 1, 1, 0, 0, 1, 0, 0},
        {128, 1, 1, 0, 1, 1, 0, 0, 1, 0, 0, 1, 0, 0, 1, 129},
        {128, 0, 1, 1, 0, 1, 1, 0, 1, 1, 0, 0, 1, 0, 0, 129},
        {128, 1, 129, 0, 0, 0, 1, 1, 1, 0, 0, 1, 1, 1, 0, 0},
        {128, 0, 129, 1, 1, 0, 0, 1, 1, 1, 0, 0, 0, 1, 1, 0},
        {128, 1, 1, 0, 1, 1, 0, 0, 1, 1, 0, 0, 1, 0, 0, 129},
        {128, 1, 1, 0, 0, 0, 1, 1, 0, 0, 1, 1, 1, 0, 0, 129},
        {128, 1, 1, 1, 1, 1, 1, 0, 1, 0, 0, 0, 0, 0, 0, 129},
        {128, 0, 0, 1, 1, 0, 0, 0, 1, 1, 1, 0, 0, 1, 1, 129},
        {128, 0, 0, 0, 1, 1, 1, 1, 0, 0, 1, 1, 0, 0, 1, 129},
        {128, 0, 129, 1, 0, 0, 1, 1, 1, 1, 1, 1, 0, 0, 0, 0},
        {128, 0, 129, 0, 0, 0, 1, 0, 1, 1, 1, 0, 1, 1, 1, 0},
        {128, 1, 0, 0, 0, 1, 0, 0, 0, 1, 1, 1, 0, 1, 1, 129},
    },
    {
        {128, 0, 1, 129, 0, 0, 1, 1, 0, 2, 2, 1, 2, 2, 2, 130},
        {128, 0, 0, 129, 0, 0, 1, 1, 130, 2, 1, 1, 2, 2, 2, 1},
        {128, 0, 0, 0, 2, 0, 0, 1, 130, 2, 1, 1, 2, 2, 1, 129},
        {128, 2, 2, 130, 0, 0, 2, 2, 0, 0, 1, 1, 0, 1, 1, 129},
        {128, 0, 0, 0, 0, 0, 0, 0, 129, 1, 2, 2, 1, 1, 2, 130},
        {128, 0, 1, 129, 0, 0, 1, 1, 0, 0, 2, 2, 0, 0, 2, 130},
        {128, 0, 2, 130, 0, 0, 2, 2, 1, 1, 1, 1, 1, 1, 1, 129},
        {128, 0, 1, 1, 0, 0, 1, 1, 130, 2, 1, 1, 2, 2, 1, 129},
        {128, 0, 0, 0, 0, 0, 0, 0, 129, 1, 1, 1, 2, 2, 2, 130},
        {128, 0, 0, 0, 1, 1, 1, 1, 129, 1, 1, 1, 2, 2, 2, 130},
        {128, 0, 0, 0, 1, 1, 129, 1, 2, 2, 2, 2, 2, 2, 2, 130},
        {128, 0, 1, 2, 0, 0, 129, 2, 0, 0, 1, 2, 0, 0, 1, 130},
        {128, 1, 1, 2, 0, 1, 129, 2, 0, 1, 1, 2, 0, 1, 1, 130},
        {128, 1, 2, 2, 0, 129, 2, 2, 0, 1, 2, 2, 0, 1, 2, 130},
        {128, 0, 1, 129, 0, 1, 1, 2, 1, 1, 2, 2, 1, 2, 2, 130},
        {128, 0, 1, 129, 2, 0, 0, 1, 130, 2, 0, 0, 2, 2, 2, 0},
        {128, 0, 0, 129, 0, 0, 1, 1, 0, 1, 1, 2, 1, 1, 2, 130},
        {128, 1, 1, 129, 0, 0, 1, 1, 130, 0, 0, 1, 2, 2, 0, 0},
        {128, 0, 0, 0, 1, 1, 2, 2, 129, 1, 2, 2, 1, 1, 2, 130},
        {128, 0, 2, 130, 0, 0, 2, 2, 0, 0, 2, 2, 1, 1, 1, 129},
        {128, 1, 1, 129, 0, 1, 1, 1, 0, 2, 2, 2, 0, 2, 2, 130},
        {128, 0, 0, 129, 0, 0, 0, 1, 130, 2, 2, 1, 2, 2, 2, 1},
        {128, 0, 0, 0, 0, 0, 129, 1, 0, 1, 2, 2, 0, 1, 2, 130},
        {128, 0, 0, 0, 1, 1, 0, 0, 130, 2, 129, 0, 2, 2, 1, 0},
        {128, 1, 2, 130, 0, 129, 2, 2, 0, 0, 1, 1, 0, 0, 0, 0},
        {128, 0, 1, 2, 0, 0, 1, 2, 129, 1, 2, 2, 2, 2, 2, 130},
        {128, 1, 1, 0, 1, 2, 130, 1, 129, 2, 2, 1, 0, 1, 1, 0},
        {128, 0, 0, 0, 0, 1, 129, 0, 1, 2, 130, 1, 1, 2, 2, 1},
        {128, 0, 2, 2, 1, 1, 0, 2, 129, 1, 0, 2, 0, 0, 2, 130},
        {128, 1, 1, 0, 0, 129, 1, 0, 2, 0, 0, 2, 2, 2, 2, 130},
        {128, 0, 1, 1, 0, 1, 2, 2, 0, 1, 130, 2, 0, 0, 1, 129},
        {128, 0, 0, 0, 2, 0, 0, 0, 130, 2, 1, 1, 2, 2, 2, 129},
        {128, 0, 0, 0, 0, 0, 0, 2, 129, 1, 2, 2, 1, 2, 2, 130},
        {128, 2, 2, 130, 0, 0, 2, 2, 0, 0, 1, 2, 0, 0, 1, 129},
        {128, 0, 1, 129, 0, 0, 1, 2, 0, 0, 2, 2, 0, 2, 2, 130},
        {128, 1, 2, 0, 0, 129, 2, 0, 0, 1, 130, 0, 0, 1, 2, 0},
        {128, 0, 0, 0, 1, 1, 129, 1, 2, 2, 130, 2, 0, 0, 0, 0},
        {128, 1, 2, 0, 1, 2, 0, 1, 130, 0, 129, 2, 0, 1, 2, 0},
        {128, 1, 2, 0, 2, 0, 1, 2, 129, 130, 0, 1, 0, 1, 2, 0},
        {128, 0, 1, 1, 2, 2, 0, 0, 1, 1, 130, 2, 0, 0, 1, 129},
        {128, 0, 1, 1, 1, 1, 130, 2, 2, 2, 0, 0, 0, 0, 1, 129},
        {128, 1, 0, 129, 0, 1, 0, 1, 2, 2, 2, 2, 2, 2, 2, 130},
        {128, 0, 0, 0, 0, 0, 0, 0, 130, 1, 2, 1, 2, 1, 2, 129},
        {128, 0, 2, 2, 1, 129, 2, 2, 0, 0, 2, 2, 1, 1, 2, 130},
        {128, 0, 2, 130, 0, 0, 1, 1, 0, 0, 2, 2, 0, 0, 1, 129},
        {128, 2, 2, 0, 1, 2, 130, 1, 0, 2, 2, 0, 1, 2, 2, 129},
        {128, 1, 0, 1, 2, 2, 130, 2, 2, 2, 2, 2, 0, 1, 0, 129},
        {128, 0, 0, 0, 2, 1, 2, 1, 130, 1, 2, 1, 2, 1, 2, 129},
        {128, 1, 0, 129, 0, 1, 0, 1, 0, 1, 0, 1, 2, 2, 2, 130},
        {128, 2, 2, 130, 0, 1, 1, 1, 0, 2, 2, 2, 0, 1, 1, 129},
        {128, 0, 0, 2, 1, 129, 1, 2, 0, 0, 0, 2, 1, 1, 1, 130},
        {128, 0, 0, 0, 2, 129, 1, 2, 2, 1, 1, 2, 2, 1, 1, 130},
        {128, 2, 2, 2, 0, 129, 1, 1, 0, 1, 1, 1, 0, 2, 2, 130},
        {128, 0, 0, 2, 1, 1, 1, 2, 129, 1, 1, 2, 0, 0, 0, 130},
        {128, 1, 1, 0, 0, 129, 1, 0, 0, 1, 1, 0, 2, 2, 2, 130},
        {128, 0, 0, 0, 0, 0, 0, 0, 2, 1, 129, 2, 2, 1, 1, 130},
        {128, 1, 1, 0, 0, 129, 1, 0, 2, 2, 2, 2, 2, 2, 2, 130},
        {128, 0, 2, 2, 0, 0, 1, 1, 0, 0, 129, 1, 0, 0, 2, 130},
        {128, 0, 2, 2, 1, 1, 2, 2, 129, 1, 2, 2, 0, 0, 2, 130},
        {128, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 2, 129, 1, 130},
        {128, 0, 0, 130, 0, 0, 0, 1, 0, 0, 0, 2, 0, 0, 0, 129},
        {128, 2, 2, 2, 1, 2, 2, 2, 0, 2, 2, 2, 129, 2, 2, 130},
        {128, 1, 0, 129, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 130},
        {128, 1, 1, 129, 2, 0, 1, 1, 130, 2, 0, 1, 2, 2, 2, 0},
    },
};

// block.rs:50-56
void decode_block_bc1(const uint8_t* in, uint8_t* out, size_t pitch) {
  decode_color_block(in, out, pitch, false);
}
// block.rs:59-71
void decode_block_bc2(const uint8_t* in, uint8_t* out, size_t pitch) {
  decode_color_block(in + 8, out, pitch, true);
  decode_sharp_alpha_block(in, out, pitch);
}
// block.rs:74-90
void decode_block_bc3(const uint8_t* in, uint8_t* out, size_t pitch) {
  decode_color_block(in + 8, out, pitch, true);
  decode_smooth_alpha_block(in, out + 3, pitch, 4);
}
// block.rs:93-100
void decode_block_bc4(const uint8_t* in, uint8_t* out, size_t pitch) {
  decode_smooth_alpha_block(in, out, pitch, 1);
}
// block.rs:103-118
void decode_block_bc5(const uint8_t* in, uint8_t* out, size_t pitch) {
  decode_smooth_alpha_block(in, out, pitch, 2);
  decode_smooth_alpha_block(in + 8, out + 1, pitch, 2);
}

// block.rs:262-769.  `pitch` is in f16 elements; RGB written at stride 3.
// `comp`, when non-null, also receives each texel's interpolated value before
// finish_unquantize (texel k channel c at comp[k*3 + c]): the value the
// encoder measures its error on.
void decode_block_bc6h(const uint8_t* in, uint16_t* out, size_t pitch, bool is_signed,
                       int32_t* comp) {
  static const uint8_t ACTUAL_BITS_COUNT[4][14] = {
      {10, 7, 11, 11, 11, 9, 8, 8, 8, 6, 10, 11, 12, 16},
      {5, 6, 5, 4, 4, 5, 6, 5, 5, 6, 10, 9, 8, 4},
      {5, 6, 4, 5, 4, 5, 5, 6, 5, 6, 10, 9, 8, 4},
      {5, 6, 4, 4, 5, 5, 5, 5, 6, 6, 10, 9, 8, 4},
  };
  static const int32_t WEIGHT3[8] = {0, 9, 18, 27, 37, 46, 55, 64};
  static const int32_t WEIGHT4[16] = {0, 4, 9, 13, 17, 21, 26, 30, 34, 38, 43, 47, 51, 55, 60, 64};

  BitStream bs(in);
  int32_t r[4] = {0}, g[4] = {0}, b[4] = {0};
  uint32_t mode = bs.read_bits(2);
  if (mode > 1) mode |= bs.read_bits(3) << 2;
  int32_t partition = 0;

#define B1(c, i, s) c[i] |= bs.read_bit_i32() << (s)
#define BN(c, i, n) c[i] |= bs.read_bits_i32(n)
  switch (mode) {
    case 0b00:
      B1(g, 2, 4); B1(b, 2, 4); B1(b, 3, 4); BN(r, 0, 10); BN(g, 0, 10); BN(b, 0, 10);
      BN(r, 1, 5); B1(g, 3, 4); BN(g, 2, 4); BN(g, 1, 5); B1(b, 3, 0); BN(g, 3, 4);
      BN(b, 1, 5); B1(b, 3, 1); BN(b, 2, 4); BN(r, 2, 5); B1(b, 3, 2); BN(r, 3, 5);
      B1(b, 3, 3); partition = bs.read_bits_i32(5); mode = 0;
      break;
    case 0b01:
      B1(g, 2, 5); B1(g, 3, 4); B1(g, 3, 5); BN(r, 0, 7); B1(b, 3, 0); B1(b, 3, 1);
      B1(b, 2, 4); BN(g, 0, 7); B1(b, 2, 5); B1(b, 3, 2); B1(g, 2, 4); BN(b, 0, 7);
      B1(b, 3, 3); B1(b, 3, 5); B1(b, 3, 4); BN(r, 1, 6); BN(g, 2, 4); BN(g, 1, 6);
      BN(g, 3, 4); BN(b, 1, 6); BN(b, 2, 4); BN(r, 2, 6); BN(r, 3, 6);
      partition = bs.read_bits_i32(5); mode = 1;
      break;
    case 0b00010:
      BN(r, 0, 10); BN(g, 0, 10); BN(b, 0, 10); BN(r, 1, 5); B1(r, 0, 10); BN(g, 2, 4);
      BN(g, 1, 4); B1(g, 0, 10); B1(b, 3, 0); BN(g, 3, 4); BN(b, 1, 4); B1(b, 0, 10);
      B1(b, 3, 1); BN(b, 2, 4); BN(r, 2, 5); B1(b, 3, 2); BN(r, 3, 5); B1(b, 3, 3);
      partition = bs.read_bits_i32(5); mode = 2;
      break;
    case 0b00110:
      BN(r, 0, 10); BN(g, 0, 10); BN(b, 0, 10); BN(r, 1, 4); B1(r, 0, 10); B1(g, 3, 4);
      BN(g, 2, 4); BN(g, 1, 5); B1(g, 0, 10); BN(g, 3, 4); BN(b, 1, 4); B1(b, 0, 10);
      B1(b, 3, 1); BN(b, 2, 4); BN(r, 2, 4); B1(b, 3, 0); B1(b, 3, 2); BN(r, 3, 4);
      B1(g, 2, 4); B1(b, 3, 3); partition = bs.read_bits_i32(5); mode = 3;
      break;
    case 0b01010:
      BN(r, 0, 10); BN(g, 0, 10); BN(b, 0, 10); BN(r, 1, 4); B1(r, 0, 10); B1(b, 2, 4);
      BN(g, 2, 4); BN(g, 1, 4); B1(g, 0, 10); B1(b, 3, 0); BN(g, 3, 4); BN(b, 1, 5);
      B1(b, 0, 10); BN(b, 2, 4); BN(r, 2, 4); B1(b, 3, 1); B1(b, 3, 2); BN(r, 3, 4);
      B1(b, 3, 4); B1(b, 3, 3); partition = bs.read_bits_i32(5); mode = 4;
      break;
    case 0b01110:
      BN(r, 0, 9); B1(b, 2, 4); BN(g, 0, 9); B1(g, 2, 4); BN(b, 0, 9); B1(b, 3, 4);
      BN(r, 1, 5); B1(g, 3, 4); BN(g, 2, 4); BN(g, 1, 5); B1(b, 3, 0); BN(g, 3, 4);
      BN(b, 1, 5); B1(b, 3, 1); BN(b, 2, 4); BN(r, 2, 5); B1(b, 3, 2); BN(r, 3, 5);
      B1(b, 3, 3); partition = bs.read_bits_i32(5); mode = 5;
      break;
    case 0b10010:
      BN(r, 0, 8); B1(g, 3, 4); B1(b, 2, 4); BN(g, 0, 8); B1(b, 3, 2); B1(g, 2, 4);
      BN(b, 0, 8); B1(b, 3, 3); B1(b, 3, 4); BN(r, 1, 6); BN(g, 2, 4); BN(g, 1, 5);
      B1(b, 3, 0); BN(g, 3, 4); BN(b, 1, 5); B1(b, 3, 1); BN(b, 2, 4); BN(r, 2, 6);
      BN(r, 3, 6); partition = bs.read_bits_i32(5); mode = 6;
      break;
    case 0b10110:
      BN(r, 0, 8); B1(b, 3, 0); B1(b, 2, 4); BN(g, 0, 8); B1(g, 2, 5); B1(g, 2, 4);
      BN(b, 0, 8); B1(g, 3, 5); B1(b, 3, 4); BN(r, 1, 5); B1(g, 3, 4); BN(g, 2, 4);
      BN(g, 1, 6); BN(g, 3, 4); BN(b, 1, 5); B1(b, 3, 1); BN(b, 2, 4); BN(r, 2, 5);
      B1(b, 3, 2); BN(r, 3, 5); B1(b, 3, 3); partition = bs.read_bits_i32(5); mode = 7;
      break;
    case 0b11010:
      BN(r, 0, 8); B1(b, 3, 1); B1(b, 2, 4); BN(g, 0, 8); B1(b, 2, 5); B1(g, 2, 4);
      BN(b, 0, 8); B1(b, 3, 5); B1(b, 3, 4); BN(r, 1, 5); B1(g, 3, 4); BN(g, 2, 4);
      BN(g, 1, 5); B1(b, 3, 0); BN(g, 3, 4); BN(b, 1, 6); BN(b, 2, 4); BN(r, 2, 5);
      B1(b, 3, 2); BN(r, 3, 5); B1(b, 3, 3); partition = bs.read_bits_i32(5); mode = 8;
      break;
    case 0b11110:
      BN(r, 0, 6); B1(g, 3, 4); B1(b, 3, 0); B1(b, 3, 1); B1(b, 2, 4); BN(g, 0, 6);
      B1(g, 2, 5); B1(b, 2, 5); B1(b, 3, 2); B1(g, 2, 4); BN(b, 0, 6); B1(g, 3, 5);
      B1(b, 3, 3); B1(b, 3, 5); B1(b, 3, 4); BN(r, 1, 6); BN(g, 2, 4); BN(g, 1, 6);
      BN(g, 3, 4); BN(b, 1, 6); BN(b, 2, 4); BN(r, 2, 6); BN(r, 3, 6);
      partition = bs.read_bits_i32(5); mode = 9;
      break;
    case 0b00011:
      BN(r, 0, 10); BN(g, 0, 10); BN(b, 0, 10); BN(r, 1, 10); BN(g, 1, 10); BN(b, 1, 10);
      mode = 10;
      break;
    case 0b00111:
      BN(r, 0, 10); BN(g, 0, 10); BN(b, 0, 10); BN(r, 1, 9); B1(r, 0, 10); BN(g, 1, 9);
      B1(g, 0, 10); BN(b, 1, 9); B1(b, 0, 10); mode = 11;
      break;
    case 0b01011:
      BN(r, 0, 10); BN(g, 0, 10); BN(b, 0, 10); BN(r, 1, 8);
      r[0] |= bs.read_bits_reversed(2) << 10;
      BN(g, 1, 8);
      g[0] |= bs.read_bits_reversed(2) << 10;
      BN(b, 1, 8);
      b[0] |= bs.read_bits_reversed(2) << 10;
      mode = 12;
      break;
    case 0b01111:
      BN(r, 0, 10); BN(g, 0, 10); BN(b, 0, 10); BN(r, 1, 4);
      r[0] |= bs.read_bits_reversed(6) << 10;
      BN(g, 1, 4);
      g[0] |= bs.read_bits_reversed(6) << 10;
      BN(b, 1, 4);
      b[0] |= bs.read_bits_reversed(6) << 10;
      mode = 13;
      break;
    default:
      // Reserved modes: RGB zero (block.rs:679-691).
      for (int i = 0; i < 4; i++)
        for (int k = 0; k < 12; k++) out[i * pitch + k] = 0;
      return;
  }
#undef B1
#undef BN

  int num_partitions = mode >= 10 ? 0 : 1;
  int32_t bits0 = ACTUAL_BITS_COUNT[0][mode];
  if (is_signed) {
    r[0] = extend_sign(r[0], bits0);
    g[0] = extend_sign(g[0], bits0);
    b[0] = extend_sign(b[0], bits0);
  }
  if ((mode != 9 && mode != 10) || is_signed) {
    for (int i = 1; i < (num_partitions + 1) * 2; i++) {
      r[i] = extend_sign(r[i], ACTUAL_BITS_COUNT[1][mode]);
      g[i] = extend_sign(g[i], ACTUAL_BITS_COUNT[2][mode]);
      b[i] = extend_sign(b[i], ACTUAL_BITS_COUNT[3][mode]);
    }
  }
  if (mode != 9 && mode != 10) {
    for (int i = 1; i < (num_partitions + 1) * 2; i++) {
      r[i] = transform_inverse(r[i], r[0], bits0, is_signed);
      g[i] = transform_inverse(g[i], g[0], bits0, is_signed);
      b[i] = transform_inverse(b[i], b[0], bits0, is_signed);
    }
  }
  for (int i = 0; i < (num_partitions + 1) * 2; i++) {
    r[i] = unquantize(r[i], bits0, is_signed);
    g[i] = unquantize(g[i], bits0, is_signed);
    b[i] = unquantize(b[i], bits0, is_signed);
  }
  const int32_t* weights = mode >= 10 ? WEIGHT4 : WEIGHT3;
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++) {
      int32_t ps = mode >= 10 ? ((i | j) == 0 ? 128 : 0) : BC7_PARTITION_SETS[0][partition][i * 4 + j];
      uint32_t index_bits = mode >= 10 ? 4 : 3;
      if (ps & 0x80) index_bits -= 1;
      ps &= 1;
      int32_t index = bs.read_bits_i32(index_bits);
      int ep = ps * 2;
      uint16_t* o = out + i * pitch + j * 3;
      int32_t v[3] = {interpolate(r[ep], r[ep + 1], weights, index),
                      interpolate(g[ep], g[ep + 1], weights, index),
                      interpolate(b[ep], b[ep + 1], weights, index)};
      for (int c = 0; c < 3; c++) {
        o[c] = finish_unquantize(v[c], is_signed);
        if (comp) comp[(i * 4 + j) * 3 + c] = v[c];
      }
    }
}

// block.rs:774-796
void decode_block_bc6h_float(const uint8_t* in, float* out, size_t pitch, bool is_signed) {
  uint16_t blk[48];
  decode_block_bc6h(in, blk, 12, is_signed);
  for (int i = 0; i < 4; i++) {
    for (int j = 0; j < 4; j++)
      for (int c = 0; c < 3; c++) out[i * pitch + j * 3 + c] = f16_bits_to_f32(blk[i * 12 + j * 3 + c]);
  }
}

// block.rs:801-1252
void decode_block_bc7(const uint8_t* in, uint8_t* out, size_t pitch) {
  static const uint8_t ACTUAL_BITS_COUNT[2][8] = {{4, 6, 5, 7, 5, 7, 7, 5},
                                                  {0, 0, 0, 0, 6, 8, 7, 5}};
  static const int32_t WEIGHT2[4] = {0, 21, 43, 64};
  static const int32_t WEIGHT3[8] = {0, 9, 18, 27, 37, 46, 55, 64};
  static const int32_t WEIGHT4[16] = {0, 4, 9, 13, 17, 21, 26, 30, 34, 38, 43, 47, 51, 55, 60, 64};
  const uint8_t MODE_HAS_P_BITS = 0b11001011;

  BitStream bs(in);
  uint32_t mode = 0;
  while (mode < 8 && bs.read_bits(1) == 0) mode++;
  if (mode >= 8) {
    for (int i = 0; i < 4; i++)
      for (int j = 0; j < 16; j++) out[i * pitch + j] = 0;
    return;
  }
  size_t partition = 0;
  int num_partitions = 1;
  uint32_t rotation = 0;
  uint32_t isb = 0;
  if (mode == 0 || mode == 1 || mode == 2 || mode == 3 || mode == 7) {
    num_partitions = (mode == 0 || mode == 2) ? 3 : 2;
    partition = bs.read_bits(mode == 0 ? 4 : 6);
  }
  int num_endpoints = num_partitions * 2;
  if (mode == 4 || mode == 5) {
    rotation = bs.read_bits(2);
    if (mode == 4) isb = bs.read_bits(1);
  }
  int32_t ep[6][4] = {{0}};
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < num_endpoints; j++) ep[j][i] = (int32_t)bs.read_bits(ACTUAL_BITS_COUNT[0][mode]);
  if (ACTUAL_BITS_COUNT[1][mode] > 0)
    for (int j = 0; j < num_endpoints; j++) ep[j][3] = (int32_t)bs.read_bits(ACTUAL_BITS_COUNT[1][mode]);

  if (mode == 0 || mode == 1 || mode == 3 || mode == 6 || mode == 7) {
    for (int e = 0; e < num_endpoints; e++)
      for (int c = 0; c < 4; c++) ep[e][c] <<= 1;
    if (mode == 1) {
      int32_t i = (int32_t)bs.read_bits(1);
      int32_t j = (int32_t)bs.read_bits(1);
      for (int k = 0; k < 3; k++) {
        ep[0][k] |= i;
        ep[1][k] |= i;
        ep[2][k] |= j;
        ep[3][k] |= j;
      }
    } else if (MODE_HAS_P_BITS & (1 << mode)) {
      for (int e = 0; e < num_endpoints; e++) {
        int32_t j = (int32_t)bs.read_bits(1);
        for (int c = 0; c < 4; c++) ep[e][c] |= j;
      }
    }
  }
  for (int i = 0; i < num_endpoints; i++) {
    int j = ACTUAL_BITS_COUNT[0][mode] + ((MODE_HAS_P_BITS >> mode) & 1);
    for (int k = 0; k < 3; k++) {
      ep[i][k] <<= 8 - j;
      ep[i][k] |= ep[i][k] >> j;
    }
    j = ACTUAL_BITS_COUNT[1][mode] + ((MODE_HAS_P_BITS >> mode) & 1);
    ep[i][3] <<= 8 - j;
    ep[i][3] |= ep[i][3] >> j;
  }
  if (ACTUAL_BITS_COUNT[1][mode] == 0)
    for (int e = 0; e < num_endpoints; e++) ep[e][3] = 0xFF;

  uint32_t index_bits = (mode == 0 || mode == 1) ? 3 : (mode == 6 ? 4 : 2);
  uint32_t index_bits2 = mode == 4 ? 3 : (mode == 5 ? 2 : 0);
  const int32_t* w = index_bits == 2 ? WEIGHT2 : (index_bits == 3 ? WEIGHT3 : WEIGHT4);
  const int32_t* w2 = index_bits2 == 2 ? WEIGHT2 : WEIGHT3;

  int32_t indices[4][4];
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++) {
      uint32_t ps = num_partitions == 1 ? ((i | j) == 0 ? 128 : 0)
                                        : BC7_PARTITION_SETS[num_partitions - 2][partition][i * 4 + j];
      uint32_t ib = index_bits;
      if (ps & 0x80) ib -= 1;
      indices[i][j] = (int32_t)bs.read_bits(ib);
    }
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++) {
      uint32_t ps = num_partitions == 1 ? ((i | j) == 0 ? 128 : 0)
                                        : BC7_PARTITION_SETS[num_partitions - 2][partition][i * 4 + j];
      int s = (int)(ps & 3);
      int32_t index = indices[i][j];
      int32_t c[4];
      if (index_bits2 == 0) {
        for (int k = 0; k < 4; k++) c[k] = interpolate(ep[s * 2][k], ep[s * 2 + 1][k], w, index);
      } else {
        int32_t index2 = (int32_t)bs.read_bits((i | j) == 0 ? index_bits2 - 1 : index_bits2);
        if (isb == 0) {
          for (int k = 0; k < 3; k++) c[k] = interpolate(ep[s * 2][k], ep[s * 2 + 1][k], w, index);
          c[3] = interpolate(ep[s * 2][3], ep[s * 2 + 1][3], w2, index2);
        } else {
          for (int k = 0; k < 3; k++)
            c[k] = interpolate(ep[s * 2][k], ep[s * 2 + 1][k], w2, index2);
          c[3] = interpolate(ep[s * 2][3], ep[s * 2 + 1][3], w, index);
        }
      }
      int32_t t;
      switch (rotation) {
        case 1: t = c[3]; c[3] = c[0]; c[0] = t; break;
        case 2: t = c[3]; c[3] = c[1]; c[1] = t; break;
        case 3: t = c[3]; c[3] = c[2]; c[2] = t; break;
        default: break;
      }
      uint8_t* o = out + i * pitch + j * 4;
      for (int k = 0; k < 4; k++) o[k] = (uint8_t)c[k];
    }
}

}  // namespace bco
