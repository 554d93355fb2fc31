// Exhaustive check that the kernels' division by a reused divisor equals
// IEEE-754 division on every input the kernels can produce.
//
// The kernels replace a / b (correctly rounded f32 division) with
//     y = RN(1 / b)           (one correctly rounded division per divisor)
//     q = RN(a * y);  r = fma(-b, q, a);  result = fma(r, y, q)
// Every step is an IEEE-754 operation (multiply, fused multiply-add, divide)
// with a single rounding, so this host program computes bit for bit what the
// GPU computes.  Each domain below is the full set of (a, b) one call site can
// see; the program counts the pairs where the result differs from a / b.
//
//   int_small   a = integer in [0, 2^24], b = integer in [1, 16]
//               (covariance and mean of masked integer moments, n = texel count)
//   const255    a = every finite f32, b = 255  (endpoint quantisation)
//   const65535  a = every finite f32, b = 65535 (BC6H endpoint quantisation)
//   proj        a = integer with |a| <= 510 * sqrt(b), b = integer in [1, 260100]
//               (index projection: a = sum (texel - e0) * (e1 - e0), b = |e1 - e0|^2,
//               channels <= 4 and |a| <= |texel - e0| |e1 - e0| by Cauchy-Schwarz)
//   channel     a = integer in [-255, 255], b = (float)k + 0.001f, k in [-255, 255]
//               (scalar-channel index projection)
//   small_f     a = every finite f32, b = integer in [1, 16]  (BC6H float moments)
//
// usage: divcheck <domain> [threads] [stride]
//   stride > 1 samples the all-float domains (every stride-th exponent/high
//   mantissa word, all 65536 low words); 1 is the exhaustive sweep.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <omp.h>

static inline float corr(float a, float b, float y) {
  float q = a * y;
  float r = std::fma(-b, q, a);
  return std::fma(r, y, q);
}
static inline uint32_t bits(float x) { uint32_t u; std::memcpy(&u, &x, 4); return u; }
static inline float from_bits(uint32_t u) { float x; std::memcpy(&x, &u, 4); return x; }
static inline bool same(float x, float y) { return bits(x) == bits(y) || (x != x && y != y); }
// -0 / b is -0 but the corrected quotient is +0 (r = fma(-b, -0, -0) = +0).
// Every site that can see a -0 numerator scales the quotient and adds 0.5
// next, where +-0 give the same value, so zeros of either sign are equal here.
static inline bool same_or_zero(float x, float y) { return same(x, y) || (x == 0.0f && y == 0.0f); }

static long long g_stride = 1;  // all-float domains: every g_stride-th high half-word

static long long check_all_floats(float b) {
  const float y = 1.0f / b;
  long long bad = 0;
#pragma omp parallel for reduction(+ : bad) schedule(static)
  for (long long hi = 0; hi < 65536; hi += g_stride) {
    for (uint32_t lo = 0; lo < 65536; lo++) {
      float a = from_bits((uint32_t)(hi << 16) | lo);
      if (!std::isfinite(a)) continue;
      if (!same_or_zero(corr(a, b, y), a / b)) bad++;
    }
  }
  return bad;
}

int main(int argc, char** argv) {
  if (argc < 2) { std::fprintf(stderr, "usage: divcheck <domain> [threads]\n"); return 2; }
  if (argc > 2 && std::atoi(argv[2]) > 0) omp_set_num_threads(std::atoi(argv[2]));
  if (argc > 3) g_stride = std::atoll(argv[3]) > 0 ? std::atoll(argv[3]) : 1;
  const char* d = argv[1];
  long long bad = 0, n = 0;
  if (!std::strcmp(d, "int_small")) {
    for (int b = 1; b <= 16; b++) {
      const float fb = (float)b, y = 1.0f / fb;
#pragma omp parallel for reduction(+ : bad) schedule(static)
      for (int a = 0; a <= (1 << 24); a++)
        if (!same(corr((float)a, fb, y), (float)a / fb)) bad++;
      n += (1 << 24) + 1;
    }
  } else if (!std::strcmp(d, "const255")) {
    bad = check_all_floats(255.0f); n = (65536 + g_stride - 1) / g_stride * 65536;
  } else if (!std::strcmp(d, "const65535")) {
    bad = check_all_floats(65535.0f); n = (65536 + g_stride - 1) / g_stride * 65536;
  } else if (!std::strcmp(d, "small_f")) {
    for (int b = 1; b <= 16; b++) bad += check_all_floats((float)b);
    n = 16 * ((65536 + g_stride - 1) / g_stride * 65536);
  } else if (!std::strcmp(d, "channel")) {
    for (int k = -255; k <= 255; k++) {
      const float fb = (float)k + 0.001f, y = 1.0f / fb;
      for (int a = -255; a <= 255; a++, n++)
        if (!same(corr((float)a, fb, y), (float)a / fb)) bad++;
    }
  } else if (!std::strcmp(d, "proj")) {
    // a >= 0 only: every step is odd in a (RN(-x) = -RN(x)), so -a gives -result.
#pragma omp parallel for reduction(+ : bad, n) schedule(dynamic, 64)
    for (int b = 1; b <= 260100; b++) {
      const float fb = (float)b, y = 1.0f / fb;
      const int amax = (int)std::ceil(510.0 * std::sqrt((double)b));
      long long lb = 0;
      for (int a = 0; a <= amax; a++) {
        float fa = (float)a;
        lb += !same(corr(fa, fb, y), fa / fb);
      }
      bad += lb;
      n += amax + 1;
    }
  } else {
    std::fprintf(stderr, "unknown domain %s\n", d);
    return 2;
  }
  std::printf("%s: %lld pairs, %lld differ\n", d, n, bad);
  return bad != 0;
}
