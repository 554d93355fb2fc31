// Block decode kernels (placeholder until the kernels land).
#include "bcn_internal.hpp"

extern "C" {
int bcn_decode_rgba8(bcn_variant, const void*, size_t, uint32_t, uint32_t, void*, uint32_t, void*) {
  return BCN_ERR_UNSUPPORTED;
}
int bcn_decode_bc6h_rgba16f(const void*, size_t, uint32_t, uint32_t, int, void*, uint32_t, void*) {
  return BCN_ERR_UNSUPPORTED;
}
}
