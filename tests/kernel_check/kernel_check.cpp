// CPU build of the product's per-block kernel code (the same headers the HIP
// kernels include), so the tests can compare the kernel math with the oracle
// without a GPU.  Built by tests/kernel_check/build.sh into
// tests/kernel_check/_build/libkernel_check.so.
#include <cstring>

#include "../../block_compression_amd/csrc/bc7_block.hpp"

extern "C" int kcheck_bc7_encode(const void* settings, const uint8_t* rgba, uint32_t width,
                                 uint32_t height, uint32_t stride, uint8_t* out) {
  bcn::Bc7Params s;
  std::memcpy(&s, settings, sizeof(s));
  uint32_t bw = width / 4, bh = height / 4;
  for (uint32_t idx = 0; idx < bw * bh; idx++) {
    uint32_t yy = idx / bw, xx = idx % bw;
    uint32_t px[16];
    for (int y = 0; y < 4; y++)
      for (int x = 0; x < 4; x++) {
        const uint8_t* p = rgba + (size_t)(yy * 4 + y) * stride + (xx * 4 + x) * 4;
        px[y * 4 + x] = (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) |
                        ((uint32_t)p[3] << 24);
      }
    int keys[64];
    uint32_t o[4];
    if (s.channels == 4) bcn::encode_bc7_block<4>(px, s, keys, 1, o);
    else bcn::encode_bc7_block<3>(px, s, keys, 1, o);
    std::memcpy(out + (size_t)idx * 16, o, 16);
  }
  return 0;
}

#include "../../block_compression_amd/csrc/bc15_block.hpp"

// BC1-BC5 through the same per-block functions as the HIP kernel.
extern "C" int kcheck_bc15_encode(int fmt, const uint8_t* rgba, uint32_t width, uint32_t height,
                                  uint32_t stride, uint8_t* out) {
  uint32_t bw = width / 4, bh = height / 4;
  int words = (fmt == 0 || fmt == 3) ? 2 : 4;
  for (uint32_t idx = 0; idx < bw * bh; idx++) {
    uint32_t yy = idx / bw, xx = idx % bw;
    uint32_t px[16];
    for (int y = 0; y < 4; y++)
      for (int x = 0; x < 4; x++) {
        const uint8_t* p = rgba + (size_t)(yy * 4 + y) * stride + (xx * 4 + x) * 4;
        px[y * 4 + x] = (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) |
                        ((uint32_t)p[3] << 24);
      }
    uint32_t o[4];
    switch (fmt) {
      case 0: bcn::bc1_core(px, o); break;
      case 1: bcn::bc2_alpha(px, o); bcn::bc1_core(px, o + 2); break;
      case 2: bcn::bc3_alpha_block(px, 3, o); bcn::bc1_core(px, o + 2); break;
      case 3: bcn::bc3_alpha_block(px, 0, o); break;
      case 4: bcn::bc3_alpha_block(px, 0, o); bcn::bc3_alpha_block(px, 1, o + 2); break;
      default: return -1;
    }
    std::memcpy(out + (size_t)idx * words * 4, o, words * 4);
  }
  return 0;
}
