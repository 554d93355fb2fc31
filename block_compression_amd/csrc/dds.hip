// DDS container for block-compressed textures (host code only).
//
// The reference CLI stores its block buffer with the `ddsfile` crate (0.5):
// Dds::new_dxgi + Dds::write (compressor/src/main.rs:313-353), the DXGI
// format chosen per variant (main.rs:355-365).  The file is
//
//   "DDS " magic | DDS_HEADER (124 bytes) | DDS_HEADER_DXT10 (20 bytes) | data
//
// with the DX10 extension always present.  Header fields as ddsfile fills
// them for a block-compressed DXGI format:
//   flags   CAPS | HEIGHT | WIDTH | PIXELFORMAT | LINEARSIZE (+ MIPMAPCOUNT
//           when mip_levels > 1)
//   pitchOrLinearSize  max(1, ceil(w/4)) * block_bytes * ceil(h/4)  (level 0)
//   depth 0, mipMapCount = mip_levels
//   pixel format: size 32, flags FOURCC, fourCC "DX10", RGBBitCount = the
//           format's bits per texel (4 for BC1/BC4, 8 otherwise), masks 0
//   caps    TEXTURE (+ COMPLEX | MIPMAP with mips, + COMPLEX with layers)
//   DXT10   dxgiFormat, resourceDimension TEXTURE2D (3), miscFlag 0,
//           arraySize, miscFlags2 = alpha mode (STRAIGHT = 1 in the CLI)
#include <cstring>

#include "bcn_internal.hpp"

namespace bcn {
namespace {

constexpr uint32_t kDdsMagic = 0x20534444u;  // "DDS "
constexpr uint32_t kFourCcDx10 = 0x30315844u;  // "DX10"

// DDS_HEADER flags / caps (DDS file format documentation).
constexpr uint32_t kFlagCaps = 0x1, kFlagHeight = 0x2, kFlagWidth = 0x4, kFlagPixelFormat = 0x1000,
                   kFlagMipMapCount = 0x20000, kFlagLinearSize = 0x80000;
constexpr uint32_t kCapsComplex = 0x8, kCapsTexture = 0x1000, kCapsMipMap = 0x400000;
constexpr uint32_t kPixelFourCc = 0x4;
constexpr uint32_t kDimTexture2D = 3;

// Bytes per 4x4 block of the block-compressed DXGI formats (BC1..BC7,
// typeless / unorm / srgb / snorm / uf16 / sf16), 0 for any other format.
uint32_t dxgi_block_bytes(uint32_t f) {
  if (f >= 70 && f <= 72) return 8;    // BC1
  if (f >= 73 && f <= 78) return 16;   // BC2, BC3
  if (f >= 79 && f <= 81) return 8;    // BC4
  if (f >= 82 && f <= 84) return 16;   // BC5
  if (f >= 94 && f <= 99) return 16;   // BC6H, BC7
  return 0;
}

// Blocks along a dimension of n texels (at least one).
uint32_t blocks_across(uint32_t n) { return n ? (n + 3) / 4 : 1u; }

void put32(uint8_t* p, uint32_t v) {
  p[0] = (uint8_t)v;
  p[1] = (uint8_t)(v >> 8);
  p[2] = (uint8_t)(v >> 16);
  p[3] = (uint8_t)(v >> 24);
}

}  // namespace
}  // namespace bcn

using namespace bcn;

extern "C" {

uint32_t bcn_dxgi_format(bcn_variant variant) {
  switch (variant) {
    case BCN_BC1: return 72;   // BC1_UNORM_SRGB
    case BCN_BC2: return 75;   // BC2_UNORM_SRGB
    case BCN_BC3: return 78;   // BC3_UNORM_SRGB
    case BCN_BC4: return 80;   // BC4_UNORM
    case BCN_BC5: return 83;   // BC5_UNORM
    case BCN_BC6H: return 95;  // BC6H_UF16
    case BCN_BC6H_SF16: return 96;  // BC6H_SF16 (the signed extension)
    case BCN_BC7: return 99;   // BC7_UNORM_SRGB
    default: return 0;
  }
}

size_t bcn_dds_data_size(uint32_t dxgi_format, uint32_t width, uint32_t height,
                         uint32_t mip_levels, uint32_t array_layers) {
  uint32_t bb = dxgi_block_bytes(dxgi_format);
  if (!bb || !width || !height || !mip_levels || !array_layers || mip_levels > 32) return 0;
  size_t per_layer = 0;
  for (uint32_t l = 0; l < mip_levels; l++) {
    uint32_t w = width >> l ? width >> l : 1, h = height >> l ? height >> l : 1;
    per_layer += (size_t)blocks_across(w) * blocks_across(h) * bb;
  }
  return per_layer * array_layers;
}

int bcn_dds_write_header(uint32_t dxgi_format, uint32_t width, uint32_t height,
                         uint32_t mip_levels, uint32_t array_layers, uint32_t alpha_mode,
                         void* out, size_t out_size) {
  uint32_t bb = dxgi_block_bytes(dxgi_format);
  if (!bb) return BCN_ERR_UNSUPPORTED;
  if (!out || !width || !height || !mip_levels || !array_layers || mip_levels > 32 ||
      alpha_mode > 4)
    return BCN_ERR_INVALID_ARGUMENT;
  if (out_size < BCN_DDS_HEADER_SIZE) return BCN_ERR_BUFFER_TOO_SMALL;
  uint64_t linear = (uint64_t)blocks_across(width) * bb * ((height + 3) / 4);
  if (linear > 0xFFFFFFFFull) return BCN_ERR_INVALID_ARGUMENT;  // does not fit the u32 field

  uint32_t flags = kFlagCaps | kFlagHeight | kFlagWidth | kFlagPixelFormat | kFlagLinearSize;
  uint32_t caps = kCapsTexture;
  if (mip_levels > 1) {
    flags |= kFlagMipMapCount;
    caps |= kCapsComplex | kCapsMipMap;
  }
  if (array_layers > 1) caps |= kCapsComplex;

  uint32_t w[37];
  std::memset(w, 0, sizeof(w));
  w[0] = kDdsMagic;
  // DDS_HEADER, words 1..31
  w[1] = 124;
  w[2] = flags;
  w[3] = height;
  w[4] = width;
  w[5] = (uint32_t)linear;
  w[6] = 0;  // depth
  w[7] = mip_levels;
  // w[8..18] reserved
  w[19] = 32;  // DDS_PIXELFORMAT size
  w[20] = kPixelFourCc;
  w[21] = kFourCcDx10;
  w[22] = bb == 8 ? 4 : 8;  // bits per texel
  // w[23..26] channel masks
  w[27] = caps;
  // w[28..31] caps2, caps3, caps4, reserved
  // DDS_HEADER_DXT10, words 32..36
  w[32] = dxgi_format;
  w[33] = kDimTexture2D;
  w[34] = 0;
  w[35] = array_layers;
  w[36] = alpha_mode;
  uint8_t* p = static_cast<uint8_t*>(out);
  for (int i = 0; i < 37; i++) put32(p + 4 * i, w[i]);
  return BCN_OK;
}

}  // extern "C"
