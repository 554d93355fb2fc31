// TEST INFRASTRUCTURE ONLY -- the DDS header the reference CLI writes.
//
// The CLI (compressor/src/main.rs:313-353) calls the `ddsfile` crate, version
// "0.5" (Cargo.toml:45), which is not vendored in the reference checkout.
// This restates that crate's published algorithm for a DXGI texture:
// Header::default, Header::new_dxgi, PixelFormat::from(DxgiFormat),
// Header10::new, and the field order of Dds::write / Header::write /
// PixelFormat::write / Header10::write (all little-endian u32), modelling the
// crate's Option<u32> fields as (present, value) pairs the way it does.
#include <cstddef>
#include <cstdint>
#include <vector>

namespace {

struct OptU32 {
  bool some = false;
  uint32_t v = 0;
};

// DxgiFormat::get_block_size / get_bits_per_pixel for the BC formats.
bool bc_info(uint32_t f, uint32_t* block_size, uint32_t* bpp) {
  struct Row {
    uint32_t lo, hi, bs, bpp;
  };
  static const Row rows[] = {{70, 72, 8, 4},   {73, 75, 16, 8}, {76, 78, 16, 8}, {79, 81, 8, 4},
                             {82, 84, 16, 8},  {94, 96, 16, 8}, {97, 99, 16, 8}};
  for (const Row& r : rows)
    if (f >= r.lo && f <= r.hi) {
      *block_size = r.bs;
      *bpp = r.bpp;
      return true;
    }
  return false;
}

struct PixelFormat {
  uint32_t size = 32;
  uint32_t flags = 0;
  OptU32 fourcc, rgb_bit_count, r_mask, g_mask, b_mask, a_mask;
};

struct Header {
  uint32_t size = 124;
  uint32_t flags = 0x1 | 0x2 | 0x4 | 0x1000;  // CAPS | HEIGHT | WIDTH | PIXELFORMAT
  uint32_t height = 0, width = 0;
  OptU32 pitch, linear_size, depth, mip_map_count;
  uint32_t reserved1[11] = {};
  PixelFormat spf;
  uint32_t caps = 0x1000;  // TEXTURE
  uint32_t caps2 = 0, caps3 = 0, caps4 = 0, reserved2 = 0;
};

struct Header10 {
  uint32_t dxgi_format = 0, resource_dimension = 0, misc_flag = 0, array_size = 0, misc_flags2 = 0;
};

void put(std::vector<uint8_t>& out, uint32_t v) {
  for (int i = 0; i < 4; i++) out.push_back((uint8_t)(v >> (8 * i)));
}
uint32_t or0(const OptU32& o) { return o.some ? o.v : 0; }

}  // namespace

extern "C" {

// Writes the 148 header bytes for new_dxgi(height, width, depth None,
// format, mipmap_levels Some(mips), array_layers Some(layers), caps2 None,
// is_cubemap false, Texture2D, alpha_mode).  Returns 0, or -1 when ddsfile
// would return UnsupportedFormat (no pitch for the format).
int oracle_dds_header(uint32_t format, uint32_t width, uint32_t height, uint32_t mips,
                      uint32_t layers, uint32_t alpha_mode, uint8_t* out148) {
  uint32_t bs, bpp;
  if (!bc_info(format, &bs, &bpp)) return -1;
  Header h;
  h.height = height;
  h.width = width;
  h.mip_map_count = {true, mips};
  // PixelFormat::from(DxgiFormat): bits per pixel, fourCC DX10, FOURCC flag.
  h.spf.rgb_bit_count = {true, bpp};
  h.spf.fourcc = {true, 0x30315844u};
  h.spf.flags |= 0x4;
  if (mips > 1) {
    h.flags |= 0x20000;            // MIPMAPCOUNT
    h.caps |= 0x8 | 0x400000;      // COMPLEX | MIPMAP
  }
  if (layers > 1) h.caps |= 0x8;   // COMPLEX
  // Compressed: LINEARSIZE = pitch * ceil(height / pitch_height) * depth(1).
  uint32_t pitch = ((width + 3) / 4 > 1 ? (width + 3) / 4 : 1) * bs;
  h.flags |= 0x80000;
  h.linear_size = {true, pitch * ((height + 3) / 4) * 1};
  Header10 x;
  x.dxgi_format = format;
  x.resource_dimension = 3;  // Texture2D
  x.misc_flag = 0;           // not a cubemap
  x.array_size = layers;
  x.misc_flags2 = alpha_mode;

  std::vector<uint8_t> o;
  put(o, 0x20534444u);  // "DDS "
  put(o, h.size);
  put(o, h.flags);
  put(o, h.height);
  put(o, h.width);
  put(o, h.pitch.some ? h.pitch.v : or0(h.linear_size));
  put(o, or0(h.depth));
  put(o, or0(h.mip_map_count));
  for (uint32_t r : h.reserved1) put(o, r);
  put(o, h.spf.size);
  put(o, h.spf.flags);
  put(o, or0(h.spf.fourcc));
  put(o, or0(h.spf.rgb_bit_count));
  put(o, or0(h.spf.r_mask));
  put(o, or0(h.spf.g_mask));
  put(o, or0(h.spf.b_mask));
  put(o, or0(h.spf.a_mask));
  put(o, h.caps);
  put(o, h.caps2);
  put(o, h.caps3);
  put(o, h.caps4);
  put(o, h.reserved2);
  put(o, x.dxgi_format);
  put(o, x.resource_dimension);
  put(o, x.misc_flag);
  put(o, x.array_size);
  put(o, x.misc_flags2);
  for (size_t i = 0; i < o.size(); i++) out148[i] = o[i];
  return (int)o.size() == 148 ? 0 : -2;
}

}  // extern "C"
