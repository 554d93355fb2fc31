// The C ABI (include/bcn_mi355x.h): a task queue mirroring GpuBlockCompressor
// (src/block_compressor.rs) on top of the per-format kernel launchers.
#include <stdio.h>
#include <string.h>

#include <new>
#include <string>
#include <vector>

#include "bcn_common.hpp"
#include "bcn_internal.hpp"
#include "texel_load.hpp"

namespace {

int fail(bcn_compressor* c, int status, const char* what) {
  if (c) c->last_error = what;
  return status;
}

uint32_t block_bytes(bcn_variant v) { return (v == BCN_BC1 || v == BCN_BC4) ? 8u : 16u; }

bool valid_variant(int v) { return v >= BCN_BC1 && v <= BCN_BC6H_SF16; }

bool is_bc6h(int v) { return v == BCN_BC6H || v == BCN_BC6H_SF16; }

}  // namespace

hipError_t bcn::launch_task(const bcn::Task& t, hipStream_t stream) {
  switch (t.variant) {
    case BCN_BC7: return bcn::launch_bc7(t, stream);
    case BCN_BC6H:
    case BCN_BC6H_SF16: return bcn::launch_bc6h(t, stream);
    default: return bcn::launch_bc15(t, stream);
  }
}

extern "C" {

const char* bcn_status_string(int status) {
  switch (status) {
    case BCN_OK: return "ok";
    case BCN_ERR_INVALID_ARGUMENT: return "invalid argument";
    case BCN_ERR_NOT_MULTIPLE_OF_4: return "width, height or texture_y_offset not a multiple of 4";
    case BCN_ERR_BUFFER_TOO_SMALL: return "buffer too small";
    case BCN_ERR_TEXTURE_TOO_SMALL: return "texture too small";
    case BCN_ERR_UNSUPPORTED: return "unsupported";
    case BCN_ERR_MISALIGNED: return "misaligned pointer or pitch";
    case BCN_ERR_HIP: return "HIP error";
    case BCN_ERR_NO_DEVICE: return "no such HIP device";
    case BCN_ERR_TOO_LARGE: return "input larger than the allocation-free entry point's scratch";
    default: return "unknown status";
  }
}

size_t bcn_blocks_byte_size(bcn_variant variant, uint32_t width, uint32_t height) {
  if (!valid_variant(variant)) return 0;
  return (size_t)((width + 3) / 4) * ((height + 3) / 4) * block_bytes(variant);
}

int bcn_compressor_create(int device, bcn_compressor** out) {
  if (!out) return BCN_ERR_INVALID_ARGUMENT;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) return BCN_ERR_NO_DEVICE;
  bcn_compressor* c = new (std::nothrow) bcn_compressor();
  if (!c) return BCN_ERR_INVALID_ARGUMENT;
  c->device = device;
  *out = c;
  return BCN_OK;
}

void bcn_compressor_destroy(bcn_compressor* c) { delete c; }

size_t bcn_task_count(const bcn_compressor* c) { return c ? c->tasks.size() : 0; }

const char* bcn_last_error(const bcn_compressor* c) { return c ? c->last_error.c_str() : ""; }

int bcn_add_compression_task(bcn_compressor* c, bcn_variant variant, const void* settings,
                             size_t settings_size, const void* texture,
                             bcn_source_format texture_format, uint32_t texture_pitch,
                             uint32_t texture_height, uint32_t width, uint32_t height,
                             void* buffer, size_t buffer_size, uint32_t texture_y_offset,
                             uint32_t blocks_offset) {
  if (!c) return BCN_ERR_INVALID_ARGUMENT;
  if (!valid_variant(variant)) return fail(c, BCN_ERR_INVALID_ARGUMENT, "unknown variant");
  // block_compressor.rs:352-357
  if (height % 4 != 0 || width % 4 != 0)
    return fail(c, BCN_ERR_NOT_MULTIPLE_OF_4, "width and height must be multiples of 4");
  if (texture_y_offset % 4 != 0)
    return fail(c, BCN_ERR_NOT_MULTIPLE_OF_4, "texture_y_offset must be a multiple of 4");
  const uint32_t texel_bytes = bcn::source_texel_bytes(texture_format);
  if (texel_bytes == 0) return fail(c, BCN_ERR_INVALID_ARGUMENT, "unknown texture format");
  const bool float_src = texture_format == BCN_SRC_RGBA16F || texture_format == BCN_SRC_RGBA32F;
  if (float_src && !is_bc6h(variant))
    return fail(c, BCN_ERR_UNSUPPORTED, "a float source is only accepted for BC6H");
  // BC6H from an 8-bit layout reads it as the replaced GPU path reads a unorm
  // view (block_compressor.rs:312, bc6h.wgsl:64-85): c/255 rounded to a half.
  // (The reference's CPU compress_rgba8 route into BC6H decodes sRGB with
  // powf instead, encode.rs:130-137, bc6h.rs:20-27; it is not this path.)
  bcn::Task t{};
  t.variant = variant;
  if (variant == BCN_BC7) {
    if (!settings || settings_size != sizeof(bcn_bc7_settings))
      return fail(c, BCN_ERR_INVALID_ARGUMENT, "BC7 needs a bcn_bc7_settings");
    bcn_bc7_settings s;
    memcpy(&s, settings, sizeof(s));
    if (s.channels != 3 && s.channels != 4)
      return fail(c, BCN_ERR_INVALID_ARGUMENT, "BC7 settings: channels must be 3 or 4");
    if (s.fast_skip_threshold_mode1 > 64 || s.fast_skip_threshold_mode3 > 64 ||
        s.fast_skip_threshold_mode7 > 64)
      return fail(c, BCN_ERR_INVALID_ARGUMENT, "BC7 settings: partition thresholds exceed 64");
    memcpy(t.settings, &s, sizeof(s));
  } else if (is_bc6h(variant)) {
    if (!settings || settings_size != sizeof(bcn_bc6h_settings))
      return fail(c, BCN_ERR_INVALID_ARGUMENT, "BC6H needs a bcn_bc6h_settings");
    bcn_bc6h_settings s;
    memcpy(&s, settings, sizeof(s));
    if (s.fast_skip_threshold > 32)
      return fail(c, BCN_ERR_INVALID_ARGUMENT, "BC6H settings: fast_skip_threshold exceeds 32");
    memcpy(t.settings, &s, sizeof(s));
  } else if (settings && settings_size != 0) {
    return fail(c, BCN_ERR_INVALID_ARGUMENT, "only BC6H and BC7 take settings");
  }
  if (width == 0 || height == 0) {
    // Nothing to encode; the reference dispatches an empty grid.
  } else if (!texture || !buffer) {
    return fail(c, BCN_ERR_INVALID_ARGUMENT, "null texture or buffer");
  }
  if ((uint64_t)texture_pitch < (uint64_t)width * texel_bytes)
    return fail(c, BCN_ERR_INVALID_ARGUMENT, "texture_pitch smaller than a row of the region");
  if ((uint64_t)texture_y_offset + height > texture_height)
    return fail(c, BCN_ERR_TEXTURE_TOO_SMALL, "texture_y_offset + height exceeds texture_height");
  // block_compressor.rs:364-373
  size_t required = bcn_blocks_byte_size(variant, width, height);
  if ((uint64_t)blocks_offset + required > buffer_size)
    return fail(c, BCN_ERR_BUFFER_TOO_SMALL, "buffer too small for the blocks at blocks_offset");
  if ((((uintptr_t)texture | (uintptr_t)buffer | texture_pitch) & 3u) != 0)
    return fail(c, BCN_ERR_MISALIGNED, "texture, buffer and texture_pitch must be 4-byte aligned");
  // Float texels are read whole (8 or 16 bytes at a time).
  if (float_src && (((uintptr_t)texture | texture_pitch) & (texel_bytes - 1)) != 0)
    return fail(c, BCN_ERR_MISALIGNED, "float texture and texture_pitch must be aligned to a texel");
  t.texture = static_cast<const uint8_t*>(texture) + (size_t)texture_y_offset * texture_pitch;
  t.format = texture_format;
  t.pitch = texture_pitch;
  t.width = width;
  t.height = height;
  // The reference hands blocks_offset / 4 to the shader as a word index
  // (block_compressor.rs:449).
  t.blocks = static_cast<uint8_t*>(buffer) + (size_t)(blocks_offset / 4) * 4;
  c->tasks.push_back(t);
  return BCN_OK;
}

int bcn_compress(bcn_compressor* c, void* hip_stream) {
  if (!c) return BCN_ERR_INVALID_ARGUMENT;
  std::vector<bcn::Task> tasks;
  tasks.swap(c->tasks);  // the queue is drained whatever happens (block_compressor.rs:550)
  bcn::DeviceGuard guard(c->device);
  if (!guard.ok) return fail(c, BCN_ERR_HIP, "hipSetDevice failed");
  hipStream_t stream = static_cast<hipStream_t>(hip_stream);
  for (const bcn::Task& t : tasks) {
    hipError_t e = bcn::launch_task(t, stream);
    if (e == hipErrorNotSupported) return fail(c, BCN_ERR_UNSUPPORTED, "format not built in this library");
    if (e != hipSuccess) {
      char buf[256];
      snprintf(buf, sizeof(buf), "kernel launch failed: %s", hipGetErrorString(e));
      return fail(c, BCN_ERR_HIP, buf);
    }
  }
  c->last_error.clear();
  return BCN_OK;
}

void bcn_partition_tables(uint32_t* pattern, uint32_t* mask01, uint8_t* anchor1,
                          uint8_t* anchor2) {
  const bcn::PartitionTables t = bcn::make_partition_tables();
  for (int i = 0; i < 128; i++) {
    if (pattern) pattern[i] = t.subset[i];
    if (mask01) mask01[i] = t.mask01[i];
    if (anchor1) anchor1[i] = t.anchor1[i];
    if (anchor2) anchor2[i] = t.anchor2[i];
  }
}

}  // extern "C"
