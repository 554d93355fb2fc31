/*
 * bcn_mi355x -- BC1-BC7 block compression on AMD MI355X (gfx950).
 *
 * C ABI of the GPU block compressor.  It replaces the reference crate's
 * `GpuBlockCompressor` (block_compression v0.3.0, src/block_compressor.rs)
 * and its WGSL kernels (src/shader/ bc1_to_5, bc6h, bc7 .wgsl): the same task API -- create,
 * add tasks, compress -- over the same formats and encoder presets, with
 * device pointers and a HIP stream where the reference takes wgpu textures,
 * buffers and a compute pass.  Every `assert!` of the reference becomes a
 * status code.  The encoded bytes equal the crate's CPU encoder output
 * (src/encode.rs) for the same input.
 *
 * Threading: one compressor per host thread (the reference's methods take
 * `&mut self`).  Work is asynchronous on the stream given to bcn_compress.
 */
#ifndef BCN_MI355X_H
#define BCN_MI355X_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* CompressionVariant (src/lib.rs:74-103), in declaration order. */
typedef enum bcn_variant {
  BCN_BC1 = 0,  /* RGB, 8 bytes per block */
  BCN_BC2 = 1,  /* RGBA with 4-bit alpha, 16 bytes */
  BCN_BC3 = 2,  /* RGBA with interpolated alpha, 16 bytes */
  BCN_BC4 = 3,  /* R, 8 bytes */
  BCN_BC5 = 4,  /* RG, 16 bytes */
  BCN_BC6H = 5, /* RGB half float (unsigned), 16 bytes; settings: bcn_bc6h_settings */
  BCN_BC7 = 6,  /* RGB(A), 16 bytes; settings: bcn_bc7_settings */
  /* RGB half float, signed (BC6H_SF16), 16 bytes; settings: bcn_bc6h_settings.
   * No counterpart in the reference's encoder (unsigned only); its decoder
   * reads these blocks (src/decode/block.rs:266, 697-719).  The unsigned
   * search over signed values: a texel +-(|h|/31)*32, endpoint magnitudes
   * quantised to 2^(b-1)-1 levels over 0..32767 and dequantised as the
   * decoder's signed unquantize, two's-complement endpoint fields. */
  BCN_BC6H_SF16 = 7
} bcn_variant;

/* Source texel layouts: the texture views the reference's shaders read with
 * textureLoad (src/block_compressor.rs:300-313).  The encoders work on the
 * raw texel values: an sRGB texture is passed as its bytes (the reference's
 * non-sRGB view of it).  A channel a layout lacks reads as textureLoad gives
 * it: G = B = 0, A = 255.  BC1-BC5 and BC7 take the 8-bit layouts.  BC6H
 * takes every layout, as the replaced shader reads them (bc6h.wgsl:59-85):
 * RGBA16F half bits as they are except -0, which reads as +0; RGBA32F and
 * the 8-bit unorm layouts (c/255 in f32) rounded to the nearest half, ties
 * to even, every zero as +0. */
typedef enum bcn_source_format {
  BCN_SRC_RGBA8 = 0,    /* Rgba8Unorm: 4 x u8 per texel, read as exact integers */
  BCN_SRC_RGBA16F = 1,  /* Rgba16Float: 4 x IEEE binary16 per texel (BC6H only) */
  BCN_SRC_R8 = 2,       /* R8Unorm: 1 byte per texel -> (r, 0, 0, 255) */
  BCN_SRC_RG8 = 3,      /* Rg8Unorm: 2 bytes per texel -> (r, g, 0, 255) */
  BCN_SRC_BGRA8 = 4,    /* Bgra8Unorm: 4 bytes per texel, B first */
  BCN_SRC_RGBA32F = 5   /* Rgba32Float: 4 x f32 per texel (BC6H only); each channel is
                           rounded to the nearest half (ties to even) and every
                           zero reads as +0, as bc6h.wgsl:59-70 converts it */
} bcn_source_format;

typedef enum bcn_status {
  BCN_OK = 0,
  BCN_ERR_INVALID_ARGUMENT = 1,     /* null handle/pointer, bad enum, bad settings size */
  BCN_ERR_NOT_MULTIPLE_OF_4 = 2,    /* width, height or texture_y_offset (block_compressor.rs:352-357) */
  BCN_ERR_BUFFER_TOO_SMALL = 3,     /* block_compressor.rs:364-373 */
  BCN_ERR_TEXTURE_TOO_SMALL = 4,    /* rows y_offset..y_offset+height exceed texture_height */
  BCN_ERR_UNSUPPORTED = 5,          /* e.g. RGBA16F source for a non-BC6H variant */
  BCN_ERR_MISALIGNED = 6,           /* source/destination not 4-byte aligned */
  BCN_ERR_HIP = 7,                  /* a HIP runtime call failed; see bcn_last_error */
  BCN_ERR_NO_DEVICE = 8,
  BCN_ERR_TOO_LARGE = 9             /* input beyond the fixed scratch of an allocation-free entry point */
} bcn_status;

/* BC6HSettings, 5 x u32 repr(C) (src/settings.rs:9-15). */
typedef struct bcn_bc6h_settings {
  uint32_t slow_mode;
  uint32_t fast_mode;
  uint32_t refine_iterations_1p;
  uint32_t refine_iterations_2p;
  uint32_t fast_skip_threshold;
} bcn_bc6h_settings;

/* BC7Settings, 19 x u32 repr(C) (src/settings.rs:80-90). */
typedef struct bcn_bc7_settings {
  uint32_t refine_iterations[8];
  uint32_t mode_selection[4];
  uint32_t skip_mode2;
  uint32_t fast_skip_threshold_mode1;
  uint32_t fast_skip_threshold_mode3;
  uint32_t fast_skip_threshold_mode7;
  uint32_t mode45_channel0;
  uint32_t refine_iterations_channel;
  uint32_t channels;
} bcn_bc7_settings;

typedef struct bcn_compressor bcn_compressor;

/* GpuBlockCompressor::new (src/block_compressor.rs:66).  `device` is a HIP
 * device ordinal. */
int bcn_compressor_create(int device, bcn_compressor** out);

/* Drop of GpuBlockCompressor. */
void bcn_compressor_destroy(bcn_compressor* c);

/* GpuBlockCompressor::add_compression_task (src/block_compressor.rs:342-387).
 *
 * texture:          device pointer to texel (0, 0) of the source texture.
 * texture_pitch:    bytes per texture row.
 * texture_height:   rows in the texture (bounds check for the y offset).
 * width, height:    size of the region to compress; multiples of 4.
 * buffer:           device pointer to the destination buffer.
 * buffer_size:      its size in bytes.
 * texture_y_offset: first texture row of the region; a multiple of 4.  It is
 *                   honoured for every format (the reference's WGSL ignores
 *                   it for BC1/BC3/BC4/BC5, bc1_to_5.wgsl:44-72).
 * blocks_offset:    byte offset of the first block in `buffer`; like the
 *                   reference (block_compressor.rs:449) it is used in whole
 *                   32-bit words, i.e. rounded down to a multiple of 4.
 * settings:         bcn_bc6h_settings for BC6H, bcn_bc7_settings for BC7,
 *                   otherwise NULL; settings_size is its sizeof.
 *
 * Blocks are written row-major, block (x, y) at
 * blocks_offset + (y * ceil(width/4) + x) * block_bytes. */
int bcn_add_compression_task(bcn_compressor* c, bcn_variant variant, const void* settings,
                             size_t settings_size, const void* texture,
                             bcn_source_format texture_format, uint32_t texture_pitch,
                             uint32_t texture_height, uint32_t width, uint32_t height,
                             void* buffer, size_t buffer_size, uint32_t texture_y_offset,
                             uint32_t blocks_offset);

/* GpuBlockCompressor::compress (src/block_compressor.rs:540-590): enqueue
 * every queued task on `hip_stream` (a hipStream_t; NULL = default stream)
 * and empty the queue.  Asynchronous: synchronise the stream before reading
 * the buffers. */
int bcn_compress(bcn_compressor* c, void* hip_stream);

/* Number of queued tasks. */
size_t bcn_task_count(const bcn_compressor* c);

/* Text of the last failure on this compressor (empty if none). */
const char* bcn_last_error(const bcn_compressor* c);

/* Name of a status code. */
const char* bcn_status_string(int status);

/* CompressionVariant::blocks_byte_size (src/lib.rs:138-144). */
size_t bcn_blocks_byte_size(bcn_variant variant, uint32_t width, uint32_t height);

/* GPU decode of a block image into RGBA8, the layout of
 * decompress_blocks_as_rgba8 (src/decode.rs:336-387): BC4 -> (R,0,0,0),
 * BC5 -> (R,G,0,0), BC6H -> (sRGB8(R), sRGB8(G), sRGB8(B), 0) with
 * linear_to_srgb (decode.rs:151-181) reproduced exactly (BC6H_SF16: the
 * same of the signed decode, negative values giving 0).
 * blocks: device pointer, (w/4)*(h/4) blocks, exactly blocks_byte_size bytes
 * (decode.rs:343-355); rgba: device pointer, rgba_pitch bytes per row, both
 * 16-byte aligned; width and height multiples of 4. */
int bcn_decode_rgba8(bcn_variant variant, const void* blocks, size_t blocks_size, uint32_t width,
                     uint32_t height, void* rgba, uint32_t rgba_pitch, void* hip_stream);

/* Flag of bcn_decode_rgba16f / _rgba32f: write the reference's layout.  The
 * reference's image-level 16F/32F drivers place each block at its RGBA
 * offset but write its texels as packed RGB (texel j at element j*3,
 * src/decode/block.rs:753 and :787-792, via src/decode.rs:271-279 and
 * :317-325): per block row, elements 0..11 hold RGB and 12..15 are left
 * untouched.  Without the flag the output is a proper RGBA image with
 * alpha 1.0. */
#define BCN_DECODE_REFERENCE_LAYOUT 1u

/* decompress_blocks_as_rgba16f (src/decode.rs:397-430) on the GPU: BC6H
 * (unsigned) and BC6H_SF16 (signed) only, any other variant is
 * BCN_ERR_UNSUPPORTED (the reference panics).  rgba_pitch in bytes,
 * >= width*8. */
int bcn_decode_rgba16f(bcn_variant variant, const void* blocks, size_t blocks_size, uint32_t width,
                       uint32_t height, void* rgba, uint32_t rgba_pitch, uint32_t flags,
                       void* hip_stream);

/* decompress_blocks_as_rgba32f (src/decode.rs:438-468): as above with f32
 * texels (exact conversions of the half results); rgba_pitch >= width*16. */
int bcn_decode_rgba32f(bcn_variant variant, const void* blocks, size_t blocks_size, uint32_t width,
                       uint32_t height, void* rgba, uint32_t rgba_pitch, uint32_t flags,
                       void* hip_stream);

/* GPU decode of BC6H blocks into RGBA16F (decode_block_bc6h, src/decode/block.rs:262-769,
 * unsigned unless is_signed) with alpha written as 1.0 (0x3C00).  rgba_pitch in bytes. */
int bcn_decode_bc6h_rgba16f(const void* blocks, size_t blocks_size, uint32_t width,
                            uint32_t height, int is_signed, void* rgba, uint32_t rgba_pitch,
                            void* hip_stream);

/* The 256 thresholds the BC6H -> RGBA8 decode uses for linear_to_srgb
 * (src/decode.rs:151-159): out[k] is the smallest half bit pattern in
 * 0..0x7BFF whose sRGB8 code is >= k (out[0] = 0; 0xFFFF if none).  For
 * table checks without a GPU. */
void bcn_bc6h_srgb8_thresholds(uint16_t* out);

/* The 3,968 bucket entries the BC6H -> RGBA8 decode reads per channel, built
 * from those thresholds: for a half bit pattern h <= 0x7BFF, the high byte of
 * out[h >> 3] + 32 * (h & 7) is its sRGB8 code.  Returns BCN_OK, or
 * BCN_ERR_UNSUPPORTED if the thresholds rise twice within 8 patterns (the
 * entries cannot express that; the decode then fails the same way).  For
 * table checks without a GPU. */
int bcn_bc6h_srgb8_buckets(uint16_t* out);

/* ---- DDS container ------------------------------------------------------------------
 *
 * The reference CLI writes its blocks as a DDS file through the `ddsfile`
 * crate 0.5 (Dds::new_dxgi + Dds::write, compressor/src/main.rs:313-353):
 * "DDS " magic, DDS_HEADER and the DX10 extension header, then the block
 * data in the buffer's order (mip levels of layer 0, then of layer 1, ...).
 * Host-only functions; no GPU needed. */

/* Bytes of the magic and both headers. */
#define BCN_DDS_HEADER_SIZE 148u

/* DDS_ALPHA_MODE values (miscFlags2); the CLI writes STRAIGHT. */
typedef enum bcn_dds_alpha_mode {
  BCN_DDS_ALPHA_UNKNOWN = 0,
  BCN_DDS_ALPHA_STRAIGHT = 1,
  BCN_DDS_ALPHA_PREMULTIPLIED = 2,
  BCN_DDS_ALPHA_OPAQUE = 3,
  BCN_DDS_ALPHA_CUSTOM = 4
} bcn_dds_alpha_mode;

/* The DXGI_FORMAT the CLI writes for a variant (dxgi_format,
 * compressor/src/main.rs:355-365): BC1/2/3/7 -> *_UNORM_SRGB (72, 75, 78,
 * 99), BC4 -> BC4_UNORM (80), BC5 -> BC5_UNORM (83), BC6H -> BC6H_UF16
 * (95).  0 for an invalid variant. */
uint32_t bcn_dxgi_format(bcn_variant variant);

/* Bytes of block data a DDS file of this shape holds: the sum over mip
 * levels of max(1, ceil(w_l/4)) * max(1, ceil(h_l/4)) * block bytes, times
 * array_layers.  0 for a format that is not block-compressed (DXGI 70-84,
 * 94-99) or a zero dimension. */
size_t bcn_dds_data_size(uint32_t dxgi_format, uint32_t width, uint32_t height,
                         uint32_t mip_levels, uint32_t array_layers);

/* Writes the BCN_DDS_HEADER_SIZE bytes that precede the block data, as
 * ddsfile's Dds::new_dxgi(width, height, depth None, format, mipmap_levels,
 * array_layers, caps2 None, not a cubemap, Texture2D, alpha_mode) followed
 * by Dds::write emits them.  The CLI's call is (format of the variant,
 * mip_levels 1, array_layers 1, BCN_DDS_ALPHA_STRAIGHT). */
int bcn_dds_write_header(uint32_t dxgi_format, uint32_t width, uint32_t height,
                         uint32_t mip_levels, uint32_t array_layers, uint32_t alpha_mode,
                         void* out, size_t out_size);

/* Copies of the encoder's partition tables (pattern: 2 bits of subset per
 * texel; mask01: subset-0 mask | subset-1 mask << 16; anchors) for table
 * checks without a GPU.  Each array has 128 entries. */
void bcn_partition_tables(uint32_t* pattern, uint32_t* mask01, uint8_t* anchor1,
                          uint8_t* anchor2);

/* ---- Batched persistent dispatch ---------------------------------------------------
 *
 * No counterpart in the reference, which records one compute dispatch per
 * queued task (GpuBlockCompressor::compress, src/block_compressor.rs:540-590).
 * A batch takes every task queued on a compressor -- any number of textures
 * and mip levels, mixed formats -- and encodes the BC1-BC5 and BC7 ones with
 * persistent kernels over a device-resident block work queue (chunks of
 * blocks of one task, most expensive first), one kernel per kernel class
 * (BC7 with alpha, BC7 opaque, BC1-BC5) so that each class keeps its own
 * register allocation.  The three class kernels run concurrently on three
 * streams (the launch stream and two of the batch's own, default priority),
 * so one class's tail is filled by the others; BC6H tasks of the batch are
 * launched per task after the BC1-BC5 class.  The output bytes are the ones
 * bcn_compress writes. */
typedef struct bcn_batch bcn_batch;

typedef struct bcn_batch_config {
  uint32_t chunk_blocks_bc7;   /* blocks per work item for BC7 tasks (multiple of 64)    */
  uint32_t chunk_blocks_bc15;  /* blocks per work item for BC1-BC5 tasks (multiple of 64) */
  uint32_t waves_per_cu;       /* persistent grid = CUs x this, one 64-lane wave per
                                  workgroup; 0 = as many as can be resident           */
  uint32_t serial_classes;     /* 0 = the class kernels run concurrently (default);
                                  1 = one after another on the launch stream          */
} bcn_batch_config;

typedef struct bcn_batch_info {
  uint32_t tasks;              /* tasks in the batch (all formats)       */
  uint32_t persistent_tasks;   /* of which BC1-BC5 / BC7 (work queue)    */
  uint64_t blocks;             /* blocks in the work queue               */
  uint32_t chunks;             /* work items in the queue                */
  uint32_t grid;               /* workgroups of the persistent launch    */
  bcn_batch_config config;     /* configuration in use                   */
} bcn_batch_info;

/* Moves every task queued on `c` into a new batch and uploads its work queue
 * (the only allocation and host-to-device copy).  config NULL = defaults.
 * On failure the tasks stay queued on `c`.  The textures and buffers the
 * tasks name must outlive the batch. */
int bcn_batch_create(bcn_compressor* c, const bcn_batch_config* config, bcn_batch** out);

/* Enqueues the batch on hip_stream: a one-thread kernel that rewinds the
 * queue, then the three class kernels forked onto hip_stream and the
 * batch's two side streams (the BC6H tasks after the BC1-BC5 class), joined
 * back into hip_stream (with serial_classes, all on hip_stream): work
 * enqueued on hip_stream afterwards sees every block.  Allocates nothing and issues no memset, so it may be captured into
 * a hipGraph (the side streams join the capture).  A batch may be launched
 * any number of times; launches of one batch must not overlap. */
int bcn_batch_launch(bcn_batch* b, void* hip_stream);

/* As bcn_batch_launch through a hipGraph: the first call captures the launch
 * on a private stream and instantiates it; every call launches the graph on
 * hip_stream. */
int bcn_batch_launch_graph(bcn_batch* b, void* hip_stream);

/* Times candidate configurations on this batch (each launched `reps` times
 * on hip_stream, best time kept), keeps the fastest and returns it in
 * *chosen (may be NULL).  Synchronises hip_stream.  The output buffers are
 * rewritten with the same bytes by every run. */
int bcn_batch_autotune(bcn_batch* b, void* hip_stream, int reps, bcn_batch_config* chosen);

int bcn_batch_get_info(const bcn_batch* b, bcn_batch_info* info);

/* Text of the last failure on this batch (empty if none). */
const char* bcn_batch_last_error(const bcn_batch* b);

/* Frees the work queue and graph.  Synchronise the launch stream first. */
void bcn_batch_destroy(bcn_batch* b);

/* ---- Decode adjoints and gradient endpoint refinement ----------------------------
 *
 * No counterpart in the reference.  The decodes of BC7 (decode_block_bc7,
 * src/decode/block.rs:801-1252) and BC1 (decode_color_block, :124-197) as
 * functions of continuous endpoints, with everything else in the block held
 * fixed: mode, partition, rotation, p-bits and indices (BC1: the colour mode
 * and indices).  Parameters per block, float, in the mode's quantised units
 * without p-bits: BC7 x[6][4] (endpoint e = 2*subset + {0,1}, channel RGBA;
 * unused entries 0), BC1 x[2][3] (endpoint, RGB in 5/6/5-bit units).
 * Texels: float[16][4] per block, row-major texels, 0..255 units.
 * Forward, smooth = 0: straight-through, the exact decode of clamp(rint(x));
 * smooth = 1: the decode's linear relaxation.  Backward: the transpose of the
 * relaxation's Jacobian (the straight-through gradient).  All device
 * pointers; blocks packed (16 or 8 bytes each). */
int bcn_decode_endpoints(bcn_variant variant, const void* blocks, uint32_t nblocks, float* x,
                         void* hip_stream);
int bcn_decode_with_endpoints(bcn_variant variant, const void* blocks, const float* x,
                              uint32_t nblocks, int smooth, float* texels, void* hip_stream);
int bcn_decode_with_endpoints_backward(bcn_variant variant, const void* blocks,
                                       const float* grad_texels, uint32_t nblocks, float* grad_x,
                                       void* hip_stream);

/* Texel-side adjoint, for texture training (no counterpart in the reference):
 * the gradient with respect to the *source* texels of the blocks' decode,
 * viewing each block as its form (as above) with endpoints re-fitted to the
 * source texels by least squares, unquantised (the fit of the reference's
 * refinement step, opt_endpoints, src/encode/common.rs:142-206).  Per subset
 * and index set that map is the orthogonal projection of the source values
 * onto span{1, interpolation weight}, which is its own adjoint; channels that
 * do not depend on the source (alpha of BC7 modes 0-3 and of BC1, BC1's
 * 3-colour black texels) get 0.  grad_texels in, grad_source out: float
 * [16][4] per block, 16-byte aligned device pointers. */
int bcn_decode_texels_backward(bcn_variant variant, const void* blocks, const float* grad_texels,
                               uint32_t nblocks, float* grad_source, void* hip_stream);

typedef struct bcn_refine_params {
  uint32_t steps;      /* Adam steps                                           */
  float lr;            /* step size, in quantised endpoint units               */
  float beta1, beta2;  /* Adam moments                                          */
  float eps;
  uint32_t channels;   /* 3: minimise RGB squared error, 4: RGBA              */
  uint32_t reassign;   /* 0: indices fixed; k > 0: after every k-th step, give
                          each texel its best index for the current endpoints
                          (mode, partition, rotation stay fixed)            */
  uint32_t pbits;      /* with reassign: also try each p-bit flipped, keep
                          flips that lower the error                         */
} bcn_refine_params;

/* Gradient endpoint refinement of a BC7 block image against a float RGBA
 * target (width x height texels, 0..1 per channel, target_pitch bytes per row,
 * 16-byte aligned): per block, Adam on the endpoints from the block's own,
 * straight-through rounding, the adjoint above for the gradient of the
 * squared error, optionally alternating with index (and p-bit) reassignment;
 * each iterate is scored by its exact decode and the best (the input
 * included) is written, with anchor indices made encodable by the usual
 * endpoint swap, so no block's error grows.  blocks_out may equal
 * blocks_in.  sse_before / sse_after (may be NULL): per-block squared error in
 * 0..255 units.  workspace: device memory of bcn_bc7_refine_workspace_size
 * bytes; calls sharing a workspace must not overlap.  Replaces, as an option, the least-squares refinement of the
 * encoder (refine_iterations) as a pass after it. */
int bcn_bc7_refine(const void* blocks_in, void* blocks_out, uint32_t width, uint32_t height,
                   const float* target, uint32_t target_pitch, const bcn_refine_params* params,
                   float* sse_before, float* sse_after, void* workspace, size_t workspace_size,
                   void* hip_stream);

/* Bytes of device workspace bcn_bc7_refine needs for nblocks blocks (a sort
 * of the blocks by mode, so that each wave refines blocks of one mode). */
size_t bcn_bc7_refine_workspace_size(uint32_t nblocks);

/* Largest block count bcn_bc7_refine_noalloc accepts: 4096 x 4096 texels. */
#define BCN_NOALLOC_MAX_BLOCKS 1048576u

/* bcn_bc7_refine without a workspace argument and without any allocation.
 * No counterpart in the reference; for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  The workspace is a fixed scratch array in the library's
 * device code, one per GPU, of 4 MiB + 160 B, held until the process ends;
 * images of more than BCN_NOALLOC_MAX_BLOCKS blocks are rejected with
 * BCN_ERR_TOO_LARGE before any device work.  All calls on one GPU share that
 * array, so calls that overlap on one GPU (on different streams, or from
 * different threads) race and must be serialised by the caller.  Results are
 * those of bcn_bc7_refine with the same arguments. */
int bcn_bc7_refine_noalloc(const void* blocks_in, void* blocks_out, uint32_t width, uint32_t height,
                           const float* target, uint32_t target_pitch, const bcn_refine_params* params,
                           float* sse_before, float* sse_after, void* hip_stream);

#ifdef __cplusplus
}
#endif

#endif /* BCN_MI355X_H */
