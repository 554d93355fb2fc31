"""Image-level GPU decoders with the reference crate's decode API.

Mirrors src/decode.rs:336-468 over the HIP decode kernels:

    decompress_blocks_as_rgba8(variant, width, height, blocks_data, rgba_data)
    decompress_blocks_as_rgba16f(variant, width, height, blocks_data, rgba_data)
    decompress_blocks_as_rgba32f(variant, width, height, blocks_data, rgba_data)

`blocks_data` is a contiguous uint8 device tensor of exactly
variant.blocks_byte_size(width, height) bytes; `rgba_data` a contiguous
device tensor of exactly width * height * 4 elements (uint8, float16 or
float32).  The reference's size asserts raise ValueError with its messages.

RGBA8: BC4 -> (R, 0, 0, 0), BC5 -> (R, G, 0, 0), BC6H -> sRGB8 of the
unsigned decode with alpha 0 (decode.rs:161-181), the rest RGBA.

RGBA16F / RGBA32F: BC6H only (other variants raise, as the reference
panics).  By default the output is a proper RGBA image with alpha 1.0.  The
reference writes each block's texels as packed RGB from the block's RGBA
offset (texel j at element j*3, src/decode/block.rs:753, :787-792), leaving
elements 12..15 of every block row untouched; reference_layout=True
reproduces that layout element for element.
"""
from __future__ import annotations

import ctypes

from . import _capi
from .settings import CompressionVariant


def _stream(tensor, stream) -> ctypes.c_void_p:
    import torch

    if stream is None:
        return ctypes.c_void_p(torch.cuda.current_stream(tensor.device).cuda_stream)
    return ctypes.c_void_p(stream.cuda_stream)


def _check_sizes(variant: CompressionVariant, width: int, height: int, blocks_data, rgba_data, dtype):
    # decode.rs:343-355 / :404-416 / :445-456
    if not blocks_data.is_contiguous() or not rgba_data.is_contiguous():
        raise ValueError("blocks_data and rgba_data must be contiguous")
    if blocks_data.numel() * blocks_data.element_size() != variant.blocks_byte_size(width, height):
        raise ValueError("the input bitstream slice has not the expected size")
    if rgba_data.dtype != dtype or rgba_data.numel() != width * height * 4:
        raise ValueError("the output slice has not the expected size")


def decompress_blocks_as_rgba8(variant: CompressionVariant, width: int, height: int, blocks_data,
                               rgba_data, stream=None) -> None:
    """decompress_blocks_as_rgba8 (src/decode.rs:336-387) on the GPU."""
    import torch

    _check_sizes(variant, width, height, blocks_data, rgba_data, torch.uint8)
    _capi.check(_capi.lib().bcn_decode_rgba8(
        variant.tag, blocks_data.data_ptr(), variant.blocks_byte_size(width, height), width, height,
        rgba_data.data_ptr(), width * 4, _stream(rgba_data, stream)))


def _decompress_float(fn, variant, width, height, blocks_data, rgba_data, dtype, texel_bytes,
                      reference_layout, stream):
    _check_sizes(variant, width, height, blocks_data, rgba_data, dtype)
    flags = _capi.DECODE_REFERENCE_LAYOUT if reference_layout else 0
    _capi.check(fn(variant.tag, blocks_data.data_ptr(), variant.blocks_byte_size(width, height),
                   width, height, rgba_data.data_ptr(), width * texel_bytes, flags,
                   _stream(rgba_data, stream)))


def decompress_blocks_as_rgba16f(variant: CompressionVariant, width: int, height: int, blocks_data,
                                 rgba_data, reference_layout: bool = False, stream=None) -> None:
    """decompress_blocks_as_rgba16f (src/decode.rs:397-430) on the GPU."""
    import torch

    _decompress_float(_capi.lib().bcn_decode_rgba16f, variant, width, height, blocks_data,
                      rgba_data, torch.float16, 8, reference_layout, stream)


def decompress_blocks_as_rgba32f(variant: CompressionVariant, width: int, height: int, blocks_data,
                                 rgba_data, reference_layout: bool = False, stream=None) -> None:
    """decompress_blocks_as_rgba32f (src/decode.rs:438-468) on the GPU."""
    import torch

    _decompress_float(_capi.lib().bcn_decode_rgba32f, variant, width, height, blocks_data,
                      rgba_data, torch.float32, 16, reference_layout, stream)


def bc6h_srgb8_thresholds():
    """The 256 sRGB8 code thresholds the BC6H -> RGBA8 kernel uses (numpy
    uint16): entry k is the smallest half bit pattern whose code is >= k."""
    import numpy as np

    out = np.zeros(256, np.uint16)
    _capi.lib().bcn_bc6h_srgb8_thresholds(out.ctypes.data_as(ctypes.c_void_p))
    return out


def bc6h_srgb8_buckets():
    """The 3,968 bucket entries the BC6H -> RGBA8 kernel reads (numpy uint16):
    the sRGB8 code of a half bit pattern h <= 0x7BFF is the high byte of
    entry[h >> 3] + 32 * (h & 7)."""
    import numpy as np

    out = np.zeros(3968, np.uint16)
    _capi.check(_capi.lib().bcn_bc6h_srgb8_buckets(out.ctypes.data_as(ctypes.c_void_p)))
    return out
