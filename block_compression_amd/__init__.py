"""BC1-BC7 texture block compression on AMD MI355X (gfx950).

A drop-in for the GPU path of the `block_compression` crate: the
`GpuBlockCompressor` task API (add_compression_task / compress) over
`CompressionVariant` with `BC6HSettings` / `BC7Settings` presets, running
hand-written HIP kernels through the C ABI in include/bcn_mi355x.h.
"""
from .settings import (  # noqa: F401
    BC6H_PRESETS,
    BC7_PRESETS,
    BC6HSettings,
    BC7Settings,
    CompressionVariant,
)
from .compressor import (  # noqa: F401
    BlockCompressionBatch,
    GpuBlockCompressor,
    compress_rgba8,
    decode_bc6h_rgba16f,
    decode_rgba8,
)
from ._capi import BcnError  # noqa: F401
from . import dds, decode  # noqa: F401
from .decode import (  # noqa: F401
    decompress_blocks_as_rgba8,
    decompress_blocks_as_rgba16f,
    decompress_blocks_as_rgba32f,
)

__all__ = [
    "BC6HSettings",
    "BC7Settings",
    "BC6H_PRESETS",
    "BC7_PRESETS",
    "CompressionVariant",
    "BlockCompressionBatch",
    "GpuBlockCompressor",
    "BcnError",
    "compress_rgba8",
    "decode_rgba8",
    "decode_bc6h_rgba16f",
    "decompress_blocks_as_rgba8",
    "decompress_blocks_as_rgba16f",
    "decompress_blocks_as_rgba32f",
    "dds",
    "decode",
]
