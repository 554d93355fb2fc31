"""Command-line compressor: an image file in, a DDS file out.

The reference's `compressor` binary (compressor/src/main.rs) on the HIP
library:

    python -m block_compression_amd.cli <compression_variant> <input_file>

reads the image (any format Pillow reads, converted to RGBA8), uploads it,
compresses it on the GPU with the CLI's preset for the variant (BC6H
very_slow, BC7 alpha_slow; main.rs:386-394), downloads the blocks and writes
<input_file with extension .dds> (DXGI format per main.rs:355-365).  For BC6H
every 8-bit channel, alpha included, goes through srgb_to_linear in float64
and is rounded to the nearest half float (main.rs:152-156, 214-222).  It
prints the timings the reference prints; the compression time is measured
with HIP events around the kernels, as the reference measures its compute
pass with timestamp queries.
"""
from __future__ import annotations

import sys
import time
from typing import List, Optional, Tuple

import numpy as np

from .settings import BC6HSettings, BC7Settings, CompressionVariant

HELP = """Usage: compressor <compression_variant> <input_file>

Compression variants:
  bc1  - BC1 compression (RGB)
  bc2  - BC2 compression with sharp alpha (RGBA)
  bc3  - BC3 compression with smooth alpha (RGBA)
  bc4  - BC4 compression (R)
  bc5  - BC5 compression (RG)
  bc6h - BC6H compression (RGB HDR)
  bc7  - BC7 compression with smooth alpha (RGBA)"""

VARIANTS = {
    "bc1": CompressionVariant.BC1,
    "bc2": CompressionVariant.BC2,
    "bc3": CompressionVariant.BC3,
    "bc4": CompressionVariant.BC4,
    "bc5": CompressionVariant.BC5,
    "bc6h": lambda: CompressionVariant.BC6H(BC6HSettings.very_slow()),
    "bc7": lambda: CompressionVariant.BC7(BC7Settings.alpha_slow()),
}


def parse_args(argv: List[str]) -> Optional[Tuple[CompressionVariant, str]]:
    """parse_args (main.rs:378-404): exactly two arguments, no --help."""
    if len(argv) != 3 or "--help" in argv:
        print(HELP)
        return None
    make = VARIANTS.get(argv[1].lower())
    if make is None:
        print("Error: Invalid compression variant")
        print(HELP)
        return None
    return make(), argv[2]


def srgb_to_linear_f16(rgba8: np.ndarray) -> np.ndarray:
    """srgb_to_linear (main.rs:214-222) in float64, then f16::from_f64
    (round to nearest even) of every channel."""
    v = rgba8.astype(np.float64) / 255.0
    lin = np.where(v <= 0.04045, v / 12.92, np.power((v + 0.055) / 1.055, 2.4))
    return lin.astype(np.float16)


def read_image(file_name: str, variant: CompressionVariant) -> np.ndarray:
    """ImageReader::open(..).decode().to_rgba8() (main.rs:131-138); RGBA16F
    for BC6H."""
    from PIL import Image

    with Image.open(file_name) as im:
        rgba = np.array(im.convert("RGBA"), dtype=np.uint8)
    if variant.is_bc6h:
        return srgb_to_linear_f16(rgba)
    return np.ascontiguousarray(rgba)


def main(argv: Optional[List[str]] = None) -> int:
    args = parse_args(sys.argv if argv is None else argv)
    if args is None:
        return 0
    variant, file_name = args

    import torch

    from . import dds
    from .compressor import GpuBlockCompressor

    device = torch.device("cuda", torch.cuda.current_device())
    compressor = GpuBlockCompressor(device.index)
    print(f"Using backend: HIP ({torch.cuda.get_device_name(device)})")

    start = time.perf_counter()
    pixels = read_image(file_name, variant)
    texture = torch.from_numpy(pixels).to(device)
    torch.cuda.synchronize(device)
    height, width = int(pixels.shape[0]), int(pixels.shape[1])
    print(f"Image read and upload took: {(time.perf_counter() - start) * 1000.0:.3f} ms")

    blocks = torch.empty(variant.blocks_byte_size(width, height), dtype=torch.uint8, device=device)
    compressor.add_compression_task(variant, texture, width, height, blocks)
    stream = torch.cuda.current_stream(device)
    t0 = torch.cuda.Event(enable_timing=True)
    t1 = torch.cuda.Event(enable_timing=True)
    t0.record(stream)
    compressor.compress(stream)
    t1.record(stream)
    t1.synchronize()
    print(f"Compression took: {t0.elapsed_time(t1):.3f} ms")

    start = time.perf_counter()
    block_data = blocks.cpu().numpy()
    print(f"Block data download took: {(time.perf_counter() - start) * 1000.0:.3f} ms")

    start = time.perf_counter()
    dds.write_dds(dds.dds_path(file_name), variant, width, height, block_data)
    print(f"DDS output to disk took: {(time.perf_counter() - start) * 1000.0:.3f} ms")
    compressor.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
