"""DDS files of block-compressed textures.

The reference CLI writes its output through the `ddsfile` crate
(compressor/src/main.rs:313-353): "DDS " magic, DDS_HEADER and the DX10
extension header, then the block data.  The header bytes come from the C ABI
(bcn_dds_write_header), so C callers and this module write identical files.

    write_dds(path, variant, width, height, block_data)     # main.rs:313-353
    dds_path(file_name)                                      # PathBuf::set_extension("dds")
    read_dds(path) -> DdsFile                                # header fields + data

Mip chains and texture arrays (layer-major, each layer's levels from the
largest) are accepted by write_dds; the CLI writes one level, one layer.
"""
from __future__ import annotations

import ctypes
import os
import struct
from dataclasses import dataclass
from typing import Union

from . import _capi
from .settings import CompressionVariant

HEADER_SIZE = 148
ALPHA_UNKNOWN, ALPHA_STRAIGHT, ALPHA_PREMULTIPLIED, ALPHA_OPAQUE, ALPHA_CUSTOM = range(5)

# The block-compressed DXGI_FORMAT values.
DXGI_NAMES = {
    70: "BC1_TYPELESS", 71: "BC1_UNORM", 72: "BC1_UNORM_SRGB",
    73: "BC2_TYPELESS", 74: "BC2_UNORM", 75: "BC2_UNORM_SRGB",
    76: "BC3_TYPELESS", 77: "BC3_UNORM", 78: "BC3_UNORM_SRGB",
    79: "BC4_TYPELESS", 80: "BC4_UNORM", 81: "BC4_SNORM",
    82: "BC5_TYPELESS", 83: "BC5_UNORM", 84: "BC5_SNORM",
    94: "BC6H_TYPELESS", 95: "BC6H_UF16", 96: "BC6H_SF16",
    97: "BC7_TYPELESS", 98: "BC7_UNORM", 99: "BC7_UNORM_SRGB",
}


def dxgi_format(variant: CompressionVariant) -> int:
    """dxgi_format (compressor/src/main.rs:355-365)."""
    return int(_capi.lib().bcn_dxgi_format(variant.tag))


def data_size(fmt: int, width: int, height: int, mip_levels: int = 1, array_layers: int = 1) -> int:
    """Bytes of block data a DDS file of this shape holds."""
    return int(_capi.lib().bcn_dds_data_size(fmt, width, height, mip_levels, array_layers))


def header(fmt: int, width: int, height: int, mip_levels: int = 1, array_layers: int = 1,
           alpha_mode: int = ALPHA_STRAIGHT) -> bytes:
    """The HEADER_SIZE bytes before the data (bcn_dds_write_header)."""
    buf = (ctypes.c_uint8 * HEADER_SIZE)()
    _capi.check(_capi.lib().bcn_dds_write_header(fmt, width, height, mip_levels, array_layers,
                                                 alpha_mode, buf, HEADER_SIZE))
    return bytes(buf)


def dds_path(file_name: Union[str, os.PathLike]) -> str:
    """PathBuf::set_extension("dds") (main.rs:347-348): the last extension of
    the file name is replaced, or ".dds" appended when there is none."""
    root, _ = os.path.splitext(os.fspath(file_name))
    return root + ".dds"


def _as_bytes(block_data) -> bytes:
    if isinstance(block_data, (bytes, bytearray, memoryview)):
        return bytes(block_data)
    if hasattr(block_data, "detach"):  # torch tensor, host or device
        block_data = block_data.detach().contiguous().cpu().numpy()
    return block_data.tobytes()


def write_dds(path: Union[str, os.PathLike], variant: Union[CompressionVariant, int], width: int,
              height: int, block_data, mip_levels: int = 1, array_layers: int = 1,
              alpha_mode: int = ALPHA_STRAIGHT) -> None:
    """Writes a DDS file the way the reference CLI does (main.rs:313-353).

    `variant` is a CompressionVariant (its DXGI format per main.rs:355-365)
    or a DXGI_FORMAT value; `block_data` holds exactly data_size(...) bytes
    (bytes, a numpy array or a torch tensor)."""
    fmt = dxgi_format(variant) if isinstance(variant, CompressionVariant) else int(variant)
    data = _as_bytes(block_data)
    want = data_size(fmt, width, height, mip_levels, array_layers)
    if want == 0:
        raise ValueError(f"not a block-compressed DXGI format or empty shape: {fmt}")
    if len(data) != want:
        raise ValueError(f"block data holds {len(data)} bytes, the DDS shape needs {want}")
    hdr = header(fmt, width, height, mip_levels, array_layers, alpha_mode)
    with open(path, "wb") as f:
        f.write(hdr)
        f.write(data)


@dataclass
class DdsFile:
    dxgi_format: int
    width: int
    height: int
    mip_levels: int
    array_layers: int
    alpha_mode: int
    data: bytes

    @property
    def format_name(self) -> str:
        return DXGI_NAMES.get(self.dxgi_format, str(self.dxgi_format))


def read_dds(path: Union[str, os.PathLike]) -> DdsFile:
    """Reads a DX10-extended DDS file of a block-compressed 2D texture."""
    with open(path, "rb") as f:
        raw = f.read()
    if len(raw) < HEADER_SIZE or raw[:4] != b"DDS ":
        raise ValueError("not a DDS file")
    w = struct.unpack_from("<37I", raw, 0)
    if w[1] != 124 or w[19] != 32 or not (w[20] & 0x4) or w[21] != 0x30315844:
        raise ValueError("DDS file without the DX10 header")
    fmt, dim, _misc, layers, alpha = w[32:37]
    if fmt not in DXGI_NAMES or dim != 3:
        raise ValueError(f"unsupported DDS texture: format {fmt}, dimension {dim}")
    mips = w[7] if w[7] else 1
    out = DdsFile(fmt, w[4], w[3], mips, layers, alpha, raw[HEADER_SIZE:])
    want = data_size(fmt, out.width, out.height, mips, layers)
    if len(out.data) != want:
        raise ValueError(f"DDS data holds {len(out.data)} bytes, its header needs {want}")
    return out
