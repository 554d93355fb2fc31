"""The batched mixed-format workload: many textures with full mip chains.

300 RGBA8 textures from the tile generator (synth.tile_image_rgba8), each
with width and height drawn independently from {64..4096} (splitmix64 keyed
by (seed, texture)), a full box-filtered mip chain, and a format drawn
uniformly from BC1, BC3, BC4, BC5, BC7 opaque_basic and BC7 alpha_basic.
Textures whose format keeps alpha (BC3, BC7 alpha_basic) take their alpha
channel from a second generator image (seed + 7919, its red channel), so the
alpha search has work to do.

Mip level l+1 averages 2x2 texels of level l, (a + b + c + d + 2) >> 2 per
channel, over floor(w/2) x floor(h/2) (an odd last row or column is dropped),
down to 1x1.  Every level is then padded to multiples of 4 by repeating its
last row and column.  Everything is built with torch on the target device.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import List

import numpy as np

from . import synth
from .settings import BC7Settings, CompressionVariant

FORMATS = ("bc1", "bc3", "bc4", "bc5", "bc7_opaque_basic", "bc7_alpha_basic")


def variant_of(name: str) -> CompressionVariant:
    return {
        "bc1": CompressionVariant.BC1(),
        "bc3": CompressionVariant.BC3(),
        "bc4": CompressionVariant.BC4(),
        "bc5": CompressionVariant.BC5(),
        "bc7_opaque_basic": CompressionVariant.BC7(BC7Settings.opaque_basic()),
        "bc7_alpha_basic": CompressionVariant.BC7(BC7Settings.alpha_basic()),
    }[name]


@dataclass
class Level:
    texture: int          # texture index
    level: int            # mip level
    fmt: str              # one of FORMATS
    width: int            # padded width (multiple of 4)
    height: int           # padded height
    texels: object        # (height, width, 4) uint8 device tensor
    offset: int = 0       # byte offset of its blocks in the output buffer
    nbytes: int = 0


def texture_specs(n: int = 300, seed: int = 4, min_size: int = 64, max_size: int = 4096):
    """(width, height, format) per texture, drawn with splitmix64."""
    idx = np.arange(n, dtype=np.uint64)
    with np.errstate(over="ignore"):
        k = synth.splitmix64(np.uint64(seed) * np.uint64(0x9E3779B97F4A7C15) + idx)
        hw = synth.splitmix64(k ^ np.uint64(0x1234567))
        hh = synth.splitmix64(k ^ np.uint64(0x89ABCDE))
        hf = synth.splitmix64(k ^ np.uint64(0x5151515))
    span = max_size - min_size + 1
    w = min_size + (hw % np.uint64(span)).astype(np.int64)
    h = min_size + (hh % np.uint64(span)).astype(np.int64)
    f = (hf % np.uint64(len(FORMATS))).astype(np.int64)
    return [(int(a), int(b), FORMATS[int(c)]) for a, b, c in zip(w, h, f)]


def _pad4(t):
    import torch

    h, w = int(t.shape[0]), int(t.shape[1])
    ph, pw = (-h) % 4, (-w) % 4
    if ph == 0 and pw == 0:
        return t.contiguous()
    out = torch.empty((h + ph, w + pw, 4), dtype=t.dtype, device=t.device)
    out[:h, :w] = t
    if pw:
        out[:h, w:] = t[:, w - 1:w]
    if ph:
        out[h:] = out[h - 1:h]
    return out


def mip_chain(base) -> list:
    """Box-filtered levels of an (h, w, 4) uint8 tensor, level 0 first (unpadded)."""
    import torch

    levels = [base]
    cur = base
    while cur.shape[0] > 1 or cur.shape[1] > 1:
        h2, w2 = max(1, cur.shape[0] // 2), max(1, cur.shape[1] // 2)
        if cur.shape[0] == 1 or cur.shape[1] == 1:
            # One dimension already 1: average pairs along the other.
            c = cur.to(torch.int32)
            if cur.shape[0] == 1:
                s = c[:, 0:2 * w2:2] + c[:, 1:2 * w2:2]
            else:
                s = c[0:2 * h2:2] + c[1:2 * h2:2]
            nxt = ((s + 1) >> 1).to(torch.uint8)
        else:
            c = cur[:2 * h2, :2 * w2].to(torch.int32)
            s = c[0::2, 0::2] + c[1::2, 0::2] + c[0::2, 1::2] + c[1::2, 1::2]
            nxt = ((s + 2) >> 2).to(torch.uint8)
        levels.append(nxt.contiguous())
        cur = nxt
    return levels


def build(n_textures: int = 300, seed: int = 4, device="cuda", max_size: int = 4096,
          min_size: int = 64) -> List[Level]:
    """Generates the textures and their padded mip levels on `device`."""
    out: List[Level] = []
    for i, (w, h, fmt) in enumerate(texture_specs(n_textures, seed, min_size, max_size)):
        w8, h8 = (w + 7) // 8 * 8, (h + 7) // 8 * 8
        img = synth.tile_image_rgba8_torch(w8, h8, seed * 1000 + i, 0.1, device)
        if fmt in ("bc3", "bc7_alpha_basic"):
            a = synth.tile_image_rgba8_torch(w8, h8, seed * 1000 + i + 7919, 0.1, device)
            img[..., 3] = a[..., 0]
        base = img[:h, :w]
        for lvl, t in enumerate(mip_chain(base)):
            p = _pad4(t)
            out.append(Level(i, lvl, fmt, int(p.shape[1]), int(p.shape[0]), p))
    off = 0
    for lv in out:
        lv.nbytes = variant_of(lv.fmt).blocks_byte_size(lv.width, lv.height)
        lv.offset = off
        off += (lv.nbytes + 15) // 16 * 16
    return out


def output_bytes(levels: List[Level]) -> int:
    return max((lv.offset + lv.nbytes for lv in levels), default=0)


def queue(compressor, levels: List[Level], buffer) -> None:
    """One compression task per level, blocks at each level's offset in `buffer`."""
    for lv in levels:
        compressor.add_compression_task(variant_of(lv.fmt), lv.texels, lv.width, lv.height, buffer,
                                        blocks_offset=lv.offset)
