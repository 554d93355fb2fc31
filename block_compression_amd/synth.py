"""Seeded synthetic textures for the measured workloads.

Every random draw comes from splitmix64 keyed by (seed, tile, purpose), so the
bytes depend only on this file's arithmetic, never on a library's random
stream.  The images are built from 8x8-pixel tiles drawn independently:

* ``tile_image_rgba8`` -- RGBA8 (sRGB-encoded values, alpha 255).  Tile kinds:
  bilinear 4-corner colour gradient (50%), 4-octave value noise with a random
  amplitude in 0..255 around a random base colour (30%), two-colour hard edge
  through a random point at an angle uniform in [0, pi) (20%).  With
  ``noise_frac`` > 0, that fraction of tiles is instead i.i.d. uniform
  per-pixel noise and the rest keep the 50/30/20 split.
* ``hdr_image_rgba16f`` -- RGBA16F HDR: luminance 2**(-8 + 18 n) with n a
  smooth 3-octave value-noise field over the image, hue uniform on the circle
  and saturation U(0,1) per tile, 5% of tiles with a x2**6 luminance step
  across a random edge (the bright side clamped to 2**10); RGB = L * (1 + s *
  cos(h + 2 pi k / 3)); alpha 1.0; rounded to float16 (round-to-nearest-even).
"""
from __future__ import annotations

import hashlib

import numpy as np

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(x: np.ndarray) -> np.ndarray:
    """splitmix64 finaliser applied elementwise to uint64 keys."""
    with np.errstate(over="ignore"):
        z = (x.astype(np.uint64) + np.uint64(0x9E3779B97F4A7C15)) & _M64
        z = ((z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & _M64
        z = ((z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & _M64
        return z ^ (z >> np.uint64(31))


def _key(seed: int, tile: np.ndarray, purpose: int) -> np.ndarray:
    with np.errstate(over="ignore"):
        base = splitmix64(np.uint64(seed) * np.uint64(0x100000001B3) + np.uint64(purpose))
        return splitmix64(base ^ (tile.astype(np.uint64) * np.uint64(0xD6E8FEB86659FD93)))


def _unit(h: np.ndarray) -> np.ndarray:
    """Top 24 bits of a hash as a float64 in [0, 1)."""
    return (h >> np.uint64(40)).astype(np.float64) / float(1 << 24)


def _draws(seed: int, tile: np.ndarray, purpose: int, n: int) -> np.ndarray:
    """n independent uniforms per tile, shape (len(tile), n)."""
    k = _key(seed, tile, purpose)
    with np.errstate(over="ignore"):
        return _unit(splitmix64(k[:, None] + np.arange(n, dtype=np.uint64)[None, :]))


def _value_noise(lattice: np.ndarray, freq: int, fx: np.ndarray, fy: np.ndarray) -> np.ndarray:
    """Bilinear value noise.  lattice: (T, C, freq+1, freq+1); fx, fy in [0,1)
    of shape (8,) -> returns (T, C, 8, 8)."""
    gx = fx * freq
    gy = fy * freq
    ix = np.floor(gx).astype(np.int64)
    iy = np.floor(gy).astype(np.int64)
    tx = gx - ix
    ty = gy - iy
    l00 = lattice[:, :, iy[:, None], ix[None, :]]
    l10 = lattice[:, :, iy[:, None], ix[None, :] + 1]
    l01 = lattice[:, :, iy[:, None] + 1, ix[None, :]]
    l11 = lattice[:, :, iy[:, None] + 1, ix[None, :] + 1]
    top = l00 + (l10 - l00) * tx[None, None, None, :]
    bot = l01 + (l11 - l01) * tx[None, None, None, :]
    return top + (bot - top) * ty[None, None, :, None]


def tile_image_rgba8(width: int, height: int, seed: int = 1, noise_frac: float = 0.0) -> np.ndarray:
    """(height, width, 4) uint8 image; width and height multiples of 8."""
    rgb = np.clip(np.floor(_tile_rgb(width, height, seed, noise_frac) + 0.5), 0, 255).astype(np.uint8)
    img = np.empty((height, width, 4), dtype=np.uint8)
    img[..., :3] = rgb
    img[..., 3] = 255
    return img


def tile_target_rgba_f32(width: int, height: int, seed: int = 1, noise_frac: float = 0.0) -> np.ndarray:
    """(height, width, 4) float32 RGBA target in [0, 1]: the tile generator's
    values before rounding to 8 bits, divided by 255 and clipped; alpha is the
    red channel of a second generator image (seed + 7919), likewise."""
    rgb = _tile_rgb(width, height, seed, noise_frac)
    a = _tile_rgb(width, height, seed + 7919, noise_frac)[..., 0]
    out = np.empty((height, width, 4), dtype=np.float32)
    out[..., :3] = np.clip(rgb / 255.0, 0.0, 1.0)
    out[..., 3] = np.clip(a / 255.0, 0.0, 1.0)
    return out


def _tile_rgb(width: int, height: int, seed: int, noise_frac: float) -> np.ndarray:
    """(height, width, 3) float64 tile-generator values (0..255 scale, before rounding)."""
    if width % 8 or height % 8:
        raise ValueError("width and height must be multiples of 8")
    tw, th = width // 8, height // 8
    nt = tw * th
    tiles = np.arange(nt, dtype=np.int64)
    kind_u = _draws(seed, tiles, 1, 2)
    is_noise_tile = kind_u[:, 0] < noise_frac
    u = kind_u[:, 1]
    kind = np.where(u < 0.5, 0, np.where(u < 0.8, 1, 2))
    kind = np.where(is_noise_tile, 3, kind)

    out = np.zeros((nt, 8, 8, 3), dtype=np.float64)
    coord = (np.arange(8, dtype=np.float64) + 0.5) / 8.0
    lin = np.arange(8, dtype=np.float64) / 7.0

    # Gradients: corners c00 c10 c01 c11, 3 channels each.
    sel = np.nonzero(kind == 0)[0]
    if sel.size:
        c = _draws(seed, sel, 2, 12).reshape(-1, 4, 3) * 255.0
        fx = lin[None, None, :, None]
        fy = lin[None, :, None, None]
        top = c[:, None, None, 0] + (c[:, None, None, 1] - c[:, None, None, 0]) * fx
        bot = c[:, None, None, 2] + (c[:, None, None, 3] - c[:, None, None, 2]) * fx
        out[sel] = top + (bot - top) * fy

    # 4-octave value noise around a base colour with amplitude in 0..255.
    sel = np.nonzero(kind == 1)[0]
    if sel.size:
        par = _draws(seed, sel, 3, 4)
        base = par[:, :3] * 255.0
        amp = par[:, 3] * 255.0
        acc = np.zeros((sel.size, 3, 8, 8))
        norm = 0.0
        for o, f in enumerate((1, 2, 4, 8)):
            lat = _draws(seed, sel, 10 + o, 3 * (f + 1) * (f + 1)).reshape(-1, 3, f + 1, f + 1)
            w = 0.5 ** o
            acc += w * _value_noise(lat, f, coord, coord)
            norm += w
        n = acc / norm - 0.5
        out[sel] = np.transpose(base[:, :, None, None] + amp[:, None, None, None] * n, (0, 2, 3, 1))

    # Two-colour hard edges.
    sel = np.nonzero(kind == 2)[0]
    if sel.size:
        par = _draws(seed, sel, 4, 9)
        c0 = par[:, 0:3] * 255.0
        c1 = par[:, 3:6] * 255.0
        px = par[:, 6] * 8.0
        py = par[:, 7] * 8.0
        th_ = par[:, 8] * np.pi
        xs = np.arange(8, dtype=np.float64) + 0.5
        d = (xs[None, None, :] - px[:, None, None]) * np.sin(th_)[:, None, None] - (
            xs[None, :, None] - py[:, None, None]
        ) * np.cos(th_)[:, None, None]
        side = (d >= 0.0)[..., None]
        out[sel] = np.where(side, c1[:, None, None, :], c0[:, None, None, :])

    # i.i.d. uniform noise tiles.
    sel = np.nonzero(kind == 3)[0]
    if sel.size:
        out[sel] = _draws(seed, sel, 5, 192).reshape(-1, 8, 8, 3) * 256.0

    return out.reshape(th, tw, 8, 8, 3).transpose(0, 2, 1, 3, 4).reshape(height, width, 3)


def hdr_image_rgba16f(width: int, height: int, seed: int = 3) -> np.ndarray:
    """(height, width, 4) float16 HDR image; width and height multiples of 8."""
    if width % 8 or height % 8:
        raise ValueError("width and height must be multiples of 8")
    tw, th = width // 8, height // 8
    nt = tw * th
    ys = (np.arange(height, dtype=np.float64) + 0.5) / height
    xs = (np.arange(width, dtype=np.float64) + 0.5) / width
    field = np.zeros((height, width))
    norm = 0.0
    for o, f in enumerate((4, 8, 16)):
        cells = np.arange((f + 1) * (f + 1), dtype=np.int64)
        lat = _unit(_key(seed, cells, 20 + o)).reshape(f + 1, f + 1)
        gx, gy = xs * f, ys * f
        ix = np.minimum(np.floor(gx).astype(np.int64), f - 1)
        iy = np.minimum(np.floor(gy).astype(np.int64), f - 1)
        tx, ty = gx - ix, gy - iy
        # Smoothstep-blended bilinear interpolation of the lattice.
        sx = tx * tx * (3 - 2 * tx)
        sy = ty * ty * (3 - 2 * ty)
        l00 = lat[iy[:, None], ix[None, :]]
        l10 = lat[iy[:, None], ix[None, :] + 1]
        l01 = lat[iy[:, None] + 1, ix[None, :]]
        l11 = lat[iy[:, None] + 1, ix[None, :] + 1]
        top = l00 + (l10 - l00) * sx[None, :]
        bot = l01 + (l11 - l01) * sx[None, :]
        w = 0.5 ** o
        field += w * (top + (bot - top) * sy[:, None])
        norm += w
    n = np.clip(field / norm, 0.0, 1.0)
    lum = np.exp2(-8.0 + 18.0 * n)

    tiles = np.arange(nt, dtype=np.int64)
    par = _draws(seed, tiles, 30, 6)
    hue = par[:, 0] * 2 * np.pi
    sat = par[:, 1]
    step = par[:, 2] < 0.05
    lum_t = lum.reshape(th, 8, tw, 8).transpose(0, 2, 1, 3).reshape(nt, 8, 8)
    if step.any():
        sel = np.nonzero(step)[0]
        px = par[sel, 3] * 8.0
        py = par[sel, 4] * 8.0
        ang = par[sel, 5] * np.pi
        c = np.arange(8, dtype=np.float64) + 0.5
        d = (c[None, None, :] - px[:, None, None]) * np.sin(ang)[:, None, None] - (
            c[None, :, None] - py[:, None, None]
        ) * np.cos(ang)[:, None, None]
        bright = np.minimum(lum_t[sel] * 64.0, 1024.0)
        lum_t[sel] = np.where(d >= 0.0, bright, lum_t[sel])
    k = np.arange(3, dtype=np.float64)
    chroma = 1.0 + sat[:, None] * np.cos(hue[:, None] + 2 * np.pi * k[None, :] / 3)
    rgb = lum_t[..., None] * chroma[:, None, None, :]
    img = np.empty((th, 8, tw, 8, 4), dtype=np.float32)
    img[..., :3] = rgb.reshape(th, tw, 8, 8, 3).transpose(0, 2, 1, 3, 4)
    img[..., 3] = 1.0
    return img.reshape(height, width, 4).astype(np.float16)


def sha256(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


# ---- the tile generator on a torch device -------------------------------------------
#
# The same arithmetic as tile_image_rgba8 on torch tensors (float64 for the
# tile maths), so that large workloads can be generated where they are used.
# torch has no uint64 arithmetic: hashes are int64 bit patterns (addition and
# multiplication wrap like uint64; right shifts are masked to be logical).


def _s64(c: int) -> int:
    c &= (1 << 64) - 1
    return c - (1 << 64) if c >= (1 << 63) else c


def _shr(z, k: int):
    return (z >> k) & ((1 << (64 - k)) - 1)


def splitmix64_t(x):
    """splitmix64 on an int64 tensor holding uint64 bit patterns."""
    z = x + _s64(0x9E3779B97F4A7C15)
    z = (z ^ _shr(z, 30)) * _s64(0xBF58476D1CE4E5B9)
    z = (z ^ _shr(z, 27)) * _s64(0x94D049BB133111EB)
    return z ^ _shr(z, 31)


def _draws_t(seed: int, tile, purpose: int, n: int):
    import torch

    base = int(splitmix64(np.uint64(seed) * np.uint64(0x100000001B3) + np.uint64(purpose)))
    k = splitmix64_t(_s64(base) ^ (tile * _s64(0xD6E8FEB86659FD93)))
    h = splitmix64_t(k[:, None] + torch.arange(n, dtype=torch.int64, device=tile.device)[None, :])
    return _shr(h, 40).to(torch.float64) / float(1 << 24)


def _value_noise_t(lattice, freq: int, coord):
    import torch

    g = coord * freq
    i = torch.floor(g).to(torch.int64)
    t = g - i
    l00 = lattice[:, :, i[:, None], i[None, :]]
    l10 = lattice[:, :, i[:, None], i[None, :] + 1]
    l01 = lattice[:, :, i[:, None] + 1, i[None, :]]
    l11 = lattice[:, :, i[:, None] + 1, i[None, :] + 1]
    top = l00 + (l10 - l00) * t[None, None, None, :]
    bot = l01 + (l11 - l01) * t[None, None, None, :]
    return top + (bot - top) * t[None, None, :, None]


def tile_image_rgba8_torch(width: int, height: int, seed: int = 1, noise_frac: float = 0.0,
                           device="cuda"):
    """tile_image_rgba8 computed with torch on `device`: a (height, width, 4)
    uint8 tensor.  Equal to the numpy version up to the last-ulp behaviour of
    the device's sin/cos on edge tiles."""
    import math

    import torch

    if width % 8 or height % 8:
        raise ValueError("width and height must be multiples of 8")
    tw, th = width // 8, height // 8
    nt = tw * th
    f64 = torch.float64
    tiles = torch.arange(nt, dtype=torch.int64, device=device)
    kind_u = _draws_t(seed, tiles, 1, 2)
    u = kind_u[:, 1]
    kind = torch.where(u < 0.5, 0, torch.where(u < 0.8, 1, 2))
    kind = torch.where(kind_u[:, 0] < noise_frac, 3, kind)
    out = torch.zeros((nt, 8, 8, 3), dtype=f64, device=device)
    coord = (torch.arange(8, dtype=f64, device=device) + 0.5) / 8.0
    lin = torch.arange(8, dtype=f64, device=device) / 7.0

    sel = torch.nonzero(kind == 0).flatten()
    if sel.numel():
        c = _draws_t(seed, sel, 2, 12).reshape(-1, 4, 3) * 255.0
        fx = lin[None, None, :, None]
        fy = lin[None, :, None, None]
        top = c[:, None, None, 0] + (c[:, None, None, 1] - c[:, None, None, 0]) * fx
        bot = c[:, None, None, 2] + (c[:, None, None, 3] - c[:, None, None, 2]) * fx
        out[sel] = top + (bot - top) * fy

    sel = torch.nonzero(kind == 1).flatten()
    if sel.numel():
        par = _draws_t(seed, sel, 3, 4)
        base = par[:, :3] * 255.0
        amp = par[:, 3] * 255.0
        acc = torch.zeros((sel.numel(), 3, 8, 8), dtype=f64, device=device)
        norm = 0.0
        for o, f in enumerate((1, 2, 4, 8)):
            lat = _draws_t(seed, sel, 10 + o, 3 * (f + 1) * (f + 1)).reshape(-1, 3, f + 1, f + 1)
            w = 0.5 ** o
            acc += w * _value_noise_t(lat, f, coord)
            norm += w
        n = acc / norm - 0.5
        out[sel] = (base[:, :, None, None] + amp[:, None, None, None] * n).permute(0, 2, 3, 1)

    sel = torch.nonzero(kind == 2).flatten()
    if sel.numel():
        par = _draws_t(seed, sel, 4, 9)
        c0 = par[:, 0:3] * 255.0
        c1 = par[:, 3:6] * 255.0
        px = par[:, 6] * 8.0
        py = par[:, 7] * 8.0
        th_ = par[:, 8] * math.pi
        xs = torch.arange(8, dtype=f64, device=device) + 0.5
        d = (xs[None, None, :] - px[:, None, None]) * torch.sin(th_)[:, None, None] - (
            xs[None, :, None] - py[:, None, None]) * torch.cos(th_)[:, None, None]
        side = (d >= 0.0)[..., None]
        out[sel] = torch.where(side, c1[:, None, None, :], c0[:, None, None, :])

    sel = torch.nonzero(kind == 3).flatten()
    if sel.numel():
        out[sel] = _draws_t(seed, sel, 5, 192).reshape(-1, 8, 8, 3) * 256.0

    rgb = torch.clamp(torch.floor(out + 0.5), 0, 255).to(torch.uint8)
    img = torch.empty((th, 8, tw, 8, 4), dtype=torch.uint8, device=device)
    img[..., :3] = rgb.reshape(th, tw, 8, 8, 3).permute(0, 2, 1, 3, 4)
    img[..., 3] = 255
    return img.reshape(height, width, 4)
