"""BC6H from every source layout on the GPU.

The replaced GPU path reads any texture view with textureLoad and converts
each channel with f32_to_f16_bits (src/block_compressor.rs:312,
bc6h.wgsl:59-85): an 8-bit unorm view gives c/255 in f32, rounded to the
nearest half (ties to even); every zero -- -0 included -- becomes +0.  The
oracle's encoder (compress_rgba16, encode.rs:211-251) is run on the half
image those rules give, and the GPU output must equal it byte for byte.
"""
import numpy as np
import pytest

import oracle_lib
from block_compression_amd import GpuBlockCompressor, synth
from block_compression_amd.settings import BC6HSettings, CompressionVariant

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

W, H = 64, 48


def _encode(variant, tex, w=W, h=H):
    buf = torch.zeros(variant.blocks_byte_size(w, h), dtype=torch.uint8, device="cuda")
    c = GpuBlockCompressor()
    c.add_compression_task(variant, tex, w, h, buf)
    c.compress()
    torch.cuda.synchronize()
    return buf.cpu().numpy()


def _oracle(signed, settings, half_img):
    if signed:
        return oracle_lib.compress_rgba16_sf16(settings.to_bytes(), half_img, threads=8)[0]
    return oracle_lib.compress_rgba16(settings.to_bytes(), half_img, threads=8)


def _unorm_to_half(img_u8):
    return (img_u8.astype(np.float32) / np.float32(255.0)).astype(np.float16)


@pytest.mark.parametrize("signed", [False, True], ids=["uf16", "sf16"])
@pytest.mark.parametrize("preset", ["fast", "basic", "slow"])
@pytest.mark.parametrize("layout", ["rgba8", "bgra8", "r8", "rg8"])
def test_bc6h_from_unorm8_layouts_matches_oracle(layout, preset, signed):
    print(f"bc6h_sources: {layout} {preset} signed={signed}", flush=True)
    img = synth.tile_image_rgba8(W, H, 71, 0.1)
    img[0, :16, :3] = np.arange(48).reshape(16, 3) * 5 + 10   # a spread of values
    img[4, :4, :3] = 0
    seen = img.copy()
    if layout == "r8":
        seen[..., 1:3] = 0
        tex = torch.from_numpy(np.ascontiguousarray(img[..., 0])).cuda()
    elif layout == "rg8":
        seen[..., 2] = 0
        tex = torch.from_numpy(np.ascontiguousarray(img[..., :2])).cuda()
    elif layout == "bgra8":
        tex = torch.from_numpy(np.ascontiguousarray(img[..., [2, 1, 0, 3]])).cuda()
    else:
        tex = torch.from_numpy(img).cuda()
    s = getattr(BC6HSettings, preset)()
    variant = CompressionVariant.BC6H_SF16(s) if signed else CompressionVariant.BC6H(s)
    fmt = {"bgra8": "bgra8"}.get(layout)
    buf = torch.zeros(variant.blocks_byte_size(W, H), dtype=torch.uint8, device="cuda")
    c = GpuBlockCompressor()
    c.add_compression_task(variant, tex, W, H, buf, texture_format=fmt)
    c.compress()
    torch.cuda.synchronize()
    assert np.array_equal(buf.cpu().numpy(), _oracle(signed, s, _unorm_to_half(seen)))


@pytest.mark.parametrize("signed", [False, True], ids=["uf16", "sf16"])
def test_bc6h_negative_zero_reads_as_plus_zero_in_every_float_layout(signed):
    print(f"bc6h_sources: -0 halves signed={signed}", flush=True)
    img = synth.hdr_image_rgba16f(W, H, 72).copy()
    bits = img.view(np.uint16)
    rng = np.random.default_rng(73)
    z = rng.random(bits.shape) < 0.2
    z[..., 3] = False
    bits[z] = 0x8000                       # -0 in RGB
    bits[:4, :4, :3] = 0x8000              # a whole block of -0
    bits[4:8, :4, :3] = np.where(rng.random((4, 4, 3)) < 0.5, 0x8000, 0)
    s = BC6HSettings.slow()
    variant = CompressionVariant.BC6H_SF16(s) if signed else CompressionVariant.BC6H(s)
    g16 = _encode(variant, torch.from_numpy(img).cuda())
    g32 = _encode(variant, torch.from_numpy(img.astype(np.float32)).cuda())
    assert np.array_equal(g16, g32)
    plus = bits.copy()
    plus[plus == 0x8000] = 0
    assert np.array_equal(g16, _oracle(signed, s, plus.view(np.float16)))
