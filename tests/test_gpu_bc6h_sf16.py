"""Signed BC6H (BC6H_SF16) on the GPU: the encoder against the oracle's
signed extension (pinned in tests/test_bc6h_sf16_cpu.py) on every preset,
from RGBA16F and RGBA32F sources, at the 2048^2 HDR size (sampled), and the
image-level decoders of signed blocks against the oracle's signed decoder."""
import numpy as np
import pytest

import oracle_lib
from block_compression_amd import synth
from block_compression_amd.settings import BC6H_PRESETS, BC6HSettings, CompressionVariant

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu


def _signed_image(w, h, seed):
    img = synth.hdr_image_rgba16f(w, h, seed).astype(np.float32)
    rng = np.random.default_rng(seed)
    # Smooth sign field: whole regions negative, plus a signed gradient channel.
    yy, xx = np.mgrid[0:h, 0:w]
    sign = np.where(np.sin(xx / 23.0) * np.cos(yy / 17.0) < 0, -1.0, 1.0).astype(np.float32)
    img[..., :3] *= sign[..., None]
    img[..., 2] = ((xx - w / 2) / w * 8.0).astype(np.float32) * (1 + 0.05 * rng.random((h, w)))
    return img.astype(np.float16)


def _encode(variant, tex, w, h):
    from block_compression_amd import GpuBlockCompressor

    buf = torch.zeros(variant.blocks_byte_size(w, h), dtype=torch.uint8, device="cuda")
    c = GpuBlockCompressor()
    c.add_compression_task(variant, tex, w, h, buf)
    c.compress()
    torch.cuda.synchronize()
    return buf


@pytest.mark.parametrize("preset", sorted(BC6H_PRESETS))
def test_sf16_presets_match_oracle(preset):
    print(f"sf16: preset {preset}", flush=True)
    img = _signed_image(128, 96, 3)
    s = BC6H_PRESETS[preset]()
    got = _encode(CompressionVariant.BC6H_SF16(s), torch.from_numpy(img).cuda(), 128, 96)
    want, _ = oracle_lib.compress_rgba16_sf16(s.to_bytes(), img, threads=8)
    assert np.array_equal(got.cpu().numpy(), want)


def test_sf16_rgba32f_source_matches_oracle():
    print("sf16: rgba32f source", flush=True)
    img32 = _signed_image(64, 64, 7).astype(np.float32) * np.float32(1.0007)
    s = BC6HSettings.basic()
    got = _encode(CompressionVariant.BC6H_SF16(s), torch.from_numpy(img32).cuda(), 64, 64)
    half = np.where(img32 == 0.0, np.float32(0.0), img32).astype(np.float16)
    want, _ = oracle_lib.compress_rgba16_sf16(s.to_bytes(), half, threads=8)
    assert np.array_equal(got.cpu().numpy(), want)


def test_sf16_2048_sampled_and_decoders():
    from block_compression_amd import decode

    print("sf16: 2048 sampled + decoders", flush=True)
    w = h = 2048
    img = _signed_image(w, h, 11)
    s = BC6HSettings.slow()
    v = CompressionVariant.BC6H_SF16(s)
    blocks = _encode(v, torch.from_numpy(img).cuda(), w, h)
    got = blocks.cpu().numpy().reshape(-1, 16)
    ids = np.random.default_rng(12).choice(got.shape[0], 1024, replace=False)
    for i in ids[:1024]:
        yy, xx = divmod(int(i), w // 4)
        tile = np.ascontiguousarray(img[yy * 4:yy * 4 + 4, xx * 4:xx * 4 + 4])
        want, _ = oracle_lib.compress_rgba16_sf16(s.to_bytes(), tile)
        assert np.array_equal(got[i], want), i
    # Image-level decodes of the signed blocks.
    f16 = torch.empty((h, w, 4), dtype=torch.float16, device="cuda")
    f32 = torch.empty((h, w, 4), dtype=torch.float32, device="cuda")
    s8 = torch.empty((h, w, 4), dtype=torch.uint8, device="cuda")
    decode.decompress_blocks_as_rgba16f(v, w, h, blocks, f16)
    decode.decompress_blocks_as_rgba32f(v, w, h, blocks, f32)
    decode.decompress_blocks_as_rgba8(v, w, h, blocks, s8)
    torch.cuda.synchronize()
    assert torch.equal(f16.float(), f32)
    f16n = f16.cpu().numpy().view(np.uint16)
    s8n = s8.cpu().numpy()
    for i in ids[:256]:
        yy, xx = divmod(int(i), w // 4)
        dec = oracle_lib.decode_block_bc6h(got[i].tobytes(), 12, True).view(np.uint16).reshape(4, 4, 3)
        assert np.array_equal(f16n[yy * 4:yy * 4 + 4, xx * 4:xx * 4 + 4, :3], dec)
        lin = dec.view(np.float16).astype(np.float32)
        want8 = oracle_lib.linear_to_srgb(lin).reshape(4, 4, 3)
        assert np.array_equal(s8n[yy * 4:yy * 4 + 4, xx * 4:xx * 4 + 4, :3], want8)
        assert np.all(s8n[yy * 4:yy * 4 + 4, xx * 4:xx * 4 + 4, 3] == 0)
