"""BC6H encode and all GPU decoders against the oracle."""
import numpy as np
import pytest

import oracle_lib
from block_compression_amd import synth
from block_compression_amd.settings import BC6H_PRESETS, BC6HSettings, BC7Settings, CompressionVariant

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu


def _gpu_bc6h(img16, settings):
    from block_compression_amd import GpuBlockCompressor

    h, w, _ = img16.shape
    v = CompressionVariant.BC6H(settings)
    tex = torch.from_numpy(img16).cuda()
    buf = torch.zeros(v.blocks_byte_size(w, h), dtype=torch.uint8, device="cuda")
    c = GpuBlockCompressor()
    c.add_compression_task(v, tex, w, h, buf)
    c.compress()
    torch.cuda.synchronize()
    return buf.cpu().numpy()


@pytest.mark.parametrize("preset", sorted(BC6H_PRESETS))
def test_bc6h_presets_match_oracle(preset):
    img = synth.hdr_image_rgba16f(128, 128, 3)
    s = BC6H_PRESETS[preset]()
    ref = oracle_lib.compress_rgba16(s.to_bytes(), img, threads=8)
    assert np.array_equal(_gpu_bc6h(img, s), ref)


def test_bc6h_slow_2048_sampled():
    """The 2048^2 HDR config (BC6H slow): every block on the GPU, a seeded
    sample of 2048 blocks checked against the oracle."""
    img = synth.hdr_image_rgba16f(2048, 2048, 3)
    s = BC6HSettings.slow()
    got = _gpu_bc6h(img, s).reshape(-1, 16)
    ids = np.random.default_rng(9).choice(got.shape[0], 2048, replace=False)
    bits = img.view(np.uint16)
    h, w, _ = bits.shape
    ref = []
    for i in ids:
        yy, xx = divmod(int(i), w // 4)
        tile = np.ascontiguousarray(img[yy * 4:yy * 4 + 4, xx * 4:xx * 4 + 4])
        ref.append(oracle_lib.compress_rgba16(s.to_bytes(), tile))
    assert np.array_equal(got[ids], np.stack(ref))


@pytest.mark.parametrize("tag,name", [(0, "bc1"), (1, "bc2"), (2, "bc3"), (3, "bc4"), (4, "bc5"), (6, "bc7")])
def test_gpu_decode_matches_oracle_random_blocks(tag, name):
    from block_compression_amd import decode_rgba8

    w, h = 64, 48
    bb = 8 if tag in (0, 3) else 16
    nb = (w // 4) * (h // 4)
    blocks = np.random.default_rng(tag).integers(0, 256, nb * bb, dtype=np.uint8)
    v = CompressionVariant(tag, BC7Settings.opaque_basic() if tag == 6 else None)
    got = decode_rgba8(v, torch.from_numpy(blocks).cuda(), w, h).cpu().numpy()
    ref = oracle_lib.decompress_rgba8(tag, w, h, blocks)
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("tag,name", [(0, "bc1"), (3, "bc4"), (6, "bc7")])
def test_gpu_decode_partial_waves(tag, name):
    """100x12 (25 blocks a row, 75 blocks): the last wave is partly empty and
    block rows end inside waves, against the oracle's decoder."""
    from block_compression_amd import decode_rgba8

    w, h = 100, 12
    bb = 8 if tag in (0, 3) else 16
    nb = (w // 4) * (h // 4)
    blocks = np.random.default_rng(90 + tag).integers(0, 256, nb * bb, dtype=np.uint8)
    v = CompressionVariant(tag, BC7Settings.opaque_basic() if tag == 6 else None)
    got = decode_rgba8(v, torch.from_numpy(blocks).cuda(), w, h).cpu().numpy()
    ref = oracle_lib.decompress_rgba8(tag, w, h, blocks)
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("mode", list(range(9)))
def test_gpu_decode_bc7_one_mode_per_image(mode):
    """Random BC7 blocks all of one mode (8: a first byte of 0), so every wave
    is uniform: the decoder's paths that skip subsets 1 and 2 for waves
    without 2- or 3-subset blocks run, against the oracle's decoder."""
    from block_compression_amd import decode_rgba8

    w, h = 256, 64  # 1,024 blocks: 16 waves
    nb = (w // 4) * (h // 4)
    blocks = np.random.default_rng(70 + mode).integers(0, 256, (nb, 16), dtype=np.uint8)
    if mode < 8:
        blocks[:, 0] = (blocks[:, 0] & np.uint8((0xFF << (mode + 1)) & 0xFF)) | np.uint8(1 << mode)
    else:
        blocks[:, 0] = 0
    blocks = blocks.reshape(-1)
    v = CompressionVariant(6, BC7Settings.opaque_basic())
    got = decode_rgba8(v, torch.from_numpy(blocks).cuda(), w, h).cpu().numpy()
    ref = oracle_lib.decompress_rgba8(6, w, h, blocks)
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("is_signed", [False, True])
def test_gpu_decode_bc6h_random_blocks(is_signed):
    from block_compression_amd import decode_bc6h_rgba16f

    w, h = 32, 32
    nb = (w // 4) * (h // 4)
    blocks = np.random.default_rng(60 + is_signed).integers(0, 256, nb * 16, dtype=np.uint8)
    got = decode_bc6h_rgba16f(torch.from_numpy(blocks).cuda(), w, h, is_signed).cpu().numpy()
    got_bits = got.view(np.uint16)
    for i, b in enumerate(blocks.reshape(-1, 16)):
        yy, xx = divmod(i, w // 4)
        ref = oracle_lib.decode_block_bc6h(b.tobytes(), 12, is_signed)[:48].reshape(4, 4, 3)
        tile = got_bits[yy * 4:yy * 4 + 4, xx * 4:xx * 4 + 4]
        assert np.array_equal(tile[..., :3], ref)
        assert (tile[..., 3] == 0x3C00).all()


@pytest.mark.parametrize("is_signed", [False, True])
def test_gpu_decode_bc6h_one_subset_waves(is_signed):
    """Random BC6H blocks all of the one-subset modes (5-bit prefixes 0x03,
    0x07, 0x0B, 0x0F), so every wave takes the decoders' one-subset path:
    RGBA16F against the oracle's decoder, bit for bit."""
    from block_compression_amd import decode_bc6h_rgba16f

    w, h = 128, 64  # 512 blocks: 8 waves
    nb = (w // 4) * (h // 4)
    rng = np.random.default_rng(80 + is_signed)
    blocks = rng.integers(0, 256, (nb, 16), dtype=np.uint8)
    prefix = rng.choice(np.array([0x03, 0x07, 0x0B, 0x0F], dtype=np.uint8), nb)
    blocks[:, 0] = (blocks[:, 0] & np.uint8(0xE0)) | prefix
    got = decode_bc6h_rgba16f(torch.from_numpy(blocks.reshape(-1)).cuda(), w, h, is_signed).cpu().numpy()
    got_bits = got.view(np.uint16)
    for i, b in enumerate(blocks):
        yy, xx = divmod(i, w // 4)
        ref = oracle_lib.decode_block_bc6h(b.tobytes(), 12, is_signed)[:48].reshape(4, 4, 3)
        assert np.array_equal(got_bits[yy * 4:yy * 4 + 4, xx * 4:xx * 4 + 4, :3], ref)


def test_gpu_encode_decode_roundtrip_bc7_4096():
    """Size-independent property at the headline size: GPU decode of the GPU
    encode equals the oracle decoder's output on the same blocks (sampled)."""
    from block_compression_amd import GpuBlockCompressor, decode_rgba8

    img = synth.tile_image_rgba8(4096, 4096, 2, 0.1)
    v = CompressionVariant.BC7(BC7Settings.opaque_basic())
    tex = torch.from_numpy(img).cuda()
    buf = torch.zeros(v.blocks_byte_size(4096, 4096), dtype=torch.uint8, device="cuda")
    c = GpuBlockCompressor()
    c.add_compression_task(v, tex, 4096, 4096, buf)
    c.compress()
    dec = decode_rgba8(v, buf, 4096, 4096).cpu().numpy()
    blocks = buf.cpu().numpy().reshape(-1, 16)
    ids = np.random.default_rng(1).choice(blocks.shape[0], 2000, replace=False)
    for i in ids:
        yy, xx = divmod(int(i), 1024)
        ref = oracle_lib.decode_block("bc7", blocks[i].tobytes(), 16, 64).reshape(4, 4, 4)
        assert np.array_equal(dec[yy * 4:yy * 4 + 4, xx * 4:xx * 4 + 4], ref)
    # Opaque preset: decoded alpha is 255 or (mode 6 quirk) 254.
    assert set(np.unique(dec[..., 3]).tolist()) <= {254, 255}
