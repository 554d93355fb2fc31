"""BC1-BC5 on the GPU vs the CPU oracle (src/encode/bc1_to_5.rs restated):
byte-identical blocks, including flat and extreme tiles, row-band tasks with
texture_y_offset / blocks_offset (tests/multi_tasks.rs) and odd widths."""
import numpy as np
import pytest

import oracle_lib
from block_compression_amd import synth
from block_compression_amd.settings import CompressionVariant

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu


def _image(w, h, seed):
    img = synth.tile_image_rgba8(w, h, seed, 0.1)
    img[..., 3] = synth.tile_image_rgba8(w, h, seed + 1, 0.2)[..., 1]
    img[:8, :8] = 255
    img[8:16, :8] = 0
    img[16:24, :8, :3] = 77
    return img


def _gpu(tag, img):
    from block_compression_amd import GpuBlockCompressor

    h, w, _ = img.shape
    v = CompressionVariant(tag)
    tex = torch.from_numpy(img).cuda()
    buf = torch.zeros(v.blocks_byte_size(w, h), dtype=torch.uint8, device="cuda")
    c = GpuBlockCompressor()
    c.add_compression_task(v, tex, w, h, buf)
    c.compress()
    torch.cuda.synchronize()
    return buf.cpu().numpy()


@pytest.mark.parametrize("tag", [0, 1, 2, 3, 4], ids=["bc1", "bc2", "bc3", "bc4", "bc5"])
def test_bc15_match_oracle(tag):
    img = _image(256, 128, 21)
    ref = oracle_lib.compress_rgba8(tag, b"", img, threads=8)
    got = _gpu(tag, img)
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("tag", [0, 3, 4], ids=["bc1", "bc4", "bc5"])
def test_bc15_full_size_sampled(tag):
    """The 4096^2 config texture: every block from the GPU, a seeded sample of
    8192 checked against the oracle."""
    img = synth.tile_image_rgba8(4096, 4096, 2, 0.1)
    got = _gpu(tag, img)
    bb = 8 if tag in (0, 3) else 16
    got = got.reshape(-1, bb)
    ids = np.random.default_rng(3).choice(got.shape[0], 8192, replace=False)
    ref = oracle_lib.compress_rgba8_sample(tag, b"", img, ids, threads=8).reshape(-1, bb)
    assert np.array_equal(got[ids], ref)


@pytest.mark.parametrize("tag", [0, 2, 3, 4, 6], ids=["bc1", "bc3", "bc4", "bc5", "bc7"])
def test_row_band_tasks(tag):
    """Two tasks per image, top and bottom halves, written at byte offsets of
    one shared buffer (tests/multi_tasks.rs:33-69): the result equals the
    whole-image encode.  The texture_y_offset is honoured for every format."""
    from block_compression_amd import GpuBlockCompressor
    from block_compression_amd.settings import BC7Settings

    img = _image(64, 64, 31)
    v = CompressionVariant(tag, BC7Settings.opaque_fast() if tag == 6 else None)
    w, h = 64, 64
    tex = torch.from_numpy(img).cuda()
    half = v.blocks_byte_size(w, h // 2)
    buf = torch.zeros(2 * half + 64, dtype=torch.uint8, device="cuda")
    c = GpuBlockCompressor()
    c.add_compression_task(v, tex, w, h // 2, buf, texture_y_offset=h // 2, blocks_offset=64 + half)
    c.add_compression_task(v, tex, w, h // 2, buf, texture_y_offset=0, blocks_offset=64)
    assert c.task_count() == 2
    c.compress()
    assert c.task_count() == 0
    torch.cuda.synchronize()
    got = buf.cpu().numpy()
    ref = oracle_lib.compress_rgba8(tag, v.settings_bytes(), img, threads=8)
    assert np.array_equal(got[64:], ref)
    assert not got[:64].any()


def test_task_preconditions_raise():
    """The reference's asserts (block_compressor.rs:352-373) as errors."""
    from block_compression_amd import BcnError, GpuBlockCompressor

    img = torch.zeros((8, 8, 4), dtype=torch.uint8, device="cuda")
    buf = torch.zeros(64, dtype=torch.uint8, device="cuda")
    c = GpuBlockCompressor()
    v = CompressionVariant.BC1()
    with pytest.raises(BcnError, match="NOT_MULTIPLE_OF_4"):
        c.add_compression_task(v, img, 6, 8, buf)
    with pytest.raises(BcnError, match="NOT_MULTIPLE_OF_4"):
        c.add_compression_task(v, img, 4, 4, buf, texture_y_offset=2)
    with pytest.raises(BcnError, match="BUFFER_TOO_SMALL"):
        c.add_compression_task(v, img, 8, 8, buf, blocks_offset=40)
    with pytest.raises(BcnError, match="TEXTURE_TOO_SMALL"):
        c.add_compression_task(v, img, 8, 8, buf, texture_y_offset=4)
    assert c.task_count() == 0
