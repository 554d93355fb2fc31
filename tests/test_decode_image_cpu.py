"""Image-level decode drivers (src/decode.rs:151-468) without a GPU.

* the oracle's linear_to_srgb (decode.rs:151-159) is pinned by known answers
  and by the sRGB transfer function evaluated in float64 on every
  non-negative finite half float (the whole input domain of the BC6H ->
  RGBA8 path), and is monotone;
* the library's threshold table (bcn_bc6h_srgb8_thresholds) and the bucket
  entries built from it (bcn_bc6h_srgb8_buckets), which the GPU kernel reads,
  reproduce the oracle's code on that whole domain;
* the oracle's RGBA16F / RGBA32F image drivers keep the reference's layout:
  each block at its RGBA offset, its texels as packed RGB (texel j at
  element j*3), elements 12..15 of every block row untouched;
* the Python wrappers raise the reference's size asserts before any device
  call.
"""
import numpy as np
import pytest

import oracle_lib

torch = pytest.importorskip("torch")

HALVES = np.arange(0x7C00, dtype=np.uint16)  # every non-negative finite half


def _srgb_f64(x):
    v = np.asarray(x, np.float64)
    return np.where(v <= 0.0031308, v * 12.92, 1.055 * np.power(v, 1 / 2.4) - 0.055)


def test_linear_to_srgb_known_answers():
    # 0.5 -> 0.735357 * 255 = 187.52; 0.18 -> 0.461369 * 255 = 117.65;
    # 0.003 is on the linear segment: 0.003 * 12.92 * 255 = 9.88.
    got = oracle_lib.linear_to_srgb([0.0, 1.0, 0.5, 0.18, 0.003, 2.0, 65504.0])
    assert got.tolist() == [0, 255, 188, 118, 10, 255, 255]


def test_linear_to_srgb_matches_float64_formula_on_every_half():
    x = HALVES.view(np.float16).astype(np.float32)
    got = oracle_lib.linear_to_srgb(x).astype(np.int64)
    f = np.clip(_srgb_f64(x), 0.0, 1.0) * 255.0
    want = np.floor(f + 0.5).astype(np.int64)
    # An f32 evaluation may round differently only next to a .5 tie.
    near_tie = np.abs(f - np.floor(f) - 0.5) < 1e-4
    assert np.all((got == want) | near_tie)
    assert np.all(np.diff(got) >= 0)


def test_library_thresholds_reproduce_oracle_codes():
    from block_compression_amd import decode

    t = decode.bc6h_srgb8_thresholds()
    assert t[0] == 0 and np.all(np.diff(t.astype(np.int64)) >= 0) and t[255] <= 0x7BFF
    code = (HALVES[:, None] >= t[None, 1:]).sum(axis=1)
    want = oracle_lib.linear_to_srgb(HALVES.view(np.float16).astype(np.float32))
    assert np.array_equal(code, want)


def test_library_buckets_reproduce_oracle_codes():
    """The bucket entries the kernel reads (bcn_bc6h_srgb8_buckets): the high
    byte of entry[h >> 3] + 32 (h & 7) is the oracle's code of every
    non-negative finite half; each entry is 256 base + 32 (8 - off) with base
    the code of the bucket's first pattern."""
    from block_compression_amd import decode

    b = decode.bc6h_srgb8_buckets().astype(np.int64)
    h = HALVES.astype(np.int64)
    got = ((b[h >> 3] + 32 * (h & 7)) >> 8) & 0xFF
    want = oracle_lib.linear_to_srgb(HALVES.view(np.float16).astype(np.float32)).astype(np.int64)
    assert np.array_equal(got, want)
    assert np.array_equal(b >> 8, want[::8])
    assert np.all((b & 0xFF) % 32 == 0) and np.all((b & 0xFF) <= 7 * 32)


def test_bucket_entries_follow_from_thresholds():
    """The same entries rebuilt from the thresholds in numpy: base = the
    count of thresholds <= 8 i, off = the first j in 1..7 whose pattern
    reaches the next threshold (8 if none)."""
    from block_compression_amd import decode

    t = decode.bc6h_srgb8_thresholds().astype(np.int64)
    start = np.arange(3968, dtype=np.int64) * 8
    base = (start[:, None] >= t[None, 1:]).sum(axis=1)
    nxt = t[np.minimum(base + 1, 255)]
    off = np.where((base < 255) & (nxt - start < 8), nxt - start, 8)
    assert np.all(off >= 1)
    assert np.array_equal(decode.bc6h_srgb8_buckets().astype(np.int64), 256 * base + 32 * (8 - off))


def _random_bc6h(nblocks, seed):
    return np.random.default_rng(seed).integers(0, 256, (nblocks, 16), dtype=np.uint8)


@pytest.mark.parametrize("kind", ["16f", "32f"])
def test_oracle_float_drivers_keep_reference_layout(kind):
    w, h = 12, 8
    bw = w // 4
    blocks = _random_bc6h(bw * (h // 4), 3)
    fill = 0x7E00 if kind == "16f" else 0x7FC00001  # NaN patterns, never decoded
    if kind == "16f":
        img = oracle_lib.decompress_rgba16f(5, w, h, blocks.reshape(-1), fill).view(np.uint16)
    else:
        img = oracle_lib.decompress_rgba32f(5, w, h, blocks.reshape(-1), fill).view(np.uint32)
    rows = img.reshape(h, w * 4)
    for i, blk in enumerate(blocks):
        by, bx = divmod(i, bw)
        if kind == "16f":
            texels = oracle_lib.decode_block_bc6h(blk.tobytes(), 12, False).view(np.uint16)
        else:
            texels = oracle_lib.decode_block_bc6h_float(blk.tobytes(), 12, False).view(np.uint32)
        texels = texels.reshape(4, 12)  # 4 rows x (4 texels x RGB)
        for y in range(4):
            seg = rows[by * 4 + y, bx * 16:bx * 16 + 16]
            assert np.array_equal(seg[:12], texels[y])
            assert np.all(seg[12:] == fill)


def test_python_wrappers_raise_reference_size_asserts():
    from block_compression_amd import decode
    from block_compression_amd.settings import BC6HSettings, CompressionVariant

    v = CompressionVariant.BC6H(BC6HSettings.basic())
    blocks = torch.zeros(16 * 4, dtype=torch.uint8)
    with pytest.raises(ValueError, match="input bitstream"):
        decode.decompress_blocks_as_rgba16f(v, 8, 12, blocks, torch.zeros(8 * 12 * 4, dtype=torch.float16))
    with pytest.raises(ValueError, match="output slice"):
        decode.decompress_blocks_as_rgba16f(v, 8, 8, blocks, torch.zeros(8 * 8 * 3, dtype=torch.float16))
    with pytest.raises(ValueError, match="output slice"):
        decode.decompress_blocks_as_rgba32f(v, 8, 8, blocks, torch.zeros(8 * 8 * 4, dtype=torch.float16))
