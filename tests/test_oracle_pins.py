"""Checks that pin the oracle itself (beyond the reference's decoder KATs).

The reference crate cannot be built or run here (no Rust toolchain), so the
oracle's restatement is pinned by independent implementations, closed
forms, brute force and exact invariants, each chosen so a plausible
transcription slip fails it:

* Pillow's BCn decoder (an independent implementation) on random blocks:
  BC7 (all modes), BC4, BC5 and the BC3 alpha block.  Pillow's BC1 colour
  rounding differs from bcdec's (block.rs:124-197), so BC1 colour is pinned
  by the KATs only.
* Encoder bitstream vs decoder: for every BC7 preset, decoding the emitted
  block with the oracle's decoder gives an SSE equal to the encoder's own
  best error for every block -- a wrong bit position, endpoint swap, anchor
  or p-bit in the packer breaks the equality.
* Closed forms: flat BC4/BC3-alpha blocks, BC2 alpha nibbles.
* partial_sort_list against a full sort; f32 -> f16 rounding against numpy.
"""
import numpy as np
import pytest

import oracle_lib
from block_compression_amd import synth
from block_compression_amd.settings import BC7_PRESETS

PIL = pytest.importorskip("PIL.Image")


def _pillow(blocks: np.ndarray, n: int, mode: str, w: int, h: int) -> np.ndarray:
    return np.array(PIL.frombytes(mode, (w, h), blocks.tobytes(), "bcn", n))


def _oracle_image(tag, blocks, w, h):
    return oracle_lib.decompress_rgba8(tag, w, h, blocks)


def test_bc7_decoder_matches_pillow_on_random_blocks_of_every_mode():
    rng = np.random.default_rng(101)
    w, h = 64, 64
    nb = (w // 4) * (h // 4)
    blocks = rng.integers(0, 256, (nb, 16), dtype=np.uint8)
    for i in range(nb):
        m = i % 8
        blocks[i, 0] = (blocks[i, 0] & (0xFF & ~((1 << (m + 1)) - 1))) | (1 << m)
    ref = _pillow(blocks, 7, "RGBA", w, h)
    assert np.array_equal(_oracle_image(6, blocks.reshape(-1), w, h), ref)


def test_bc4_bc5_decoders_match_pillow():
    rng = np.random.default_rng(102)
    w, h = 64, 32
    nb = (w // 4) * (h // 4)
    b4 = rng.integers(0, 256, nb * 8, dtype=np.uint8)
    assert np.array_equal(_oracle_image(3, b4, w, h)[..., 0], _pillow(b4, 4, "L", w, h))
    b5 = rng.integers(0, 256, nb * 16, dtype=np.uint8)
    got = _oracle_image(4, b5, w, h)
    ref = _pillow(b5, 5, "RGB", w, h)
    assert np.array_equal(got[..., :2], ref[..., :2])


def test_bc3_alpha_matches_pillow():
    rng = np.random.default_rng(103)
    w, h = 32, 32
    nb = (w // 4) * (h // 4)
    b3 = rng.integers(0, 256, nb * 16, dtype=np.uint8)
    assert np.array_equal(_oracle_image(2, b3, w, h)[..., 3], _pillow(b3, 3, "RGBA", w, h)[..., 3])


def _block_sse(img, dec, ch):
    h, w, _ = img.shape
    d = (dec.astype(np.int64) - img.astype(np.int64))[..., :ch]
    return (d * d).reshape(h // 4, 4, w // 4, 4, ch).transpose(0, 2, 1, 3, 4).reshape(-1, 16 * ch).sum(1)


@pytest.mark.parametrize("preset", sorted(BC7_PRESETS))
def test_bc7_encoder_error_equals_decoded_sse(preset):
    s = BC7_PRESETS[preset]()
    img = synth.tile_image_rgba8(64, 64, 21, 0.1)
    if s.channels == 4:
        img[..., 3] = synth.tile_image_rgba8(64, 64, 22, 0.2)[..., 0]
    blocks, errs = oracle_lib.bc7_block_errors(s.to_bytes(), img)
    dec = oracle_lib.decompress_rgba8(6, 64, 64, blocks)
    sse = _block_sse(img, dec, s.channels)
    assert np.array_equal(sse.astype(np.float32), errs)


@pytest.mark.parametrize("v", [0, 1, 77, 128, 254, 255])
def test_flat_bc4_block_closed_form(v):
    img = np.full((4, 4, 4), v, np.uint8)
    out = oracle_lib.compress_rgba8(3, b"", img)
    # ep1 = v + 0.1 truncates to v; every texel gets index 1 (bc1_to_5.rs:441-470).
    assert out.tolist() == [v, v, 0x49, 0x92, 0x24, 0x49, 0x92, 0x24]


def test_bc2_alpha_closed_form():
    a = np.arange(256, dtype=np.uint8)
    img = np.zeros((4 * 16, 4 * 4, 4), np.uint8)
    img[..., 3] = np.resize(a, (64, 16))
    out = oracle_lib.compress_rgba8(1, b"", img).reshape(-1, 16)
    nib = np.array([[(row[i // 2] >> (4 * (i % 2))) & 15 for i in range(16)] for row in out[:, :8]])
    expect = (img[..., 3].reshape(16, 4, 4, 4).transpose(0, 2, 1, 3).reshape(-1, 16).astype(np.float32)
              / np.float32(255.0) * np.float32(15.0)).astype(np.int64)
    assert np.array_equal(nib, expect)


def test_partial_sort_list_brute_force():
    rng = np.random.default_rng(7)
    for _ in range(200):
        n = int(rng.integers(1, 65))
        keys = rng.choice(100000, n, replace=False).astype(np.int32)
        k = int(rng.integers(0, n + 1))
        got = oracle_lib.partial_sort_list(keys, k)
        assert np.array_equal(got[:k], np.sort(keys)[:k])
        assert sorted(got.tolist()) == sorted(keys.tolist())


def test_f32_to_f16_matches_numpy_rne():
    rng = np.random.default_rng(8)
    vals = np.concatenate([
        rng.standard_normal(20000).astype(np.float32) * np.float32(1000.0),
        np.exp2(rng.uniform(-30, 17, 20000)).astype(np.float32),
        np.array([0.0, -0.0, 65504.0, 65520.0, 1e-8, 5.96e-8, 2.98e-8, 6.1e-5], np.float32),
    ])
    lib = oracle_lib.lib()
    got = np.array([lib.oracle_f32_to_f16(float(v)) for v in vals], np.uint16)
    with np.errstate(over="ignore"):  # values past 65504 round to +-inf on purpose
        want = vals.astype(np.float16)
    assert np.array_equal(got, want.view(np.uint16))
    back = np.array([lib.oracle_f16_to_f32(int(h)) for h in got[:5000]], np.float32)
    assert np.array_equal(back, got[:5000].view(np.float16).astype(np.float32))
