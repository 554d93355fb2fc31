"""Signed BC6H (BC6H_SF16) encode without a GPU.

The reference's encoder is unsigned only; its decoder reads signed blocks
(src/decode/block.rs:266, 697-719), and the oracle's decoder is pinned to the
reference's signed known-answer tests.  The signed encoder (the oracle's
extension, oracle/bc6h.cpp `sgn`, and the kernels' encode_bc6h_block<true>)
is pinned by:

* decode(encode(x)) reproducing the encoder's own error on every block of
  every preset, for positive, negative, smooth signed and mixed-sign images:
  the decoder's interpolated values before finish_unquantize against the
  texel values +-(|h|/31)*32 (a wrong sign, field, delta or rounding in the
  packing breaks the equality);
* exact round trips of constant blocks whose |h| is a multiple of 31 (their
  value is on the 16-bit endpoint grid, so the zero-error mode 13 is found);
* quality: on non-negative input the signed encoding loses at most one bit
  of endpoint precision, so its error stays close to the unsigned one; signs
  of a smooth signed field survive;
* the host build of the kernels' code equals the oracle byte for byte.
"""
import numpy as np
import pytest

import kernel_check_lib as kc
import oracle_lib
from block_compression_amd import synth
from block_compression_amd.settings import BC6H_PRESETS, BC6HSettings


def _images():
    img = synth.hdr_image_rgba16f(64, 64, 3)
    x = np.linspace(-4, 4, 64, dtype=np.float32)
    smooth = np.zeros((64, 64, 4), np.float32)
    smooth[..., 0] = x[None, :]
    smooth[..., 1] = x[:, None]
    smooth[..., 2] = np.sin(3 * x)[None, :] * x[:, None]
    rng = np.random.default_rng(5)
    mixed = (img.astype(np.float32) * np.where(rng.random((64, 64, 1)) < 0.5, -1, 1)).astype(np.float16)
    return {"positive": img, "negative": (-img.astype(np.float32)).astype(np.float16),
            "smooth_signed": smooth.astype(np.float16), "mixed": mixed}


def _blocks_of(img):
    h, w, _ = img.shape
    bits = img.view(np.uint16).astype(np.int64)
    return bits.reshape(h // 4, 4, w // 4, 4, 4).transpose(0, 2, 1, 3, 4).reshape(-1, 16, 4)[..., :3]


def _signed_values(bits):
    mag = (bits & 0x7FFF).astype(np.float32)
    v = (mag / np.float32(31)) * np.float32(32)
    return np.where(bits & 0x8000, -v, v).astype(np.float64)


@pytest.mark.parametrize("name", ["positive", "negative", "smooth_signed", "mixed"])
@pytest.mark.parametrize("preset", sorted(BC6H_PRESETS))
def test_decoder_reproduces_encoder_error(name, preset):
    img = _images()[name]
    blocks, errs = oracle_lib.compress_rgba16_sf16(BC6H_PRESETS[preset]().to_bytes(), img, threads=8)
    comp = np.stack([oracle_lib.decode_block_bc6h_comp(b.tobytes(), True)
                     for b in blocks.reshape(-1, 16)]).astype(np.float64)
    err = ((comp - _signed_values(_blocks_of(img))) ** 2).sum(axis=(1, 2))
    # f32 accumulation in the encoder: relative 1e-5.
    assert np.all(np.abs(err - errs) <= 1e-5 * np.maximum(errs, 1.0))


def test_constant_blocks_on_the_grid_round_trip_exactly():
    rng = np.random.default_rng(8)
    n = 256
    mags = rng.integers(0, 0x7BFF // 31 + 1, (n, 3)) * 31
    signs = rng.integers(0, 2, (n, 3)) << 15
    vals = (mags | signs).astype(np.uint16)
    img = np.zeros((4, 4 * n, 4), np.uint16)
    for i in range(n):
        img[:, 4 * i:4 * i + 4, :3] = vals[i]
    blocks, errs = oracle_lib.compress_rgba16_sf16(BC6HSettings.slow().to_bytes(), img.view(np.float16))
    assert np.all(errs == 0)
    for i, b in enumerate(blocks.reshape(-1, 16)):
        dec = oracle_lib.decode_block_bc6h(b.tobytes(), 12, True).view(np.uint16).reshape(16, 3)
        # -0 decodes as +0 (a zero magnitude has no sign in the encoding).
        want = np.where((vals[i] & 0x7FFF) == 0, 0, vals[i])
        assert np.all(dec == want[None, :]), i


def _decoded_f64(blocks, signed):
    return np.stack([oracle_lib.decode_block_bc6h(b.tobytes(), 12, signed).view(np.float16).reshape(16, 3)
                     for b in blocks.reshape(-1, 16)]).astype(np.float64)


def test_quality_against_unsigned_and_sign_survival():
    imgs = _images()
    s = BC6HSettings.slow().to_bytes()
    pos = imgs["positive"]
    ref = pos.reshape(16, 4, 16, 4, 4).transpose(0, 2, 1, 3, 4).reshape(-1, 16, 4)[..., :3].astype(np.float64)
    sb, _ = oracle_lib.compress_rgba16_sf16(s, pos, threads=8)
    ub = oracle_lib.compress_rgba16(s, pos, threads=8)
    rel_s = np.median(np.abs(_decoded_f64(sb, True) - ref) / (np.abs(ref) + 1e-3))
    rel_u = np.median(np.abs(_decoded_f64(ub, False) - ref) / (np.abs(ref) + 1e-3))
    assert rel_s <= 1.1 * rel_u + 1e-4
    sm = imgs["smooth_signed"]
    ref = sm.reshape(16, 4, 16, 4, 4).transpose(0, 2, 1, 3, 4).reshape(-1, 16, 4)[..., :3].astype(np.float64)
    b, _ = oracle_lib.compress_rgba16_sf16(s, sm, threads=8)
    dec = _decoded_f64(b, True)
    big = np.abs(ref) > 0.05
    assert np.all(np.sign(dec[big]) == np.sign(ref[big]))


@pytest.mark.parametrize("name", ["positive", "negative", "smooth_signed", "mixed"])
def test_host_build_of_kernel_code_equals_oracle(name):
    img = _images()[name]
    for preset, make in sorted(BC6H_PRESETS.items()):
        want, _ = oracle_lib.compress_rgba16_sf16(make().to_bytes(), img, threads=8)
        assert np.array_equal(kc.bc6h_encode(make().to_bytes(), img, True), want), preset
