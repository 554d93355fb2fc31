"""The image-level GPU decoders (decompress_blocks_as_rgba8 / _rgba16f /
_rgba32f, src/decode.rs:336-468) against the oracle.

* the BC6H -> RGBA8 sRGB8 code equals the oracle's linear_to_srgb on every
  non-negative finite half (exhaustive, through the kernel's own device
  function and an LDS copy of the library's bucket table);
* BC6H -> RGBA8 images (random blocks and a GPU-encoded HDR image) equal the
  oracle's decompress_blocks_as_rgba8 byte for byte;
* RGBA16F / RGBA32F with the reference layout equal the oracle element for
  element, including the untouched elements; the default RGBA layout holds
  the same RGB values with alpha 1.0;
* non-BC6H variants are rejected for the float outputs.
"""
import ctypes
import os

import numpy as np
import pytest

import oracle_lib
from block_compression_amd import synth
from block_compression_amd.settings import BC6HSettings, BC7Settings, CompressionVariant

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

FASTMATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fastmath", "_build",
                        "libfastmath_check.so")
V6 = CompressionVariant.BC6H(BC6HSettings.basic())


def test_srgb8_code_exhaustive_on_every_half():
    from block_compression_amd import decode

    print("decode-image: sRGB8 code sweep", flush=True)
    lib = ctypes.CDLL(FASTMATH)
    b = decode.bc6h_srgb8_buckets()
    n = 0x7C00
    code = np.zeros(n, np.uint8)
    rc = lib.srgb8_check(b.ctypes.data_as(ctypes.c_void_p), ctypes.c_uint32(n),
                         code.ctypes.data_as(ctypes.c_void_p))
    assert rc == 0
    want = oracle_lib.linear_to_srgb(np.arange(n, dtype=np.uint16).view(np.float16).astype(np.float32))
    assert np.array_equal(code, want)


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _encoded_hdr(w, h):
    from block_compression_amd import GpuBlockCompressor

    img = synth.hdr_image_rgba16f(w, h, 5)
    buf = torch.zeros(V6.blocks_byte_size(w, h), dtype=torch.uint8, device="cuda")
    c = GpuBlockCompressor()
    c.add_compression_task(V6, _dev(img), w, h, buf)
    c.compress()
    return buf


@pytest.mark.parametrize("source", ["random", "encoded"])
def test_bc6h_rgba8_matches_oracle(source):
    from block_compression_amd import decode

    print(f"decode-image: BC6H rgba8 {source}", flush=True)
    w, h = (256, 128) if source == "random" else (128, 128)
    if source == "random":
        blocks = _dev(np.random.default_rng(31).integers(0, 256, V6.blocks_byte_size(w, h), dtype=np.uint8))
    else:
        blocks = _encoded_hdr(w, h)
    out = torch.full((h, w, 4), 0xAB, dtype=torch.uint8, device="cuda")
    decode.decompress_blocks_as_rgba8(V6, w, h, blocks, out)
    torch.cuda.synchronize()
    want = oracle_lib.decompress_rgba8(5, w, h, blocks.cpu().numpy())
    assert np.array_equal(out.cpu().numpy(), want)


# (96, 64): 24 blocks per row, every wave spans rows (per-lane stores);
# (512, 16): whole waves inside one block row (row-run stores through LDS);
# (320, 8): 80 blocks per row, one wave of each kind, and a partial last wave.
@pytest.mark.parametrize("size", [(96, 64), (512, 16), (320, 8)])
@pytest.mark.parametrize("kind", ["16f", "32f"])
@pytest.mark.parametrize("source", ["random", "encoded"])
def test_bc6h_float_images_match_oracle(kind, source, size):
    from block_compression_amd import decode

    print(f"decode-image: BC6H rgba{kind} {source} {size}", flush=True)
    w, h = size
    if source == "random":
        blocks = _dev(np.random.default_rng(37).integers(0, 256, V6.blocks_byte_size(w, h), dtype=np.uint8))
    else:
        blocks = _encoded_hdr(w, h)
    nb = blocks.cpu().numpy()
    if kind == "16f":
        dt, it, fill, fn = torch.float16, np.uint16, 0x7E00, decode.decompress_blocks_as_rgba16f
        want = oracle_lib.decompress_rgba16f(5, w, h, nb, fill).view(it)
        one = 0x3C00
    else:
        dt, it, fill, fn = torch.float32, np.uint32, 0x7FC00001, decode.decompress_blocks_as_rgba32f
        want = oracle_lib.decompress_rgba32f(5, w, h, nb, fill).view(it)
        one = 0x3F800000
    tdt = torch.int16 if kind == "16f" else torch.int32
    # Reference layout over a NaN-filled buffer: element for element.
    out = torch.full((h, w, 4), fill, dtype=tdt, device="cuda").view(dt)
    fn(V6, w, h, blocks, out, reference_layout=True)
    torch.cuda.synchronize()
    assert np.array_equal(out.cpu().numpy().view(it), want)
    # Default RGBA layout: the same RGB per texel, alpha 1.0.
    rgba = torch.empty((h, w, 4), dtype=dt, device="cuda")
    fn(V6, w, h, blocks, rgba)
    torch.cuda.synchronize()
    got = rgba.cpu().numpy().view(it)
    rows = want.reshape(h, w // 4, 16)[:, :, :12].reshape(h, w, 3)
    assert np.array_equal(got[..., :3], rows)
    assert np.all(got[..., 3] == one)


def test_float_outputs_reject_other_variants():
    from block_compression_amd import BcnError, decode

    v = CompressionVariant.BC7(BC7Settings.alpha_basic())
    blocks = torch.zeros(v.blocks_byte_size(8, 8), dtype=torch.uint8, device="cuda")
    with pytest.raises(BcnError, match="UNSUPPORTED"):
        decode.decompress_blocks_as_rgba16f(v, 8, 8, blocks, torch.zeros((8, 8, 4), dtype=torch.float16, device="cuda"))
    with pytest.raises(BcnError, match="UNSUPPORTED"):
        decode.decompress_blocks_as_rgba32f(v, 8, 8, blocks, torch.zeros((8, 8, 4), dtype=torch.float32, device="cuda"))


def test_bc6h_2048_float_and_srgb_consistent():
    """At the 2048^2 HDR config: the RGBA32F decode is the exact widening of
    the RGBA16F decode, and a seeded sample of RGBA8 rows equals the oracle."""
    from block_compression_amd import decode

    print("decode-image: BC6H 2048 consistency", flush=True)
    w = h = 2048
    blocks = _encoded_hdr(w, h)
    f16 = torch.empty((h, w, 4), dtype=torch.float16, device="cuda")
    f32 = torch.empty((h, w, 4), dtype=torch.float32, device="cuda")
    s8 = torch.empty((h, w, 4), dtype=torch.uint8, device="cuda")
    decode.decompress_blocks_as_rgba16f(V6, w, h, blocks, f16)
    decode.decompress_blocks_as_rgba32f(V6, w, h, blocks, f32)
    decode.decompress_blocks_as_rgba8(V6, w, h, blocks, s8)
    torch.cuda.synchronize()
    assert torch.equal(f16.float(), f32)
    rows = np.random.default_rng(41).choice(h // 4, 16, replace=False)
    nb = blocks.cpu().numpy().reshape(h // 4, -1)
    got = s8.cpu().numpy()
    for by in rows:
        want = oracle_lib.decompress_rgba8(5, w, 4, nb[by])
        assert np.array_equal(got[by * 4:by * 4 + 4], want)


@pytest.mark.parametrize("is_signed", [False, True])
def test_bc6h_one_subset_waves_match_oracle(is_signed):
    """Waves of only one-subset blocks take the decoder's one-subset path,
    mixed waves the general one: random blocks with one-subset prefixes
    (modes 11-14 and a reserved prefix) in the first 4 waves of each row,
    mixed blocks after, decoded to RGBA16F, RGBA32F and RGBA8 and compared
    with the oracle texel for texel."""
    from block_compression_amd import decode

    print(f"decode-image: BC6H one-subset waves signed={is_signed}", flush=True)
    v = CompressionVariant.BC6H_SF16(BC6HSettings.basic()) if is_signed else V6
    w, h = 1024, 8  # 256 blocks per row: 4 waves
    rng = np.random.default_rng(43 + is_signed)
    blocks = rng.integers(0, 256, (h // 4, w // 4, 16), dtype=np.uint8)
    one = np.array([3, 7, 11, 15, 19], np.uint8)  # 19: a reserved prefix
    blocks[:, :192, 0] = (blocks[:, :192, 0] & ~np.uint8(31)) | one[rng.integers(0, 5, (h // 4, 192))]
    nb = blocks.reshape(-1)
    want16 = np.stack([oracle_lib.decode_block_bc6h(b.tobytes(), 12, is_signed)[:48]
                       for b in blocks.reshape(-1, 16)]).reshape(h // 4, w // 4, 4, 4, 3)
    want16 = want16.transpose(0, 2, 1, 3, 4).reshape(h, w, 3)
    dev = _dev(nb)
    f16 = torch.empty((h, w, 4), dtype=torch.float16, device="cuda")
    f32 = torch.empty((h, w, 4), dtype=torch.float32, device="cuda")
    s8 = torch.empty((h, w, 4), dtype=torch.uint8, device="cuda")
    decode.decompress_blocks_as_rgba16f(v, w, h, dev, f16)
    decode.decompress_blocks_as_rgba32f(v, w, h, dev, f32)
    decode.decompress_blocks_as_rgba8(v, w, h, dev, s8)
    torch.cuda.synchronize()
    g16 = f16.cpu().numpy().view(np.uint16)
    assert np.array_equal(g16[..., :3], want16)
    assert torch.equal(f16.float(), f32)
    # linear_to_srgb of each channel (negative values clamp to 0), alpha 0.
    lin = want16.view(np.float16).astype(np.float32)
    want8 = np.zeros((h, w, 4), np.uint8)
    want8[..., :3] = oracle_lib.linear_to_srgb(lin.reshape(-1)).reshape(h, w, 3)
    assert np.array_equal(s8.cpu().numpy(), want8)
    if not is_signed:
        assert np.array_equal(want8, oracle_lib.decompress_rgba8(5, w, h, nb))
