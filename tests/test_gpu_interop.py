"""Texture interop on the GPU: the source layouts a texture view can carry
(src/block_compressor.rs:300-313), row-strided views, DLPack import, and the
batched queue over mixed layouts.

Each layout is checked against the oracle's encoder on the RGBA8 image the
reference's shaders would see through textureLoad: R8 -> (r, 0, 0, 255),
RG8 -> (r, g, 0, 255), BGRA8 -> channels swapped back; RGBA32F for BC6H ->
each channel rounded to the nearest half (ties to even), zeros as +0.
"""
import numpy as np
import pytest

import oracle_lib
from block_compression_amd import synth
from block_compression_amd.settings import BC6HSettings, BC7Settings, CompressionVariant

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

W, H = 64, 48


def _rgba(seed):
    img = synth.tile_image_rgba8(W, H, seed, 0.1)
    img[..., 3] = synth.tile_image_rgba8(W, H, seed + 1, 0.2)[..., 0]
    return img


def _variant(tag):
    if tag == 6:
        return CompressionVariant.BC7(BC7Settings.alpha_basic())
    return CompressionVariant(tag, None)


def _settings(tag):
    return BC7Settings.alpha_basic().to_bytes() if tag == 6 else b""


def _as_source(img, layout):
    """(device tensor in `layout`, the RGBA8 image the shaders would see)."""
    if layout == "r8":
        seen = np.zeros_like(img)
        seen[..., 0] = img[..., 0]
        seen[..., 3] = 255
        return torch.from_numpy(np.ascontiguousarray(img[..., 0])).cuda(), seen
    if layout == "rg8":
        seen = np.zeros_like(img)
        seen[..., :2] = img[..., :2]
        seen[..., 3] = 255
        return torch.from_numpy(np.ascontiguousarray(img[..., :2])).cuda(), seen
    if layout == "bgra8":
        bgra = np.ascontiguousarray(img[..., [2, 1, 0, 3]])
        return torch.from_numpy(bgra).cuda(), img
    return torch.from_numpy(img).cuda(), img


def _encode(variant, tex, w, h, **kw):
    from block_compression_amd import GpuBlockCompressor

    buf = torch.zeros(variant.blocks_byte_size(w, h), dtype=torch.uint8, device="cuda")
    c = GpuBlockCompressor()
    c.add_compression_task(variant, tex, w, h, buf, **kw)
    c.compress()
    torch.cuda.synchronize()
    return buf.cpu().numpy()


@pytest.mark.parametrize("layout", ["r8", "rg8", "bgra8"])
@pytest.mark.parametrize("tag", [0, 1, 2, 3, 4, 6], ids=["bc1", "bc2", "bc3", "bc4", "bc5", "bc7"])
def test_source_layouts_match_oracle(tag, layout):
    print(f"interop: tag {tag} {layout}", flush=True)
    tex, seen = _as_source(_rgba(3 + tag), layout)
    kw = {"texture_format": "bgra8"} if layout == "bgra8" else {}
    got = _encode(_variant(tag), tex, W, H, **kw)
    assert np.array_equal(got, oracle_lib.compress_rgba8(tag, _settings(tag), seen, threads=8))


@pytest.mark.parametrize("layout", ["rgba8", "r8", "rg8"])
def test_row_strided_view_with_y_offset(layout):
    """A sub-rectangle of a larger texture (row pitch of the parent) with a
    texture_y_offset of 8 rows."""
    print(f"interop: strided {layout}", flush=True)
    big = synth.tile_image_rgba8(160, 72, 9, 0.1)
    tex_big, seen_big = _as_source(big, layout)
    view = tex_big[:, 32:32 + W]
    assert not view.is_contiguous()
    got = _encode(CompressionVariant.BC7(BC7Settings.opaque_basic()), view, W, 48, texture_y_offset=8)
    want = oracle_lib.compress_rgba8(6, BC7Settings.opaque_basic().to_bytes(),
                                     np.ascontiguousarray(seen_big[8:56, 32:32 + W]), threads=8)
    assert np.array_equal(got, want)


def test_bc6h_rgba32f_source_matches_oracle():
    print("interop: BC6H rgba32f", flush=True)
    rng = np.random.default_rng(11)
    img = synth.hdr_image_rgba16f(W, H, 7).astype(np.float32)
    img *= rng.uniform(0.5, 2.0, img.shape).astype(np.float32)  # values between halves
    flat = img.reshape(-1, 4)
    flat[:40, :3] = 0.0
    flat[40:80, :3] = -0.0
    flat[80:120, :3] = rng.uniform(1e-7, 6e-5, (40, 3))     # half subnormals
    flat[120:140, :3] = rng.uniform(7e4, 1e6, (20, 3))      # beyond the half range
    flat[140:160, :3] = -rng.uniform(0.1, 4.0, (20, 3))     # negative values
    s = BC6HSettings.basic()
    got = _encode(CompressionVariant.BC6H(s), torch.from_numpy(img).cuda(), W, H)
    with np.errstate(over="ignore"):  # values beyond the half range become inf, as on the GPU
        half = np.where(img == 0.0, np.float32(0.0), img).astype(np.float16)
    assert np.array_equal(got, oracle_lib.compress_rgba16(s.to_bytes(), half, threads=8))


class _Foreign:
    """A device array from another framework: only the DLPack protocol."""

    def __init__(self, t):
        self._t = t

    def __dlpack__(self, *args, **kwargs):
        return self._t.__dlpack__(*args, **kwargs)

    def __dlpack_device__(self):
        return self._t.__dlpack_device__()


def test_dlpack_texture_and_buffer():
    print("interop: dlpack", flush=True)
    img = _rgba(21)
    v = CompressionVariant.BC7(BC7Settings.alpha_basic())
    from block_compression_amd import GpuBlockCompressor

    tex = torch.from_numpy(img).cuda()
    buf = torch.zeros(v.blocks_byte_size(W, H), dtype=torch.uint8, device="cuda")
    c = GpuBlockCompressor()
    c.add_compression_task(v, _Foreign(tex), W, H, _Foreign(buf))
    c.compress()
    torch.cuda.synchronize()
    assert np.array_equal(buf.cpu().numpy(), oracle_lib.compress_rgba8(6, _settings(6), img, threads=8))


def test_batch_over_mixed_layouts_matches_per_task():
    from block_compression_amd import GpuBlockCompressor

    print("interop: batch mixed layouts", flush=True)
    jobs = [(3, "r8"), (4, "rg8"), (6, "bgra8"), (0, "bgra8"), (6, "r8")]
    srcs, outs_q, outs_t = [], [], []
    cq, ct = GpuBlockCompressor(), GpuBlockCompressor()
    for i, (tag, layout) in enumerate(jobs):
        tex, _ = _as_source(_rgba(30 + i), layout)
        kw = {"texture_format": "bgra8"} if layout == "bgra8" else {}
        v = _variant(tag)
        bq = torch.zeros(v.blocks_byte_size(W, H), dtype=torch.uint8, device="cuda")
        bt = torch.zeros_like(bq)
        cq.add_compression_task(v, tex, W, H, bq, **kw)
        ct.add_compression_task(v, tex, W, H, bt, **kw)
        srcs.append(tex)
        outs_q.append(bq)
        outs_t.append(bt)
    batch = cq.create_batch()
    batch.launch()
    ct.compress()
    torch.cuda.synchronize()
    for bq, bt in zip(outs_q, outs_t):
        assert torch.equal(bq, bt)
    batch.close()


def test_layout_variant_checks():
    from block_compression_amd import BcnError, GpuBlockCompressor

    c = GpuBlockCompressor()
    buf = torch.zeros(4096, dtype=torch.uint8, device="cuda")
    with pytest.raises(BcnError, match="UNSUPPORTED"):
        c.add_compression_task(_variant(6), torch.zeros((8, 8, 4), dtype=torch.float32, device="cuda"), 8, 8, buf)
    # RGBA32F rows must keep 16-byte texels aligned: a view starting one
    # float into the allocation is refused.
    raw = torch.zeros(8 * 8 * 4 + 1, dtype=torch.float32, device="cuda")
    with pytest.raises(BcnError, match="MISALIGNED"):
        c.add_compression_task(CompressionVariant.BC6H(BC6HSettings.basic()),
                               raw[1:].view(8, 8, 4), 8, 8, buf)
    assert c.task_count() == 0
    # A view narrower than the region is refused before any launch (the C ABI
    # sees only a pointer and a pitch).
    big = torch.zeros((16, 64, 4), dtype=torch.uint8, device="cuda")
    with pytest.raises(ValueError, match="width"):
        c.add_compression_task(_variant(6), big[:, 16:32], 32, 16, buf)
    assert c.task_count() == 0
