"""The benchmark configurations at their full sizes on the GPU.

C1: BC1 512x512 (tile generator seed 1, no noise tiles): every block equals
the oracle, the GPU decode equals the oracle's decoder, and the reference
harness's PSNR (tests/metrics.rs:40-114) of the two is the same.

C2 (BC4 from R, BC5 from RG, 4096x4096, seed 2 with 10% noise tiles): every
block equals the oracle.  (BC7 opaque_basic on the same texture is checked
block by block in test_gpu_bc7.py.)

C3: BC6H slow on the 2048x2048 HDR texture: every one of the 262,144 blocks
equals the oracle.

C4: 300 tile-generator textures with full box-filtered mip chains in mixed
formats (block_compression_amd.workloads, 3,513 levels, about 1.7 Gpx): the
batched persistent launch equals one launch per level on every byte, and a
seeded sample of small and large levels equals the oracle.

C5: 1024x1024 gradient-refined BC7: BC7 alpha_slow on the target rounded to
RGBA8 (a sample checked against the oracle), then 64 Adam steps (lr 0.5)
with index reassignment and p-bit trials every 4 steps.  No block gets
worse; the reported errors equal a recount from the decoded blocks; a
sample of blocks equals the host build of the same refinement code.
"""
import numpy as np
import pytest

import kernel_check_lib as kc
import oracle_lib
from block_compression_amd import GpuBlockCompressor, grad, synth, workloads
from block_compression_amd.settings import BC7Settings, CompressionVariant

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu


def _encode(variant, tex, w, h):
    buf = torch.zeros(variant.blocks_byte_size(w, h), dtype=torch.uint8, device="cuda")
    c = GpuBlockCompressor()
    c.add_compression_task(variant, tex, w, h, buf)
    c.compress()
    torch.cuda.synchronize()
    return buf


def test_c1_bc1_512_every_block_decode_and_psnr():
    from block_compression_amd import decode_rgba8
    from metrics_ref import calculate_image_metrics_rgba8

    print("configs: C1 BC1 512^2", flush=True)
    img = synth.tile_image_rgba8(512, 512, 1, 0.0)
    v = CompressionVariant.BC1()
    buf = _encode(v, torch.from_numpy(img).cuda(), 512, 512)
    got = buf.cpu().numpy()
    ref = oracle_lib.compress_rgba8(v.tag, v.settings_bytes(), img, threads=8)
    assert np.array_equal(got, ref)
    dec = decode_rgba8(v, buf, 512, 512).cpu().numpy()
    want = oracle_lib.decompress_rgba8(v.tag, 512, 512, ref)
    assert np.array_equal(dec, want)
    p_gpu = calculate_image_metrics_rgba8(img, dec, 3)["overall_psnr"]
    p_ref = calculate_image_metrics_rgba8(img, want, 3)["overall_psnr"]
    assert p_gpu == p_ref and p_ref > 25.0


@pytest.mark.parametrize("fmt", ["bc4", "bc5"])
def test_c2_bc4_bc5_4096_every_block(fmt):
    print(f"configs: C2 {fmt} 4096^2", flush=True)
    img = synth.tile_image_rgba8(4096, 4096, 2, 0.1)
    v = CompressionVariant.BC4() if fmt == "bc4" else CompressionVariant.BC5()
    got = _encode(v, torch.from_numpy(img).cuda(), 4096, 4096).cpu().numpy()
    assert np.array_equal(got, oracle_lib.compress_rgba8(v.tag, v.settings_bytes(), img, threads=16))


def test_c3_bc6h_slow_2048_every_block():
    from block_compression_amd.settings import BC6HSettings

    print("configs: C3 BC6H slow 2048^2 (oracle on every block)", flush=True)
    img = synth.hdr_image_rgba16f(2048, 2048, 3)
    s = BC6HSettings.slow()
    v = CompressionVariant.BC6H(s)
    tex = torch.from_numpy(img.view(np.uint16)).cuda().view(torch.float16)
    got = _encode(v, tex, 2048, 2048).cpu().numpy()
    assert np.array_equal(got, oracle_lib.compress_rgba16(s.to_bytes(), img, threads=16))


def test_c4_batched_mip_chains_full_size():
    print("configs: C4 build 300 textures", flush=True)
    levels = workloads.build(300, seed=4, device="cuda")
    assert len(levels) == 3513
    total = workloads.output_bytes(levels)
    plain = torch.full((total,), 0xAB, dtype=torch.uint8, device="cuda")
    batched = torch.full((total,), 0xCD, dtype=torch.uint8, device="cuda")
    print("configs: C4 per-level launches", flush=True)
    c = GpuBlockCompressor()
    workloads.queue(c, levels, plain)
    c.compress()
    print("configs: C4 batched launch (graph)", flush=True)
    workloads.queue(c, levels, batched)
    batch = c.create_batch()
    batch.launch(graph=True)
    batch.launch(graph=True)
    torch.cuda.synchronize()
    out = batched.cpu().numpy()
    ref = plain.cpu().numpy()
    batch.close()
    # Levels sit at 16-byte aligned offsets; the padding after an 8-byte-block
    # level is written by neither path and keeps its fill.
    for lv in levels:
        a, b = out[lv.offset:lv.offset + lv.nbytes], ref[lv.offset:lv.offset + lv.nbytes]
        assert np.array_equal(a, b), (lv.texture, lv.level, lv.fmt)
    assert not (out[levels[-1].offset:levels[-1].offset + levels[-1].nbytes] == 0xCD).all()
    rng = np.random.default_rng(11)
    small = [i for i, lv in enumerate(levels) if lv.width * lv.height <= 64 * 64]
    big = [i for i, lv in enumerate(levels) if lv.width * lv.height > 64 * 64]
    pick = list(rng.choice(small, 48, replace=False)) + list(rng.choice(big, 6, replace=False))
    fmts = set()
    for i in pick:
        lv = levels[int(i)]
        v = workloads.variant_of(lv.fmt)
        fmts.add(lv.fmt)
        img = lv.texels.cpu().numpy()
        nb = (lv.width // 4) * (lv.height // 4)
        ids = np.arange(nb) if nb <= 8192 else np.sort(rng.choice(nb, 8192, replace=False))
        ref = oracle_lib.compress_rgba8_sample(v.tag, v.settings_bytes(), img, ids, threads=16)
        bb = v.block_byte_size()
        got = out[lv.offset:lv.offset + lv.nbytes].reshape(-1, bb)[ids]
        assert np.array_equal(got, ref.reshape(-1, bb)), (lv.texture, lv.level, lv.fmt)
    assert len(fmts) >= 5


def test_c5_refine_1024_alpha_slow():
    n = 1024
    print("configs: C5 plain alpha_slow 1024^2", flush=True)
    target = synth.tile_target_rgba_f32(n, n, 5, 0.0)
    img = np.clip(np.floor(target * 255 + 0.5), 0, 255).astype(np.uint8)
    s = BC7Settings.alpha_slow()
    v = CompressionVariant.BC7(s)
    blocks = torch.zeros(v.blocks_byte_size(n, n), dtype=torch.uint8, device="cuda")
    c = GpuBlockCompressor()
    c.add_compression_task(v, torch.from_numpy(img).cuda(), n, n, blocks)
    c.compress()
    torch.cuda.synchronize()
    plain = blocks.cpu().numpy().reshape(-1, 16)
    rng = np.random.default_rng(12)
    ids = np.sort(rng.choice(plain.shape[0], 1024, replace=False))
    ref = oracle_lib.compress_rgba8_sample(6, s.to_bytes(), img, ids, threads=16)
    assert np.array_equal(plain[ids], ref.reshape(-1, 16))
    print("configs: C5 refine 64 steps, reassign/p-bits every 4", flush=True)
    out, s0, s1 = grad.refine_bc7(blocks, torch.from_numpy(target).cuda(), n, n, steps=64, lr=0.5,
                                  channels=4, reassign=4, pbits=4)
    torch.cuda.synchronize()
    out = out.cpu().numpy()
    s0, s1 = s0.cpu().numpy(), s1.cpu().numpy()
    assert np.all(s1 <= s0) and s1.sum() < 0.98 * s0.sum()
    # Recount from the decoded blocks (0..255 units, RGBA).
    tb = (target.astype(np.float64) * 255).reshape(n // 4, 4, n // 4, 4, 4).transpose(0, 2, 1, 3, 4).reshape(-1, 16, 4)
    for blk, want in ((plain.reshape(-1), s0), (out, s1)):
        dec = oracle_lib.decompress_rgba8(6, n, n, blk).astype(np.float64)
        db = dec.reshape(n // 4, 4, n // 4, 4, 4).transpose(0, 2, 1, 3, 4).reshape(-1, 16, 4)
        sse = ((db - tb) ** 2).sum(axis=(1, 2))
        assert np.allclose(sse, want, rtol=1e-5, atol=1e-3)
    # Host build of the same code on a sample of blocks.
    hs = ids[:512]
    hout, h0, h1 = kc.bc7_refine(plain[hs], (tb[hs]).astype(np.float32), steps=64, lr=0.5, channels=4,
                                 reassign=4 | (4 << 16))
    assert np.array_equal(out.reshape(-1, 16)[hs], hout.reshape(-1, 16))
    assert np.array_equal(s0[hs], h0) and np.array_equal(s1[hs], h1)


def test_c4_textures_generated_on_the_gpu_match_the_committed_digests():
    """The workload builds its textures with torch on the device; the first
    four equal the CPU-generated bytes whose digests are committed."""
    import hashlib
    import json
    import os

    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "config_digests.json")) as f:
        want = json.load(f)
    for i, (w, h, fmt) in enumerate(workloads.texture_specs(300, 4)[:4]):
        w8, h8 = (w + 7) // 8 * 8, (h + 7) // 8 * 8
        img = synth.tile_image_rgba8_torch(w8, h8, 4 * 1000 + i, 0.1, "cuda")
        if fmt in ("bc3", "bc7_alpha_basic"):
            img[..., 3] = synth.tile_image_rgba8_torch(w8, h8, 4 * 1000 + i + 7919, 0.1, "cuda")[..., 0]
        got = hashlib.sha256(img[:h, :w].contiguous().cpu().numpy().tobytes()).hexdigest()
        assert got == want[f"C4_texture{i}_{w}x{h}_{fmt}"], i
