"""Batched persistent dispatch (bcn_batch_*): every task of a batch -- mixed
formats, sizes, alignments, BC6H among them -- comes out byte-identical to the
per-task path (bcn_compress), which the other GPU tests pin to the CPU oracle;
a sample is also checked against the oracle directly.  Graph replays and the
autotuner must leave the bytes unchanged."""
import numpy as np
import pytest

import oracle_lib
from block_compression_amd import synth
from block_compression_amd.settings import BC6HSettings, BC7Settings, CompressionVariant

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

# (variant, width, height, texel offset into a wider texture, seed)
CASES = [
    (CompressionVariant.BC7(BC7Settings.alpha_basic()), 96, 64, 0, 1),
    (CompressionVariant.BC7(BC7Settings.opaque_basic()), 200, 136, 0, 2),
    (CompressionVariant.BC7(BC7Settings.opaque_fast()), 4, 4, 1, 3),
    (CompressionVariant.BC1(), 520, 260, 1, 4),
    (CompressionVariant.BC2(), 64, 8, 0, 5),
    (CompressionVariant.BC3(), 132, 100, 0, 6),
    (CompressionVariant.BC4(), 1024, 516, 0, 7),
    (CompressionVariant.BC5(), 300, 12, 1, 8),
    (CompressionVariant.BC7(BC7Settings.opaque_basic()), 8, 12, 0, 9),
    (CompressionVariant.BC1(), 4, 8, 0, 10),
]


def _texture(w, h, off, seed):
    """(texture tensor view, numpy copy of the region): the region sits at texel
    column `off` of a wider texture, so off=1 gives rows that are 4-byte but
    not 16-byte aligned."""
    ww = w + 8
    full = synth.tile_image_rgba8(((ww + 7) // 8) * 8, ((h + 7) // 8) * 8, seed, 0.1)
    full[..., 3] = synth.tile_image_rgba8(full.shape[1], full.shape[0], seed + 100, 0.2)[..., 2]
    dev = torch.from_numpy(full).cuda()
    view = dev[:h, off:off + w]
    return view, full[:h, off:off + w].copy()


def _queue(comp, cases, bufs, texs):
    for (v, w, h, off, seed), buf in zip(cases, bufs):
        tex, _ = texs[(w, h, off, seed)]
        comp.add_compression_task(v, tex, w, h, buf)


def _setup(cases):
    texs = {}
    for v, w, h, off, seed in cases:
        texs[(w, h, off, seed)] = _texture(w, h, off, seed)
    return texs


def _bufs(cases):
    return [torch.full((v.blocks_byte_size(w, h),), 0xAB, dtype=torch.uint8, device="cuda")
            for v, w, h, _, _ in cases]


def _per_task(cases, texs):
    from block_compression_amd import GpuBlockCompressor

    bufs = _bufs(cases)
    c = GpuBlockCompressor()
    _queue(c, cases, bufs, texs)
    c.compress()
    torch.cuda.synchronize()
    return [b.cpu().numpy() for b in bufs]


def _hdr_case():
    img = synth.hdr_image_rgba16f(64, 32, 5)
    return torch.from_numpy(img.view(np.uint16)).cuda().view(torch.float16), img


def test_batch_matches_per_task_and_oracle():
    from block_compression_amd import GpuBlockCompressor

    print("batch: mixed formats, 10 tasks + BC6H", flush=True)
    texs = _setup(CASES)
    ref = _per_task(CASES, texs)
    bufs = _bufs(CASES)
    hdr_tex, hdr_img = _hdr_case()
    v6 = CompressionVariant.BC6H(BC6HSettings.basic())
    b6 = torch.zeros(v6.blocks_byte_size(64, 32), dtype=torch.uint8, device="cuda")
    c = GpuBlockCompressor()
    _queue(c, CASES, bufs, texs)
    c.add_compression_task(v6, hdr_tex, 64, 32, b6)
    batch = c.create_batch()
    assert c.task_count() == 0
    info = batch.info()
    assert info["tasks"] == len(CASES) + 1 and info["persistent_tasks"] == len(CASES)
    batch.launch()
    torch.cuda.synchronize()
    for (v, w, h, off, seed), b, r in zip(CASES, bufs, ref):
        assert np.array_equal(b.cpu().numpy(), r), (v, w, h, off)
    # The oracle directly, on two of the tasks.
    for i in (0, 7):
        v, w, h, off, seed = CASES[i]
        img = texs[(w, h, off, seed)][1]
        want = oracle_lib.compress_rgba8(v.tag, v.settings_bytes(), img, threads=8)
        assert np.array_equal(bufs[i].cpu().numpy(), want), v
    want6 = oracle_lib.compress_rgba16(v6.settings_bytes(), hdr_img, threads=8)
    assert np.array_equal(b6.cpu().numpy(), want6)
    batch.close()


def test_batch_graph_replays_and_autotune_keep_bytes():
    from block_compression_amd import GpuBlockCompressor

    print("batch: graph replays + autotune", flush=True)
    texs = _setup(CASES)
    ref = _per_task(CASES, texs)
    bufs = _bufs(CASES)
    c = GpuBlockCompressor()
    _queue(c, CASES, bufs, texs)
    batch = c.create_batch({"chunk_blocks_bc7": 128, "chunk_blocks_bc15": 64, "waves_per_cu": 2})
    for rep in range(3):
        for b in bufs:
            b.fill_(0x5A)  # eager work between replays
        torch.cuda.synchronize()
        batch.launch(graph=True)
        torch.cuda.synchronize()
        for b, r in zip(bufs, ref):
            assert np.array_equal(b.cpu().numpy(), r), rep
    cfg = batch.autotune(reps=1)
    assert cfg["chunk_blocks_bc7"] % 64 == 0 and cfg["chunk_blocks_bc15"] % 64 == 0
    assert cfg["serial_classes"] in (0, 1)
    assert batch.info()["config"] == cfg
    for b in bufs:
        b.fill_(0)
    batch.launch(graph=True)
    torch.cuda.synchronize()
    for b, r in zip(bufs, ref):
        assert np.array_equal(b.cpu().numpy(), r)
    batch.close()


@pytest.mark.parametrize("serial", [0, 1])
def test_batch_class_schedules_write_the_same_bytes(serial):
    """The class kernels concurrent (three streams, joined) or one after
    another: the same bytes as per-task launches, eagerly and through the
    graph, and work queued on the launch stream afterwards sees them."""
    from block_compression_amd import GpuBlockCompressor

    print(f"batch: serial_classes={serial}", flush=True)
    texs = _setup(CASES)
    ref = _per_task(CASES, texs)
    bufs = _bufs(CASES)
    c = GpuBlockCompressor()
    _queue(c, CASES, bufs, texs)
    batch = c.create_batch({"chunk_blocks_bc7": 64, "chunk_blocks_bc15": 256, "waves_per_cu": 0,
                            "serial_classes": serial})
    assert batch.info()["config"]["serial_classes"] == serial
    for graph in (False, True, True):
        for b in bufs:
            b.fill_(0x5A)
        batch.launch(graph=graph)
        # Same-stream consumers after the join: copies enqueued right after the launch.
        got = [b.clone() for b in bufs]
        torch.cuda.synchronize()
        for g, r in zip(got, ref):
            assert np.array_equal(g.cpu().numpy(), r), graph
    batch.close()


def test_empty_batch_and_bad_config():
    from block_compression_amd import BcnError, GpuBlockCompressor

    c = GpuBlockCompressor()
    batch = c.create_batch()
    assert batch.info()["tasks"] == 0
    batch.launch()
    batch.launch(graph=True)
    torch.cuda.synchronize()
    batch.close()
    with pytest.raises(BcnError):
        c.create_batch({"chunk_blocks_bc7": 100, "chunk_blocks_bc15": 64, "waves_per_cu": 0})
    with pytest.raises(BcnError):
        c.create_batch({"chunk_blocks_bc7": 64, "chunk_blocks_bc15": 64, "waves_per_cu": 0,
                        "serial_classes": 2})


def test_mip_chain_workload_batched_vs_plain_and_oracle():
    """A small instance of the batched workload (12 textures up to 192 texels,
    full mip chains, six formats): batched == per-task, and every block of
    every level equals the oracle."""
    from block_compression_amd import GpuBlockCompressor, workloads

    print("batch: mip-chain workload, 12 textures", flush=True)
    levels = workloads.build(12, seed=11, device="cuda", max_size=192, min_size=4)
    total = workloads.output_bytes(levels)
    plain = torch.zeros(total, dtype=torch.uint8, device="cuda")
    batched = torch.full((total,), 0xCD, dtype=torch.uint8, device="cuda")
    c = GpuBlockCompressor()
    workloads.queue(c, levels, plain)
    c.compress()
    workloads.queue(c, levels, batched)
    batch = c.create_batch()
    batch.launch(graph=True)
    torch.cuda.synchronize()
    p, b = plain.cpu().numpy(), batched.cpu().numpy()
    for lv in levels:
        sl = slice(lv.offset, lv.offset + lv.nbytes)
        assert np.array_equal(b[sl], p[sl]), (lv.texture, lv.level, lv.fmt)
        v = workloads.variant_of(lv.fmt)
        want = oracle_lib.compress_rgba8(v.tag, v.settings_bytes(), lv.texels.cpu().numpy(), threads=8)
        assert np.array_equal(b[sl], want), (lv.texture, lv.level, lv.fmt)
    batch.close()


def test_failed_batch_creation_keeps_the_tasks_queued():
    """A batch that cannot be created (here: a chunk size that is not a
    multiple of 64) leaves every task queued on the compressor, which can
    then still compress them."""
    from block_compression_amd import BcnError, GpuBlockCompressor

    cases = CASES[:3]
    texs = _setup(cases)
    bufs = _bufs(cases)
    c = GpuBlockCompressor()
    _queue(c, cases, bufs, texs)
    with pytest.raises(BcnError):
        c.create_batch({"chunk_blocks_bc7": 65, "chunk_blocks_bc15": 1024, "waves_per_cu": 0})
    assert c.task_count() == len(cases)
    c.compress()
    torch.cuda.synchronize()
    for got, want in zip([b.cpu().numpy() for b in bufs], _per_task(cases, texs)):
        assert np.array_equal(got, want)
