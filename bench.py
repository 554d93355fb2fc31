#!/usr/bin/env python3
"""Benchmark: BC7 'opaque_basic' encode of one 4096x4096 RGBA8 texture on one MI355X.

A step is one pass of the GPU block compressor over the texture: queue the
task (GpuBlockCompressor.add_compression_task) and enqueue it
(GpuBlockCompressor.compress), the texture already resident in HBM.  After W
warm-up steps, K steps are timed between two device synchronisations; each
step is also bracketed by HIP events on the launch stream, which gives the
kernel's average launch duration for the roofline.

Besides the JSON line's throughput, the run checks its own output against the
CPU oracle (oracle/, a C++ restatement of the reference's CPU encoder) on
every block -- byte identity and the reference harness's per-image PSNR
(tests/metrics.rs:40-114) -- and times the oracle on the host's CPUs and on
one core as the CPU baseline.

    python bench.py [--steps K] [--warmup W] [--gpus 1]
"""
from __future__ import annotations

import argparse
import ctypes
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "BC7-basic 4096² encode Mpix/s at ref PSNR±0.01dB; batched & grad-refined vs plain"
PEAK_FP32_TFLOPS = 157.3        # MI355X published FP32 vector peak (packed FMA)
ISSUE_CEILING_TOPS = 39.3       # 256 CU x 64 lanes x 2.4 GHz, one op per lane per cycle
PACKED_CEILING_TOPS = 78.6      # the same with two ops per lane per cycle (packed, no FMA)
PEAK_HBM_GBS = 8000.0


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_threads(requested: int):
    """Oracle threads and a description of the host: every CPU this process
    may run on (its affinity set), at most 16 -- a GPU box gives one GPU's
    job a share of 16 CPUs whatever nproc shows."""
    visible = os.cpu_count() or 1
    try:
        allowed = len(os.sched_getaffinity(0))
    except AttributeError:
        allowed = visible
    threads = requested or max(1, min(16, allowed))
    return threads, {"cpus_visible": visible, "cpus_allowed": allowed, "cpu_model": cpu_model()}


def cpu_baseline_leg(img, settings, nblocks: int, threads: int, host: dict) -> dict:
    """The oracle (C++ restatement of the reference's CPU encoder, OpenMP over
    blocks as the crate's rayon feature spreads it) on the same image, timed
    as BASELINE.md plans it: 1 warm-up, then the median of 5 runs, on all the
    host CPUs this job may use and on one core.  Bounded seeded samples of
    the workload's blocks keep the leg to about 10 s."""
    import numpy as np

    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib

    def median_rate(ids, nthreads):
        oracle_lib.compress_rgba8_sample(6, settings.to_bytes(), img, ids[: min(ids.size, 4096)], nthreads)
        times = []
        for _ in range(5):
            t = time.perf_counter()
            oracle_lib.compress_rgba8_sample(6, settings.to_bytes(), img, ids, nthreads)
            times.append(time.perf_counter() - t)
        t = float(np.median(times))
        return ids.size * 16 / t / 1e6, t, [round(x, 4) for x in times]

    rng = np.random.default_rng(7)
    many = np.sort(rng.choice(nblocks, 65536, replace=False))
    one = np.sort(rng.choice(nblocks, 16384, replace=False))
    log(f"bench: CPU baseline, {threads} threads (65,536 blocks) and 1 thread (16,384 blocks), median of 5")
    v_all, t_all, runs_all = median_rate(many, threads)
    v_one, t_one, runs_one = median_rate(one, 1)
    return {
        "value": round(v_all, 3),
        "unit": "Mpix/s",
        "cores": threads,
        "kind": "port",
        "sample": f"65,536 seeded random blocks (1,048,576 px) of the same 4096x4096 opaque_basic "
                  f"workload, C++ oracle, OpenMP over blocks on {threads} threads; 1 warm-up, median of 5 "
                  f"({t_all:.3f} s)",
        "runs_s": runs_all,
        "one_core": {"value": round(v_one, 3), "unit": "Mpix/s", "cores": 1,
                     "sample": f"16,384 seeded random blocks (262,144 px) of the same workload; 1 warm-up, "
                               f"median of 5 ({t_one:.3f} s)", "runs_s": runs_one},
        "host": host,
        "build": "g++ -O3 -ffp-contract=off -fopenmp, baseline x86-64 (no -march)",
    }


def bench_formats(img, reps: int = 5, launches: int = 20) -> dict:
    """The other kernels on their configurations: BC1/BC3/BC4/BC5 encode of the
    same 4096x4096 texture (BC4 also from its R8 channel, BC5 from RG8), BC6H
    'slow' encode of the 2048x2048 HDR texture, the GPU decode of the
    4096x4096 BC7 output and the RGBA16F / RGBA32F / sRGB8 decodes of the BC6H
    output.  Each figure is the best of
    `reps` runs of `launches` back-to-back launches between two HIP events,
    divided by `launches` (one launch of BC4 is ~10 us, comparable to launch
    gaps).  HBM roofline against 8 TB/s with the algorithmic bytes per block
    (64 B read + block out for RGBA8 sources, 128 B + 16 B for BC6H, 16 B +
    64 B for decode)."""
    import numpy as np
    import torch

    from block_compression_amd import GpuBlockCompressor, synth
    from block_compression_amd._capi import check, lib
    from block_compression_amd.settings import BC6HSettings, BC7Settings, CompressionVariant

    stream = torch.cuda.current_stream()
    comp = GpuBlockCompressor(0)
    out = {}

    def timed(fn):
        best = None
        for _ in range(reps):
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            torch.cuda.synchronize()
            a.record(stream)
            for _ in range(launches):
                fn()
            b.record(stream)
            torch.cuda.synchronize()
            ms = a.elapsed_time(b) / launches
            best = ms if best is None else min(best, ms)
        return best

    def entry(ms, w, h, bytes_per_block, bound="hbm"):
        blocks = (w // 4) * (h // 4)
        gbs = blocks * bytes_per_block / (ms * 1e-3) / 1e9
        return {"ms": round(ms, 4), "mpix_s": round(w * h / (ms * 1e-3) / 1e6, 1),
                "bytes_per_block": bytes_per_block, "achieved_gbs": round(gbs, 1),
                "hbm_frac": round(gbs / PEAK_HBM_GBS, 4), "bound": bound}

    W, H = img.shape[1], img.shape[0]
    tex = torch.from_numpy(img).cuda()
    for name, v in [("bc1", CompressionVariant.BC1()), ("bc3", CompressionVariant.BC3()),
                    ("bc4", CompressionVariant.BC4()), ("bc5", CompressionVariant.BC5())]:
        buf = torch.empty(v.blocks_byte_size(W, H), dtype=torch.uint8, device="cuda")

        def run(v=v, buf=buf):
            comp.add_compression_task(v, tex, W, H, buf)
            comp.compress()

        run()
        out[f"{name}_encode_{W}"] = entry(timed(run), W, H, 64 + v.block_byte_size())
    # BC6H slow, 2048x2048 RGBA16F.
    hdr = synth.hdr_image_rgba16f(2048, 2048, 3)
    htex = torch.from_numpy(hdr.view(np.uint16)).cuda().view(torch.float16)
    v6 = CompressionVariant.BC6H(BC6HSettings.slow())
    b6 = torch.empty(v6.blocks_byte_size(2048, 2048), dtype=torch.uint8, device="cuda")

    def run6():
        comp.add_compression_task(v6, htex, 2048, 2048, b6)
        comp.compress()

    run6()
    e6 = entry(timed(run6), 2048, 2048, 144, bound="valu")
    out["bc6h_slow_encode_2048"] = e6
    # Decode of the BC7 opaque_basic blocks of the 4096^2 texture.
    v7 = CompressionVariant.BC7(BC7Settings.opaque_basic())
    b7 = torch.empty(v7.blocks_byte_size(W, H), dtype=torch.uint8, device="cuda")
    comp.add_compression_task(v7, tex, W, H, b7)
    comp.compress()
    rgba = torch.empty((H, W, 4), dtype=torch.uint8, device="cuda")

    def rund():
        check(lib().bcn_decode_rgba8(6, b7.data_ptr(), b7.numel(), W, H, rgba.data_ptr(), W * 4,
                                     ctypes.c_void_p(stream.cuda_stream)))

    rund()
    out[f"bc7_decode_{W}"] = entry(timed(rund), W, H, 80)
    # BC4 / BC5 from single- and two-channel sources (R8, RG8): 16 / 32 bytes
    # of texels per block instead of 64.
    for name, v, src in [("bc4_encode_r8", CompressionVariant.BC4(), tex[..., 0].contiguous()),
                         ("bc5_encode_rg8", CompressionVariant.BC5(), tex[..., :2].contiguous())]:
        buf = torch.empty(v.blocks_byte_size(W, H), dtype=torch.uint8, device="cuda")

        def runs(v=v, buf=buf, src=src):
            comp.add_compression_task(v, src, W, H, buf)
            comp.compress()

        runs()
        texel_bytes = 16 * src.element_size() * (1 if src.dim() == 2 else src.shape[2])
        out[f"{name}_{W}"] = entry(timed(runs), W, H, texel_bytes + v.block_byte_size())
    # Image-level decodes of the BC6H blocks (2048^2): RGBA16F, RGBA32F and
    # sRGB8 (exact linear_to_srgb).
    for name, fn, dt, tb in [("rgba16f", "bcn_decode_rgba16f", torch.float16, 8),
                             ("rgba32f", "bcn_decode_rgba32f", torch.float32, 16),
                             ("srgb8", "bcn_decode_rgba8", torch.uint8, 4)]:
        img_out = torch.empty((2048, 2048, 4), dtype=dt, device="cuda")
        f = getattr(lib(), fn)

        def rundh(f=f, img_out=img_out, tb=tb, fn=fn):
            args = [5, b6.data_ptr(), b6.numel(), 2048, 2048, img_out.data_ptr(), 2048 * tb]
            if fn != "bcn_decode_rgba8":
                args.append(0)
            check(f(*args, ctypes.c_void_p(stream.cuda_stream)))

        rundh()
        out[f"bc6h_decode_{name}_2048"] = entry(timed(rundh), 2048, 2048, 16 + 16 * tb)
    torch.cuda.synchronize()
    return out


def count_format_ops(formats: dict, img) -> None:
    """Adds the as-written f32 op count per block (instrumented oracle, 1024
    seeded blocks of each kernel's own input) and the FP32 roofline fraction
    to the encode kernels of the other-kernel leg.  The BC1-BC5 counts are on
    the same 4096x4096 texture, BC6H 'slow' on the 2048x2048 HDR texture."""
    import numpy as np

    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib
    from block_compression_amd import synth
    from block_compression_amd.settings import BC6HSettings

    W = img.shape[1]
    ids = np.sort(np.random.default_rng(7).choice((W // 4) ** 2, 1024, replace=False))
    counts = {f"{n}_encode_{W}": oracle_lib.count_ops_rgba8(t, b"", img, ids) / 1024.0
              for t, n in [(0, "bc1"), (2, "bc3"), (3, "bc4"), (4, "bc5")]}
    hdr = synth.hdr_image_rgba16f(2048, 2048, 3)
    ids6 = np.sort(np.random.default_rng(7).choice(512 * 512, 1024, replace=False))
    counts["bc6h_slow_encode_2048"] = oracle_lib.count_ops_rgba16(BC6HSettings.slow().to_bytes(), hdr,
                                                                  ids6) / 1024.0
    for key, ops in counts.items():
        e = formats.get(key)
        if e is None:
            continue
        side = 2048 if key.startswith("bc6h") else W
        tops = ops * (side // 4) ** 2 / (e["ms"] * 1e-3) / 1e12
        e["ops_per_block"] = round(ops, 1)
        e["achieved_tops"] = round(tops, 3)
        e["fp32_frac"] = round(tops / PEAK_FP32_TFLOPS, 4)


def bench_batched(n_textures: int, reps: int, oracle_levels: int, threads: int) -> dict:
    """The batched mixed mip-chain workload (block_compression_amd.workloads):
    the same tasks encoded by per-task launches (bcn_compress, one kernel per
    mip level, as the reference dispatches) and by one batched persistent
    launch (bcn_batch, autotuned, hipGraph).  Both outputs must be identical;
    a seeded sample of levels is checked against the oracle."""
    import numpy as np
    import torch

    from block_compression_amd import GpuBlockCompressor, workloads

    t0 = time.perf_counter()
    levels = workloads.build(n_textures, seed=4, device="cuda")
    torch.cuda.synchronize()
    gen_s = time.perf_counter() - t0
    total = workloads.output_bytes(levels)
    texels = sum(lv.width * lv.height for lv in levels)
    blocks = texels // 16
    log(f"bench: batched workload {n_textures} textures, {len(levels)} levels, "
        f"{texels / 1e9:.3f} Gpx, generated in {gen_s:.1f} s")
    plain = torch.zeros(total, dtype=torch.uint8, device="cuda")
    batched = torch.zeros(total, dtype=torch.uint8, device="cuda")
    stream = torch.cuda.current_stream()
    comp = GpuBlockCompressor(0)

    def timed(fn, n, prep=None):
        """Best of n: HIP events on the launch stream around fn(); prep()
        (host-side queuing) runs before the first event, outside the timing."""
        best = None
        for _ in range(n):
            if prep is not None:
                prep()
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            torch.cuda.synchronize()
            a.record(stream)
            fn()
            b.record(stream)
            torch.cuda.synchronize()
            ms = a.elapsed_time(b)
            best = ms if best is None else min(best, ms)
        return best

    def wall(fn, n):
        """Best of n wall-clock seconds of fn() to a device synchronisation."""
        best = None
        for _ in range(n):
            torch.cuda.synchronize()
            t = time.perf_counter()
            fn()
            torch.cuda.synchronize()
            ms = (time.perf_counter() - t) * 1e3
            best = ms if best is None else min(best, ms)
        return best

    def plain_step():
        workloads.queue(comp, levels, plain)
        comp.compress()

    plain_step()
    torch.cuda.synchronize()
    # Kernels only: the 3,513 tasks are queued before the first event.
    plain_ms = timed(comp.compress, reps, prep=lambda: workloads.queue(comp, levels, plain))
    # Host set-up included: queuing every task, then compress().
    plain_setup_ms = wall(plain_step, reps)
    log(f"bench: per-task launches {plain_ms:.2f} ms (kernels), {plain_setup_ms:.2f} ms with queuing")
    workloads.queue(comp, levels, batched)
    batch = comp.create_batch()
    cfg = batch.autotune(reps=1)
    batch.launch(graph=True)
    torch.cuda.synchronize()
    batched_ms = timed(lambda: batch.launch(graph=True), reps)

    def batched_setup_step():
        workloads.queue(comp, levels, batched)
        b2 = comp.create_batch(cfg)
        b2.launch()
        torch.cuda.synchronize()
        b2.close()

    batched_setup_ms = wall(batched_setup_step, reps)
    # The other class schedule (concurrent class kernels vs one class after
    # another) with the autotuned chunking, through its own graph.
    other = dict(cfg, serial_classes=1 - cfg["serial_classes"])
    workloads.queue(comp, levels, batched)
    b3 = comp.create_batch(other)
    b3.launch(graph=True)
    torch.cuda.synchronize()
    other_ms = timed(lambda: b3.launch(graph=True), reps)
    b3.close()
    info = batch.info()
    log(f"bench: batched persistent launch {batched_ms:.2f} ms (graph replay), {batched_setup_ms:.2f} ms "
        f"with queuing and batch creation, config {cfg}")
    same = bool(torch.equal(plain, batched))
    # Oracle on a seeded sample of levels (every level of the sample, all blocks).
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib

    rng = np.random.default_rng(5)
    small = [i for i, lv in enumerate(levels) if lv.width * lv.height <= 64 * 64]
    big = [i for i, lv in enumerate(levels) if lv.width * lv.height > 64 * 64]
    pick = list(rng.choice(small, min(oracle_levels, len(small)), replace=False))
    pick += list(rng.choice(big, min(max(1, oracle_levels // 16), len(big)), replace=False))
    out = batched.cpu().numpy()
    checked = identical = 0
    for i in pick:
        lv = levels[int(i)]
        v = workloads.variant_of(lv.fmt)
        img = lv.texels.cpu().numpy()
        nb = (lv.width // 4) * (lv.height // 4)
        ids = np.arange(nb) if nb <= 4096 else np.sort(rng.choice(nb, 4096, replace=False))
        ref = oracle_lib.compress_rgba8_sample(v.tag, v.settings_bytes(), img, ids, threads)
        bb = v.block_byte_size()
        got = out[lv.offset:lv.offset + lv.nbytes].reshape(-1, bb)[ids]
        identical += int((got == ref.reshape(-1, bb)).all(axis=1).sum())
        checked += len(ids)
    batch.close()
    return {
        "workload": f"{n_textures} tile-generator RGBA8 textures, w,h ~ U{{64..4096}}, full box-filtered "
                    "mip chains padded to 4, formats ~ U{BC1,BC3,BC4,BC5,BC7 opaque_basic,BC7 alpha_basic}",
        "levels": len(levels),
        "texels": int(texels),
        "blocks": int(blocks),
        "plain_ms": round(plain_ms, 3),
        "plain_mpix_s": round(texels / (plain_ms * 1e-3) / 1e6, 1),
        "batched_ms": round(batched_ms, 3),
        "batched_mpix_s": round(texels / (batched_ms * 1e-3) / 1e6, 1),
        "speedup_vs_plain": round(plain_ms / batched_ms, 3),
        "plain_with_setup_ms": round(plain_setup_ms, 3),
        "batched_with_setup_ms": round(batched_setup_ms, 3),
        "speedup_with_setup": round(plain_setup_ms / batched_setup_ms, 3),
        "autotuned_config": cfg,
        "schedules_ms": {"concurrent" if cfg["serial_classes"] == 0 else "serial": round(batched_ms, 3),
                         "serial" if cfg["serial_classes"] == 0 else "concurrent": round(other_ms, 3)},
        "grid": info["grid"],
        "chunks": info["chunks"],
        "identical_to_plain": same,
        "oracle_blocks_checked": checked,
        "oracle_identical_frac": round(identical / checked, 6) if checked else None,
        "timing": f"best of {reps} each.  Kernels only (HIP events on the launch stream): plain = "
                  "compress() of the already-queued tasks, one kernel per mip level as the reference "
                  "dispatches; batched = one replay of the batch's hipGraph.  With set-up (wall clock "
                  "to a device sync): plain = queue every level + compress(); batched = queue every "
                  "level + create_batch (queue upload) + launch.  schedules_ms: the autotuned "
                  "chunking with the three class kernels concurrent (three streams) and one after "
                  "another.",
    }


def bench_grad(size: int, preset: str, reps: int) -> dict:
    """Gradient-refined BC7: an FP32 RGBA target from the tile generator; the
    plain GPU encoder (the preset, on the target rounded to RGBA8) picks mode,
    partition and indices; then 64 Adam steps (lr 0.5 in endpoint units,
    straight-through rounding) through the HIP decode adjoint; every 4th step
    the indices are reassigned and each p-bit flip is tried.  PSNR over RGBA against the FP32 target.
    The adjoint is checked against FP64 central differences (step 1e-3) on 256
    random blocks of the plain output."""
    import numpy as np
    import torch

    from block_compression_amd import GpuBlockCompressor, decode_rgba8, grad, synth
    from block_compression_amd.settings import BC7_PRESETS, CompressionVariant

    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import adjoint_ref

    n = size
    target = synth.tile_target_rgba_f32(n, n, 5, 0.0)
    img = np.clip(np.floor(target * 255 + 0.5), 0, 255).astype(np.uint8)
    v = CompressionVariant.BC7(BC7_PRESETS[preset]())
    tex = torch.from_numpy(img).cuda()
    td = torch.from_numpy(target).cuda()
    blocks = torch.zeros(v.blocks_byte_size(n, n), dtype=torch.uint8, device="cuda")
    refined = torch.zeros_like(blocks)
    comp = GpuBlockCompressor(0)
    stream = torch.cuda.current_stream()

    def plain():
        comp.add_compression_task(v, tex, n, n, blocks)
        comp.compress()

    steps, lr = 64, 0.5
    res = {}

    def refine():
        res["out"] = grad.refine_bc7(blocks, td, n, n, steps=steps, lr=lr, channels=4, reassign=4,
                                     pbits=4, out=refined)

    def timed(fn):
        best = None
        for _ in range(reps):
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            torch.cuda.synchronize()
            a.record(stream)
            fn()
            b.record(stream)
            torch.cuda.synchronize()
            ms = a.elapsed_time(b)
            best = ms if best is None else min(best, ms)
        return best

    plain()
    refine()
    plain_ms = timed(plain)
    refine_ms = timed(refine)
    _, s0, s1 = res["out"]
    s0 = s0.double().sum().item()
    s1 = s1.double().sum().item()
    denom = float(n * n * 4) * 255.0 ** 2

    def psnr(sse):
        return 10.0 * math.log10(denom / sse) if sse > 0 else 0.0

    # Ablations on the same blocks: the same schedule with lr = 0 (only the
    # index reassignment and p-bit flips act), and Adam with indices and
    # p-bits fixed.
    def variant(src, **kw):
        o = torch.zeros_like(src)
        fn = lambda: grad.refine_bc7(src, td, n, n, steps=steps, channels=4, out=o, **kw)
        r = fn()
        ms = timed(fn)
        return r[2].double().sum().item(), ms

    sse_lr0, ms_lr0 = variant(blocks, lr=0.0, reassign=4, pbits=4)
    sse_fixed, ms_fixed = variant(blocks, lr=lr, reassign=0, pbits=0)
    # The gradient pass in place of the encoder's own refinement: the preset
    # with every refine_iterations 0, then the same 64 steps.
    import dataclasses

    s_noref = dataclasses.replace(BC7_PRESETS[preset](), refine_iterations=(0,) * 8,
                                  refine_iterations_channel=0)
    v0 = CompressionVariant.BC7(s_noref)
    blocks0 = torch.zeros_like(blocks)

    def plain0():
        comp.add_compression_task(v0, tex, n, n, blocks0)
        comp.compress()

    plain0()
    plain0_ms = timed(plain0)
    dec_n = decode_rgba8(v0, blocks0, n, n).double()
    sse_noref = ((dec_n - td.double() * 255.0) ** 2).sum().item()
    sse_repl, ms_repl = variant(blocks0, lr=lr, reassign=4, pbits=4)
    # The reported errors are those of the blocks written: decode and recount.
    dec = decode_rgba8(v, refined, n, n).double()
    s_dec = ((dec - td.double() * 255.0) ** 2).sum().item()
    dec0 = decode_rgba8(v, blocks, n, n).double()
    s_dec0 = ((dec0 - td.double() * 255.0) ** 2).sum().item()
    # Adjoint vs FP64 finite differences on 256 random blocks.
    rng = np.random.default_rng(6)
    ids = rng.choice(blocks.numel() // 16, 256, replace=False)
    bsub = blocks.view(-1, 16)[torch.from_numpy(ids).cuda()].contiguous()
    x0 = grad.endpoints(6, bsub.view(-1)).double().cpu().numpy()
    xc = x0 + rng.uniform(-0.4, 0.4, x0.shape)
    tgt = rng.uniform(0, 255, (256, 16, 4))
    x = torch.tensor(xc, dtype=torch.float32, device="cuda", requires_grad=True)
    out = grad.DecodeWithEndpoints.apply(x, bsub.view(-1), 6, True)
    ((out - torch.tensor(tgt, dtype=torch.float32, device="cuda")) ** 2).sum().backward()
    g = x.grad.double().cpu().numpy()
    bnp = bsub.cpu().numpy()
    fd = np.zeros_like(xc)
    h = 1e-3
    with torch.no_grad():
        for i in range(256):
            f = adjoint_ref.parse_bc7(bnp[i].tobytes())
            xi = torch.tensor(xc[i].reshape(6, 4), dtype=torch.float64)
            tt = torch.tensor(tgt[i], dtype=torch.float64)
            for j in range(24):
                d = torch.zeros(24, dtype=torch.float64)
                d[j] = h
                d = d.reshape(6, 4)
                lp = ((adjoint_ref.smooth_bc7(f, xi + d) - tt) ** 2).sum()
                lm = ((adjoint_ref.smooth_bc7(f, xi - d) - tt) ** 2).sum()
                fd[i, j] = float((lp - lm) / (2 * h))
    fd_err = float(np.abs(g - fd).max() / max(np.abs(fd).max(), 1e-12))
    return {
        "workload": f"{n}x{n} FP32 RGBA target (tile generator, values before 8-bit rounding; alpha "
                    f"from a second generator image), BC7 {preset}",
        "plain_ms": round(plain_ms, 3),
        "refine_ms": round(refine_ms, 3),
        "time_ratio_vs_plain": round((plain_ms + refine_ms) / plain_ms, 3),
        "psnr_plain_db": round(psnr(s0), 4),
        "psnr_refined_db": round(psnr(s1), 4),
        "gain_db": round(psnr(s1) - psnr(s0), 4),
        "decoded_error_matches": bool(abs(s_dec - s1) <= 1e-6 * s1 + 1.0 and abs(s_dec0 - s0) <= 1e-6 * s0 + 1.0),
        "adam": {"steps": steps, "lr": lr, "beta1": 0.9, "beta2": 0.999, "reassign_every": 4,
                 "pbit_flips_every": 4},
        "ablation": {
            "lr0_reassign_pbits": {"gain_db": round(psnr(sse_lr0) - psnr(s0), 4), "refine_ms": round(ms_lr0, 3)},
            "adam_fixed_indices": {"gain_db": round(psnr(sse_fixed) - psnr(s0), 4),
                                   "refine_ms": round(ms_fixed, 3)},
            "note": "gains over the plain encode: the full schedule minus lr (only index reassignment "
                    "and p-bit flips), and Adam with indices and p-bits fixed",
        },
        "replace_refine_iterations": {
            "plain_no_refine_ms": round(plain0_ms, 3),
            "psnr_plain_no_refine_db": round(psnr(sse_noref), 4),
            "psnr_gradient_db": round(psnr(sse_repl), 4),
            "gain_vs_plain_db": round(psnr(sse_repl) - psnr(s0), 4),
            "time_ratio_vs_plain": round((plain0_ms + ms_repl) / plain_ms, 3),
            "note": f"the {preset} preset with refine_iterations 0 (least-squares refinement off), "
                    "then the 64-step gradient pass, against the plain preset with its own refinement",
        },
        "fd_check": {"blocks": 256, "step": h, "dtype": "f64", "max_rel_err": fd_err},
        "psnr": "10 log10(255^2 / MSE) over R, G, B, A of the decoded blocks vs the FP32 target x 255",
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=10,
                    help="untimed steps before the timed ones (the engine clock ramps up over the first ~6)")
    ap.add_argument("--parity-blocks", type=int, default=-1,
                    help="blocks checked against the CPU oracle: -1 = every block (default), "
                         "n > 0 = a seeded sample of n, 0 = none")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="oracle threads (0 = the host CPUs this process may use, at most 16)")
    ap.add_argument("--no-cpu-baseline", action="store_true", help="skip the CPU-baseline leg")
    ap.add_argument("--batched-textures", type=int, default=300,
                    help="textures in the batched mip-chain workload (0 = skip it)")
    ap.add_argument("--batched-reps", type=int, default=3)
    ap.add_argument("--grad-size", type=int, default=1024,
                    help="side of the gradient-refinement target (0 = skip it)")
    ap.add_argument("--grad-preset", default="alpha_slow")
    ap.add_argument("--no-formats", action="store_true", help="skip the other-format kernels")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a plain device-to-device copy of MIB MiB five times "
                         "(reference kernel for scaling HBM byte counters) and exit")
    args = ap.parse_args()

    if args.calibrate_copy:
        import torch

        n = args.calibrate_copy << 20
        a = torch.empty(n, dtype=torch.uint8, device="cuda")
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"calibrate_copy_bytes": n, "copies": 5}), flush=True)
        return 0

    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); the path is single-GPU "
              "(blocks are independent; see DESIGN.md for how it would shard).", file=sys.stderr)
        return 2

    import numpy as np
    import torch

    from block_compression_amd import GpuBlockCompressor, synth
    from block_compression_amd.settings import BC7Settings, CompressionVariant

    W = H = 4096
    seed, noise = 2, 0.1
    settings = BC7Settings.opaque_basic()
    variant = CompressionVariant.BC7(settings)
    nblocks = (W // 4) * (H // 4)

    log("bench: generating the 4096x4096 tile texture")
    img = synth.tile_image_rgba8(W, H, seed, noise)
    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    tex = torch.from_numpy(img).to(dev)
    buf = torch.zeros(variant.blocks_byte_size(W, H), dtype=torch.uint8, device=dev)
    comp = GpuBlockCompressor(0)

    def step():
        comp.add_compression_task(variant, tex, W, H, buf)
        comp.compress()

    log(f"bench: warmup {args.warmup}")
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()

    log(f"bench: timing {args.steps} steps")
    stream = torch.cuda.current_stream(dev)
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
           for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for a, b in evs:
        a.record(stream)
        step()
        b.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    launch_ms = [a.elapsed_time(b) for a, b in evs]
    ms_per_step = wall * 1e3 / args.steps
    kernel_ms = float(np.mean(launch_ms))
    mpix = W * H / (ms_per_step * 1e-3) / 1e6
    log(f"bench: {ms_per_step:.3f} ms/step wall, {kernel_ms:.3f} ms/launch (events), {mpix:.1f} Mpix/s")

    out = buf.cpu().numpy().reshape(-1, 16)

    # ---- parity (every block, or a seeded sample) and the CPU baseline ----------
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib

    threads, host = cpu_threads(args.cpu_threads)
    parity = None
    cpu_baseline = None
    ops_per_block = None
    if args.parity_blocks != 0:
        from metrics_ref import calculate_image_metrics_rgba8

        if args.parity_blocks < 0 or args.parity_blocks >= nblocks:
            log(f"bench: oracle on all {nblocks} blocks with {threads} threads")
            t1 = time.perf_counter()
            ref = oracle_lib.compress_rgba8(6, settings.to_bytes(), img, threads).reshape(-1, 16)
            full_s = time.perf_counter() - t1
            same = (out == ref).all(axis=1)
            p_gpu = calculate_image_metrics_rgba8(img, oracle_lib.decompress_rgba8(6, W, H, out), 3)
            p_ref = calculate_image_metrics_rgba8(img, oracle_lib.decompress_rgba8(6, W, H, ref), 3)
            checked = nblocks
        else:
            full_s = None
            ids = np.sort(np.random.default_rng(7).choice(nblocks, args.parity_blocks, replace=False))
            ref = oracle_lib.compress_rgba8_sample(6, settings.to_bytes(), img, ids, threads).reshape(-1, 16)
            same = (out[ids] == ref).all(axis=1)
            bw = W // 4
            sub = np.stack([img[(i // bw) * 4:(i // bw) * 4 + 4, (i % bw) * 4:(i % bw) * 4 + 4] for i in ids])
            dec = lambda blocks: np.stack([oracle_lib.decode_block("bc7", b.tobytes(), 16, 64).reshape(4, 4, 4)
                                           for b in blocks])
            p_gpu = calculate_image_metrics_rgba8(sub, dec(out[ids]), 3)
            p_ref = calculate_image_metrics_rgba8(sub, dec(ref), 3)
            checked = int(ids.size)
        parity = {
            "blocks_checked": int(checked),
            "identical_frac": float(same.mean()),
            "psnr_gpu_db": round(p_gpu["overall_psnr"], 6),
            "psnr_ref_db": round(p_ref["overall_psnr"], 6),
            "psnr_delta_db": round(p_gpu["overall_psnr"] - p_ref["overall_psnr"], 9),
            "psnr": "calculate_image_metrics_rgba8 (tests/metrics.rs:40-114), RGB in linear light, "
                    "both block images decoded by the oracle's decoder",
        }
        if full_s is not None:
            parity["oracle_full_image_s"] = round(full_s, 3)
            parity["oracle_full_image_threads"] = threads

    if not args.no_cpu_baseline:
        cpu_baseline = cpu_baseline_leg(img, settings, nblocks, threads, host)
        ops = oracle_lib.count_ops_rgba8(6, settings.to_bytes(), img,
                                         np.sort(np.random.default_rng(7).choice(nblocks, 1024, replace=False)))
        ops_per_block = ops / 1024.0

    roofline = None
    if ops_per_block:
        achieved = ops_per_block * nblocks / (kernel_ms * 1e-3) / 1e12
        traffic = None
        tpath = os.path.join(ROOT, "profiles", "bc7_opaque_basic_4096_traffic.json")
        if os.path.exists(tpath):
            with open(tpath) as f:
                traffic = json.load(f).get("hbm_bytes_per_launch")
        roofline = {
            "bound": "valu",
            "achieved": round(achieved, 3),
            "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s",
            "frac": round(achieved / PEAK_FP32_TFLOPS, 4),
            "traffic": traffic,
            "ops_per_block": round(ops_per_block, 1),
            "issue_ceiling_no_fma": ISSUE_CEILING_TOPS,
            "frac_of_issue_ceiling": round(achieved / ISSUE_CEILING_TOPS, 4),
            "packed_ceiling_no_fma": PACKED_CEILING_TOPS,
            "frac_of_packed_ceiling": round(achieved / PACKED_CEILING_TOPS, 4),
            "algorithmic_bytes_per_launch": nblocks * 80,
            "hbm_frac": round(nblocks * 80 / (kernel_ms * 1e-3) / 1e9 / PEAK_HBM_GBS, 5),
        }

    formats = None
    if not args.no_formats:
        log("bench: other formats (BC1/BC3/BC4/BC5 4096^2, BC6H slow 2048^2, BC7 decode)")
        formats = bench_formats(img)
        if not args.no_cpu_baseline:
            count_format_ops(formats, img)

    batched = None
    if args.batched_textures > 0:
        batched = bench_batched(args.batched_textures, args.batched_reps, 24, threads)

    grad_refined = None
    if args.grad_size > 0:
        log(f"bench: gradient refinement {args.grad_size}^2 {args.grad_preset}")
        grad_refined = bench_grad(args.grad_size, args.grad_preset, 3)
        log(f"bench: refined {grad_refined['gain_db']:+.4f} dB at {grad_refined['time_ratio_vs_plain']:.3f}x time")

    line = {
        "metric": METRIC,
        "value": round(mpix, 2),
        "unit": "Mpix/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "bc7_opaque_basic_4096x4096",
            "format": "BC7",
            "preset": "opaque_basic",
            "width": W,
            "height": H,
            "texture": f"tile generator seed={seed} noise_frac={noise}",
            "parallelism": "single-gpu",
        },
        "kernel_ms": round(kernel_ms, 4),
        "roofline": roofline,
        "cpu_baseline": cpu_baseline,
        "parity": parity,
        "other_kernels": formats,
        "batched": batched,
        "grad_refined": grad_refined,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
