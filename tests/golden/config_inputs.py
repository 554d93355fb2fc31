"""The benchmark workloads' seeded input buffers and their SHA-256 digests.

C1: BC1 512x512 RGBA8, tile generator seed 1 (no noise tiles), alpha 255.
C2: BC7 opaque_basic / BC4 / BC5 4096x4096 RGBA8, seed 2, 10% noise tiles
    (the bench's headline texture).
C3: BC6H slow 2048x2048 RGBA16F, HDR generator seed 3.
C4: the 300-texture mip-chain workload (seed 4): its texture specs (w, h,
    format) and the level-0 texels of its first 4 textures, generated on the
    CPU with the same torch generator the workload uses on the GPU.
C5: 1024x1024 float32 RGBA target, seed 5, no noise tiles.
"""
import hashlib
import json

import numpy as np


def _sha(a) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def digests() -> dict:
    import torch

    from block_compression_amd import synth, workloads

    c1 = synth.tile_image_rgba8(512, 512, 1, 0.0)
    c1[..., 3] = 255
    out = {
        "C1_bc1_512_rgba8_seed1": _sha(c1),
        "C2_bc7_4096_rgba8_seed2_noise0.1": _sha(synth.tile_image_rgba8(4096, 4096, 2, 0.1)),
        "C3_bc6h_2048_rgba16f_seed3": _sha(synth.hdr_image_rgba16f(2048, 2048, 3).view(np.uint16)),
        "C5_target_1024_rgba32f_seed5": _sha(synth.tile_target_rgba_f32(1024, 1024, 5, 0.0)),
    }
    specs = workloads.texture_specs(300, 4)
    out["C4_specs_300_seed4"] = hashlib.sha256(json.dumps(specs).encode()).hexdigest()
    for i, (w, h, fmt) in enumerate(specs[:4]):
        w8, h8 = (w + 7) // 8 * 8, (h + 7) // 8 * 8
        img = synth.tile_image_rgba8_torch(w8, h8, 4 * 1000 + i, 0.1, "cpu")
        if fmt in ("bc3", "bc7_alpha_basic"):
            img[..., 3] = synth.tile_image_rgba8_torch(w8, h8, 4 * 1000 + i + 7919, 0.1, "cpu")[..., 0]
        out[f"C4_texture{i}_{w}x{h}_{fmt}"] = _sha(img[:h, :w].contiguous().numpy())
    return out
