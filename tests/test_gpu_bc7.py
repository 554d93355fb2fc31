"""BC7 on the GPU vs the CPU oracle: byte-identical blocks.

Parity cases run the HIP kernels through the C ABI (GpuBlockCompressor) and
compare with oracle/ (the C++ restatement of src/encode/bc7.rs) on the same
seeded synthetic textures.
"""
import numpy as np
import pytest

import oracle_lib
from block_compression_amd import synth
from block_compression_amd.settings import BC7_PRESETS, BC7Settings, CompressionVariant

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu


def _gpu_bc7(img: np.ndarray, settings: BC7Settings) -> np.ndarray:
    from block_compression_amd import GpuBlockCompressor

    h, w, _ = img.shape
    v = CompressionVariant.BC7(settings)
    tex = torch.from_numpy(img).cuda()
    buf = torch.zeros(v.blocks_byte_size(w, h), dtype=torch.uint8, device="cuda")
    c = GpuBlockCompressor()
    c.add_compression_task(v, tex, w, h, buf)
    c.compress()
    torch.cuda.synchronize()
    return buf.cpu().numpy()


def _alpha_image(w, h, seed):
    img = synth.tile_image_rgba8(w, h, seed, 0.1)
    a = synth.tile_image_rgba8(w, h, seed + 100, 0.1)[..., 0]
    img[..., 3] = a
    return img


OPAQUE = sorted(p for p in BC7_PRESETS if p.startswith("opaque"))
ALPHA = sorted(p for p in BC7_PRESETS if p.startswith("alpha"))


def test_every_preset_is_covered():
    assert len(OPAQUE) + len(ALPHA) == len(BC7_PRESETS) == 10


@pytest.mark.parametrize("preset", OPAQUE)
def test_bc7_opaque_presets_match_oracle(preset):
    img = synth.tile_image_rgba8(128, 128, 11, 0.1)
    s = getattr(BC7Settings, preset)()
    ref = oracle_lib.compress_rgba8(6, s.to_bytes(), img, threads=8)
    got = _gpu_bc7(img, s)
    assert np.array_equal(got, ref), f"{(got.reshape(-1,16) != ref.reshape(-1,16)).any(1).sum()} blocks differ"


@pytest.mark.parametrize("preset", ALPHA)
def test_bc7_alpha_presets_match_oracle(preset):
    img = _alpha_image(96, 64, 12)
    s = getattr(BC7Settings, preset)()
    ref = oracle_lib.compress_rgba8(6, s.to_bytes(), img, threads=8)
    got = _gpu_bc7(img, s)
    assert np.array_equal(got, ref), f"{(got.reshape(-1,16) != ref.reshape(-1,16)).any(1).sum()} blocks differ"


def test_bc7_opaque_basic_4096_every_block():
    """Headline workload (4096^2, opaque_basic): all 1,048,576 blocks
    byte-for-byte against the oracle, and the per-image PSNR
    (tests/metrics.rs:40-114) of both decodes equal."""
    from metrics_ref import calculate_image_metrics_rgba8

    print("bc7: 4096^2 opaque_basic, every block", flush=True)
    img = synth.tile_image_rgba8(4096, 4096, 2, 0.1)
    s = BC7Settings.opaque_basic()
    got = _gpu_bc7(img, s)
    ref = oracle_lib.compress_rgba8(6, s.to_bytes(), img, threads=16)
    same = (got.reshape(-1, 16) == ref.reshape(-1, 16)).all(axis=1)
    assert same.all(), f"{int((~same).sum())} of {same.size} blocks differ"
    dec = oracle_lib.decompress_rgba8(6, 4096, 4096, got)
    p = calculate_image_metrics_rgba8(img, dec, 3)["overall_psnr"]
    assert p > 20.0
