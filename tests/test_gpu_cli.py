"""The command-line compressor end to end on the GPU (block_compression_amd/
cli.py, the reference's compressor/src/main.rs): PNG in, DDS out.  The DDS
header equals the oracle's ddsfile restatement and the block data equals the
oracle's encoder with the CLI's preset for the variant (BC6H very_slow on the
sRGB-to-linear half-float image, BC7 alpha_slow)."""
import numpy as np
import pytest

import oracle_lib
from block_compression_amd import cli, dds, synth
from block_compression_amd.settings import BC6HSettings, BC7Settings

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["bc1", "bc3", "bc4", "bc5", "bc6h", "bc7"])
def test_cli_png_to_dds_matches_oracle(tmp_path, capsys, name):
    from PIL import Image

    print(f"cli: {name} 64x48", flush=True)
    w, h = 64, 48
    img = synth.tile_image_rgba8(w, h, 17, 0.1)
    img[..., 3] = synth.tile_image_rgba8(w, h, 18, 0.2)[..., 0]
    png = tmp_path / "in.png"
    Image.fromarray(img, "RGBA").save(png)
    assert cli.main(["compressor", name, str(png)]) == 0
    out = capsys.readouterr().out
    for line in ("Image read and upload took:", "Compression took:", "Block data download took:",
                 "DDS output to disk took:"):
        assert line in out
    f = dds.read_dds(tmp_path / "in.dds")
    tag = ["bc1", "bc2", "bc3", "bc4", "bc5", "bc6h", "bc7"].index(name)
    fmt = [72, 75, 78, 80, 83, 95, 99][tag]
    raw = (tmp_path / "in.dds").read_bytes()
    assert raw[:dds.HEADER_SIZE] == oracle_lib.dds_header(fmt, w, h)
    if name == "bc6h":
        half = cli.srgb_to_linear_f16(img)
        want = oracle_lib.compress_rgba16(BC6HSettings.very_slow().to_bytes(), half, threads=8)
    else:
        settings = BC7Settings.alpha_slow().to_bytes() if name == "bc7" else b""
        want = oracle_lib.compress_rgba8(tag, settings, img, threads=8)
    assert np.array_equal(np.frombuffer(f.data, np.uint8), want.reshape(-1))
