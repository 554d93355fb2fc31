"""DDS container and CLI host logic (block_compression_amd/dds.py, cli.py,
bcn_dds_* in the C ABI) without a GPU.

The reference CLI writes through the `ddsfile` crate 0.5, which the
reference checkout does not vendor; oracle/dds.cpp restates it.  Pins:
* a known-answer header for the CLI's BC7 4096x4096 call, field by field
  from the DDS file format documentation (DDS_HEADER + DDS_HEADER_DXT10);
  one word is parity unpinned: the pixel format's RGBBitCount (4 / 8 here,
  from ddsfile's bits-per-pixel), which readers ignore when the FOURCC flag
  is set;
* the C ABI's header == the oracle's over every BC format, shapes, mip
  counts, array sizes and alpha modes;
* Pillow's independent DDS reader opens the files and decodes BC4/BC5/BC7
  to exactly the oracle's decode, BC1-BC3 to within 1 (Pillow expands 565
  and interpolates with its own rounding) with alpha exact;
* read_dds(write_dds(x)) == x, with mip chains and arrays.
"""
import struct

import numpy as np
import pytest

import oracle_lib

from block_compression_amd import BcnError, cli, dds, synth
from block_compression_amd.settings import BC6HSettings, BC7Settings, CompressionVariant

SHAPES = [(4, 4, 1, 1), (64, 32, 1, 1), (4096, 4096, 1, 1), (13, 7, 3, 2), (1, 1, 1, 1),
          (2048, 1024, 12, 1), (100, 60, 7, 6)]


def test_known_answer_header_bc7_4096():
    fourcc = lambda s: int.from_bytes(s, "little")  # noqa: E731
    words = [fourcc(b"DDS "), 124,
             0x1 | 0x2 | 0x4 | 0x1000 | 0x80000,  # CAPS HEIGHT WIDTH PIXELFORMAT LINEARSIZE
             4096, 4096,
             1024 * 16 * 1024,  # linear size: 1024 blocks per row x 16 bytes x 1024 rows
             0, 1,              # depth, mip count
             *([0] * 11),
             32, 0x4, fourcc(b"DX10"), 8, 0, 0, 0, 0,  # pixel format: FOURCC "DX10"
             0x1000, 0, 0, 0, 0,  # caps TEXTURE
             99, 3, 0, 1, 1]      # BC7_UNORM_SRGB, TEXTURE2D, misc 0, 1 layer, STRAIGHT
    want = struct.pack("<37I", *words)
    v = CompressionVariant.BC7(BC7Settings.alpha_slow())
    assert len(want) == dds.HEADER_SIZE == 148
    assert dds.header(dds.dxgi_format(v), 4096, 4096) == want
    assert oracle_lib.dds_header(99, 4096, 4096) == want


def test_dxgi_format_mapping():
    # compressor/src/main.rs:355-365
    got = [dds.dxgi_format(CompressionVariant(t, s)) for t, s in
           [(0, None), (1, None), (2, None), (3, None), (4, None),
            (5, BC6HSettings.basic()), (6, BC7Settings.alpha_basic())]]
    assert [dds.DXGI_NAMES[f] for f in got] == [
        "BC1_UNORM_SRGB", "BC2_UNORM_SRGB", "BC3_UNORM_SRGB", "BC4_UNORM", "BC5_UNORM",
        "BC6H_UF16", "BC7_UNORM_SRGB"]
    # The signed BC6H extension.
    assert dds.DXGI_NAMES[dds.dxgi_format(CompressionVariant.BC6H_SF16(BC6HSettings.basic()))] == "BC6H_SF16"


def test_header_matches_oracle_everywhere():
    for fmt in dds.DXGI_NAMES:
        for w, h, m, l in SHAPES:
            for a in range(5):
                assert dds.header(fmt, w, h, m, l, a) == oracle_lib.dds_header(fmt, w, h, m, l, a), \
                    (fmt, w, h, m, l, a)


def test_header_rejects_bad_arguments():
    with pytest.raises(BcnError, match="UNSUPPORTED"):
        dds.header(28, 4, 4)  # R8G8B8A8_UNORM is not block-compressed
    with pytest.raises(BcnError, match="INVALID"):
        dds.header(99, 0, 4)
    with pytest.raises(BcnError, match="INVALID"):
        dds.header(99, 4, 4, alpha_mode=5)
    assert dds.data_size(28, 4, 4) == 0


def test_data_size():
    # 13x7, 3 levels: 4x2 + 2x1 + 1x1 blocks; 2 layers; 8 bytes (BC1).
    assert dds.data_size(71, 13, 7, 3, 2) == (8 + 2 + 1) * 8 * 2
    assert dds.data_size(99, 4096, 4096) == 4096 * 4096
    # Levels below 4 texels still take one block.
    assert dds.data_size(98, 4, 4, 3, 1) == 3 * 16


def _img(n):
    img = synth.tile_image_rgba8(n, n, 12, 0.1)
    img[..., 3] = synth.tile_image_rgba8(n, n, 13, 0.2)[..., 0]
    return img


@pytest.mark.parametrize("tag,fmt,exact", [(0, 71, False), (1, 74, False), (2, 77, False),
                                           (3, 80, True), (4, 83, True), (6, 99, True)])
def test_pillow_reads_and_decodes_our_files(tmp_path, tag, fmt, exact):
    from PIL import Image

    img = _img(64)
    settings = BC7Settings.alpha_basic().to_bytes() if tag == 6 else b""
    blocks = oracle_lib.compress_rgba8(tag, settings, img, threads=4)
    path = tmp_path / "t.dds"
    dds.write_dds(path, fmt, 64, 64, blocks)
    with Image.open(path) as im:
        im.load()
        assert im.size == (64, 64)
        got = np.asarray(im)
    got = got[..., None] if got.ndim == 2 else got
    want = oracle_lib.decompress_rgba8(tag, 64, 64, blocks)[..., :got.shape[2]]
    d = np.abs(got.astype(int) - want.astype(int))
    if exact:
        assert d.max() == 0
    else:
        assert d.max() <= 1 and d[..., 3].max() == 0


def test_write_read_round_trip(tmp_path):
    rng = np.random.default_rng(5)
    for fmt, (w, h, m, l) in [(72, (64, 32, 1, 1)), (95, (13, 7, 3, 2)), (80, (100, 60, 7, 6))]:
        data = rng.integers(0, 256, dds.data_size(fmt, w, h, m, l), dtype=np.uint8)
        p = tmp_path / f"{fmt}.dds"
        dds.write_dds(p, fmt, w, h, data, m, l, dds.ALPHA_PREMULTIPLIED)
        f = dds.read_dds(p)
        assert (f.dxgi_format, f.width, f.height, f.mip_levels, f.array_layers, f.alpha_mode) == \
            (fmt, w, h, m, l, dds.ALPHA_PREMULTIPLIED)
        assert f.data == data.tobytes()
    with pytest.raises(ValueError, match="needs"):
        dds.write_dds(tmp_path / "x.dds", 99, 8, 8, np.zeros(63, np.uint8))


def test_dds_path_is_set_extension():
    assert dds.dds_path("a/b/marble.png") == "a/b/marble.dds"
    assert dds.dds_path("noext") == "noext.dds"
    assert dds.dds_path("x.tar.gz") == "x.tar.dds"
    assert dds.dds_path(".hidden") == ".hidden.dds"


def test_cli_argument_parsing(capsys):
    assert cli.parse_args(["compressor"]) is None
    assert cli.parse_args(["compressor", "bc7", "--help"]) is None
    assert cli.parse_args(["compressor", "bc9", "x.png"]) is None
    assert "Invalid compression variant" in capsys.readouterr().out
    v, f = cli.parse_args(["compressor", "BC7", "x.png"])
    assert v == CompressionVariant.BC7(BC7Settings.alpha_slow()) and f == "x.png"
    v, _ = cli.parse_args(["compressor", "bc6h", "x.png"])
    assert v == CompressionVariant.BC6H(BC6HSettings.very_slow())
    assert cli.parse_args(["compressor", "bc4", "x.png"])[0] == CompressionVariant.BC4()


def test_cli_srgb_to_linear_half():
    v = np.arange(256, dtype=np.uint8)
    got = cli.srgb_to_linear_f16(v)
    assert got[0] == 0 and got[255] == 1.0
    # 128/255 = 0.50196 -> ((0.50196 + 0.055) / 1.055)^2.4 = 0.2158605 -> half 0.2158203
    assert got[128] == np.float16(0.2158605)
    # the linear segment: 10/255 / 12.92 = 0.0030353
    assert got[10] == np.float16(10 / 255 / 12.92)
    assert np.all(np.diff(got.astype(np.float64)) > 0)


def test_cli_reads_png_as_rgba8(tmp_path):
    from PIL import Image

    img = _img(16)
    Image.fromarray(img[..., :3], "RGB").save(tmp_path / "rgb.png")
    got = cli.read_image(str(tmp_path / "rgb.png"), CompressionVariant.BC1())
    assert got.dtype == np.uint8 and np.array_equal(got[..., :3], img[..., :3]) and np.all(got[..., 3] == 255)
    h16 = cli.read_image(str(tmp_path / "rgb.png"), CompressionVariant.BC6H(BC6HSettings.basic()))
    assert h16.dtype == np.float16 and np.all(h16[..., 3] == 1.0)
