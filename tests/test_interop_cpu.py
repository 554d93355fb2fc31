"""Texture layout inference and the argument checks of the interop path
(block_compression_amd/compressor.py texture_layout / import_texture and
bcn_add_compression_task) without a GPU."""
import numpy as np
import pytest

from block_compression_amd import _capi
from block_compression_amd.compressor import import_texture, texture_layout

torch = pytest.importorskip("torch")


def test_layout_inferred_from_dtype_and_channels():
    assert texture_layout(torch.zeros((8, 8), dtype=torch.uint8)) == _capi.SRC_R8
    assert texture_layout(torch.zeros((8, 8, 1), dtype=torch.uint8)) == _capi.SRC_R8
    assert texture_layout(torch.zeros((8, 8, 2), dtype=torch.uint8)) == _capi.SRC_RG8
    assert texture_layout(torch.zeros((8, 8, 4), dtype=torch.uint8)) == _capi.SRC_RGBA8
    assert texture_layout(torch.zeros((8, 8, 4), dtype=torch.uint8), "BGRA8") == _capi.SRC_BGRA8
    assert texture_layout(torch.zeros((8, 8, 4), dtype=torch.float16)) == _capi.SRC_RGBA16F
    assert texture_layout(torch.zeros((8, 8, 4), dtype=torch.float32)) == _capi.SRC_RGBA32F


def test_layout_rejections():
    with pytest.raises(TypeError, match="unsupported"):
        texture_layout(torch.zeros((8, 8, 3), dtype=torch.uint8))
    with pytest.raises(ValueError, match="does not match"):
        texture_layout(torch.zeros((8, 8, 2), dtype=torch.uint8), "rgba8")
    with pytest.raises(ValueError, match="unknown"):
        texture_layout(torch.zeros((8, 8, 4), dtype=torch.uint8), "rgb565")
    # Texels must be packed within a row; rows may have any pitch.
    big = torch.zeros((8, 16, 4), dtype=torch.uint8)
    assert texture_layout(big[:, 4:12]) == _capi.SRC_RGBA8
    with pytest.raises(ValueError, match="packed"):
        texture_layout(big[:, ::2])


def test_dlpack_import_shares_memory():
    a = np.arange(64, dtype=np.uint8).reshape(4, 4, 4)
    t = import_texture(a)  # numpy exports __dlpack__
    assert t.data_ptr() == a.ctypes.data
    with pytest.raises(TypeError, match="DLPack"):
        import_texture([1, 2, 3])


def test_region_wider_than_a_strided_view_is_rejected():
    """A row-strided view of a larger texture reports the large pitch; the
    width check uses the view's own width."""
    import torch

    from block_compression_amd.compressor import check_region

    big = torch.zeros((16, 256, 4), dtype=torch.uint8)
    view = big[:, 192:256]
    check_region(view, 64)
    with pytest.raises(ValueError, match="width"):
        check_region(view, 68)
    with pytest.raises(ValueError, match="width"):
        check_region(big[:, 200:204], 8)


def test_unorm8_to_half_is_the_nearest_half_of_the_exact_quotient():
    """The BC6H unorm8 source rule, c/255 in f32 then round to the nearest
    half (ties to even), equals the half nearest to the exact rational c/255
    for every c: there is no double-rounding case, so the rule has one
    meaning."""
    from fractions import Fraction

    halves = np.arange(0, 0x3C01, dtype=np.uint16).view(np.float16).astype(np.float64)
    for c in range(256):
        got = float((np.float32(c) / np.float32(255.0)).astype(np.float16))
        q = Fraction(c, 255)
        i = int(np.searchsorted(halves, float(q)))
        cands = [h for h in halves[max(i - 2, 0):i + 2]]
        best = min(cands, key=lambda h: abs(Fraction(h) - q))
        assert got == best, c
