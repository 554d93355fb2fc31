"""The C-ABI library without a GPU: it loads, exports every symbol that
include/bcn_mi355x.h declares, and its partition tables equal the oracle's
(which equal the reference's decoder tables, src/decode/block.rs:806-952)."""
import ctypes
import os
import re

import numpy as np
import pytest

import oracle_lib
from block_compression_amd import _capi
from block_compression_amd.settings import BC7Settings, CompressionVariant

HEADER = os.path.join(os.path.dirname(os.path.dirname(__file__)), "include", "bcn_mi355x.h")


def declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|void|size_t|uint32_t|const char\*)\s+(bcn_\w+)\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    lib = _capi.lib()
    names = declared_functions()
    assert len(names) >= 10
    for n in names:
        assert hasattr(lib, n), n
    assert sorted(n for n, _, _ in _capi.SIGNATURES) == names


def test_partition_tables_match_oracle():
    pattern = np.zeros(128, np.uint32)
    mask01 = np.zeros(128, np.uint32)
    a1 = np.zeros(128, np.uint8)
    a2 = np.zeros(128, np.uint8)
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    _capi.lib().bcn_partition_tables(p(pattern), p(mask01), p(a1), p(a2))
    opat, m0, m1, m2, skips = oracle_lib.tables()
    assert np.array_equal(pattern, opat)
    assert np.array_equal(mask01 & 0xFFFF, m0)
    assert np.array_equal(mask01 >> 16, m1)
    assert np.array_equal(m2 & 0xFFFF, ~(m0 | m1) & 0xFFFF)
    assert np.array_equal(a1, skips[:, 1])
    assert np.array_equal(a2, skips[:, 2])


def test_blocks_byte_size_matches_lib_rs():
    lib = _capi.lib()
    for tag in range(7):
        for w, h in [(4, 4), (5, 9), (4096, 4096), (1, 1), (0, 0)]:
            v = CompressionVariant(tag, BC7Settings.opaque_basic() if tag == 6 else
                                   (__import__("block_compression_amd").BC6HSettings.slow() if tag == 5 else None))
            assert lib.bcn_blocks_byte_size(tag, w, h) == v.blocks_byte_size(w, h)


def test_status_strings():
    lib = _capi.lib()
    for s in range(10):
        assert lib.bcn_status_string(s)
        assert lib.bcn_status_string(s) != b"unknown status"
    assert lib.bcn_status_string(10) == b"unknown status"


def test_add_task_preconditions_without_gpu():
    """The argument checks run before any device work, so a null compressor
    handle is rejected without a GPU."""
    lib = _capi.lib()
    rc = lib.bcn_add_compression_task(None, 6, None, 0, None, 0, 16, 4, 4, 4, None, 16, 0, 0)
    assert rc == 1
    assert lib.bcn_compress(None, None) == 1


def test_noalloc_refine_size_and_argument_checks_without_gpu():
    """bcn_bc7_refine_noalloc rejects an image beyond its fixed scratch
    (BCN_ERR_TOO_LARGE) and bad arguments before any device work."""
    lib = _capi.lib()
    p = _capi.RefineParams(1, 0.5, 0.9, 0.999, 1e-8, 4, 0, 0)
    big = 4 * int(np.sqrt(_capi.NOALLOC_MAX_BLOCKS))
    assert lib.bcn_bc7_refine_noalloc(None, None, big, big + 4, None, big * 16, ctypes.byref(p),
                                      None, None, None) == 9
    assert lib.bcn_bc7_refine_noalloc(None, None, big, big, None, big * 16, ctypes.byref(p),
                                      None, None, None) == 1
    assert lib.bcn_bc7_refine_noalloc(None, None, 6, 4, None, 96, ctypes.byref(p), None, None, None) == 1
    assert lib.bcn_bc7_refine_noalloc(None, None, 8, 8, None, 128, None, None, None, None) == 1


def test_texel_adjoint_argument_checks_without_gpu():
    """bcn_decode_texels_backward: BC1 and BC7 only, null or misaligned
    pointers rejected, zero blocks a no-op -- all before any device work; the
    image <-> block layout reshapes are exact inverses."""
    import torch

    from block_compression_amd import grad

    lib = _capi.lib()
    assert lib.bcn_decode_texels_backward(3, None, None, 0, None, None) == 1  # BC4: no adjoint
    assert lib.bcn_decode_texels_backward(6, None, None, 4, None, None) == 1
    assert lib.bcn_decode_texels_backward(6, 16, 8, 4, 16, None) == 6  # grad not 16-byte aligned
    assert lib.bcn_decode_texels_backward(0, 16, 16, 0, 16, None) == 0
    img = torch.arange(8 * 12 * 4, dtype=torch.float32).reshape(8, 12, 4)
    blocks = grad.image_to_blocks(img)
    assert blocks.shape == (6, 16, 4)
    assert torch.equal(blocks[1, 5], img[1, 5])  # block (0, 1), texel (1, 1)
    assert torch.equal(grad.blocks_to_image(blocks, 8, 12), img)
