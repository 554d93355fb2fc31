"""The oracle's decoders against the reference's 18 known-answer tests.

Fixture: tests/golden/decode_kats.json, extracted from
src/decode/block.rs:1405-1841 by tests/golden/make_decode_kats.py.  Like the
reference's `test_block` (block.rs:1409-1430), only bytes [y*pitch, y*pitch +
pitch) for y in 0..4 are compared; the BC7 KATs use pitch 8 with 16-byte rows,
so rows overlap exactly as in the reference.
"""
import json
import os

import numpy as np
import pytest

import oracle_lib

FIXTURE = os.path.join(os.path.dirname(__file__), "golden", "decode_kats.json")
KATS = json.load(open(FIXTURE))["kats"]


@pytest.mark.parametrize("kat", KATS, ids=[k["name"] for k in KATS])
def test_decode_kat(kat):
    comp = bytes(kat["compressed"])
    if kat["decoder"] == "bc6h":
        out = oracle_lib.decode_block_bc6h(comp, kat["pitch"], kat["signed"])
        assert out[:48].tolist() == kat["expected_f16_bits"]
    elif kat["decoder"] == "bc6h_float":
        out = oracle_lib.decode_block_bc6h_float(comp, kat["pitch"], kat["signed"])
        exp = np.array(kat["expected_f32"], dtype=np.float32)
        assert np.array_equal(out[:48], exp)
    else:
        pitch = kat["pitch"]
        out = oracle_lib.decode_block(kat["decoder"], comp, pitch, 64)
        exp = np.array(kat["expected"], dtype=np.uint8)
        for y in range(4):
            assert out[y * pitch:(y + 1) * pitch].tolist() == exp[y * pitch:(y + 1) * pitch].tolist(), y


def test_kat_count():
    assert len(KATS) == 18
