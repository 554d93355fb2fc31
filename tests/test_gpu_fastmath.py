"""The kernels' square root, reciprocal and reciprocal square root (sqrt_cr,
rcp_cr, rsqrt_cr in
block_compression_amd/csrc/bcn_common.hpp) equal the compiler's correctly
rounded sqrtf and 1.0f / x on every finite float, swept exhaustively on the
GPU, and the unguarded reciprocal (rcp_normal) equals 1.0f / x on every normal
|x| < 2^126 (tests/fastmath/fastmath_check.hip, built with the library's flags by
__graft_entry__.build()).  The unguarded fast paths are swept too, to show
that the guards cover exactly where they differ: square roots of inputs below
2^-102 (exponent field <= 24) and reciprocals of zero, denormals and |x| >= 2^126."""
import ctypes
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

LIB = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fastmath", "_build", "libfastmath_check.so")


def _sweep(lo, hi, mode):
    lib = ctypes.CDLL(LIB)
    bs = (ctypes.c_ulonglong * 256)()
    br = (ctypes.c_ulonglong * 256)()
    rc = lib.fastmath_check(ctypes.c_uint32(lo), ctypes.c_uint32(hi), ctypes.c_int(mode), bs, br)
    assert rc == 0
    return np.array(bs[:], np.int64), np.array(br[:], np.int64)


def test_guarded_sqrt_and_reciprocal_exact_on_every_finite_float():
    for lo, hi in [(0, 0x7F800000), (0x80000000, 0xFF800000)]:
        print(f"fastmath: sweep {lo:#x}..{hi:#x}", flush=True)
        s, r = _sweep(lo, hi, 0)
        assert s.sum() == 0 and r.sum() == 0


def test_rsqrt_exact_on_every_non_negative_float():
    """rsqrt_cr (the power iteration's renormalisation) equals 1.0f / sqrtf(x)
    bit for bit on every non-negative float, +inf and NaN patterns included."""
    print("fastmath: rsqrt sweep", flush=True)
    lib = ctypes.CDLL(LIB)
    bad = (ctypes.c_ulonglong * 256)()
    assert lib.rsqrt_check(ctypes.c_uint32(0), ctypes.c_uint32(0x80000000), bad) == 0
    assert sum(bad[:]) == 0


def test_unguarded_fast_paths_differ_only_where_guarded():
    print("fastmath: unguarded sweep", flush=True)
    s, r = _sweep(0, 0x7F800000, 1)
    assert np.all(s[25:] == 0) and s[:25].sum() > 0
    assert set(np.nonzero(r)[0].tolist()) <= {0, 253, 254}
    # Negative divisors (the scalar-channel denominator can be negative).
    _, r = _sweep(0x80000000, 0xFF800000, 1)
    assert set(np.nonzero(r)[0].tolist()) <= {0, 253, 254}
