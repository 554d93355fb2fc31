"""ctypes bridge to the CPU oracle (oracle/_build/liboracle_bc.so).

Test infrastructure: the oracle is the checker the GPU path is compared
against, and the bench's CPU baseline.  Nothing in block_compression_amd/
imports this module.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
# BCO_ORACLE_LIB names another build of the oracle's sources (the planted-slip
# builds of test_oracle_pins_catch_slips.py); default: the oracle itself.
LIB_PATH = os.environ.get("BCO_ORACLE_LIB") or os.path.join(ORACLE_DIR, "_build", "liboracle_bc.so")
OPCOUNT_PATH = os.path.join(ORACLE_DIR, "_build", "liboracle_bc_opcount.so")

_lib = None
_oplib = None

_vp = ctypes.c_void_p
_sz = ctypes.c_size_t
_u32 = ctypes.c_uint32
_i32 = ctypes.c_int


def build() -> None:
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)


def _declare(lib):
    lib.oracle_compress_rgba8.argtypes = [_i32, _vp, _vp, _vp, _sz, _u32, _u32, _u32, _i32]
    lib.oracle_compress_rgba8.restype = _i32
    lib.oracle_compress_rgba16.argtypes = [_vp, _vp, _vp, _sz, _u32, _u32, _u32, _i32]
    lib.oracle_compress_rgba16.restype = _i32
    lib.oracle_compress_rgba8_sample.argtypes = [_i32, _vp, _vp, _u32, _u32, _vp, _sz, _vp, _i32]
    lib.oracle_compress_rgba8_sample.restype = _i32
    lib.oracle_bc7_block_errors.argtypes = [_vp, _vp, _u32, _u32, _u32, _vp, _vp]
    lib.oracle_bc6h_block_errors.argtypes = [_vp, _vp, _u32, _u32, _u32, _vp, _vp]
    for n in ("bc1", "bc2", "bc3", "bc4", "bc5", "bc7"):
        getattr(lib, f"oracle_decode_block_{n}").argtypes = [_vp, _vp, _sz]
        getattr(lib, f"oracle_decode_block_{n}").restype = None
    lib.oracle_decode_block_bc6h.argtypes = [_vp, _vp, _sz, _i32]
    lib.oracle_decode_block_bc6h_float.argtypes = [_vp, _vp, _sz, _i32]
    lib.oracle_decode_block_bc6h_comp.argtypes = [_vp, _vp, _i32]
    lib.oracle_decode_block_bc6h_comp.restype = None
    lib.oracle_decompress_rgba8.argtypes = [_i32, _u32, _u32, _vp, _sz, _vp, _sz]
    lib.oracle_decompress_rgba8.restype = _i32
    for n in ("16f", "32f"):
        getattr(lib, f"oracle_decompress_rgba{n}").argtypes = [_i32, _u32, _u32, _vp, _sz, _vp, _sz]
        getattr(lib, f"oracle_decompress_rgba{n}").restype = _i32
    lib.oracle_linear_to_srgb.argtypes = [ctypes.c_float]
    lib.oracle_compress_rgba16_sf16.argtypes = [_vp, _vp, _vp, _sz, _u32, _u32, _u32, _vp, _i32]
    lib.oracle_compress_rgba16_sf16.restype = _i32
    lib.oracle_dds_header.argtypes = [_u32, _u32, _u32, _u32, _u32, _u32, _vp]
    lib.oracle_dds_header.restype = _i32
    lib.oracle_linear_to_srgb.restype = ctypes.c_uint8
    lib.oracle_partial_sort_list.argtypes = [_vp, _sz, _u32]
    lib.oracle_get_tables.argtypes = [_vp, _vp, _vp, _vp, _vp]
    lib.oracle_get_partition_sets.argtypes = [_vp]
    lib.oracle_f32_to_f16.argtypes = [ctypes.c_float]
    lib.oracle_f32_to_f16.restype = ctypes.c_uint16
    lib.oracle_f16_to_f32.argtypes = [ctypes.c_uint16]
    lib.oracle_f16_to_f32.restype = ctypes.c_float
    # Single encoder primitives (pin tests).
    lib.oracle_compute_axis.argtypes = [_vp, _u32, _u32, _vp]
    lib.oracle_get_pca_bound.argtypes = [_vp, _u32]
    lib.oracle_get_pca_bound.restype = ctypes.c_float
    lib.oracle_block_pca_axis.argtypes = [_vp, _u32, _u32, _vp, _vp]
    lib.oracle_block_segment_core.argtypes = [_vp, _u32, _u32, _vp]
    lib.oracle_opt_endpoints.argtypes = [_vp, _u32, _vp, _u32, _u32, _vp]
    lib.oracle_block_quant.argtypes = [_vp, _u32, _vp, _u32, _vp]
    lib.oracle_block_quant.restype = ctypes.c_float
    lib.oracle_compute_stats_masked.argtypes = [_vp, _u32, _u32, _vp]
    lib.oracle_block_pca_bound_split.argtypes = [_vp, _u32, _u32]
    lib.oracle_block_pca_bound_split.restype = ctypes.c_float
    lib.oracle_bc7_ep_quant_pair.argtypes = [_i32, _vp, _i32, _vp]
    lib.oracle_bc7_ep_dequant_pair.argtypes = [_i32, _vp, _vp]
    lib.oracle_bc7_channel_opt_endpoints.argtypes = [_vp, _u32, _vp, _vp]
    lib.oracle_bc7_channel_opt_quant.argtypes = [_vp, _u32, _vp, _vp]
    lib.oracle_bc7_channel_opt_quant.restype = ctypes.c_float
    lib.oracle_bc1_enc_rgb565.argtypes = [ctypes.c_float, ctypes.c_float, ctypes.c_float]
    lib.oracle_bc1_enc_rgb565.restype = ctypes.c_int32
    lib.oracle_bc1_dec_rgb565.argtypes = [ctypes.c_int32, _vp]
    lib.oracle_bc1_fast_quant.argtypes = [_vp, ctypes.c_int32, ctypes.c_int32]
    lib.oracle_bc1_fast_quant.restype = ctypes.c_uint32
    lib.oracle_bc1_refine.argtypes = [_vp, _u32, _vp, _vp]
    lib.oracle_bc1_axis.argtypes = [_vp, _vp, _vp]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = ctypes.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


def oplib():
    global _oplib
    if _oplib is None:
        if not os.path.exists(OPCOUNT_PATH):
            build()
        _oplib = ctypes.CDLL(OPCOUNT_PATH)
        _declare(_oplib)
        _oplib.oracle_count_ops_rgba8.argtypes = [_i32, _vp, _vp, _u32, _u32, _vp, _sz]
        _oplib.oracle_count_ops_rgba8.restype = ctypes.c_uint64
        _oplib.oracle_count_ops_rgba16.argtypes = [_vp, _vp, _u32, _u32, _vp, _sz]
        _oplib.oracle_count_ops_rgba16.restype = ctypes.c_uint64
    return _oplib


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


def block_bytes(tag: int) -> int:
    return 8 if tag in (0, 3) else 16


def compress_rgba8(tag: int, settings: bytes, rgba: np.ndarray, threads: int = 1) -> np.ndarray:
    """compress_rgba8 (encode.rs:83-151) over a (h, w, 4) uint8 image."""
    rgba = np.ascontiguousarray(rgba, dtype=np.uint8)
    h, w, _ = rgba.shape
    out = np.zeros(((w + 3) // 4) * ((h + 3) // 4) * block_bytes(tag), np.uint8)
    sbuf = ctypes.create_string_buffer(settings or b"\0" * 4)
    rc = lib().oracle_compress_rgba8(tag, sbuf, _ptr(rgba), _ptr(out), out.size, w, h, w * 4, threads)
    if rc != 0:
        raise ValueError(f"oracle_compress_rgba8 failed: {rc}")
    return out


def compress_rgba16(settings: bytes, rgba16: np.ndarray, threads: int = 1) -> np.ndarray:
    """compress_rgba16 (encode.rs:211-251) over a (h, w, 4) float16 image."""
    bits = np.ascontiguousarray(rgba16).view(np.uint16)
    h, w, _ = bits.shape
    out = np.zeros((w // 4) * (h // 4) * 16, np.uint8)
    sbuf = ctypes.create_string_buffer(settings)
    rc = lib().oracle_compress_rgba16(sbuf, _ptr(bits), _ptr(out), out.size, w, h, w * 4, threads)
    if rc != 0:
        raise ValueError(f"oracle_compress_rgba16 failed: {rc}")
    return out


def compress_rgba16_sf16(settings: bytes, rgba16: np.ndarray, threads: int = 1):
    """Signed (SF16) BC6H encode of a (h, w, 4) float16 image (the oracle's
    extension of the reference's unsigned encoder); returns (blocks, the
    search error of each block)."""
    bits = np.ascontiguousarray(rgba16).view(np.uint16)
    h, w, _ = bits.shape
    out = np.zeros((w // 4) * (h // 4) * 16, np.uint8)
    errs = np.zeros((w // 4) * (h // 4), np.float32)
    sbuf = ctypes.create_string_buffer(settings)
    rc = lib().oracle_compress_rgba16_sf16(sbuf, _ptr(bits), _ptr(out), out.size, w, h, w * 4,
                                           _ptr(errs), threads)
    if rc != 0:
        raise ValueError(f"oracle_compress_rgba16_sf16 failed: {rc}")
    return out, errs


def compress_rgba8_sample(tag: int, settings: bytes, rgba: np.ndarray, block_ids: np.ndarray,
                          threads: int = 1) -> np.ndarray:
    rgba = np.ascontiguousarray(rgba, dtype=np.uint8)
    h, w, _ = rgba.shape
    ids = np.ascontiguousarray(block_ids, dtype=np.int64)
    out = np.zeros(ids.size * block_bytes(tag), np.uint8)
    sbuf = ctypes.create_string_buffer(settings or b"\0" * 4)
    lib().oracle_compress_rgba8_sample(tag, sbuf, _ptr(rgba), w, w * 4, _ptr(ids), ids.size,
                                       _ptr(out), threads)
    return out


def bc7_block_errors(settings: bytes, rgba: np.ndarray):
    rgba = np.ascontiguousarray(rgba, dtype=np.uint8)
    h, w, _ = rgba.shape
    n = (w // 4) * (h // 4)
    out = np.zeros(n * 16, np.uint8)
    errs = np.zeros(n, np.float32)
    sbuf = ctypes.create_string_buffer(settings)
    lib().oracle_bc7_block_errors(sbuf, _ptr(rgba), w, h, w * 4, _ptr(out), _ptr(errs))
    return out, errs


def bc6h_block_errors(settings: bytes, rgba16: np.ndarray):
    bits = np.ascontiguousarray(rgba16).view(np.uint16)
    h, w, _ = bits.shape
    n = (w // 4) * (h // 4)
    out = np.zeros(n * 16, np.uint8)
    errs = np.zeros(n, np.float32)
    sbuf = ctypes.create_string_buffer(settings)
    lib().oracle_bc6h_block_errors(sbuf, _ptr(bits), w, h, w * 4, _ptr(out), _ptr(errs))
    return out, errs


def decode_block(name: str, block: bytes, pitch: int, out_len: int) -> np.ndarray:
    """One block through decode_block_<name> with the given pitch (block.rs)."""
    src = np.frombuffer(bytes(block), np.uint8).copy()
    out = np.zeros(out_len, np.uint8)
    getattr(lib(), f"oracle_decode_block_{name}")(_ptr(src), _ptr(out), pitch)
    return out


def decode_block_bc6h(block: bytes, pitch: int, is_signed: bool) -> np.ndarray:
    src = np.frombuffer(bytes(block), np.uint8).copy()
    out = np.zeros(4 * pitch, np.uint16)
    lib().oracle_decode_block_bc6h(_ptr(src), _ptr(out), pitch, int(is_signed))
    return out


def decode_block_bc6h_float(block: bytes, pitch: int, is_signed: bool) -> np.ndarray:
    src = np.frombuffer(bytes(block), np.uint8).copy()
    out = np.zeros(4 * pitch, np.float32)
    lib().oracle_decode_block_bc6h_float(_ptr(src), _ptr(out), pitch, int(is_signed))
    return out


def decode_block_bc6h_comp(block: bytes, is_signed: bool) -> np.ndarray:
    """The 16 x 3 interpolated BC6H values before finish_unquantize."""
    src = np.frombuffer(bytes(block), np.uint8).copy()
    out = np.zeros(48, np.int32)
    lib().oracle_decode_block_bc6h_comp(_ptr(src), _ptr(out), int(is_signed))
    return out.reshape(16, 3)


def decompress_rgba8(tag: int, width: int, height: int, blocks: np.ndarray) -> np.ndarray:
    """decompress_blocks_as_rgba8 (decode.rs:336-387)."""
    blocks = np.ascontiguousarray(blocks, dtype=np.uint8)
    out = np.zeros(width * height * 4, np.uint8)
    rc = lib().oracle_decompress_rgba8(tag, width, height, _ptr(blocks), blocks.size, _ptr(out), out.size)
    if rc != 0:
        raise ValueError(f"oracle_decompress_rgba8 failed: {rc}")
    return out.reshape(height, width, 4)


def _decompress_float(name, dtype, tag, width, height, blocks, fill):
    blocks = np.ascontiguousarray(blocks, dtype=np.uint8)
    out = np.empty(width * height * 4, dtype)
    out.view(np.uint16 if dtype == np.float16 else np.uint32)[:] = fill
    rc = getattr(lib(), name)(tag, width, height, _ptr(blocks), blocks.size, _ptr(out), out.size)
    if rc != 0:
        raise ValueError(f"{name} failed: {rc}")
    return out.reshape(height, width, 4)


def decompress_rgba16f(tag: int, width: int, height: int, blocks: np.ndarray, fill: int = 0) -> np.ndarray:
    """decompress_blocks_as_rgba16f (decode.rs:397-430) into a buffer whose
    every element starts as the bit pattern `fill`; the reference's layout
    (RGB at stride 3 inside each block's RGBA span) is kept."""
    return _decompress_float("oracle_decompress_rgba16f", np.float16, tag, width, height, blocks, fill)


def decompress_rgba32f(tag: int, width: int, height: int, blocks: np.ndarray, fill: int = 0) -> np.ndarray:
    """decompress_blocks_as_rgba32f (decode.rs:438-468), as above."""
    return _decompress_float("oracle_decompress_rgba32f", np.float32, tag, width, height, blocks, fill)


def linear_to_srgb(values) -> np.ndarray:
    """linear_to_srgb (decode.rs:151-159) of each f32 value."""
    f = lib().oracle_linear_to_srgb
    return np.array([f(float(v)) for v in np.asarray(values, np.float32).ravel()], np.uint8)


def dds_header(dxgi_format: int, width: int, height: int, mips: int = 1, layers: int = 1,
               alpha_mode: int = 1) -> bytes:
    """The 148 bytes ddsfile 0.5 writes before the data (oracle/dds.cpp)."""
    out = np.zeros(148, np.uint8)
    rc = lib().oracle_dds_header(dxgi_format, width, height, mips, layers, alpha_mode, _ptr(out))
    if rc != 0:
        raise ValueError(f"oracle_dds_header failed: {rc}")
    return out.tobytes()


def tables():
    n = 128
    pattern = np.zeros(n, np.uint32)
    m0 = np.zeros(n, np.uint32)
    m1 = np.zeros(n, np.uint32)
    m2 = np.zeros(n, np.uint32)
    skips = np.zeros(n * 3, np.uint32)
    lib().oracle_get_tables(_ptr(pattern), _ptr(m0), _ptr(m1), _ptr(m2), _ptr(skips))
    return pattern, m0, m1, m2, skips.reshape(n, 3)


def partition_sets() -> np.ndarray:
    out = np.zeros((2, 64, 16), np.uint8)
    lib().oracle_get_partition_sets(_ptr(out))
    return out


def partial_sort_list(values, partial_count: int) -> np.ndarray:
    a = np.ascontiguousarray(values, dtype=np.int32).copy()
    lib().oracle_partial_sort_list(_ptr(a), a.size, partial_count)
    return a


def count_ops_rgba8(tag: int, settings: bytes, rgba: np.ndarray, block_ids: np.ndarray) -> int:
    rgba = np.ascontiguousarray(rgba, dtype=np.uint8)
    h, w, _ = rgba.shape
    ids = np.ascontiguousarray(block_ids, dtype=np.int64)
    sbuf = ctypes.create_string_buffer(settings or b"\0" * 4)
    return int(oplib().oracle_count_ops_rgba8(tag, sbuf, _ptr(rgba), w, w * 4, _ptr(ids), ids.size))


def count_ops_rgba16(settings: bytes, rgba16: np.ndarray, block_ids: np.ndarray) -> int:
    bits = np.ascontiguousarray(rgba16).view(np.uint16)
    h, w, _ = bits.shape
    ids = np.ascontiguousarray(block_ids, dtype=np.int64)
    sbuf = ctypes.create_string_buffer(settings)
    return int(oplib().oracle_count_ops_rgba16(sbuf, _ptr(bits), w, w * 4, _ptr(ids), ids.size))


# ---- single encoder primitives (the oracle's pin tests) ---------------------
#
# `block` arguments are 64 floats, channel-major: R[16], G[16], B[16], A[16].

def _f32(a, n=None):
    a = np.ascontiguousarray(a, dtype=np.float32)
    if n is not None:
        assert a.size == n
    return a


def compute_axis(covar10, iterations: int, channels: int) -> np.ndarray:
    """compute_axis (common.rs:260-292)."""
    c = _f32(covar10, 10)
    out = np.zeros(4, np.float32)
    lib().oracle_compute_axis(_ptr(c), iterations, channels, _ptr(out))
    return out


def get_pca_bound(covar10, channels: int) -> float:
    """get_pca_bound (common.rs:208-245)."""
    c = _f32(covar10, 10)
    return float(lib().oracle_get_pca_bound(_ptr(c), channels))


def block_pca_axis(block64, mask: int, channels: int):
    """block_pca_axis (common.rs:370-394): (axis, dc)."""
    b = _f32(block64, 64)
    axis = np.zeros(4, np.float32)
    dc = np.zeros(4, np.float32)
    lib().oracle_block_pca_axis(_ptr(b), mask, channels, _ptr(axis), _ptr(dc))
    return axis, dc


def block_segment_core(block64, mask: int, channels: int) -> np.ndarray:
    """block_segment_core (common.rs:483-518): endpoints A (0..3), B (4..7)."""
    b = _f32(block64, 64)
    ep = np.zeros(8, np.float32)
    lib().oracle_block_segment_core(_ptr(b), mask, channels, _ptr(ep))
    return ep


def opt_endpoints(block64, bits: int, qblock2, mask: int, channels: int) -> np.ndarray:
    """opt_endpoints (common.rs:142-205)."""
    b = _f32(block64, 64)
    q = np.ascontiguousarray(qblock2, dtype=np.uint32)
    ep = np.zeros(8, np.float32)
    lib().oracle_opt_endpoints(_ptr(b), bits, _ptr(q), mask, channels, _ptr(ep))
    return ep


def block_quant(block64, bits: int, ep8, channels: int):
    """block_quant (common.rs:422-481) with every texel in subset 0: (err, qblock)."""
    b = _f32(block64, 64)
    e = _f32(ep8, 8)
    q = np.zeros(2, np.uint32)
    err = lib().oracle_block_quant(_ptr(b), bits, _ptr(e), channels, _ptr(q))
    return float(err), q


def compute_stats_masked(block64, mask: int, channels: int) -> np.ndarray:
    b = _f32(block64, 64)
    st = np.zeros(15, np.float32)
    lib().oracle_compute_stats_masked(_ptr(b), mask, channels, _ptr(st))
    return st


def block_pca_bound_split(block64, mask: int, channels: int) -> float:
    """block_pca_bound_split (common.rs:396-420) with the block's full stats."""
    b = _f32(block64, 64)
    return float(lib().oracle_block_pca_bound_split(_ptr(b), mask, channels))


def bc7_ep_quant_pair(mode: int, ep8, channels: int) -> np.ndarray:
    """ep_quant0367 / ep_quant1 / ep_quant245 of one endpoint pair (bc7.rs:102-189)."""
    e = _f32(ep8, 8)
    q = np.zeros(8, np.int32)
    lib().oracle_bc7_ep_quant_pair(mode, _ptr(e), channels, _ptr(q))
    return q


def bc7_channel_opt_endpoints(channel16, bits: int, qblock2) -> np.ndarray:
    """channel_opt_endpoints (bc7.rs:322-369)."""
    c = _f32(channel16, 16)
    q = np.ascontiguousarray(qblock2, dtype=np.uint32)
    ep = np.zeros(2, np.float32)
    lib().oracle_bc7_channel_opt_endpoints(_ptr(c), bits, _ptr(q), _ptr(ep))
    return ep


def bc7_channel_opt_quant(channel16, bits: int, ep2):
    """channel_opt_quant (bc7.rs:278-320): (err, qblock)."""
    c = _f32(channel16, 16)
    e = _f32(ep2, 2)
    q = np.zeros(2, np.uint32)
    err = lib().oracle_bc7_channel_opt_quant(_ptr(c), bits, _ptr(e), _ptr(q))
    return float(err), q


def bc1_enc_rgb565(r: float, g: float, b: float) -> int:
    """enc_rgb565 (bc1_to_5.rs:268-278)."""
    return int(lib().oracle_bc1_enc_rgb565(r, g, b))


def bc1_dec_rgb565(p: int) -> np.ndarray:
    out = np.zeros(3, np.float32)
    lib().oracle_bc1_dec_rgb565(p, _ptr(out))
    return out


def bc1_fast_quant(block64, p0: int, p1: int) -> int:
    """fast_quant (bc1_to_5.rs:280-321): 2-bit indices, texel k at bits 2k."""
    b = _f32(block64, 64)
    return int(lib().oracle_bc1_fast_quant(_ptr(b), p0, p1))


def bc1_refine(block64, bits: int):
    """bc1_refine (bc1_to_5.rs:323-374): (colours c0 c1 before 565 rounding, codes)."""
    b = _f32(block64, 64)
    col = np.zeros(6, np.float32)
    pe = np.zeros(2, np.int32)
    lib().oracle_bc1_refine(_ptr(b), bits, _ptr(col), _ptr(pe))
    return col.reshape(2, 3), pe


def bc1_axis(block64):
    """BC1's principal axis and mean (bc1_to_5.rs:157-223, 386-397)."""
    b = _f32(block64, 64)
    axis = np.zeros(3, np.float32)
    dc = np.zeros(3, np.float32)
    lib().oracle_bc1_axis(_ptr(b), _ptr(axis), _ptr(dc))
    return axis, dc
