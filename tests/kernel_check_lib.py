"""ctypes bridge to tests/kernel_check/_build/libkernel_check.so: the
product's per-block kernel code (block_compression_amd/csrc/*_block.hpp)
compiled for the host, so its arithmetic can be compared with the oracle
without a GPU."""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# KC_LIB names another build of kernel_check.cpp (the sanitizer build of
# test_sanitizers_cpu.py); default: the plain build, made on first use.
LIB = os.environ.get("KC_LIB") or os.path.join(HERE, "kernel_check", "_build", "libkernel_check.so")
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.environ.get("KC_LIB"):
            subprocess.run([os.path.join(HERE, "kernel_check", "build.sh")], check=True)
        _lib = ctypes.CDLL(LIB)
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def bc7_encode(settings: bytes, img: np.ndarray) -> np.ndarray:
    img = np.ascontiguousarray(img)
    h, w, _ = img.shape
    out = np.zeros((w // 4) * (h // 4) * 16, np.uint8)
    lib().kcheck_bc7_encode(ctypes.create_string_buffer(settings), _p(img), w, h, w * 4, _p(out))
    return out


def pbit_mismatches(levels: int) -> int:
    f = lib().kcheck_pbit_mismatches
    f.restype = ctypes.c_long
    return f(levels)


def quant4_mismatches() -> int:
    f = lib().kcheck_quant4_mismatches
    f.restype = ctypes.c_long
    return f()


def chan_q4_mismatches() -> int:
    f = lib().kcheck_chan_q4_mismatches
    f.restype = ctypes.c_long
    return f()


def chan_q8_mismatches() -> int:
    f = lib().kcheck_chan_q8_mismatches
    f.restype = ctypes.c_long
    return f()


def quant8_mismatches() -> int:
    f = lib().kcheck_quant8_mismatches
    f.restype = ctypes.c_long
    return f()


def bc15_encode(fmt: int, img: np.ndarray) -> np.ndarray:
    img = np.ascontiguousarray(img)
    h, w, _ = img.shape
    bb = 8 if fmt in (0, 3) else 16
    out = np.zeros((w // 4) * (h // 4) * bb, np.uint8)
    lib().kcheck_bc15_encode(fmt, _p(img), w, h, w * 4, _p(out))
    return out


def bc6h_encode(settings: bytes, img16: np.ndarray, is_signed: bool = False) -> np.ndarray:
    bits = np.ascontiguousarray(img16).view(np.uint16)
    h, w, _ = bits.shape
    out = np.zeros((w // 4) * (h // 4) * 16, np.uint8)
    lib().kcheck_bc6h_encode(ctypes.create_string_buffer(settings), _p(bits), w, h, w * 4, _p(out),
                             int(is_signed))
    return out


def decode(fmt: int, blocks: np.ndarray, is_signed: bool = False, one_path: bool = False) -> np.ndarray:
    """one_path (BC6H): one-subset blocks through the decoder's one-subset
    path, as in a wave of only such blocks."""
    blocks = np.ascontiguousarray(blocks, dtype=np.uint8)
    bb = 8 if fmt in (0, 3) else 16
    n = blocks.size // bb
    out = np.zeros(n * 48, np.uint16) if fmt == 5 else np.zeros(n * 16, np.uint32)
    lib().kcheck_decode(fmt, int(is_signed) | (2 if one_path else 0), _p(blocks), n, _p(out))
    return out


def bc7_repack(blocks: np.ndarray) -> np.ndarray:
    blocks = np.ascontiguousarray(blocks, dtype=np.uint8)
    out = np.zeros_like(blocks)
    lib().kcheck_bc7_repack(_p(blocks), blocks.size // 16, _p(out))
    return out


def _nparams(fmt):
    return 24 if fmt == 6 else 6


def endpoints(fmt: int, blocks: np.ndarray) -> np.ndarray:
    blocks = np.ascontiguousarray(blocks, dtype=np.uint8)
    n = blocks.size // (16 if fmt == 6 else 8)
    x = np.zeros(n * _nparams(fmt), np.float32)
    lib().kcheck_endpoints(fmt, _p(blocks), n, _p(x))
    return x.reshape(n, -1)


def forward(fmt: int, blocks: np.ndarray, x: np.ndarray, smooth: bool) -> np.ndarray:
    blocks = np.ascontiguousarray(blocks, dtype=np.uint8)
    x = np.ascontiguousarray(x, dtype=np.float32)
    n = x.shape[0]
    t = np.zeros((n, 16, 4), np.float32)
    lib().kcheck_forward(fmt, _p(blocks), _p(x), n, int(smooth), _p(t))
    return t


def backward(fmt: int, blocks: np.ndarray, g: np.ndarray) -> np.ndarray:
    blocks = np.ascontiguousarray(blocks, dtype=np.uint8)
    g = np.ascontiguousarray(g, dtype=np.float32)
    n = g.shape[0]
    gx = np.zeros((n, _nparams(fmt)), np.float32)
    lib().kcheck_backward(fmt, _p(blocks), _p(g), n, _p(gx))
    return gx


def texel_backward(fmt: int, blocks: np.ndarray, g: np.ndarray) -> np.ndarray:
    """Host build of the texel-side adjoint: g (n, 16, 4) -> (n, 16, 4)."""
    blocks = np.ascontiguousarray(blocks, dtype=np.uint8)
    g = np.ascontiguousarray(g, dtype=np.float32)
    out = np.zeros_like(g)
    lib().kcheck_texel_backward(fmt, _p(blocks), _p(g), g.shape[0], _p(out))
    return out


def bc7_refine(blocks: np.ndarray, target: np.ndarray, steps: int, lr: float, channels: int,
               reassign: int = 0):
    blocks = np.ascontiguousarray(blocks, dtype=np.uint8)
    target = np.ascontiguousarray(target, dtype=np.float32)
    n = target.shape[0]
    out = np.zeros_like(blocks)
    s0 = np.zeros(n, np.float32)
    s1 = np.zeros(n, np.float32)
    lib().kcheck_bc7_refine(_p(blocks), _p(target), n, steps, ctypes.c_float(lr), channels, reassign,
                            _p(out), _p(s0), _p(s1))
    return out, s0, s1


def block_pos(bw: np.ndarray, idx: np.ndarray):
    bw = np.ascontiguousarray(bw, np.uint32)
    idx = np.ascontiguousarray(idx, np.uint32)
    rows = np.zeros(len(bw), np.uint32)
    cols = np.zeros(len(bw), np.uint32)
    lib().kcheck_block_pos(_p(bw), _p(idx), ctypes.c_size_t(len(bw)), _p(rows), _p(cols))
    return rows, cols
