"""ctypes binding of the HIP library's C ABI (include/bcn_mi355x.h).

The library is block_compression_amd/_lib/libbcn_mi355x.so, built by
block_compression_amd/_build.py.  Loading fails loudly if it is missing:
there is no CPU fallback behind this package.
"""
from __future__ import annotations

import ctypes
import os

from ._build import LIB_PATH

_vp = ctypes.c_void_p
_sz = ctypes.c_size_t
_u32 = ctypes.c_uint32
_i32 = ctypes.c_int

BCN_OK = 0
STATUS_NAMES = {
    0: "BCN_OK",
    1: "BCN_ERR_INVALID_ARGUMENT",
    2: "BCN_ERR_NOT_MULTIPLE_OF_4",
    3: "BCN_ERR_BUFFER_TOO_SMALL",
    4: "BCN_ERR_TEXTURE_TOO_SMALL",
    5: "BCN_ERR_UNSUPPORTED",
    6: "BCN_ERR_MISALIGNED",
    7: "BCN_ERR_HIP",
    8: "BCN_ERR_NO_DEVICE",
    9: "BCN_ERR_TOO_LARGE",
}
NOALLOC_MAX_BLOCKS = 1 << 20

SRC_RGBA8 = 0
SRC_RGBA16F = 1
SRC_R8 = 2
SRC_RG8 = 3
SRC_BGRA8 = 4
SRC_RGBA32F = 5
DECODE_REFERENCE_LAYOUT = 1

# (name, restype, argtypes) of every exported symbol.
SIGNATURES = [
    ("bcn_compressor_create", _i32, [_i32, ctypes.POINTER(_vp)]),
    ("bcn_compressor_destroy", None, [_vp]),
    ("bcn_add_compression_task", _i32,
     [_vp, _i32, _vp, _sz, _vp, _i32, _u32, _u32, _u32, _u32, _vp, _sz, _u32, _u32]),
    ("bcn_compress", _i32, [_vp, _vp]),
    ("bcn_task_count", _sz, [_vp]),
    ("bcn_last_error", ctypes.c_char_p, [_vp]),
    ("bcn_status_string", ctypes.c_char_p, [_i32]),
    ("bcn_blocks_byte_size", _sz, [_i32, _u32, _u32]),
    ("bcn_decode_rgba8", _i32, [_i32, _vp, _sz, _u32, _u32, _vp, _u32, _vp]),
    ("bcn_decode_bc6h_rgba16f", _i32, [_vp, _sz, _u32, _u32, _i32, _vp, _u32, _vp]),
    ("bcn_decode_rgba16f", _i32, [_i32, _vp, _sz, _u32, _u32, _vp, _u32, _u32, _vp]),
    ("bcn_decode_rgba32f", _i32, [_i32, _vp, _sz, _u32, _u32, _vp, _u32, _u32, _vp]),
    ("bcn_bc6h_srgb8_thresholds", None, [_vp]),
    ("bcn_bc6h_srgb8_buckets", _i32, [_vp]),
    ("bcn_dxgi_format", _u32, [_i32]),
    ("bcn_dds_data_size", _sz, [_u32, _u32, _u32, _u32, _u32]),
    ("bcn_dds_write_header", _i32, [_u32, _u32, _u32, _u32, _u32, _u32, _vp, _sz]),
    ("bcn_partition_tables", None, [_vp, _vp, _vp, _vp]),
    ("bcn_batch_create", _i32, [_vp, _vp, ctypes.POINTER(_vp)]),
    ("bcn_batch_launch", _i32, [_vp, _vp]),
    ("bcn_batch_launch_graph", _i32, [_vp, _vp]),
    ("bcn_batch_autotune", _i32, [_vp, _vp, _i32, _vp]),
    ("bcn_batch_get_info", _i32, [_vp, _vp]),
    ("bcn_batch_last_error", ctypes.c_char_p, [_vp]),
    ("bcn_batch_destroy", None, [_vp]),
    ("bcn_decode_endpoints", _i32, [_i32, _vp, _u32, _vp, _vp]),
    ("bcn_decode_with_endpoints", _i32, [_i32, _vp, _vp, _u32, _i32, _vp, _vp]),
    ("bcn_decode_with_endpoints_backward", _i32, [_i32, _vp, _vp, _u32, _vp, _vp]),
    ("bcn_decode_texels_backward", _i32, [_i32, _vp, _vp, _u32, _vp, _vp]),
    ("bcn_bc7_refine", _i32, [_vp, _vp, _u32, _u32, _vp, _u32, _vp, _vp, _vp, _vp, _sz, _vp]),
    ("bcn_bc7_refine_workspace_size", _sz, [_u32]),
    ("bcn_bc7_refine_noalloc", _i32, [_vp, _vp, _u32, _u32, _vp, _u32, _vp, _vp, _vp, _vp]),
]


class RefineParams(ctypes.Structure):
    """bcn_refine_params."""

    _fields_ = [("steps", _u32), ("lr", ctypes.c_float), ("beta1", ctypes.c_float),
                ("beta2", ctypes.c_float), ("eps", ctypes.c_float), ("channels", _u32),
                ("reassign", _u32), ("pbits", _u32)]


class BatchConfig(ctypes.Structure):
    """bcn_batch_config."""

    _fields_ = [("chunk_blocks_bc7", _u32), ("chunk_blocks_bc15", _u32), ("waves_per_cu", _u32),
                ("serial_classes", _u32)]

    def as_dict(self) -> dict:
        return {k: int(getattr(self, k)) for k, _ in self._fields_}


class BatchInfo(ctypes.Structure):
    """bcn_batch_info."""

    _fields_ = [("tasks", _u32), ("persistent_tasks", _u32), ("blocks", ctypes.c_uint64),
                ("chunks", _u32), ("grid", _u32), ("config", BatchConfig)]


class BcnError(RuntimeError):
    """A non-zero bcn_status from the C ABI."""

    def __init__(self, status: int, message: str = ""):
        self.status = status
        name = STATUS_NAMES.get(status, str(status))
        super().__init__(f"{name}: {message}" if message else name)


_lib = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python -m block_compression_amd._build` "
                "(or __graft_entry__.build()); there is no CPU fallback"
            )
        l = ctypes.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            fn = getattr(l, name)
            fn.restype = res
            fn.argtypes = args
        _lib = l
    return _lib


def check(status: int, handle=None, batch=None) -> None:
    if status != BCN_OK:
        msg = ""
        if handle is not None:
            msg = (lib().bcn_last_error(handle) or b"").decode()
        elif batch is not None:
            msg = (lib().bcn_batch_last_error(batch) or b"").decode()
        raise BcnError(status, msg)
