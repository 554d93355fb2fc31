"""GpuBlockCompressor: the reference crate's GPU task API over the HIP library.

Mirrors `GpuBlockCompressor` (src/block_compressor.rs:45-62):

    compressor = GpuBlockCompressor()                          # ::new  (:66)
    compressor.add_compression_task(variant, texture, width, height, buffer,
                                    texture_y_offset=None, blocks_offset=None)  # (:342)
    compressor.compress(stream=None)                            # (:540)

`texture` is a CUDA/HIP torch tensor of shape (H, W, 4): uint8 (Rgba8Unorm)
or float16 (Rgba16Float, BC6H only); `buffer` is a contiguous uint8 torch
tensor on the same device.  The argument checks are the reference's asserts
(block_compressor.rs:352-373), raised as BcnError.  compress() enqueues the
kernels on `stream` (default: torch's current stream) and empties the queue.
torch is used only to hand over device pointers and streams.
"""
from __future__ import annotations

import ctypes
from typing import List, Optional

from . import _capi
from .settings import BC6H, BC7, CompressionVariant


def _current_stream_handle(device) -> int:
    import torch

    return torch.cuda.current_stream(device).cuda_stream


class GpuBlockCompressor:
    def __init__(self, device: Optional[int] = None):
        import torch

        if device is None:
            device = torch.cuda.current_device()
        self.device = int(device)
        handle = ctypes.c_void_p()
        _capi.check(_capi.lib().bcn_compressor_create(self.device, ctypes.byref(handle)))
        self._handle = handle
        # Keep tensors alive until compress() has enqueued their work.
        self._keepalive: List[object] = []

    def close(self) -> None:
        if getattr(self, "_handle", None):
            _capi.lib().bcn_compressor_destroy(self._handle)
            self._handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_compression_task(
        self,
        variant: CompressionVariant,
        texture,
        width: int,
        height: int,
        buffer,
        texture_y_offset: Optional[int] = None,
        blocks_offset: Optional[int] = None,
    ) -> None:
        import torch

        if not isinstance(variant, CompressionVariant):
            raise TypeError("variant must be a CompressionVariant")
        if texture.dim() != 3 or texture.shape[2] != 4:
            raise ValueError("texture must have shape (height, width, 4)")
        if texture.dtype == torch.uint8:
            fmt = _capi.SRC_RGBA8
        elif texture.dtype == torch.float16:
            fmt = _capi.SRC_RGBA16F
        else:
            raise TypeError("texture must be uint8 (RGBA8) or float16 (RGBA16F)")
        if texture.stride(2) != 1 or texture.stride(1) != 4:
            raise ValueError("texture texels must be packed (stride (pitch, 4, 1))")
        if buffer.dtype != torch.uint8 or not buffer.is_contiguous():
            raise TypeError("buffer must be a contiguous uint8 tensor")
        if not texture.is_cuda or not buffer.is_cuda:
            raise ValueError("texture and buffer must be device tensors")
        pitch = texture.stride(0) * texture.element_size()
        settings = variant.settings_bytes()
        sbuf = ctypes.create_string_buffer(settings) if settings else None
        status = _capi.lib().bcn_add_compression_task(
            self._handle,
            variant.tag,
            ctypes.cast(sbuf, ctypes.c_void_p) if sbuf is not None else None,
            len(settings),
            texture.data_ptr(),
            fmt,
            pitch,
            texture.shape[0],
            width,
            height,
            buffer.data_ptr(),
            buffer.numel(),
            0 if texture_y_offset is None else texture_y_offset,
            0 if blocks_offset is None else blocks_offset,
        )
        _capi.check(status, self._handle)
        self._keepalive.extend((texture, buffer))

    def task_count(self) -> int:
        return int(_capi.lib().bcn_task_count(self._handle))

    def compress(self, stream=None) -> None:
        if stream is None:
            handle = _current_stream_handle(self.device)
        elif isinstance(stream, int):
            handle = stream
        else:
            handle = stream.cuda_stream
        status = _capi.lib().bcn_compress(self._handle, ctypes.c_void_p(handle))
        self._keepalive.clear()
        _capi.check(status, self._handle)


    def create_batch(self, config: Optional[dict] = None) -> "BlockCompressionBatch":
        """Moves every queued task into a BlockCompressionBatch (one persistent
        launch over all of them; see bcn_batch_create)."""
        return BlockCompressionBatch(self, config)


def _stream_handle(device, stream) -> int:
    if stream is None:
        return _current_stream_handle(device)
    if isinstance(stream, int):
        return stream
    return stream.cuda_stream


class BlockCompressionBatch:
    """Every task queued on a GpuBlockCompressor, encoded by one persistent
    kernel over a block work queue (no counterpart in the reference, which
    dispatches per task).  launch() may be called any number of times; the
    textures and buffers of the tasks must stay alive while the batch does."""

    def __init__(self, compressor: GpuBlockCompressor, config: Optional[dict] = None):
        self.device = compressor.device
        cfg = None
        if config is not None:
            cfg = _capi.BatchConfig(**config)
        handle = ctypes.c_void_p()
        status = _capi.lib().bcn_batch_create(
            compressor._handle, ctypes.byref(cfg) if cfg is not None else None, ctypes.byref(handle))
        _capi.check(status, compressor._handle)
        self._handle = handle
        self._keepalive = list(compressor._keepalive)
        compressor._keepalive.clear()

    def launch(self, stream=None, graph: bool = False) -> None:
        fn = _capi.lib().bcn_batch_launch_graph if graph else _capi.lib().bcn_batch_launch
        _capi.check(fn(self._handle, ctypes.c_void_p(_stream_handle(self.device, stream))),
                    batch=self._handle)

    def autotune(self, reps: int = 2, stream=None) -> dict:
        cfg = _capi.BatchConfig()
        _capi.check(_capi.lib().bcn_batch_autotune(
            self._handle, ctypes.c_void_p(_stream_handle(self.device, stream)), reps,
            ctypes.byref(cfg)), batch=self._handle)
        return cfg.as_dict()

    def info(self) -> dict:
        inf = _capi.BatchInfo()
        _capi.check(_capi.lib().bcn_batch_get_info(self._handle, ctypes.byref(inf)), batch=self._handle)
        return {"tasks": inf.tasks, "persistent_tasks": inf.persistent_tasks, "blocks": inf.blocks,
                "chunks": inf.chunks, "grid": inf.grid, "config": inf.config.as_dict()}

    def close(self) -> None:
        if getattr(self, "_handle", None):
            _capi.lib().bcn_batch_destroy(self._handle)
            self._handle = None
            self._keepalive = []

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def compress_rgba8(variant: CompressionVariant, texture, stream=None):
    """One-shot helper: encode a whole (H, W, 4) device texture, return the
    block buffer (a uint8 device tensor of variant.blocks_byte_size bytes)."""
    import torch

    h, w = int(texture.shape[0]), int(texture.shape[1])
    out = torch.empty(variant.blocks_byte_size(w, h), dtype=torch.uint8, device=texture.device)
    c = GpuBlockCompressor(texture.device.index)
    c.add_compression_task(variant, texture, w, h, out)
    c.compress(stream)
    return out


def decode_rgba8(variant: CompressionVariant, blocks, width: int, height: int, stream=None):
    """GPU decode of a block image into a (height, width, 4) uint8 device tensor
    in the layout of decompress_blocks_as_rgba8 (src/decode.rs:336-387)."""
    import torch

    out = torch.empty((height, width, 4), dtype=torch.uint8, device=blocks.device)
    handle = _current_stream_handle(blocks.device) if stream is None else stream.cuda_stream
    status = _capi.lib().bcn_decode_rgba8(
        variant.tag, blocks.data_ptr(), blocks.numel(), width, height, out.data_ptr(), width * 4,
        ctypes.c_void_p(handle))
    _capi.check(status)
    return out


def decode_bc6h_rgba16f(blocks, width: int, height: int, is_signed: bool = False, stream=None):
    """GPU decode of BC6H blocks into a (height, width, 4) float16 device tensor."""
    import torch

    out = torch.empty((height, width, 4), dtype=torch.float16, device=blocks.device)
    handle = _current_stream_handle(blocks.device) if stream is None else stream.cuda_stream
    status = _capi.lib().bcn_decode_bc6h_rgba16f(
        blocks.data_ptr(), blocks.numel(), width, height, int(is_signed), out.data_ptr(),
        width * 8, ctypes.c_void_p(handle))
    _capi.check(status)
    return out
