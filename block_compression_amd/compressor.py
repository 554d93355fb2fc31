"""GpuBlockCompressor: the reference crate's GPU task API over the HIP library.

Mirrors `GpuBlockCompressor` (src/block_compressor.rs:45-62):

    compressor = GpuBlockCompressor()                          # ::new  (:66)
    compressor.add_compression_task(variant, texture, width, height, buffer,
                                    texture_y_offset=None, blocks_offset=None)  # (:342)
    compressor.compress(stream=None)                            # (:540)

`texture` is a device tensor (or a DLPack-exporting array) whose dtype and
channel count give the texture view's format: uint8 RGBA8 / BGRA8 / RG8 / R8
for every format, float16 or float32 RGBA for BC6H only; `buffer` is a
contiguous uint8 torch tensor on the same device.  The argument checks are the reference's asserts
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
        texture_format: Optional[str] = None,
    ) -> None:
        """add_compression_task (src/block_compressor.rs:342-387).

        `texture` is a device tensor, or any object with `__dlpack__` (imported
        without a copy), of shape (H, W, C) or (H, W): its layout is the
        texture view's format (see texture_layout).  Rows may have any pitch
        (a row-strided view of a larger texture is fine); texels within a row
        must be packed.  `buffer` is a contiguous uint8 device tensor."""
        import torch

        if not isinstance(variant, CompressionVariant):
            raise TypeError("variant must be a CompressionVariant")
        texture = import_texture(texture)
        buffer = import_texture(buffer)
        fmt = texture_layout(texture, texture_format)
        if buffer.dtype != torch.uint8 or not buffer.is_contiguous():
            raise TypeError("buffer must be a contiguous uint8 tensor")
        if not texture.is_cuda or not buffer.is_cuda:
            raise ValueError("texture and buffer must be device tensors")
        check_region(texture, width)
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
        """compress (src/block_compressor.rs:540-590): enqueues every queued
        task on `stream` (a torch stream, a raw hipStream_t handle, or None for
        torch's current stream) and empties the queue.  Each texture and
        buffer is marked as used on that stream (record_stream), so torch's
        caching allocator does not hand their memory to a new tensor before
        the kernels that read and write it have run."""
        import torch

        if stream is None:
            stream = torch.cuda.current_stream(self.device)
        elif isinstance(stream, int):
            stream = torch.cuda.ExternalStream(stream, device=self.device)
        status = _capi.lib().bcn_compress(self._handle, ctypes.c_void_p(stream.cuda_stream))
        if status == _capi.BCN_OK:
            for t in self._keepalive:
                t.record_stream(stream)
        self._keepalive.clear()
        _capi.check(status, self._handle)


    def create_batch(self, config: Optional[dict] = None) -> "BlockCompressionBatch":
        """Moves every queued task into a BlockCompressionBatch (one launch
        over all of them: a persistent kernel per kernel class, the classes
        running concurrently; see bcn_batch_create)."""
        return BlockCompressionBatch(self, config)


# Source layouts by name (bcn_source_format).
SOURCE_FORMATS = {
    "rgba8": _capi.SRC_RGBA8,
    "bgra8": _capi.SRC_BGRA8,
    "r8": _capi.SRC_R8,
    "rg8": _capi.SRC_RG8,
    "rgba16f": _capi.SRC_RGBA16F,
    "rgba32f": _capi.SRC_RGBA32F,
}


def check_region(texture, width: int) -> None:
    """The region's width fits the texture view.  The C ABI sees a pointer
    and a row pitch only, so a row-strided view of a larger texture (e.g.
    big[:, 192:256]) would otherwise be read past its right edge, and on the
    last rows past the allocation; the reference's textureLoad stays inside
    the view.  Rows (height, texture_y_offset) are checked by the C ABI
    against the view's height."""
    if width > int(texture.shape[1]):
        raise ValueError(f"width {width} exceeds the texture view's width {int(texture.shape[1])}")


def import_texture(obj):
    """A torch tensor as is; any other object exporting `__dlpack__` (CuPy,
    JAX, another framework's device array) through torch.from_dlpack, which
    shares its memory."""
    import torch

    if isinstance(obj, torch.Tensor):
        return obj
    if hasattr(obj, "__dlpack__"):
        return torch.from_dlpack(obj)
    raise TypeError(f"expected a torch tensor or a DLPack-exporting array, got {type(obj).__name__}")


def texture_layout(texture, texture_format: Optional[str] = None) -> int:
    """The bcn_source_format of a texture tensor, from its dtype and channel
    count, the formats of the views the reference accepts
    (src/block_compressor.rs:300-313):

        uint8   (H, W) or (H, W, 1) -> R8      uint8   (H, W, 2) -> RG8
        uint8   (H, W, 4)           -> RGBA8   (or "bgra8" when named)
        float16 (H, W, 4)           -> RGBA16F float32 (H, W, 4) -> RGBA32F

    `texture_format` names the layout explicitly (a key of SOURCE_FORMATS);
    it must agree with the tensor's dtype and channel count."""
    import torch

    if texture.dim() == 2:
        channels = 1
    elif texture.dim() == 3:
        channels = int(texture.shape[2])
    else:
        raise ValueError("texture must have shape (height, width[, channels])")
    inferred = {(torch.uint8, 1): "r8", (torch.uint8, 2): "rg8", (torch.uint8, 4): "rgba8",
                (torch.float16, 4): "rgba16f", (torch.float32, 4): "rgba32f"}.get((texture.dtype, channels))
    if inferred is None:
        raise TypeError(f"unsupported texture: dtype {texture.dtype} with {channels} channels")
    name = inferred if texture_format is None else texture_format.lower()
    if name not in SOURCE_FORMATS:
        raise ValueError(f"unknown texture_format {texture_format!r}")
    if name != inferred and not (name == "bgra8" and inferred == "rgba8"):
        raise ValueError(f"texture_format {name!r} does not match a {inferred} tensor")
    # Texels packed within a row; rows at any pitch.
    if texture.dim() == 3 and texture.stride(2) != 1:
        raise ValueError("texture channels must be packed")
    if texture.stride(1) != channels:
        raise ValueError("texture texels must be packed within a row")
    return SOURCE_FORMATS[name]


def _stream_handle(device, stream) -> int:
    if stream is None:
        return _current_stream_handle(device)
    if isinstance(stream, int):
        return stream
    return stream.cuda_stream


class BlockCompressionBatch:
    """Every task queued on a GpuBlockCompressor, encoded by persistent
    kernels over a block work queue, one per kernel class, running
    concurrently (no counterpart in the reference, which dispatches per
    task).  launch() may be called any number of times; the
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
        import torch

        fn = _capi.lib().bcn_batch_launch_graph if graph else _capi.lib().bcn_batch_launch
        if stream is None:
            stream = torch.cuda.current_stream(self.device)
        elif isinstance(stream, int):
            stream = torch.cuda.ExternalStream(stream, device=self.device)
        _capi.check(fn(self._handle, ctypes.c_void_p(stream.cuda_stream)), batch=self._handle)
        # As in GpuBlockCompressor.compress: the memory stays reserved for the
        # launch stream's work even after the batch lets go of the tensors.
        for t in self._keepalive:
            t.record_stream(stream)

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
