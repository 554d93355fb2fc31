"""Differentiable BC7 / BC1 decode and gradient endpoint refinement on the GPU.

No counterpart in the reference crate.  The decode of a block is viewed as a
function of continuous endpoints with the rest of the block held fixed (mode,
partition, rotation, p-bits, indices); see include/bcn_mi355x.h for the
parameter layout and the straight-through / smooth forwards.

    x = endpoints(variant, blocks)                    # (n, 24) BC7, (n, 6) BC1
    texels = DecodeWithEndpoints.apply(x, blocks, variant_tag, smooth)   # (n, 16, 4)
    loss(texels).backward()                           # x.grad via the HIP adjoint

    # texture training: gradients to the source texels through the blocks
    src = image_to_blocks(texture)                    # (n, 16, 4), 0..255 units
    decoded = DecodeTexels.apply(src, blocks, variant_tag)
    loss(decoded).backward()                          # texture.grad via the texel-side adjoint

    refined, sse0, sse1 = refine_bc7(blocks, target, width, height)

torch supplies device memory and streams; the arithmetic is the HIP kernels'.
"""
from __future__ import annotations

import ctypes

from . import _capi

BC1, BC7 = 0, 6


def _nparams(tag: int) -> int:
    if tag == BC7:
        return 24
    if tag == BC1:
        return 6
    raise ValueError("the decode adjoint covers BC1 (tag 0) and BC7 (tag 6)")


def _block_bytes(tag: int) -> int:
    return 16 if tag == BC7 else 8


def _stream(dev) -> ctypes.c_void_p:
    import torch

    return ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)


def _check_blocks(blocks, tag):
    import torch

    if blocks.dtype != torch.uint8 or not blocks.is_cuda or not blocks.is_contiguous():
        raise TypeError("blocks must be a contiguous uint8 device tensor")
    if blocks.numel() % _block_bytes(tag):
        raise ValueError("blocks size is not a whole number of blocks")
    return blocks.numel() // _block_bytes(tag)


def endpoints(tag: int, blocks):
    """The blocks' own endpoints as float parameters, shape (n, 24) or (n, 6)."""
    import torch

    n = _check_blocks(blocks, tag)
    x = torch.empty((n, _nparams(tag)), dtype=torch.float32, device=blocks.device)
    _capi.check(_capi.lib().bcn_decode_endpoints(tag, blocks.data_ptr(), n, x.data_ptr(),
                                                  _stream(blocks.device)))
    return x


def decode_with_endpoints(tag: int, blocks, x, smooth: bool = False):
    import torch

    n = _check_blocks(blocks, tag)
    if x.shape != (n, _nparams(tag)) or x.dtype != torch.float32 or not x.is_contiguous():
        raise ValueError(f"x must be contiguous float32 of shape ({n}, {_nparams(tag)})")
    out = torch.empty((n, 16, 4), dtype=torch.float32, device=blocks.device)
    _capi.check(_capi.lib().bcn_decode_with_endpoints(tag, blocks.data_ptr(), x.data_ptr(), n,
                                                       int(smooth), out.data_ptr(), _stream(blocks.device)))
    return out


def decode_with_endpoints_backward(tag: int, blocks, grad_texels):
    import torch

    n = _check_blocks(blocks, tag)
    g = grad_texels.contiguous().to(torch.float32)
    if g.shape != (n, 16, 4):
        raise ValueError(f"grad_texels must have shape ({n}, 16, 4)")
    gx = torch.empty((n, _nparams(tag)), dtype=torch.float32, device=blocks.device)
    _capi.check(_capi.lib().bcn_decode_with_endpoints_backward(tag, blocks.data_ptr(), g.data_ptr(), n,
                                                                gx.data_ptr(), _stream(blocks.device)))
    return gx


def decode_texels_backward(tag: int, blocks, grad_texels):
    """Texel-side adjoint (bcn_decode_texels_backward): dL/d source texels
    (n, 16, 4) for dL/d decoded texels (n, 16, 4), the decode viewed as the
    least-squares re-fit of the source with each block's form held fixed."""
    import torch

    n = _check_blocks(blocks, tag)
    _nparams(tag)
    g = grad_texels.contiguous().to(torch.float32)
    if g.shape != (n, 16, 4) or g.device != blocks.device:
        raise ValueError(f"grad_texels must have shape ({n}, 16, 4) on the blocks' device")
    gs = torch.empty((n, 16, 4), dtype=torch.float32, device=blocks.device)
    _capi.check(_capi.lib().bcn_decode_texels_backward(tag, blocks.data_ptr(), g.data_ptr(), n,
                                                        gs.data_ptr(), _stream(blocks.device)))
    return gs


def image_to_blocks(img):
    """(H, W, C) -> (H/4 * W/4, 16, C): each block's texels row by row, blocks
    in row-major order (the layout of the adjoint functions).  A view-based
    torch reshape, so it is differentiable."""
    h, w, c = img.shape
    return img.reshape(h // 4, 4, w // 4, 4, c).permute(0, 2, 1, 3, 4).reshape(-1, 16, c)


def blocks_to_image(x, height: int, width: int):
    """Inverse of image_to_blocks."""
    c = x.shape[-1]
    return x.reshape(height // 4, width // 4, 4, 4, c).permute(0, 2, 1, 3, 4).reshape(height, width, c)


class DecodeTexels:
    """torch.autograd.Function for texture training: the decoded texels
    (n, 16, 4) of `blocks`, the encoding of the source texels `texels`
    (n, 16, 4).  The value is the exact decode of `blocks` (0..255 units);
    the backward is the texel-side adjoint (decode_texels_backward), so the
    gradient reaches the source texture through the block structure instead
    of as the identity.  Re-encode the source (GpuBlockCompressor) whenever
    it has changed enough for the blocks' form to change."""

    @staticmethod
    def apply(texels, blocks, tag: int):
        return _texels_function().apply(texels, blocks, tag)


_FN_TEXELS = None


def _texels_function():
    global _FN_TEXELS
    if _FN_TEXELS is None:
        import torch

        class _DecodeTexels(torch.autograd.Function):
            @staticmethod
            def forward(ctx, texels, blocks, tag):
                n = _check_blocks(blocks, tag)
                if tuple(texels.shape) != (n, 16, 4) or texels.device != blocks.device:
                    raise ValueError(f"texels must have shape ({n}, 16, 4), on the blocks' device")
                ctx.save_for_backward(blocks)
                ctx.tag = tag
                return decode_with_endpoints(tag, blocks, endpoints(tag, blocks), False)

            @staticmethod
            def backward(ctx, grad_out):
                (blocks,) = ctx.saved_tensors
                return decode_texels_backward(ctx.tag, blocks, grad_out), None, None

        _FN_TEXELS = _DecodeTexels
    return _FN_TEXELS


class DecodeWithEndpoints:
    """torch.autograd.Function: texels (n, 16, 4) of `blocks` decoded with
    endpoints x; the backward is the HIP adjoint (straight-through through
    the rounding when smooth is False)."""

    @staticmethod
    def apply(x, blocks, tag: int, smooth: bool = False):
        return _function().apply(x, blocks, tag, smooth)


_FN = None


def _function():
    global _FN
    if _FN is None:
        import torch

        class _Decode(torch.autograd.Function):
            @staticmethod
            def forward(ctx, x, blocks, tag, smooth):
                ctx.save_for_backward(blocks)
                ctx.tag = tag
                return decode_with_endpoints(tag, blocks, x.detach().contiguous(), smooth)

            @staticmethod
            def backward(ctx, grad_out):
                (blocks,) = ctx.saved_tensors
                return decode_with_endpoints_backward(ctx.tag, blocks, grad_out), None, None, None

        _FN = _Decode
    return _FN


def refine_bc7(blocks, target, width: int, height: int, steps: int = 64, lr: float = 0.5,
               channels: int = 4, reassign: int = 4, pbits: int = 4, beta1: float = 0.9,
               beta2: float = 0.999, eps: float = 1e-8, out=None):
    """Gradient endpoint refinement of a BC7 block image (bcn_bc7_refine).
    target: (height, width, 4) float32 device tensor in [0, 1].  reassign /
    pbits: every that many steps, re-pick each texel's index / try each p-bit
    flipped (0 = never).  Returns (refined blocks, per-block squared error
    before, after), errors in 0..255 units over `channels` channels."""
    import torch

    n = _check_blocks(blocks, BC7)
    if n != (width // 4) * (height // 4):
        raise ValueError("blocks do not cover width x height")
    if target.dtype != torch.float32 or tuple(target.shape) != (height, width, 4):
        raise ValueError("target must be float32 of shape (height, width, 4)")
    if target.stride(2) != 1 or target.stride(1) != 4:
        raise ValueError("target texels must be packed")
    if out is None:
        out = torch.empty_like(blocks)
    elif _check_blocks(out, BC7) != n or out.device != blocks.device:
        raise ValueError("out must hold as many blocks as `blocks`, on the same device")
    if (blocks.data_ptr() | out.data_ptr() | target.data_ptr()) & 15:
        raise ValueError("blocks, out and target must be 16-byte aligned")
    if target.device != blocks.device:
        raise ValueError("target must be on the blocks' device")
    s0 = torch.empty(n, dtype=torch.float32, device=blocks.device)
    s1 = torch.empty(n, dtype=torch.float32, device=blocks.device)
    wsize = int(_capi.lib().bcn_bc7_refine_workspace_size(n))
    ws = torch.empty(wsize, dtype=torch.uint8, device=blocks.device)
    p = _capi.RefineParams(steps, lr, beta1, beta2, eps, channels, reassign, int(pbits))
    _capi.check(_capi.lib().bcn_bc7_refine(
        blocks.data_ptr(), out.data_ptr(), width, height, target.data_ptr(),
        target.stride(0) * 4, ctypes.byref(p), s0.data_ptr(), s1.data_ptr(), ws.data_ptr(), wsize,
        _stream(blocks.device)))
    return out, s0, s1
