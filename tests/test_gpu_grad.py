"""The HIP decode adjoint and BC7 refinement (grad.hip) on the GPU.

* forward / backward kernels == the host build of the same code (which the
  CPU suite checks against the FP64 reference, torch.autograd and FP64
  finite differences), and the straight-through forward == the oracle decode;
* the torch.autograd.Function's gradient == the FP64 reference gradient and
  FP64 central differences (step 1e-3) on 256 random blocks;
* refinement is byte-identical to the host build (every step is IEEE
  single-precision arithmetic with one rounding) and never raises an error.
"""
import numpy as np
import pytest

import adjoint_ref as ref
import kernel_check_lib as kc
import oracle_lib

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu


def _blocks(fmt, n, seed):
    if fmt == 6:
        return ref.random_bc7_blocks(n, seed)
    return np.random.default_rng(seed).integers(0, 256, (n, 8), dtype=np.uint8)


@pytest.mark.parametrize("fmt", [6, 0], ids=["bc7", "bc1"])
def test_forward_backward_match_host_build(fmt):
    from block_compression_amd import grad

    print(f"grad: forward/backward fmt {fmt}", flush=True)
    b = _blocks(fmt, 4096, 21)
    bd = torch.from_numpy(b.reshape(-1).copy()).cuda()
    x = grad.endpoints(fmt, bd)
    assert np.array_equal(x.cpu().numpy(), kc.endpoints(fmt, b.reshape(-1)))
    rng = np.random.default_rng(22)
    xc = (x.cpu().numpy() + rng.uniform(-0.6, 0.6, x.shape)).astype(np.float32)
    xcd = torch.from_numpy(xc).cuda()
    for smooth in (False, True):
        got = grad.decode_with_endpoints(fmt, bd, xcd, smooth).cpu().numpy()
        assert np.array_equal(got, kc.forward(fmt, b.reshape(-1), xc, smooth)), smooth
    st = grad.decode_with_endpoints(fmt, bd, x, False).cpu().numpy()
    dec = np.stack([oracle_lib.decode_block("bc7" if fmt == 6 else "bc1", blk.tobytes(), 16, 64).reshape(16, 4)
                    for blk in b[:512]])
    assert np.array_equal(st[:512], dec.astype(np.float32))
    g = rng.standard_normal((len(b), 16, 4)).astype(np.float32)
    gx = grad.decode_with_endpoints_backward(fmt, bd, torch.from_numpy(g).cuda()).cpu().numpy()
    assert np.array_equal(gx, kc.backward(fmt, b.reshape(-1), g))


@pytest.mark.parametrize("fmt", [6, 0], ids=["bc7", "bc1"])
def test_autograd_function_vs_fp64_reference_and_finite_differences(fmt):
    from block_compression_amd import grad

    print(f"grad: autograd vs FP64 fmt {fmt}", flush=True)
    b = _blocks(fmt, 256, 23)
    bd = torch.from_numpy(b.reshape(-1).copy()).cuda()
    rng = np.random.default_rng(24)
    x0 = grad.endpoints(fmt, bd).cpu().numpy().astype(np.float64)
    xc = x0 + rng.uniform(-0.4, 0.4, x0.shape)
    target = rng.uniform(0, 255, (len(b), 16, 4))
    # Smooth forward through the Function; loss in float32 on the GPU.
    x = torch.tensor(xc, dtype=torch.float32, device="cuda", requires_grad=True)
    tex = grad.DecodeWithEndpoints.apply(x, bd, fmt, True)
    loss = ((tex - torch.tensor(target, dtype=torch.float32, device="cuda")) ** 2).sum()
    loss.backward()
    got = x.grad.double().cpu().numpy()
    shape = (6, 4) if fmt == 6 else (2, 3)
    smooth = ref.smooth_bc7 if fmt == 6 else ref.smooth_bc1
    parse = ref.parse_bc7 if fmt == 6 else ref.parse_bc1
    want = np.zeros_like(xc)
    fd = np.zeros_like(xc)
    h = 1e-3
    for i in range(len(b)):
        f = parse(b[i].tobytes())
        xi = torch.tensor(xc[i].reshape(shape), dtype=torch.float64, requires_grad=True)
        tt = torch.tensor(target[i], dtype=torch.float64)
        ((smooth(f, xi) - tt) ** 2).sum().backward()
        want[i] = xi.grad.numpy().reshape(-1)
        with torch.no_grad():
            for j in range(xc.shape[1]):
                d = torch.zeros(xc.shape[1], dtype=torch.float64)
                d[j] = h
                d = d.reshape(shape)
                lp = ((smooth(f, xi.detach() + d) - tt) ** 2).sum()
                lm = ((smooth(f, xi.detach() - d) - tt) ** 2).sum()
                fd[i, j] = float((lp - lm) / (2 * h))
    scale = np.abs(want).max()
    # float32 forward and loss on the GPU: relative 1e-4 of the largest gradient.
    assert np.abs(got - want).max() <= 1e-4 * scale
    assert np.abs(got - fd).max() <= 1e-4 * scale


def test_refine_matches_host_build_and_never_worse():
    from block_compression_amd import GpuBlockCompressor, grad, synth
    from block_compression_amd.settings import BC7Settings, CompressionVariant

    print("grad: refine 128x128 alpha_basic", flush=True)
    n = 128
    target = synth.tile_target_rgba_f32(n, n, 41, 0.0)
    img = np.clip(np.floor(target * 255 + 0.5), 0, 255).astype(np.uint8)
    v = CompressionVariant.BC7(BC7Settings.alpha_basic())
    tex = torch.from_numpy(img).cuda()
    blocks = torch.zeros(v.blocks_byte_size(n, n), dtype=torch.uint8, device="cuda")
    c = GpuBlockCompressor()
    c.add_compression_task(v, tex, n, n, blocks)
    c.compress()
    td = torch.from_numpy(target).cuda()
    out, s0, s1 = grad.refine_bc7(blocks, td, n, n, steps=64, lr=0.5, channels=4, reassign=1, pbits=True)
    torch.cuda.synchronize()
    tb = (target * 255).reshape(n // 4, 4, n // 4, 4, 4).transpose(0, 2, 1, 3, 4).reshape(-1, 16, 4)
    hout, h0, h1 = kc.bc7_refine(blocks.cpu().numpy(), tb, steps=64, lr=0.5, channels=4,
                                 reassign=1 | (1 << 16))
    assert np.array_equal(out.cpu().numpy(), hout)
    assert np.array_equal(s0.cpu().numpy(), h0) and np.array_equal(s1.cpu().numpy(), h1)
    assert np.all(h1 <= h0) and h1.sum() < h0.sum()
    # In place gives the same bytes.
    inplace = blocks.clone()
    grad.refine_bc7(inplace, td, n, n, steps=64, lr=0.5, channels=4, reassign=1, pbits=True, out=inplace)
    torch.cuda.synchronize()
    assert torch.equal(inplace, out)


@pytest.mark.parametrize("fmt", [6, 0], ids=["bc7", "bc1"])
def test_texel_adjoint_gpu_equals_host_build_and_fp64(fmt):
    """bcn_decode_texels_backward == the host build of the same code (bit for
    bit) on 4096 random blocks of every mode, and == the float64 hat matrix
    (numpy pinv) on 256 of them."""
    import kernel_check_lib as kcl
    from block_compression_amd import grad

    print(f"grad: texel adjoint fmt {fmt}", flush=True)
    b = _blocks(fmt, 4096, 31)
    bd = torch.from_numpy(b.reshape(-1).copy()).cuda()
    g = np.random.default_rng(32).standard_normal((len(b), 16, 4)).astype(np.float32)
    got = grad.decode_texels_backward(fmt, bd, torch.from_numpy(g).cuda()).cpu().numpy()
    assert np.array_equal(got, kcl.texel_backward(fmt, b.reshape(-1), g))
    for i in range(256):
        if fmt == 6:
            want = ref.texel_adjoint_bc7(ref.parse_bc7(b[i].tobytes()), g[i].astype(np.float64))
        else:
            want = ref.texel_adjoint_bc1(ref.parse_bc1(b[i].tobytes()), g[i].astype(np.float64))
        np.testing.assert_allclose(got[i], want, rtol=1e-5, atol=2e-5)


def test_decode_texels_autograd_on_an_encoded_texture():
    """Texture training through DecodeTexels: the value is the exact decode
    of the GPU encoding of the texture (== the oracle decoder), and the
    gradient reaching the (H, W, 4) texture through image_to_blocks is the
    texel-side adjoint of the loss gradient, per block == float64."""
    from block_compression_amd import GpuBlockCompressor, grad, synth
    from block_compression_amd.settings import BC7Settings, CompressionVariant

    print("grad: DecodeTexels on a 64x64 BC7 texture", flush=True)
    n = 64
    img = synth.tile_image_rgba8(n, n, 47, 0.2)
    v = CompressionVariant.BC7(BC7Settings.alpha_basic())
    blocks = torch.zeros(v.blocks_byte_size(n, n), dtype=torch.uint8, device="cuda")
    c = GpuBlockCompressor()
    c.add_compression_task(v, torch.from_numpy(img).cuda(), n, n, blocks)
    c.compress()
    tex = torch.tensor(img, dtype=torch.float32, device="cuda", requires_grad=True)
    src = grad.image_to_blocks(tex)
    assert torch.equal(grad.blocks_to_image(src, n, n), tex)
    dec = grad.DecodeTexels.apply(src, blocks, 6)
    b = blocks.cpu().numpy().reshape(-1, 16)
    want_dec = oracle_lib.decompress_rgba8(6, n, n, blocks.cpu().numpy())
    assert np.array_equal(grad.blocks_to_image(dec, n, n).detach().cpu().numpy(), want_dec.astype(np.float32))
    target = torch.from_numpy(synth.tile_image_rgba8(n, n, 48, 0.0).astype(np.float32)).cuda()
    loss = ((dec - grad.image_to_blocks(target)) ** 2).sum()
    loss.backward()
    gsrc = grad.image_to_blocks(tex.grad).cpu().numpy()
    gdec = (2.0 * (dec - grad.image_to_blocks(target))).detach().cpu().numpy().astype(np.float64)
    for i in range(len(b)):
        want = ref.texel_adjoint_bc7(ref.parse_bc7(b[i].tobytes()), gdec[i])
        np.testing.assert_allclose(gsrc[i], want, rtol=1e-4, atol=1e-3 * (1 + np.abs(gdec[i]).max()))
