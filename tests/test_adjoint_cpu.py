"""The decode adjoint and the refinement (block_compression_amd/csrc/
decode_adjoint.hpp, host build) against an independent FP64 reference
(tests/adjoint_ref.py), which is itself pinned to the oracle decoder.

* the reference parser + exact decode == the oracle's decoder, every mode;
* parse -> pack is the identity on valid blocks;
* the kernels' straight-through forward at their own endpoints == the
  oracle's decode; the smooth forward == the FP64 relaxation;
* the relaxation stays within 1.22 code values of the exact decode (bit
  replication vs exact scaling, plus the interpolation's rounding);
* the adjoint == torch.autograd of the FP64 relaxation (reference gradient)
  and == FP64 central finite differences (step 1e-3) of the squared error on
  256 random blocks;
* refinement never raises a block's error, and the error it reports is the
  exact error of the block it writes.
"""
import numpy as np
import pytest
import torch

import adjoint_ref as ref
import kernel_check_lib as kc
import oracle_lib


def _oracle_decode_bc7(blocks):
    return np.stack([oracle_lib.decode_block("bc7", b.tobytes(), 16, 64).reshape(16, 4) for b in blocks])


def _oracle_decode_bc1(blocks):
    return np.stack([oracle_lib.decode_block("bc1", b.tobytes(), 16, 64).reshape(16, 4) for b in blocks])


def _random_bc1(n, seed):
    rng = np.random.default_rng(seed)
    return rng.integers(0, 256, (n, 8), dtype=np.uint8)


def test_reference_parser_matches_oracle_decoder():
    blocks = ref.random_bc7_blocks(400, 1)
    want = _oracle_decode_bc7(blocks)
    got = np.stack([ref.exact_bc7(ref.parse_bc7(b.tobytes())) for b in blocks])
    assert np.array_equal(got, want)
    b1 = _random_bc1(300, 2)
    assert np.array_equal(np.stack([ref.exact_bc1(ref.parse_bc1(b.tobytes())) for b in b1]),
                          _oracle_decode_bc1(b1))


def test_parse_pack_identity():
    blocks = ref.random_bc7_blocks(2000, 3)
    assert np.array_equal(kc.bc7_repack(blocks.reshape(-1)).reshape(-1, 16), blocks)


@pytest.mark.parametrize("fmt", [6, 0], ids=["bc7", "bc1"])
def test_straight_through_forward_is_exact_decode(fmt):
    blocks = ref.random_bc7_blocks(400, 4) if fmt == 6 else _random_bc1(400, 4)
    x = kc.endpoints(fmt, blocks.reshape(-1))
    got = kc.forward(fmt, blocks.reshape(-1), x, smooth=False)
    want = _oracle_decode_bc7(blocks) if fmt == 6 else _oracle_decode_bc1(blocks)
    assert np.array_equal(got, want.astype(np.float32))
    # Off-integer parameters round to the nearest code (straight-through).
    rng = np.random.default_rng(5)
    got2 = kc.forward(fmt, blocks.reshape(-1), x + rng.uniform(-0.45, 0.45, x.shape).astype(np.float32),
                      smooth=False)
    assert np.array_equal(got2, got)


def _forms(fmt, blocks):
    if fmt == 6:
        return [ref.parse_bc7(b.tobytes()) for b in blocks]
    return [ref.parse_bc1(b.tobytes()) for b in blocks]


def _smooth(fmt, f, x64):
    return ref.smooth_bc7(f, x64) if fmt == 6 else ref.smooth_bc1(f, x64)


def _xshape(fmt):
    return (6, 4) if fmt == 6 else (2, 3)


@pytest.mark.parametrize("fmt", [6, 0], ids=["bc7", "bc1"])
def test_smooth_forward_matches_fp64_relaxation_and_stays_near_exact(fmt):
    blocks = ref.random_bc7_blocks(256, 6) if fmt == 6 else _random_bc1(256, 6)
    forms = _forms(fmt, blocks)
    x = kc.endpoints(fmt, blocks.reshape(-1))
    rng = np.random.default_rng(7)
    xc = (x + rng.uniform(-0.5, 0.5, x.shape)).astype(np.float32)
    got = kc.forward(fmt, blocks.reshape(-1), xc, smooth=True)
    at_int = kc.forward(fmt, blocks.reshape(-1), x, smooth=True)
    exact = kc.forward(fmt, blocks.reshape(-1), x, smooth=False)
    for i, f in enumerate(forms):
        want = _smooth(fmt, f, torch.tensor(xc[i].reshape(_xshape(fmt)), dtype=torch.float64)).numpy()
        np.testing.assert_allclose(got[i], want, rtol=0, atol=2e-4)
    # Relaxation vs the integer decode at integer endpoints: bit replication
    # differs from v * 255 / (2^n - 1) by less than 0.72 for n <= 8 bits, and
    # the interpolation's +32 >> 6 rounds by at most 0.5.
    bits = np.arange(4, 9)
    rep_err = max(np.abs(((v << (8 - n)) | ((v << (8 - n)) >> n)) - v * 255 / (2 ** n - 1)).max()
                  for n in bits for v in [np.arange(2 ** n)])
    assert rep_err < 0.72
    assert np.abs(at_int - exact).max() <= rep_err + 0.5


@pytest.mark.parametrize("fmt", [6, 0], ids=["bc7", "bc1"])
def test_adjoint_matches_autograd_and_fp64_finite_differences(fmt):
    """256 random blocks, loss = sum (smooth(x) - T)^2 in FP64.  The kernel's
    adjoint of g = 2 (smooth(x) - T) must equal torch.autograd's gradient of
    the FP64 relaxation and central differences with step 1e-3."""
    blocks = ref.random_bc7_blocks(256, 8) if fmt == 6 else _random_bc1(256, 8)
    forms = _forms(fmt, blocks)
    x = kc.endpoints(fmt, blocks.reshape(-1))
    rng = np.random.default_rng(9)
    xc = x + rng.uniform(-0.4, 0.4, x.shape)
    target = rng.uniform(0, 255, (len(forms), 16, 4))
    g = np.zeros((len(forms), 16, 4))
    grad_ref = np.zeros_like(xc)
    fd = np.zeros_like(xc)
    h = 1e-3
    for i, f in enumerate(forms):
        xi = torch.tensor(xc[i].reshape(_xshape(fmt)), dtype=torch.float64, requires_grad=True)
        tt = torch.tensor(target[i], dtype=torch.float64)
        loss = ((_smooth(fmt, f, xi) - tt) ** 2).sum()
        loss.backward()
        grad_ref[i] = xi.grad.numpy().reshape(-1)
        g[i] = (2.0 * (_smooth(fmt, f, xi.detach()) - tt)).numpy()
        with torch.no_grad():
            for j in range(xc.shape[1]):
                d = torch.zeros(xc.shape[1], dtype=torch.float64)
                d[j] = h
                d = d.reshape(_xshape(fmt))
                lp = ((_smooth(fmt, f, xi.detach() + d) - tt) ** 2).sum()
                lm = ((_smooth(fmt, f, xi.detach() - d) - tt) ** 2).sum()
                fd[i, j] = float((lp - lm) / (2 * h))
    got = kc.backward(fmt, blocks.reshape(-1), g.astype(np.float32)).astype(np.float64)
    scale = np.abs(grad_ref).max() + 1e-9
    assert np.abs(got - grad_ref).max() <= 1e-5 * scale
    assert np.abs(got - fd).max() <= 1e-4 * scale
    # Parameters the decode does not read (unused subsets, absent alpha) get 0.
    assert np.all(got[np.abs(grad_ref) == 0] == 0)


def test_refine_never_worse_and_reports_exact_error():
    from block_compression_amd import synth
    from block_compression_amd.settings import BC7Settings

    img = synth.tile_image_rgba8(64, 64, 12, 0.1)
    img[..., 3] = synth.tile_image_rgba8(64, 64, 13, 0.2)[..., 0]
    s = BC7Settings.alpha_basic()
    blocks = oracle_lib.compress_rgba8(6, s.to_bytes(), img, threads=4).reshape(-1, 16)
    # Target: the image plus a sub-code offset, in 0..255 units, block-major.
    rng = np.random.default_rng(14)
    tgt = img.astype(np.float32) + rng.uniform(-0.5, 0.5, img.shape).astype(np.float32)
    tgt = tgt.reshape(16, 4, 16, 4, 4).transpose(0, 2, 1, 3, 4).reshape(-1, 16, 4)
    out, s0, s1 = kc.bc7_refine(blocks.reshape(-1), tgt, steps=64, lr=0.5, channels=4)
    out = out.reshape(-1, 16)
    assert np.all(s1 <= s0)
    dec = _oracle_decode_bc7(out).astype(np.float64)
    sse = ((dec - tgt) ** 2).sum(axis=(1, 2))
    np.testing.assert_allclose(sse, s1, rtol=1e-5, atol=1e-2)
    dec0 = _oracle_decode_bc7(blocks).astype(np.float64)
    np.testing.assert_allclose(((dec0 - tgt) ** 2).sum(axis=(1, 2)), s0, rtol=1e-5, atol=1e-2)
    assert s1.sum() < s0.sum()
    # Blocks it leaves alone stay byte-identical.
    same = s1 == s0
    assert np.array_equal(out[same], blocks[same])


# ---- texel-side adjoint (bc7_texel_backward / bc1_texel_backward) -----------


def _bc1_blocks(n, seed):
    return np.random.default_rng(seed).integers(0, 256, (n, 8), dtype=np.uint8)


@pytest.mark.parametrize("fmt", [6, 0], ids=["bc7", "bc1"])
def test_texel_adjoint_host_build_matches_fp64_hat_matrix(fmt):
    """The kernels' closed form (host build) equals A pinv(A) from numpy in
    float64, per subset and index set, on random blocks of every BC7 mode
    (and blocks with no valid mode) and of both BC1 colour modes."""
    import kernel_check_lib as kc

    b = ref.random_bc7_blocks(800, 41) if fmt == 6 else _bc1_blocks(800, 41)
    if fmt == 6:
        b[-8:, 0] = 0  # no valid mode
    g = np.random.default_rng(42).standard_normal((len(b), 16, 4)).astype(np.float32)
    got = kc.texel_backward(fmt, b.reshape(-1), g)
    for i in range(len(b)):
        if fmt == 6:
            want = ref.texel_adjoint_bc7(ref.parse_bc7(b[i].tobytes()), g[i].astype(np.float64))
        else:
            want = ref.texel_adjoint_bc1(ref.parse_bc1(b[i].tobytes()), g[i].astype(np.float64))
        np.testing.assert_allclose(got[i], want, rtol=1e-5, atol=2e-5, err_msg=f"block {i}")


def test_texel_adjoint_fp64_reference_is_the_refit_jacobian():
    """The reference hat matrix is the Jacobian of the re-fit decode: central
    differences of refit_decode_bc7 (a linear map of the source) reproduce
    the adjoint's action, <J^T g, e_j> = <g, J e_j>."""
    b = ref.random_bc7_blocks(64, 43)
    rng = np.random.default_rng(44)
    for i in range(len(b)):
        f = ref.parse_bc7(b[i].tobytes())
        t = rng.uniform(0, 255, (16, 4))
        g = rng.standard_normal((16, 4))
        adj = ref.texel_adjoint_bc7(f, g)
        for k, o in [(rng.integers(16), rng.integers(4)) for _ in range(6)]:
            e = np.zeros((16, 4))
            e[k, o] = 1.0
            h = 1e-3
            jd = (ref.refit_decode_bc7(f, t + h * e) - ref.refit_decode_bc7(f, t - h * e)) / (2 * h)
            assert abs(np.sum(g * jd) - adj[k, o]) < 1e-6 * (1 + abs(adj[k, o])), (i, k, o)


@pytest.mark.parametrize("fmt", [6, 0], ids=["bc7", "bc1"])
def test_texel_adjoint_is_a_symmetric_projection(fmt):
    """H H g = H g and <H g1, g2> = <g1, H g2> on the host build."""
    import kernel_check_lib as kc

    b = ref.random_bc7_blocks(512, 45) if fmt == 6 else _bc1_blocks(512, 45)
    rng = np.random.default_rng(46)
    g1 = rng.standard_normal((len(b), 16, 4)).astype(np.float32)
    g2 = rng.standard_normal((len(b), 16, 4)).astype(np.float32)
    h1 = kc.texel_backward(fmt, b.reshape(-1), g1)
    hh1 = kc.texel_backward(fmt, b.reshape(-1), h1)
    np.testing.assert_allclose(hh1, h1, rtol=1e-4, atol=1e-4)
    h2 = kc.texel_backward(fmt, b.reshape(-1), g2)
    lhs = np.sum(h1.astype(np.float64) * g2, axis=(1, 2))
    rhs = np.sum(g1.astype(np.float64) * h2, axis=(1, 2))
    np.testing.assert_allclose(lhs, rhs, rtol=1e-4, atol=1e-4)
