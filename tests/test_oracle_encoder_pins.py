"""Pins for the oracle's encoder functions (not only its decoders).

The reference crate cannot run here (no Rust toolchain), so each encoder
function of the oracle is checked against something the mathematics fixes,
computed independently of the oracle's code: exact rational rounding, brute
force over every lattice point, numpy's least squares and eigenvalues,
rank-1 and diagonal closed forms, decoder round trips through the decoder the
reference's own known-answer vectors pin, and consistency of the BC7 mode
search with restricted runs of itself.  Each test names the reference lines
it pins and the slips it is meant to catch.
"""
import dataclasses
from fractions import Fraction

import numpy as np
import pytest

import oracle_lib
from block_compression_amd import synth
from block_compression_amd.settings import BC6H_PRESETS, BC7_PRESETS, BC6HSettings, BC7Settings


def _round_half_up(num: int, den: int) -> int:
    """floor(num/den + 1/2) for non-negative integers, exactly."""
    return (2 * num + den) // (2 * den)


def _block64(texels: np.ndarray) -> np.ndarray:
    """(16, C) texels -> the 64-float channel-major block (missing channels 0)."""
    b = np.zeros((4, 16), np.float32)
    b[: texels.shape[1]] = texels.T
    return b.reshape(64)


def _mask_bits(mask: int) -> np.ndarray:
    return np.array([(mask >> k) & 1 for k in range(16)], bool)


def _qblock(idx: np.ndarray) -> np.ndarray:
    """16 indices -> the encoder's qblock[2] (4 bits per texel)."""
    q = np.zeros(2, np.uint32)
    for k, v in enumerate(idx):
        q[k // 8] |= np.uint32(int(v) << (4 * (k % 8)))
    return q


# ---------------------------------------------------------------------------
# BC1: 565 quantisation, decoder expansion, fast_quant, refinement, axis
# ---------------------------------------------------------------------------

def test_enc_rgb565_is_exact_rounding_of_the_truncated_channel():
    """enc_rgb565 (bc1_to_5.rs:268-278): r5 = round(trunc(r)*31/255), g6 =
    round(trunc(g)*63/255), packed r<<11 | g<<5 | b, for every 8-bit value
    and fractional inputs.  Catches a wrong constant, shift, rounding bias,
    channel order or truncation."""
    for v in range(256):
        want5 = _round_half_up(v * 31, 255)
        want6 = _round_half_up(v * 63, 255)
        for frac in (0.0, 0.25, 0.999):
            x = float(np.float32(v + frac))
            assert oracle_lib.bc1_enc_rgb565(x, 0.0, 0.0) == want5 << 11, (v, frac)
            assert oracle_lib.bc1_enc_rgb565(0.0, x, 0.0) == want6 << 5, (v, frac)
            assert oracle_lib.bc1_enc_rgb565(0.0, 0.0, x) == want5, (v, frac)
    rng = np.random.default_rng(1)
    for r, g, b in rng.integers(0, 256, (500, 3)):
        want = (_round_half_up(int(r) * 31, 255) << 11) | (_round_half_up(int(g) * 63, 255) << 5) \
            | _round_half_up(int(b) * 31, 255)
        assert oracle_lib.bc1_enc_rgb565(float(r), float(g), float(b)) == want


def _bc1_palette(c0: int, c1: int, opaque: bool):
    """The BC1 palette from exact rationals: endpoints round(c*255/(2^n-1));
    4-colour mode round((2a+b)/3 ...); 3-colour mode round((a+b)/2 ...) and
    transparent black (the bcdec formulas the decoder restates,
    block.rs:124-197, are these roundings)."""
    def chans(c):
        return (c >> 11) & 31, (c >> 5) & 63, c & 31
    a, b = chans(c0), chans(c1)
    m = (31, 63, 31)
    pal = np.zeros((4, 4), np.int64)
    for ch in range(3):
        pal[0, ch] = _round_half_up(a[ch] * 255, m[ch])
        pal[1, ch] = _round_half_up(b[ch] * 255, m[ch])
        if c0 > c1 or opaque:
            pal[2, ch] = _round_half_up((2 * a[ch] + b[ch]) * 255, 3 * m[ch])
            pal[3, ch] = _round_half_up((a[ch] + 2 * b[ch]) * 255, 3 * m[ch])
        else:
            pal[2, ch] = _round_half_up((a[ch] + b[ch]) * 255, 2 * m[ch])
    pal[:3, 3] = 255
    pal[3, 3] = 255 if (c0 > c1 or opaque) else 0
    return pal


def test_bc1_decoder_palette_is_exact_rational_rounding_for_every_channel_pair():
    """decode_color_block (block.rs:124-197): for every 5-bit (and 6-bit)
    channel pair, in both the 4-colour and the 3-colour mode, the decoded
    palette equals round(exact interpolation of the expanded endpoints).
    BC3's colour block is always 4-colour (OPAQUE_MODE, block.rs:87-100)."""
    blocks = []
    for a in range(64):
        for b in range(64):
            r0, r1 = a % 32, b % 32
            for order in (0, 1):
                c0 = (r0 << 11) | (a << 5) | (31 - r0)
                c1 = (r1 << 11) | (b << 5) | (31 - r1)
                if order:
                    c0, c1 = c1, c0
                blocks.append((c0, c1))
    idx = 0b11100100_11100100_11100100_11100100  # texel k uses palette entry k % 4
    for c0, c1 in blocks:
        raw = np.array([c0 & 255, c0 >> 8, c1 & 255, c1 >> 8, idx & 255, (idx >> 8) & 255,
                        (idx >> 16) & 255, idx >> 24], np.uint8)
        dec = oracle_lib.decode_block("bc1", raw.tobytes(), 16, 64).reshape(16, 4)
        pal = _bc1_palette(c0, c1, False)
        assert np.array_equal(dec[:4], pal), (c0, c1)
    for c0, c1 in blocks[:: 37]:
        raw = np.zeros(16, np.uint8)
        raw[8:16] = [c0 & 255, c0 >> 8, c1 & 255, c1 >> 8, idx & 255, (idx >> 8) & 255,
                     (idx >> 16) & 255, idx >> 24]
        dec = oracle_lib.decode_block("bc3", raw.tobytes(), 16, 64).reshape(16, 4)
        assert np.array_equal(dec[:4, :3], _bc1_palette(c0, c1, True)[:, :3]), (c0, c1)


def _expand565(c: int) -> np.ndarray:
    return _bc1_palette(c, c, True)[0, :3]


def test_bc1_flat_block_closed_form():
    """A flat block of colour c encodes as two copies of enc565(c) with
    all-zero indices (bc1_to_5.rs:386-429: the refinement of a uniform index
    set returns the mean, fast_quant of equal endpoints gives index 0)."""
    rng = np.random.default_rng(2)
    cols = np.concatenate([rng.integers(0, 256, (300, 3)), [[0, 0, 0], [255, 255, 255], [255, 0, 128]]])
    for c in cols:
        img = np.zeros((4, 4, 4), np.uint8)
        img[..., :3] = c
        img[..., 3] = 255
        out = oracle_lib.compress_rgba8(0, b"", img)
        e = (_round_half_up(int(c[0]) * 31, 255) << 11) | (_round_half_up(int(c[1]) * 63, 255) << 5) \
            | _round_half_up(int(c[2]) * 31, 255)
        assert out.tolist() == [e & 255, e >> 8, e & 255, e >> 8, 0, 0, 0, 0], c


def _random_codes(rng, n):
    out = []
    while len(out) < n:
        a, b = (int(x) for x in rng.integers(0, 65536, 2))
        if a != b:
            out.append((a, b))
    return out


@pytest.mark.parametrize("palette_entries", [2, 4])
def test_bc1_blocks_of_palette_colours_round_trip_exactly(palette_entries):
    """Blocks whose texels are entries of a decodable BC1 palette (two 565
    endpoints and, with 4 entries, their decoded interpolants) decode back
    exactly after encoding: the PCA, the endpoint pick, fast_quant, the
    refinement and the 565 packing all have to be right for that."""
    rng = np.random.default_rng(3 + palette_entries)
    for c0, c1 in _random_codes(rng, 300):
        hi, lo = max(c0, c1), min(c0, c1)
        pal = _bc1_palette(hi, lo, True)[:, :3]
        if palette_entries == 2:
            pal = pal[:2]
        if len({tuple(p) for p in pal}) < len(pal):
            continue
        # A colour difference orthogonal to the power iteration's start vector
        # (1,1,1) leaves the axis at the start (bc1_to_5.rs:200-223): such a
        # block collapses to its mean in the reference too.
        d = (pal[1] - pal[0]).astype(np.float64)
        if abs(d.sum()) < 0.1 * np.linalg.norm(d):
            continue
        sel = rng.permutation(np.resize(np.arange(len(pal)), 16))
        img = np.zeros((4, 4, 4), np.uint8)
        img[..., :3] = pal[sel].reshape(4, 4, 3)
        img[..., 3] = 255
        out = oracle_lib.compress_rgba8(0, b"", img)
        dec = oracle_lib.decode_block("bc1", out.tobytes(), 16, 64).reshape(4, 4, 4)
        assert np.array_equal(dec[..., :3], img[..., :3]), (hi, lo)


def test_bc1_fast_quant_picks_the_palette_position_of_texels_on_the_segment():
    """fast_quant (bc1_to_5.rs:280-321): a texel at c0 + i/3 (c1 - c0) of the
    expanded endpoints gets index i; texels beyond the ends clamp to 0 and 3.
    Catches a wrong bias, scale or reversed direction."""
    rng = np.random.default_rng(4)
    for p0, p1 in _random_codes(rng, 400):
        e0 = oracle_lib.bc1_dec_rgb565(p0).astype(np.float64)
        e1 = oracle_lib.bc1_dec_rgb565(p1).astype(np.float64)
        if np.sum((e1 - e0) ** 2) < 64:
            continue
        idx = rng.integers(0, 4, 16)
        tex = e0[None, :] + idx[:, None] / 3.0 * (e1 - e0)[None, :]
        out = rng.random(16) < 0.15
        idx = idx.copy()
        tex[out & (idx == 0)] -= (e1 - e0) * 0.4
        tex[out & (idx == 3)] += (e1 - e0) * 0.4
        got = oracle_lib.bc1_fast_quant(_block64(tex.astype(np.float32)), p0, p1)
        want = sum(int(i) << (2 * k) for k, i in enumerate(idx))
        assert got == want, (p0, p1)
    # dec_rgb565 is the bit replication of the 565 fields.
    for p in range(0, 65536, 97):
        r, g, b = (p >> 11) & 31, (p >> 5) & 63, p & 31
        want = [(r << 3) | (r >> 2), (g << 2) | (g >> 4), (b << 3) | (b >> 2)]
        assert oracle_lib.bc1_dec_rgb565(p).tolist() == want


def test_bc1_refine_is_the_least_squares_fit_of_the_index_weights():
    """bc1_refine (bc1_to_5.rs:323-374) solves min sum |x_k - ((3-q_k) c0 +
    q_k c1)/3|^2 over c0, c1 per channel (then clamps to 0..255); with one
    index value everywhere it returns the block's mean for both."""
    rng = np.random.default_rng(5)
    for _ in range(400):
        tex = rng.integers(0, 256, (16, 3)).astype(np.float32)
        q = rng.integers(0, 4, 16)
        if len(set(q.tolist())) < 2 or np.bincount(q, minlength=4).max() > 13:
            continue
        bits = sum(int(v) << (2 * k) for k, v in enumerate(q))
        col, _ = oracle_lib.bc1_refine(_block64(tex), bits)
        A = np.stack([(3 - q) / 3.0, q / 3.0], 1)
        sol = np.clip(np.linalg.lstsq(A, tex.astype(np.float64), rcond=None)[0], 0, 255)
        cond = np.linalg.cond(A)
        assert np.allclose(col, sol, atol=2e-5 * 255 * cond, rtol=0), (q, col, sol)
    for v in range(4):
        tex = rng.integers(0, 256, (16, 3)).astype(np.float32)
        col, _ = oracle_lib.bc1_refine(_block64(tex), v * 0x55555555)
        mean = tex.astype(np.float64).mean(0)
        assert np.allclose(col[0], mean, atol=1e-4) and np.allclose(col[1], mean, atol=1e-4)


def _line_texels(rng, channels, lo=0, hi=255):
    """Integer texels on one line: base + t*d with integer t, d (rank 1)."""
    while True:
        d = rng.integers(-6, 7, channels)
        if np.all(d == 0) or d.sum() == 0:
            continue
        t = rng.integers(0, 9, 16)
        if t.max() - t.min() < 3:
            continue
        span = (t.max() - t.min()) * np.abs(d)
        if np.any(span > hi - lo):
            continue
        base = np.array([rng.integers(lo - min(0, dd * t.min()) - min(0, dd * t.max()),
                                      hi - max(0, dd * t.max()) - max(0, dd * t.min()) + 1)
                         for dd in d])
        tex = base[None, :] + t[:, None] * d[None, :]
        if tex.min() < lo or tex.max() > hi:
            continue
        return tex.astype(np.float32), d.astype(np.float64), t


def test_bc1_axis_is_the_line_direction_of_rank1_blocks_and_the_dominant_channel():
    """compute_covar_dc + compute_axis3 (bc1_to_5.rs:157-223): on a block
    whose texels lie on a line the axis is that line's unit direction (its
    sign set by the start vector (1,1,1)); with uncorrelated channel
    variances the axis points along the largest one.  The mean is exact.
    Catches a wrong covariance entry, ssymv index or normalisation."""
    rng = np.random.default_rng(6)
    for _ in range(300):
        tex, d, _ = _line_texels(rng, 3)
        axis, dc = oracle_lib.bc1_axis(_block64(tex))
        u = d / np.linalg.norm(d) * np.sign(d.sum())
        assert np.allclose(axis, u, atol=2e-5), (d, axis)
        assert np.allclose(dc, tex.astype(np.float64).mean(0), atol=1e-4)
    # Orthogonal +-1 patterns make the covariance exactly diagonal.
    h = np.array([[1, 1, 1, 1], [1, -1, 1, -1], [1, 1, -1, -1], [1, -1, -1, 1]])
    pats = np.kron(h, h)[1:4]
    for amp in ([40, 10, 5], [5, 50, 12], [3, 7, 60]):
        tex = (128 + pats.T * np.array(amp)[None, :]).astype(np.float32)
        axis, _ = oracle_lib.bc1_axis(_block64(tex))
        j = int(np.argmax(amp))
        assert abs(axis[j]) > 0.99 and np.all(np.abs(np.delete(axis, j)) < 0.05), (amp, axis)


# ---------------------------------------------------------------------------
# Shared primitives (common.rs): PCA, bounds, least squares, segments
# ---------------------------------------------------------------------------

def _sym10(m: np.ndarray) -> np.ndarray:
    """4x4 symmetric -> the reference's covar[10] packing (00 01 02 03 11 12
    13 22 23 33)."""
    return np.array([m[0, 0], m[0, 1], m[0, 2], m[0, 3], m[1, 1], m[1, 2], m[1, 3], m[2, 2],
                     m[2, 3], m[3, 3]], np.float32)


@pytest.mark.parametrize("channels", [3, 4])
def test_compute_axis_converges_to_the_dominant_eigenvector(channels):
    """compute_axis (common.rs:247-292): power iteration from (1,1,1[,1]),
    renormalised every second step, ends at the unit dominant eigenvector of
    a symmetric matrix with a large eigen-gap (numpy.linalg.eigh)."""
    rng = np.random.default_rng(10 + channels)
    n = 0
    while n < 200:
        q, _ = np.linalg.qr(rng.standard_normal((channels, channels)))
        if abs(q[:, 0].sum()) < 0.5:
            continue
        lam = np.sort(rng.uniform(0.1, 1.0, channels))[::-1] * np.r_[40.0, np.ones(channels - 1)]
        m = np.zeros((4, 4))
        m[:channels, :channels] = (q * lam) @ q.T
        axis = oracle_lib.compute_axis(_sym10(m), 8, channels)[:channels]
        w, v = np.linalg.eigh(m[:channels, :channels])
        u = v[:, -1] * np.sign(v[:, -1].sum())
        assert np.allclose(axis, u, atol=1e-5), (axis, u)
        n += 1


@pytest.mark.parametrize("channels", [3, 4])
def test_get_pca_bound_is_trace_minus_the_top_eigenvalue(channels):
    """get_pca_bound (common.rs:208-245): trace - lambda_max of covar/65536
    with 1e-6 added to the first three diagonal entries, clamped at 0; 4
    power iterations reach lambda_max when the eigen-gap is large.  Within
    2e-6 of the trace (f32), against numpy.linalg.eigvalsh in float64."""
    rng = np.random.default_rng(20 + channels)
    n = 0
    while n < 300:
        q, _ = np.linalg.qr(rng.standard_normal((channels, channels)))
        if abs(q[:, 0].sum()) < 0.5:
            continue
        scale = 10.0 ** rng.uniform(0, 5)
        lam = np.r_[1.0, rng.uniform(0, 0.04, channels - 1)] * scale
        m = np.zeros((4, 4))
        m[:channels, :channels] = (q * lam) @ q.T
        c10 = _sym10(m)
        got = oracle_lib.get_pca_bound(c10, channels)
        s = _sym10(m).astype(np.float64)
        full = np.array([[s[0], s[1], s[2], s[3]], [s[1], s[4], s[5], s[6]],
                         [s[2], s[5], s[7], s[8]], [s[3], s[6], s[8], s[9]]])[:channels, :channels]
        full = full / 65536.0 + np.diag([1e-6, 1e-6, 1e-6, 0.0][:channels])
        w = np.linalg.eigvalsh(full)
        want = max(np.trace(full) - w[-1], 0.0)
        assert abs(got - want) <= 2e-6 * np.trace(full) + 1e-12, (got, want)
        n += 1


@pytest.mark.parametrize("channels", [3, 4])
def test_block_pca_axis_and_segment_on_rank1_subsets(channels):
    """block_pca_axis / block_segment_core (common.rs:352-394, 483-518) on a
    masked subset whose texels lie on a line: dc is the subset mean, the
    axis the line's unit direction, and the endpoints the two extreme
    texels of the subset.  Texels outside the mask are random and must not
    matter.  Catches a mask, stats, covariance or extent slip."""
    rng = np.random.default_rng(30 + channels)
    pattern, m0, m1, m2, _ = oracle_lib.tables()
    masks = [int(x) for x in np.concatenate([m0[:64], m1[:64], m0[64:], m1[64:], m2[64:]])
             if bin(int(x)).count("1") >= 3] + [0xFFFF]
    for i in range(300):
        mask = masks[i % len(masks)]
        sel = _mask_bits(mask)
        tex = rng.integers(0, 256, (16, channels)).astype(np.float32)
        line, d, t = _line_texels(rng, channels)
        if len(set(t[sel].tolist())) < 2 or (t[sel].max() - t[sel].min()) * np.linalg.norm(d) < 2:
            continue
        tex[sel] = line[sel]
        b = _block64(tex)
        axis, dc = oracle_lib.block_pca_axis(b, mask, channels)
        assert np.allclose(dc[:channels], tex[sel].astype(np.float64).mean(0), atol=1e-4)
        u = d / np.linalg.norm(d) * np.sign(d.sum())
        # The covariance is sum(x^2) - sum(x)^2/n in f32: its entries carry an
        # absolute error of a few ulp of sum(x^2), relative to the variance
        # along the line.
        x = tex[sel].astype(np.float64)
        var = np.sum((x - x.mean(0)) ** 2)
        tol = 2e-5 + 64 * 2.0 ** -24 * np.sum(x * x) / var
        assert np.allclose(axis[:channels], u, atol=tol), (mask, axis, u)
        ep = oracle_lib.block_segment_core(b, mask, channels)
        lo = tex[sel][np.argmin(t[sel])]
        hi = tex[sel][np.argmax(t[sel])]
        if d.sum() < 0:
            lo, hi = hi, lo
        etol = 2e-3 + 300 * tol
        assert np.allclose(ep[:channels], lo, atol=etol) and np.allclose(ep[4:4 + channels], hi, atol=etol)


def test_compute_stats_masked_are_the_masked_moments():
    """compute_stats_masked (common.rs:294-332): stats[10..14] are the
    masked channel sums, stats[14] the count, and the upper-triangle
    entries the masked second moments in the covar[10] packing."""
    rng = np.random.default_rng(40)
    for _ in range(200):
        ch = int(rng.integers(3, 5))
        tex = rng.integers(0, 256, (16, ch)).astype(np.float64)
        mask = int(rng.integers(1, 1 << 16))
        sel = _mask_bits(mask)
        st = oracle_lib.compute_stats_masked(_block64(tex.astype(np.float32)), mask, ch)
        x = np.zeros((16, 4))
        x[:, :ch] = tex
        x = x[sel]
        mom = x.T @ x
        want = np.zeros(15)
        want[:10] = _sym10(mom)
        if ch == 3:
            want[[3, 6, 8, 9]] = 0
        want[10:10 + ch] = x[:, :ch].sum(0)
        want[14] = sel.sum()
        assert np.array_equal(st.astype(np.float64), want)


def test_block_pca_bound_split_on_subsets_with_a_dominant_axis():
    """block_pca_bound_split (common.rs:396-420): sqrt(b0 + b1) * 256 with b_i
    get_pca_bound of subset i's covariance (subset 1 from the complement of
    the full stats).  On subsets with one dominant direction each, against
    the eigenvalues of both subsets' covariances (numpy, float64)."""
    rng = np.random.default_rng(41)
    _, m0, _, _, _ = oracle_lib.tables()
    n = 0
    while n < 200:
        part = int(rng.integers(0, 64))
        mask = int(m0[part])
        sel = _mask_bits(mask)
        if sel.sum() < 3 or (~sel).sum() < 3:
            continue
        tex = np.zeros((16, 3))
        for s in (sel, ~sel):
            d = rng.standard_normal(3)
            d /= np.linalg.norm(d)
            if abs(d.sum()) < 0.5:
                break
            t = rng.uniform(-60, 60, s.sum())
            noise = rng.uniform(-1.5, 1.5, (s.sum(), 3))
            tex[s] = np.clip(np.round(128 + t[:, None] * d[None, :] + noise), 0, 255)
        else:
            got = oracle_lib.block_pca_bound_split(_block64(tex.astype(np.float32)), mask, 3)
            want = 0.0
            for s in (sel, ~sel):
                x = tex[s]
                cov = (x.T @ x - np.outer(x.sum(0), x.sum(0)) / s.sum()) / 65536.0 + 1e-6 * np.eye(3)
                w = np.linalg.eigvalsh(cov)
                want += max(np.trace(cov) - w[-1], 0.0)
            want = np.sqrt(want) * 256
            assert abs(got - want) <= 2e-3 * max(want, 1.0), (got, want)
            n += 1


@pytest.mark.parametrize("bits", [2, 3, 4])
def test_opt_endpoints_is_the_masked_least_squares_fit(bits):
    """opt_endpoints (common.rs:142-205): over the masked texels, minimise
    sum |x_k - ((L-1-q_k) e0 + q_k e1)/(L-1)|^2 (numpy.linalg.lstsq, no
    clamp); with one index value on every masked texel (|det| < 0.001) both
    endpoints are the masked mean."""
    rng = np.random.default_rng(50 + bits)
    levels = 1 << bits
    n = 0
    while n < 300:
        ch = int(rng.integers(3, 5))
        mask = int(rng.integers(1, 1 << 16))
        sel = _mask_bits(mask)
        if sel.sum() < 4:
            continue
        tex = rng.integers(0, 256, (16, ch)).astype(np.float32)
        q = rng.integers(0, levels, 16)
        qs = q[sel]
        if len(set(qs.tolist())) < 2 or np.bincount(qs).max() > sel.sum() - 2:
            continue
        ep = oracle_lib.opt_endpoints(_block64(tex), bits, _qblock(q), mask, ch)
        A = np.stack([(levels - 1 - qs) / (levels - 1), qs / (levels - 1)], 1)
        sol = np.linalg.lstsq(A, tex[sel].astype(np.float64), rcond=None)[0]
        cond = np.linalg.cond(A)
        tol = 3e-6 * cond * cond * 255
        assert np.allclose(ep[:ch], sol[0], atol=tol) and np.allclose(ep[4:4 + ch], sol[1], atol=tol), \
            (ep, sol, cond)
        n += 1
    for v in range(levels):
        tex = rng.integers(0, 256, (16, 4)).astype(np.float32)
        mask = 0xF0F0
        ep = oracle_lib.opt_endpoints(_block64(tex), bits, _qblock(np.full(16, v)), mask, 4)
        mean = tex[_mask_bits(mask)].astype(np.float64).mean(0)
        assert np.allclose(ep[:4], mean, atol=1e-4) and np.allclose(ep[4:], mean, atol=1e-4)


def _bptc_weights(bits):
    # The BPTC index weights of the format: round(64 i / (2^bits - 1)).
    m = (1 << bits) - 1
    return [_round_half_up(64 * i, m) for i in range(m + 1)]


@pytest.mark.parametrize("bits", [2, 3, 4])
def test_block_quant_error_is_the_decoded_error_and_no_neighbour_index_is_better(bits):
    """block_quant (common.rs:422-481): the returned error is the SSE of the
    BPTC interpolation (64-w)*a + w*b + 32 >> 6 at the chosen indices, and
    for each texel the chosen index is at least as good as its neighbours
    on the line (the search compares q1-1 and q1 around the projection)."""
    rng = np.random.default_rng(60 + bits)
    w = _bptc_weights(bits)
    for _ in range(300):
        ch = int(rng.integers(3, 5))
        ea = rng.integers(0, 256, ch)
        eb = rng.integers(0, 256, ch)
        if np.all(ea == eb):
            continue
        t = rng.uniform(-0.1, 1.1, 16)
        tex = np.clip(np.round(ea[None, :] + t[:, None] * (eb - ea)[None, :] + rng.uniform(-8, 8, (16, ch))),
                      0, 255).astype(np.float32)
        ep8 = np.zeros(8, np.float32)
        ep8[:ch] = ea
        ep8[4:4 + ch] = eb
        err, qb = oracle_lib.block_quant(_block64(tex), bits, ep8, ch)
        idx = [(int(qb[k // 8]) >> (4 * (k % 8))) & 15 for k in range(16)]
        dec = lambda i: (((64 - w[i]) * ea + w[i] * eb + 32) >> 6)
        sse = [float(np.sum((dec(i) - tex[k]) ** 2)) for k, i in enumerate(idx)]
        assert err == pytest.approx(sum(sse), abs=0, rel=0)
        for k, i in enumerate(idx):
            proj = np.dot(tex[k] - ea, eb - ea) / np.dot(eb - ea, eb - ea)
            q1 = min(max(int(np.floor(proj * (1 << bits) + 0.5)), 1), (1 << bits) - 1)
            assert i in (q1 - 1, q1)
            other = q1 - 1 if i == q1 else q1
            e_other = float(np.sum((dec(other) - tex[k]) ** 2))
            assert sse[k] <= e_other
            if sse[k] == e_other:  # a tie keeps q1 (strict <, common.rs:467-471)
                assert i == q1


@pytest.mark.parametrize("bits", [2, 3, 4])
def test_block_quant_ties_keep_the_rounded_projection(bits):
    """block_quant (common.rs:463-471): a texel exactly midway between two
    adjacent decoded levels has equal errors for both; the reference keeps
    q1 (the rounded projection) because only a strictly smaller error
    switches to q1 - 1."""
    rng = np.random.default_rng(65 + bits)
    w = _bptc_weights(bits)
    L = 1 << bits
    ties = 0
    for _ in range(2000):
        ea = rng.integers(0, 256, 3)
        eb = ea.copy()
        eb[0] = rng.integers(0, 256)
        if abs(int(eb[0]) - int(ea[0])) < 20:
            continue
        lv = [((64 - w[i]) * int(ea[0]) + w[i] * int(eb[0]) + 32) >> 6 for i in range(L)]
        i = int(rng.integers(1, L))
        if (lv[i] + lv[i - 1]) % 2:
            continue
        tex = np.tile(ea, (16, 1)).astype(np.float32)
        tex[:, 0] = (lv[i] + lv[i - 1]) // 2
        ep8 = np.zeros(8, np.float32)
        ep8[:3], ep8[4:7] = ea, eb
        _, qb = oracle_lib.block_quant(_block64(tex), bits, ep8, 3)
        proj = (float(tex[0, 0]) - ea[0]) / float(eb[0] - ea[0])
        q1 = min(max(int(np.floor(proj * L + 0.5)), 1), L - 1)
        if (q1 - 1, q1) != (i - 1, i):
            continue
        assert (int(qb[0]) & 15) == q1
        ties += 1
    assert ties > 50


# ---------------------------------------------------------------------------
# BC7 endpoint quantisers and the scalar channel
# ---------------------------------------------------------------------------

def _unpack(v, n):
    return (v << (8 - n)) | (v >> (2 * n - 8)) if n < 8 else v


def _nearest(y: float, cands):
    """The candidate nearest to y (exact arithmetic), ties to the larger."""
    fy = Fraction(y)
    return min(cands, key=lambda c: (abs(fy - c), -c))


# (mode) -> (code bits incl. p-bit, p-bit kind)
_PBIT_MODES = {0: 5, 3: 8, 6: 8, 7: 6}


def _pbit_codes(e, levels2):
    """For each p-bit b, the nearest code 2k+b to y = f32(f32(e/255)*levels2)
    among b, b+2, ..., levels2-1+b (brute force over the lattice)."""
    y = float(np.float32(np.float32(np.float32(e) / np.float32(255.0)) * np.float32(levels2)))
    return [_nearest(y, range(b, levels2 + b, 2)) for b in (0, 1)]


@pytest.mark.parametrize("mode", [0, 1, 2, 3, 4, 5, 6, 7])
def test_bc7_endpoint_quantisers_are_nearest_lattice_points(mode):
    """ep_quant0367 / ep_quant1 / ep_quant245 (bc7.rs:102-189).  Without a
    p-bit (modes 2, 4, 5) each code is the nearest integer to e/255*(L-1),
    clamped.  With p-bits, for each b the code is the nearest point of the
    lattice 2k+b to e/255*(2L-1), and the p-bit kept is the one whose
    (dequantised, for modes 0 and 1) codes are closer to the endpoint in
    squared error -- b=1 on a tie -- per endpoint (modes 0/3/6/7) or shared
    by both endpoints of the pair (mode 1).  Brute force over the lattice;
    inputs include out-of-range endpoints from the least-squares refits."""
    rng = np.random.default_rng(70 + mode)
    vals = np.concatenate([rng.uniform(-30, 285, 4000), rng.uniform(0, 255, 4000),
                           np.arange(256, dtype=np.float64)]).astype(np.float32)
    ints = np.arange(256, dtype=np.float32)
    for it in range(1200):
        channels = 3 if mode in (1, 2, 4, 5) else int(rng.integers(3, 5))
        # Every third pair is integer-valued: there the reference's f32 error
        # sums are exact, so exact p-bit ties occur and must go to b = 1.
        ep = ints[rng.integers(0, 256, 8)] if it % 3 == 0 else vals[rng.integers(0, vals.size, 8)]
        got = oracle_lib.bc7_ep_quant_pair(mode, ep, channels)
        if mode in (2, 4, 5):
            L = 1 << (7 if mode == 5 else 5)
            for i in range(8):
                y = float(np.float32(np.float32(ep[i] / np.float32(255.0)) * np.float32(L - 1)))
                want = _nearest(y, range(L)) if 0 <= y <= L - 1 else (0 if y < 0 else L - 1)
                assert got[i] == want, (mode, ep[i], got[i], want)
            continue
        if mode == 1:
            codes = [[_pbit_codes(e, 127)[b] for e in ep] for b in (0, 1)]
            err = [sum((Fraction(float(ep[j])) - _unpack(codes[b][j], 7)) ** 2 for j in range(8) if j % 4 < 3)
                   for b in (0, 1)]
            _check_pbit_choice(got.tolist(), codes, err, ep)
            continue
        n = _PBIT_MODES[mode]
        levels2 = (1 << (n - 1)) * 2 - 1
        for i in range(2):
            e = ep[4 * i:4 * i + 4]
            codes = [[_pbit_codes(x, levels2)[b] for x in e] for b in (0, 1)]
            val = (lambda c: _unpack(c, 5)) if mode == 0 else (lambda c: c)
            err = [sum((Fraction(float(e[p])) - val(codes[b][p])) ** 2 for p in range(channels)) for b in (0, 1)]
            _check_pbit_choice(got[4 * i:4 * i + 4].tolist(), codes, err, e)


def _check_pbit_choice(got, codes, err, ep):
    """The p-bit with the strictly smaller squared error wins, b = 1 on a
    tie (bc7.rs:139-146, 169-177).  Errors are exact rationals; the
    reference sums them in f32, which is exact for integer endpoints, so
    ties and near-ties of non-integer endpoints accept either code."""
    integral = all(float(x).is_integer() for x in ep)
    gap = abs(err[0] - err[1])
    if err[0] != err[1] and gap > Fraction(1, 10 ** 6) * (err[0] + err[1] + 1):
        assert got == codes[0 if err[0] < err[1] else 1], (ep, got, codes, err)
    elif err[0] == err[1] and integral:
        assert got == codes[1], (ep, got, codes)
    else:
        assert got in codes


@pytest.mark.parametrize("bits", [2, 3])
def test_bc7_channel_least_squares_and_index_choice(bits):
    """channel_opt_endpoints (bc7.rs:322-369) is the 16-texel scalar least
    squares fit with weights (L-1-q)/(L-1), q/(L-1) clamped to 0..255 (mean
    for a single index value); channel_opt_quant (bc7.rs:278-320) returns
    the SSE of the BPTC interpolation at its indices, each of which is no
    worse than the other candidate around the projection."""
    rng = np.random.default_rng(80 + bits)
    L = 1 << bits
    w = _bptc_weights(bits)
    n = 0
    while n < 300:
        x = rng.integers(0, 256, 16).astype(np.float32)
        q = rng.integers(0, L, 16)
        if len(set(q.tolist())) < 2 or np.bincount(q).max() > 13:
            continue
        ep = oracle_lib.bc7_channel_opt_endpoints(x, bits, _qblock(q))
        A = np.stack([(L - 1 - q) / (L - 1), q / (L - 1)], 1)
        sol = np.clip(np.linalg.lstsq(A, x.astype(np.float64), rcond=None)[0], 0, 255)
        cond = np.linalg.cond(A)
        assert np.allclose(ep, sol, atol=3e-6 * cond * cond * 255), (ep, sol)
        n += 1
    x = rng.integers(0, 256, 16).astype(np.float32)
    ep = oracle_lib.bc7_channel_opt_endpoints(x, bits, _qblock(np.full(16, 1)))
    assert np.allclose(ep, x.astype(np.float64).mean(), atol=1e-4)
    for _ in range(300):
        x = rng.integers(0, 256, 16).astype(np.float32)
        a, b = (int(v) for v in rng.integers(0, 256, 2))
        err, qb = oracle_lib.bc7_channel_opt_quant(x, bits, np.array([a, b], np.float32))
        idx = [(int(qb[k // 8]) >> (4 * (k % 8))) & 15 for k in range(16)]
        dec = lambda i: ((64 - w[i]) * a + w[i] * b + 32) >> 6
        assert err == float(sum((dec(i) - float(x[k])) ** 2 for k, i in enumerate(idx)))
        for k, i in enumerate(idx):
            proj = (float(x[k]) - a) / (b - a + 0.001)
            q1 = min(max(int(np.floor(proj * L + 0.5)), 1), L - 1)
            assert i in (q1 - 1, q1)


# ---------------------------------------------------------------------------
# BC7 mode search: group order and ties
# ---------------------------------------------------------------------------

def _mode_of(block: np.ndarray) -> int:
    b0 = int(block[0])
    return (b0 & -b0).bit_length() - 1 if b0 else 8


def _search_blocks():
    rng = np.random.default_rng(90)
    img = synth.tile_image_rgba8(32, 32, 91, 0.1)
    a = synth.tile_image_rgba8(32, 32, 92, 0.2)[..., 0]
    imgs = [img]
    # Flat and two-colour tiles: many modes reach the same error (ties).
    flat = np.zeros((32, 32, 4), np.uint8)
    for by in range(8):
        for bx in range(8):
            c = rng.integers(0, 256, 4)
            if (bx + by) % 3 == 0:
                c[:3] = [0, 0, 0] if bx % 2 else [255, 255, 255]
            c[3] = 255 if (bx + by) % 2 else rng.integers(0, 256)
            flat[4 * by:4 * by + 4, 4 * bx:4 * bx + 4] = c
            if (bx * 7 + by) % 4 == 1:
                flat[4 * by + 2:4 * by + 4, 4 * bx:4 * bx + 4, :3] = 255 - c[:3]
    imgs.append(flat)
    return imgs, a


@pytest.mark.parametrize("preset", sorted(BC7_PRESETS))
def test_bc7_mode_groups_min_and_first_wins(preset):
    """compress_block_bc7_core (bc7.rs:853-867): the full search's error is
    the minimum of the errors of runs restricted to one mode group
    (mode_selection), and the winning block's mode is from the first group,
    in the order 0/2 -> 1/3/7 -> 4/5 -> 6, that reaches the minimum (strict
    < between groups).  Within group 1 the same holds for modes 1 -> 3 -> 7
    (fast_skip thresholds).  Flat and two-colour blocks create ties, so a <=
    or a reordered group fails."""
    s = BC7_PRESETS[preset]()
    imgs, alpha = _search_blocks()
    groups = {0: (0, 2), 1: (1, 3, 7), 2: (4, 5), 3: (6,)}
    for img in imgs:
        img = img.copy()
        if s.channels == 4:
            img[..., 3] = np.where(img[..., 3] == 255, alpha, img[..., 3])
        full_blocks, full_err = oracle_lib.bc7_block_errors(s.to_bytes(), img)
        runs = {}
        for g in range(4):
            if s.mode_selection[g] == 0:
                continue
            r = dataclasses.replace(s, mode_selection=tuple(1 if i == g else 0 for i in range(4)))
            runs[g] = oracle_lib.bc7_block_errors(r.to_bytes(), img)[1]
        errs = np.stack([runs[g] for g in sorted(runs)])
        order = sorted(runs)
        assert np.array_equal(full_err, errs.min(0))
        first = np.array([order[int(np.flatnonzero(errs[:, i] == full_err[i])[0])] for i in range(errs.shape[1])])
        modes = np.array([_mode_of(b) for b in full_blocks.reshape(-1, 16)])
        for i, (m, g) in enumerate(zip(modes, first)):
            assert m in groups[g], (i, m, g, errs[:, i])
        if 1 in runs and (s.fast_skip_threshold_mode1 or s.fast_skip_threshold_mode3):
            sub = {}
            for m, (t1, t3, t7) in {1: (s.fast_skip_threshold_mode1, 0, 0),
                                    3: (0, s.fast_skip_threshold_mode3, 0),
                                    7: (0, 0, s.fast_skip_threshold_mode7)}.items():
                if (t1, t3, t7) == (0, 0, 0):
                    continue
                r = dataclasses.replace(s, mode_selection=(0, 1, 0, 0), fast_skip_threshold_mode1=t1,
                                        fast_skip_threshold_mode3=t3, fast_skip_threshold_mode7=t7)
                sub[m] = oracle_lib.bc7_block_errors(r.to_bytes(), img)[1]
            e = np.stack([sub[m] for m in sorted(sub)])
            assert np.array_equal(runs[1], e.min(0))
            g1 = dataclasses.replace(s, mode_selection=(0, 1, 0, 0))
            gb = oracle_lib.bc7_block_errors(g1.to_bytes(), img)[0].reshape(-1, 16)
            ms = sorted(sub)
            for i in range(e.shape[1]):
                want = ms[int(np.flatnonzero(e[:, i] == runs[1][i])[0])]
                assert _mode_of(gb[i]) == want, (i, e[:, i])


# ---------------------------------------------------------------------------
# BC3 alpha / BC4 / BC5: two-level and ramp closed forms
# ---------------------------------------------------------------------------

def _alpha_codes(raw8: np.ndarray):
    bits = int.from_bytes(bytes(raw8[2:8]), "little")
    return [(bits >> (3 * k)) & 7 for k in range(16)]


def _ramp_code(i):
    return 1 if i == 0 else (0 if i == 7 else 8 - i)


def test_alpha_block_two_level_and_ramp_closed_forms():
    """compress_block_bc3_alpha (bc1_to_5.rs:431-472), through BC4 (R), BC5
    (R and G) and BC3 (A): a block of two values a < b encodes as [b, a] with
    index 1 on a and 0 on b; a ramp a + i*m (i = 0..7, b = a + 7m) gets the
    codes 1, 7, 6, ..., 2, 0 (level i -> 8 - i) and decodes back exactly."""
    rng = np.random.default_rng(100)
    for _ in range(300):
        a, b = sorted(int(v) for v in rng.choice(256, 2, replace=False))
        lv = rng.integers(0, 2, 16)
        lv[rng.integers(0, 16)] = 0
        lv[(rng.integers(1, 16) + int(np.argmin(lv))) % 16] = 1
        if lv.min() == lv.max():
            continue
        m = int(rng.integers(1, 37))
        a2 = int(rng.integers(0, 256 - 7 * m))
        i = rng.integers(0, 8, 16)
        i[:2] = [0, 7]
        i = rng.permutation(i)
        two = np.where(lv == 1, b, a).astype(np.uint8)
        ramp = (a2 + i * m).astype(np.uint8)
        img = np.zeros((4, 4, 4), np.uint8)
        img[..., 0] = two.reshape(4, 4)
        img[..., 1] = ramp.reshape(4, 4)
        img[..., 3] = ramp.reshape(4, 4)
        bc4 = oracle_lib.compress_rgba8(3, b"", img)
        assert bc4[:2].tolist() == [b, a] and _alpha_codes(bc4) == [1 if v == 0 else 0 for v in lv]
        bc5 = oracle_lib.compress_rgba8(4, b"", img)
        assert np.array_equal(bc5[:8], bc4)
        assert bc5[8:10].tolist() == [a2 + 7 * m, a2]
        assert _alpha_codes(bc5[8:]) == [_ramp_code(int(v)) for v in i]
        bc3 = oracle_lib.compress_rgba8(2, b"", img)
        assert np.array_equal(bc3[:8], bc5[8:])
        dec = oracle_lib.decode_block("bc4", bc5[8:].tobytes(), 4, 16)
        assert np.array_equal(dec, ramp)


def test_bc2_and_bc3_colour_halves_are_the_bc1_block():
    """encode.rs:322-337, 384-398: the colour half of BC2 and BC3 is
    compress_block_bc1_core of the same texels (BC1's 8 bytes)."""
    img = synth.tile_image_rgba8(64, 64, 101, 0.1)
    img[..., 3] = synth.tile_image_rgba8(64, 64, 102, 0.2)[..., 1]
    bc1 = oracle_lib.compress_rgba8(0, b"", img).reshape(-1, 8)
    for tag in (1, 2):
        out = oracle_lib.compress_rgba8(tag, b"", img).reshape(-1, 16)
        assert np.array_equal(out[:, 8:], bc1)


# ---------------------------------------------------------------------------
# BC6H (unsigned): decode reproduces the encoder's error; modes; grid values
# ---------------------------------------------------------------------------

def _bc6h_images():
    c3 = synth.hdr_image_rgba16f(64, 64, 3)
    rng = np.random.default_rng(110)
    ext = np.zeros((16, 16, 4), np.uint16)
    choices = np.array([0, 1, 2, 0x03FF, 0x0400, 0x3C00, 0x7BFF, 0x7BFE, 0x5000, 0x0001], np.uint16)
    for by in range(4):
        for bx in range(4):
            k = by * 4 + bx
            if k % 4 == 0:
                blk = np.full((4, 4, 3), choices[k % choices.size], np.uint16)
            elif k % 4 == 1:
                blk = choices[rng.integers(0, choices.size, (4, 4, 3))]
            elif k % 4 == 2:
                blk = rng.integers(0, 0x7C00, (4, 4, 3)).astype(np.uint16)
            else:
                blk = np.where(rng.random((4, 4, 3)) < 0.5, 0, 0x7BFF).astype(np.uint16)
            ext[4 * by:4 * by + 4, 4 * bx:4 * bx + 4, :3] = blk
    ext[..., 3] = 0x3C00
    return {"c3_generator": c3, "extremes": ext.view(np.float16)}


def _bc6h_texel_values(img):
    h, w, _ = img.shape
    bits = img.view(np.uint16).reshape(h // 4, 4, w // 4, 4, 4).transpose(0, 2, 1, 3, 4).reshape(-1, 16, 4)
    return (bits[..., :3].astype(np.float32) / np.float32(31.0)) * np.float32(64.0)


_RESERVED_BC6H = {0b10011, 0b10111, 0b11011, 0b11111}


@pytest.mark.parametrize("name", ["c3_generator", "extremes"])
@pytest.mark.parametrize("preset", sorted(BC6H_PRESETS))
def test_bc6h_decoder_reproduces_the_unsigned_encoders_error(name, preset):
    """compress_bc6h_core and bc6h_pack (bc6h.rs:166-860): decoding every
    emitted block with the decoder the reference's known-answer vectors pin
    (block.rs:262-769) gives interpolated values whose squared error against
    the texel values bits/31*64 equals the encoder's own error (f32
    accumulation: relative 1e-5); no block uses a reserved mode.  A wrong
    field, delta, transform, sign extension or index packing breaks it."""
    img = _bc6h_images()[name]
    blocks, errs = oracle_lib.bc6h_block_errors(BC6H_PRESETS[preset]().to_bytes(), img)
    blocks = blocks.reshape(-1, 16)
    for b in blocks:
        m = int(b[0]) & 3 if (int(b[0]) & 2) == 0 else int(b[0]) & 31
        assert m not in _RESERVED_BC6H
    comp = np.stack([oracle_lib.decode_block_bc6h_comp(b.tobytes(), False) for b in blocks]).astype(np.float64)
    err = ((comp - _bc6h_texel_values(img).astype(np.float64)) ** 2).sum(axis=(1, 2))
    assert np.all(np.abs(err - errs) <= 1e-5 * np.maximum(errs, 1.0)), np.max(np.abs(err - errs))


def test_bc6h_constant_blocks_on_the_grid_round_trip_exactly():
    """A constant block whose half bits h are a multiple of 31 has value
    h/31*64 on the 16-bit endpoint grid, so the slow search finds a
    zero-error encoding and the decoder (finish_unquantize: x*31 >> 6)
    returns h exactly (bc6h.rs:793-860, block.rs:1391-1397)."""
    rng = np.random.default_rng(111)
    n = 256
    vals = (rng.integers(0, 0x7BFF // 31 + 1, (n, 3)) * 31).astype(np.uint16)
    vals[:4] = [[0, 0, 0], [31, 0, 62], [0x7BFF // 31 * 31] * 3, [31 * 500, 31 * 7, 31 * 900]]
    img = np.zeros((4, 4 * n, 4), np.uint16)
    for i in range(n):
        img[:, 4 * i:4 * i + 4, :3] = vals[i]
    blocks, errs = oracle_lib.bc6h_block_errors(BC6HSettings.slow().to_bytes(), img.view(np.float16))
    assert np.all(errs == 0)
    for i, b in enumerate(blocks.reshape(-1, 16)):
        dec = oracle_lib.decode_block_bc6h(b.tobytes(), 12, False).reshape(16, 3)
        assert np.all(dec == vals[i][None, :]), i
