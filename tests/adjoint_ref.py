"""FP64 reference of the differentiable BC7 / BC1 decode, for the adjoint tests.

Written from the BC7 format description and the reference decoder
(src/decode/block.rs:801-1252 for BC7, :124-197 for BC1), independently of
the kernels' parser (block_compression_amd/csrc/decode_adjoint.hpp): its own
bit reader, mode table and field order, and the oracle's copy of the
reference decoder's partition table (anchors flagged by bit 7).

`exact_*` decode integer endpoints exactly as the reference decoder does,
which pins this parser to the oracle; `smooth_*` are the linear relaxation
the kernels differentiate, written as torch float64 operations on the
endpoint tensor so that torch.autograd gives a reference gradient.
"""
import numpy as np
import torch

import oracle_lib

# ns, partition bits, rotation bits, index-selector bits, colour bits, alpha
# bits, p-bit kind (0 none, 1 per endpoint, 2 per subset), index bits, second index bits
MODES = [
    (3, 4, 0, 0, 4, 0, 1, 3, 0), (2, 6, 0, 0, 6, 0, 2, 3, 0), (3, 6, 0, 0, 5, 0, 0, 2, 0),
    (2, 6, 0, 0, 7, 0, 1, 2, 0), (1, 0, 2, 1, 5, 6, 0, 2, 3), (1, 0, 2, 0, 7, 8, 0, 2, 2),
    (1, 0, 0, 0, 7, 7, 1, 4, 0), (2, 6, 0, 0, 5, 5, 1, 2, 0),
]
WEIGHTS = {2: [0, 21, 43, 64], 3: [0, 9, 18, 27, 37, 46, 55, 64],
           4: [0, 4, 9, 13, 17, 21, 26, 30, 34, 38, 43, 47, 51, 55, 60, 64]}
_SETS = None


def _sets():
    global _SETS
    if _SETS is None:
        _SETS = oracle_lib.partition_sets()  # [2|3 subsets][64][16], bit 7 = anchor
    return _SETS


class _Bits:
    def __init__(self, block: bytes):
        self.v = int.from_bytes(bytes(block), "little")
        self.pos = 0

    def read(self, n):
        r = (self.v >> self.pos) & ((1 << n) - 1)
        self.pos += n
        return r


def parse_bc7(block: bytes) -> dict:
    b = _Bits(block)
    first = block[0]
    if first == 0:
        return {"mode": 8}
    mode = (first & -first).bit_length() - 1
    ns, pb, rb, isb_bits, cb, ab, pkind, ib, ib2 = MODES[mode]
    b.read(mode + 1)
    part, rot, isb = b.read(pb), b.read(rb), b.read(isb_bits)
    ne = 2 * ns
    q = np.zeros((6, 4), np.int64)
    for c in range(3):
        for e in range(ne):
            q[e, c] = b.read(cb)
    if ab:
        for e in range(ne):
            q[e, 3] = b.read(ab)
    p = np.zeros(6, np.int64)
    if pkind == 1:
        for e in range(ne):
            p[e] = b.read(1)
    elif pkind == 2:
        for s in range(2):
            p[2 * s] = p[2 * s + 1] = b.read(1)
    if ns == 1:
        subset = np.zeros(16, np.int64)
        anchors = {0}
    else:
        row = _sets()[ns - 2][part]
        subset = (row & 3).astype(np.int64)
        anchors = {k for k in range(16) if row[k] & 0x80} | {0}
    idx = np.array([b.read(ib - 1 if k in anchors else ib) for k in range(16)])
    idx2 = np.array([b.read(ib2 - 1 if k == 0 else ib2) for k in range(16)]) if ib2 else None
    return dict(mode=mode, ns=ns, cb=cb, ab=ab, pkind=pkind, ib=ib, ib2=ib2, part=part, rot=rot,
                isb=isb, q=q, p=p, subset=subset, idx=idx, idx2=idx2)


def _weights(f):
    w1 = np.array([WEIGHTS[f["ib"]][i] for i in f["idx"]])
    if not f["ib2"]:
        return w1, w1
    w2 = np.array([WEIGHTS[f["ib2"]][i] for i in f["idx2"]])
    return (w2, w1) if f["isb"] else (w1, w2)


def _has_p(f, c):
    return f["pkind"] == 1 or (f["pkind"] == 2 and c < 3)


def _bits(f, c):
    return f["cb"] if c < 3 else f["ab"]


def _out_channel(rot, c):
    if rot == 0:
        return c
    if c == 3:
        return rot - 1
    return 3 if c == rot - 1 else c


def exact_bc7(f, q=None) -> np.ndarray:
    """Decode with integer endpoints q (default: the block's own)."""
    out = np.zeros((16, 4), np.int64)
    if f["mode"] == 8:
        return out
    q = f["q"] if q is None else q
    wc, wa = _weights(f)
    for c in range(4):
        n = _bits(f, c)
        for k in range(16):
            s = f["subset"][k]
            vals = []
            for e in (2 * s, 2 * s + 1):
                if n == 0:
                    vals.append(255)
                    continue
                v, m = int(q[e, c]), n
                if _has_p(f, c):
                    v, m = (v << 1) | int(f["p"][e]), n + 1
                v <<= 8 - m
                vals.append(v | (v >> m))
            w = wc[k] if c < 3 else wa[k]
            out[k, _out_channel(f["rot"], c)] = (vals[0] * (64 - w) + vals[1] * w + 32) >> 6
    return out


def smooth_bc7(f, x: torch.Tensor) -> torch.Tensor:
    """Linear relaxation (float64 torch ops on x[6][4])."""
    if f["mode"] == 8:
        return torch.zeros((16, 4), dtype=x.dtype)
    wc, wa = _weights(f)
    cols = [None] * 4
    for c in range(4):
        n = _bits(f, c)
        texels = []
        for k in range(16):
            s = int(f["subset"][k])
            ends = []
            for e in (2 * s, 2 * s + 1):
                if n == 0:
                    ends.append(torch.tensor(255.0, dtype=x.dtype))
                elif _has_p(f, c):
                    ends.append((2.0 * x[e, c] + float(f["p"][e])) * (255.0 / ((1 << (n + 1)) - 1)))
                else:
                    ends.append(x[e, c] * (255.0 / ((1 << n) - 1)))
            w = float(wc[k] if c < 3 else wa[k])
            texels.append(((64.0 - w) * ends[0] + w * ends[1]) / 64.0)
        cols[_out_channel(f["rot"], c)] = torch.stack(texels)
    return torch.stack(cols, dim=1)


# ---- BC1 ---------------------------------------------------------------------


def parse_bc1(block: bytes) -> dict:
    w0 = int.from_bytes(bytes(block[:4]), "little")
    w1 = int.from_bytes(bytes(block[4:8]), "little")
    c0, c1 = w0 & 0xFFFF, w0 >> 16
    q = np.array([[(c >> 11) & 31, (c >> 5) & 63, c & 31] for c in (c0, c1)], np.int64)
    return dict(four=c0 > c1, q=q, idx=np.array([(w1 >> (2 * k)) & 3 for k in range(16)]))


def exact_bc1(f, q=None) -> np.ndarray:
    q = f["q"] if q is None else q
    (r0, g0, b0), (r1, g1, b1) = [tuple(int(v) for v in row) for row in q]
    pal = [((r0 * 527 + 23) >> 6, (g0 * 259 + 33) >> 6, (b0 * 527 + 23) >> 6, 255),
           ((r1 * 527 + 23) >> 6, (g1 * 259 + 33) >> 6, (b1 * 527 + 23) >> 6, 255)]
    if f["four"]:
        pal.append((((2 * r0 + r1) * 351 + 61) >> 7, ((2 * g0 + g1) * 2763 + 1039) >> 11,
                    ((2 * b0 + b1) * 351 + 61) >> 7, 255))
        pal.append((((r0 + 2 * r1) * 351 + 61) >> 7, ((g0 + 2 * g1) * 2763 + 1039) >> 11,
                    ((b0 + 2 * b1) * 351 + 61) >> 7, 255))
    else:
        pal.append((((r0 + r1) * 1053 + 125) >> 8, ((g0 + g1) * 4145 + 1019) >> 11,
                    ((b0 + b1) * 1053 + 125) >> 8, 255))
        pal.append((0, 0, 0, 0))
    return np.array([pal[i] for i in f["idx"]], np.int64)


def smooth_bc1(f, x: torch.Tensor) -> torch.Tensor:
    scale = torch.tensor([255.0 / 31, 255.0 / 63, 255.0 / 31], dtype=x.dtype)
    b0, b1 = x[0] * scale, x[1] * scale
    if f["four"]:
        pal = [b0, b1, (2.0 * b0 + b1) / 3.0, (b0 + 2.0 * b1) / 3.0]
        alpha = [255.0] * 4
    else:
        pal = [b0, b1, (b0 + b1) / 2.0, torch.zeros(3, dtype=x.dtype)]
        alpha = [255.0, 255.0, 255.0, 0.0]
    rows = [torch.cat([pal[i], torch.tensor([alpha[i]], dtype=x.dtype)]) for i in f["idx"]]
    return torch.stack(rows)


def random_bc7_blocks(n: int, seed: int) -> np.ndarray:
    """Random 128-bit blocks spread over modes 0-7 (every bit pattern after the
    mode prefix is a valid block)."""
    rng = np.random.default_rng(seed)
    blocks = rng.integers(0, 256, (n, 16), dtype=np.uint8)
    for i in range(n):
        m = i % 8
        blocks[i, 0] = (blocks[i, 0] & (0xFF & ~((1 << (m + 1)) - 1))) | (1 << m)
    return blocks


# ---- texel-side adjoint ------------------------------------------------------
#
# The decode as a function of the source texels: each subset's endpoints
# re-fitted to the source by least squares (one index set at a time), then
# the relaxed decode.  Its Jacobian per subset and channel is the hat matrix
# A pinv(A) of the design A = [1 - w, w] (w: the texels' weights in [0, 1]);
# computed here with numpy's pinv in float64, independently of the kernels'
# closed form.


def _hat(w: np.ndarray) -> np.ndarray:
    a = np.stack([1.0 - w, w], axis=1)
    return a @ np.linalg.pinv(a)


def texel_adjoint_bc7(f, g: np.ndarray) -> np.ndarray:
    """dL/d source (16, 4) for dL/d decoded g (16, 4), float64."""
    out = np.zeros((16, 4))
    if f["mode"] == 8:
        return out
    wc, wa = _weights(f)
    for c in range(4):
        if _bits(f, c) == 0:
            continue  # alpha decodes as 255 whatever the source
        o = _out_channel(f["rot"], c)
        w = (wc if c < 3 else wa) / 64.0
        for s in range(f["ns"]):
            m = f["subset"] == s
            if m.any():
                out[m, o] = _hat(w[m]).T @ g[m, o]
    return out


def texel_adjoint_bc1(f, g: np.ndarray) -> np.ndarray:
    out = np.zeros((16, 4))
    if f["four"]:
        wt = np.array([0.0, 1.0, 1.0 / 3.0, 2.0 / 3.0])
        m = np.ones(16, bool)
    else:
        wt = np.array([0.0, 1.0, 0.5, 0.0])
        m = f["idx"] != 3  # black whatever the source
    w = wt[f["idx"]]
    if m.any():
        h = _hat(w[m])
        for c in range(3):
            out[m, c] = h.T @ g[m, c]
    return out


def refit_decode_bc7(f, t: np.ndarray) -> np.ndarray:
    """The map the texel adjoint differentiates, float64: per subset and
    channel the least-squares endpoints of the source t (16, 4), then the
    relaxed interpolation; channels without stored bits decode to 255."""
    out = np.full((16, 4), 255.0)
    if f["mode"] == 8:
        return np.zeros((16, 4))
    wc, wa = _weights(f)
    for c in range(4):
        o = _out_channel(f["rot"], c)
        if _bits(f, c) == 0:
            continue
        w = (wc if c < 3 else wa) / 64.0
        for s in range(f["ns"]):
            m = f["subset"] == s
            if m.any():
                a = np.stack([1.0 - w[m], w[m]], axis=1)
                e = np.linalg.lstsq(a, t[m, o], rcond=None)[0]
                out[m, o] = a @ e
    return out
