"""The oracle's pin tests fail on plausible transcription slips.

Each case copies the oracle's sources to a temporary directory, plants one
slip of the kind a restatement of the reference could contain (a swapped
index, a wrong sign, a dropped rounding term, <= for <, a wrong bit), builds
that copy, and runs the pin tests that are meant to catch it against the
copy (BCO_ORACLE_LIB).  The case passes when the pin tests fail.  The planted
slips live only in the temporary copies; the oracle itself is untouched.
"""
import os
import shutil
import subprocess
import sys
import tempfile
from concurrent.futures import ThreadPoolExecutor

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")
SOURCES = ["common.cpp", "bc7.cpp", "bc15.cpp", "bc6h.cpp", "decode.cpp", "capi.cpp", "dds.cpp"]

PINS = "tests/test_oracle_encoder_pins.py"
OLD_PINS = "tests/test_oracle_pins.py"

# (name, file, original text, slip, pin file, -k selector)
SLIPS = [
    ("enc565_channel_order", "bc15.cpp", "return (r5 << 11) + (g6 << 5) + b5;",
     "return (b5 << 11) + (g6 << 5) + r5;", PINS, "enc_rgb565"),
    ("fast_quant_bias", "bc15.cpp", "F bias = 0.5f;", "F bias = 0.0f;", PINS, "fast_quant"),
    ("bc1_refine_sign", "bc15.cpp", "atb2[p] = F(3.0f) * sum[p] - atb1[p];",
     "atb2[p] = F(3.0f) * sum[p] + atb1[p];", PINS, "bc1_refine"),
    ("bc1_ssymv_index", "bc15.cpp", "r[1] = c[1] * a[0] + c[3] * a[1] + c[4] * a[2];",
     "r[1] = c[1] * a[0] + c[3] * a[1] + c[2] * a[2];", PINS, "bc1_axis"),
    ("alpha_code_wrap", "bc15.cpp", "if (q == 8) q = 1;", "if (q == 8) q = 0;", PINS, "alpha_block"),
    ("ssymv3_index", "common.cpp", "a[2] = covar[2] * b[0] + covar[5] * b[1] + covar[7] * b[2];",
     "a[2] = covar[2] * b[0] + covar[5] * b[1] + covar[8] * b[2];", PINS, "compute_axis"),
    ("pca_bound_eps_slot", "common.cpp", "covar_scaled[7] += EPS;", "covar_scaled[9] += EPS;", PINS,
     "get_pca_bound"),
    ("opt_endpoints_cxy", "common.cpp", "F cxy = i2f(levels - 1) * sum_q - sum_qq;",
     "F cxy = i2f(levels - 1) * sum_q + sum_qq;", PINS, "opt_endpoints"),
    ("stats_moment", "common.cpp", "stats[5] += rgba[1] * rgba[2];", "stats[5] += rgba[1] * rgba[1];",
     PINS, "compute_stats_masked"),
    ("segment_extents_swapped", "common.cpp",
     "for (size_t p = 0; p < channels; p++) ep[4 * i + p] = ext[i] * axis[p] + dc[p];",
     "for (size_t p = 0; p < channels; p++) ep[4 * i + p] = ext[1 - i] * axis[p] + dc[p];", PINS,
     "segment"),
    ("block_quant_tie", "common.cpp", "if (err0 < err1) {", "if (err0 <= err1) {", PINS, "block_quant"),
    ("pbit_tie", "bc7.cpp", "bool take0 = err0 < err1;", "bool take0 = err0 <= err1;", PINS,
     "endpoint_quantisers"),
    ("mode1_code_clamp", "bc7.cpp", "qep_b[b * 8 + i] = iclamp(v, b, 126 + b);",
     "qep_b[b * 8 + i] = iclamp(v, b, 127 + b);", PINS, "endpoint_quantisers"),
    ("quant245_truncates", "bc7.cpp",
     "int32_t v = f2i(ep[i] / F(255.0f) * i2f(levels - 1) + F(0.5f));",
     "int32_t v = f2i(ep[i] / F(255.0f) * i2f(levels - 1));", PINS, "endpoint_quantisers"),
    ("mode_group_tie", "bc7.cpp", "if (best < best_err) {", "if (best <= best_err) {", PINS,
     "mode_groups"),
    ("channel_lsq_cxx", "bc7.cpp", "F cxx = F(16.0f) * sq(i2f(levels - 1))",
     "F cxx = F(15.0f) * sq(i2f(levels - 1))", PINS, "channel_least_squares"),
    ("mode6_pbit_source", "bc7.cpp", "put_bits(data, &pos, 1, (uint32_t)qep[4] & 1);",
     "put_bits(data, &pos, 1, (uint32_t)qep[0] & 1);", OLD_PINS, "encoder_error_equals_decoded_sse"),
    ("bc6h_pack_bit", "bc6h.cpp", "pq[5] += bit_at(pr[12 + 2], 0) << 5;",
     "pq[5] += bit_at(pr[12 + 2], 1) << 5;", PINS, "unsigned_encoders_error"),
    ("bc6h_unquant_top", "bc6h.cpp", "if (v == (1u << bits) - 1) return 0xFFFF;",
     "if (v == (1u << bits) - 1) return 0xFFFE;", PINS, "bc6h"),
]


def _build(tmp: str, name: str, fname: str, old: str, new: str) -> str:
    d = os.path.join(tmp, name)
    os.makedirs(d)
    for f in SOURCES + ["bc_oracle.hpp"]:
        shutil.copy(os.path.join(ORACLE, f), d)
    path = os.path.join(d, fname)
    with open(path) as fh:
        text = fh.read()
    assert old in text, (name, old)
    with open(path, "w") as fh:
        fh.write(text.replace(old, new, 1))
    lib = os.path.join(d, "liboracle_slip.so")
    subprocess.run(["g++", "-std=c++17", "-O1", "-fPIC", "-shared", "-fopenmp", "-ffp-contract=off",
                    "-fno-fast-math", "-w", *[os.path.join(d, s) for s in SOURCES], "-o", lib],
                   check=True)
    return lib


def _run_pins(lib: str, pin_file: str, selector: str) -> subprocess.CompletedProcess:
    env = dict(os.environ, BCO_ORACLE_LIB=lib)
    return subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider",
                           "-p", "no:xdist", pin_file, "-k", selector],
                          cwd=ROOT, env=env, capture_output=True, text=True, timeout=600)


@pytest.mark.slow
def test_every_planted_slip_fails_its_pins():
    with tempfile.TemporaryDirectory(prefix="bco_slips_") as tmp:
        with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as pool:
            libs = list(pool.map(lambda s: _build(tmp, *s[:4]), SLIPS))
            runs = list(pool.map(lambda a: _run_pins(a[0], a[1][4], a[1][5]), zip(libs, SLIPS)))
        caught = {s[0]: r.returncode == 1 for s, r in zip(SLIPS, runs)}
        # A selector that matches nothing (returncode 5) or an error (2+) is
        # not a catch.
        missed = [n for n, ok in caught.items() if not ok]
        assert not missed, {n: runs[[s[0] for s in SLIPS].index(n)].stdout[-400:] for n in missed}


def test_the_pin_selectors_pass_on_the_oracle_itself():
    """The same selectors select tests that pass on the real oracle."""
    for pin_file, sel in sorted({(s[4], s[5]) for s in SLIPS}):
        r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-p", "no:cacheprovider", "-p", "no:xdist",
                            "--collect-only", pin_file, "-k", sel], cwd=ROOT, capture_output=True, text=True)
        assert r.returncode == 0, (pin_file, sel, r.stdout[-300:])
