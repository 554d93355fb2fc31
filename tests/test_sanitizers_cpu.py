"""The oracle (oracle/*.cpp) and the kernels' host build (tests/kernel_check)
under AddressSanitizer and UBSan.

Both are rebuilt with -fsanitize=address,undefined (oracle/Makefile's
`sanitize` target, tests/kernel_check/build.sh with SANITIZE=1) and a
representative set of their CPU tests runs against those builds in a child
Python with libasan preloaded (Python itself is not instrumented).  Any
out-of-bounds access, use after free or undefined behaviour (shifts past the
width, signed overflow, misaligned loads) aborts the child, and the test
fails with its report.
"""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HERE = os.path.join(ROOT, "tests")

# (test file, -k selection): decoders, encoders of every format on small
# images, the quantiser and PCA pins, the kernels' decoders and encoders.
RUNS = [
    ("tests/test_oracle_pins.py", None),
    ("tests/test_oracle_encoder_pins.py",
     "enc_rgb565 or bc1_decoder or fast_quant or bc1_refine or compute_axis or endpoint_quantisers "
     "or alpha_block or bc6h"),
    ("tests/test_kernel_check_cpu.py",
     "decoder or block_pos or channel or bc15 or bc6h_kernel or opaque_basic or alpha_slow or pbit"),
]


def _asan_runtime() -> str:
    path = subprocess.run(["g++", "-print-file-name=libasan.so"], capture_output=True, text=True,
                          check=True).stdout.strip()
    if not os.path.isabs(path) or not os.path.exists(path):
        pytest.skip("no libasan for the host compiler")
    return path


@pytest.mark.slow
def test_oracle_and_kernel_host_build_are_clean_under_asan_and_ubsan():
    asan = _asan_runtime()
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "sanitize"], check=True)
    subprocess.run([os.path.join(HERE, "kernel_check", "build.sh")], check=True,
                   env=dict(os.environ, SANITIZE="1"))
    env = dict(os.environ,
               LD_PRELOAD=asan,
               ASAN_OPTIONS="detect_leaks=0:abort_on_error=1:halt_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1",
               BCO_ORACLE_LIB=os.path.join(ROOT, "oracle", "_build", "liboracle_bc_san.so"),
               KC_LIB=os.path.join(HERE, "kernel_check", "_build", "libkernel_check_san.so"),
               OMP_NUM_THREADS="4")
    for path, sel in RUNS:
        # -s: a sanitizer report goes to the child's stderr, not into a
        # capture buffer that an abort would discard.
        cmd = [sys.executable, "-m", "pytest", "-q", "-x", "-s", "-p", "no:cacheprovider", "-p", "no:xdist",
               "-m", "not gpu", path]
        if sel:
            cmd += ["-k", sel]
        r = subprocess.run(cmd, cwd=ROOT, env=env, capture_output=True, text=True, timeout=900)
        out = r.stdout + r.stderr
        assert r.returncode == 0, (path, out[-3000:])
        assert "runtime error" not in out and "AddressSanitizer" not in out, (path, out[-3000:])
        assert " passed" in r.stdout, (path, r.stdout[-500:])
