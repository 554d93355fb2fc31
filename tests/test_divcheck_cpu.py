"""The kernels' reciprocal-and-correct division equals IEEE division on every
input its call sites can see (tests/divcheck/divcheck.cpp states the domains).

The integer domains are swept exhaustively here.  The all-float domains
(divisors 255 and 65535) are sampled on every 16th high half-word by default;
BCN_DIVCHECK_FULL=1 runs the full 2^32 sweep (about a minute each).
"""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
EXE = os.path.join(HERE, "divcheck", "_build", "divcheck")


@pytest.fixture(scope="module")
def divcheck():
    subprocess.run([os.path.join(HERE, "divcheck", "build.sh")], check=True)
    return EXE


@pytest.mark.parametrize("domain", ["int_small", "channel", "proj"])
def test_integer_domains_exhaustive(divcheck, domain):
    r = subprocess.run([divcheck, domain], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert ", 0 differ" in r.stdout


@pytest.mark.parametrize("domain", ["const255", "const65535"])
def test_constant_divisors_all_floats(divcheck, domain):
    stride = "1" if os.environ.get("BCN_DIVCHECK_FULL") == "1" else "16"
    r = subprocess.run([divcheck, domain, "0", stride], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert ", 0 differ" in r.stdout


def test_checker_detects_an_uncorrected_quotient(divcheck):
    # BC6H's float moments over a texel count are outside the exact domains;
    # the checker must report differences there (this is why BC6H keeps IEEE
    # division for them).
    r = subprocess.run([divcheck, "small_f", "0", "4096"], capture_output=True, text=True)
    assert r.returncode == 1
    assert ", 0 differ" not in r.stdout
