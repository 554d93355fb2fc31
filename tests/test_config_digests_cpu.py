"""The seeded benchmark inputs (C1-C5) hash to the committed SHA-256 digests
(tests/golden/config_digests.json, written by make_config_digests.py)."""
import json
import os
import sys

import pytest

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
sys.path.insert(0, GOLDEN)


@pytest.mark.slow
def test_config_inputs_match_committed_digests():
    from config_inputs import digests

    with open(os.path.join(GOLDEN, "config_digests.json")) as f:
        want = json.load(f)
    assert digests() == want
