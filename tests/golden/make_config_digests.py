"""Writes tests/golden/config_digests.json: the SHA-256 of every seeded input
buffer of the benchmark workloads, so that a change of generator, seed or
layout shows up as a failing test (tests/test_config_digests_cpu.py).

    python tests/golden/make_config_digests.py
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
sys.path.insert(0, HERE)

from config_inputs import digests  # noqa: E402

if __name__ == "__main__":
    out = os.path.join(HERE, "config_digests.json")
    with open(out, "w") as f:
        json.dump(digests(), f, indent=1, sort_keys=True)
        f.write("\n")
    print(out)
