"""Extract the reference's decoder known-answer tests into a JSON fixture.

Reads the 18 #[test] functions of src/decode/block.rs (lines 1405-1841 of the
reference crate) as text and writes, for each, the decoder it calls, the
destination pitch, the compressed bytes and the expected output values, to
tests/golden/decode_kats.json.  Run once in a container where the reference
checkout exists:

    python tests/golden/make_decode_kats.py /root/reference
"""
import json
import os
import re
import sys


def parse_ints(text):
    return [int(x, 0) for x in re.findall(r"0x[0-9A-Fa-f]+|\b\d+\b", text)]


def main(ref_root):
    src = open(os.path.join(ref_root, "src/decode/block.rs")).read()
    tests = src[src.index("#[cfg(test)]"):]
    kats = []
    for m in re.finditer(r"fn (test_\w+)\(\) \{(.*?)\n    \}\n", tests, re.S):
        name, body = m.group(1), m.group(2)
        cb = re.search(r"let compressed_block = \[(.*?)\];", body, re.S)
        rep = re.fullmatch(r"\s*(\d+)u8;\s*(\d+)\s*", cb.group(1))
        if rep:  # `[0u8; 8]`
            compressed = [int(rep.group(1))] * int(rep.group(2))
        else:
            compressed = parse_ints(cb.group(1))
        if "decode_block_bc6h_float" in body:
            exp = re.search(r"let expected_output: \[f32; 48\] = \[(.*?)\];", body, re.S)
            values = [float(x) for x in re.findall(r"-?\d+\.\d+(?:e-?\d+)?", exp.group(1))]
            kats.append({"name": name, "decoder": "bc6h_float", "pitch": 12, "signed": True,
                         "compressed": compressed, "expected_f32": values})
        elif "decode_block_bc6h(" in body:
            bits = [int(x, 16) for x in re.findall(r"f16::from_bits\((0x[0-9A-Fa-f]+)\)", body)]
            kats.append({"name": name, "decoder": "bc6h", "pitch": 12, "signed": True,
                         "compressed": compressed, "expected_f16_bits": bits})
        else:
            call = re.search(r"test_block\(\s*decode_block_(\w+),\s*(\d+),", body, re.S)
            exp = re.search(r"let expected_output = \[(.*?)\];", body, re.S)
            kats.append({"name": name, "decoder": call.group(1), "pitch": int(call.group(2)),
                         "compressed": compressed, "expected": parse_ints(exp.group(1))})
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "decode_kats.json")
    with open(out, "w") as f:
        json.dump({"source": "src/decode/block.rs:1405-1841", "kats": kats}, f, indent=1)
    print(f"wrote {len(kats)} KATs to {out}")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
