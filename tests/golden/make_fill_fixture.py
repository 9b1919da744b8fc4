"""Copies the reference's Smart Blend / hole fill model fixtures into tests/golden. Data only.

* plugins/retime-engine/tests/fixtures/f1_cases.hex (F1C1: 36 Smart Blend composites and 73 hole
  fills of a numpy float32 model, read by holefill/tests.rs:90-160, bit-exact there) becomes
  f1_cases.bin.gz: the same bytes, decoded from hex and gzip-compressed.
* The HF1 probe of holefill/tests.rs:588-628 (a 12x9 plane, its hole mask, the 48 values the
  Vectors In plugin filled with the row above first, and the 48 values of the below-first order)
  becomes hf1.json.

Run in a container that has the reference checked out at /root/reference."""
import gzip
import json
import os
import re

REF = "/root/reference/plugins/retime-engine"
HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    with open(os.path.join(REF, "tests/fixtures/f1_cases.hex")) as fh:
        blob = bytes.fromhex("".join(fh.read().split()))
    assert blob[:4] == b"F1C1"
    with gzip.open(os.path.join(HERE, "f1_cases.bin.gz"), "wb", compresslevel=9) as fh:
        fh.write(blob)

    with open(os.path.join(REF, "src/holefill/tests.rs")) as fh:
        text = fh.read()

    def usize(name):
        return int(re.search(rf"const {name}: usize = (\d+);", text).group(1))

    def words(name):
        body = re.search(rf"const {name}: \[u32; \d+\] = \[(.*?)\];", text, re.S).group(1)
        return [int(v, 16) for v in re.findall(r"0x[0-9a-fA-F]+", body)]

    mask = re.search(r'const HF1_MASK: &str = "([01]+)";', text).group(1)
    hf1 = dict(w=usize("HF1_W"), h=usize("HF1_H"), input=words("HF1_INPUT"), mask=mask,
               plugin=words("HF1_PLUGIN"), below_first=words("HF1_BELOW_FIRST"))
    assert len(hf1["input"]) == hf1["w"] * hf1["h"] == len(mask)
    assert len(hf1["plugin"]) == len(hf1["below_first"]) == mask.count("1")
    with open(os.path.join(HERE, "hf1.json"), "w") as fh:
        json.dump(hf1, fh)
        fh.write("\n")
    print(len(blob), "bytes of F1C1,", len(hf1["plugin"]), "HF1 values")


if __name__ == "__main__":
    main()
