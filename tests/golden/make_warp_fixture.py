"""Copies the reference's inverse-warp model fixture (WRP1, 60 cases, bit-exact:
plugins/retime-engine/tests/fixtures/warp_cases.hex, read by warp/tests.rs:590-635) into
tests/golden/warp_cases.bin.gz: the same bytes, decoded from hex and gzip-compressed. Data only.

Run in a container that has the reference checked out at /root/reference."""
import gzip
import os

SRC = "/root/reference/plugins/retime-engine/tests/fixtures/warp_cases.hex"
HERE = os.path.dirname(os.path.abspath(__file__))

with open(SRC) as fh:
    digits = "".join(fh.read().split())
blob = bytes.fromhex(digits)
assert blob[:4] == b"WRP1"
with gzip.open(os.path.join(HERE, "warp_cases.bin.gz"), "wb", compresslevel=9) as fh:
    fh.write(blob)
print(len(blob), "bytes")
