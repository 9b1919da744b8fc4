"""Copies the reference's forward-warp model fixture into tests/golden. Data only.

plugins/retime-engine/tests/fixtures/f2_cases.hex (F2C1: 26 cases of the textured forward warp
with coverage from a numpy float32 model over the three depths: translations, shears, squeezes,
a swirl, random flows, a half resolution flow, negative, tiny and zero amounts; read by
forward/tests.rs:552-600, where u8, u16 and the coverage are bit-exact and f32 is within 2e-6)
becomes f2_cases.bin.gz: the same bytes, decoded from hex and gzip-compressed.

Run in a container that has the reference checked out at /root/reference."""
import gzip
import os

REF = "/root/reference/plugins/retime-engine"
HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    with open(os.path.join(REF, "tests/fixtures/f2_cases.hex")) as fh:
        blob = bytes.fromhex("".join(fh.read().split()))
    assert blob[:4] == b"F2C1"
    with gzip.open(os.path.join(HERE, "f2_cases.bin.gz"), "wb", compresslevel=9) as fh:
        fh.write(blob)


if __name__ == "__main__":
    main()
