"""Copies the reference's vector-coding model fixture into tests/golden. Data only.

plugins/retime-engine/tests/fixtures/vectors_cases.hex (VEC1: 12 encode and 9 decode cases of
a numpy model over the three depths, read by vectors/tests.rs:340-393, bit-exact there) becomes
vectors_cases.bin.gz: the same bytes, decoded from hex and gzip-compressed.

Run in a container that has the reference checked out at /root/reference."""
import gzip
import os

REF = "/root/reference/plugins/retime-engine"
HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    with open(os.path.join(REF, "tests/fixtures/vectors_cases.hex")) as fh:
        blob = bytes.fromhex("".join(fh.read().split()))
    assert blob[:4] == b"VEC1"
    with gzip.open(os.path.join(HERE, "vectors_cases.bin.gz"), "wb", compresslevel=9) as fh:
        fh.write(blob)


if __name__ == "__main__":
    main()
