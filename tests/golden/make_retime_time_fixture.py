"""Extracts the in-scope cases of the reference's retime-time parity grid with the SHA-256 of the
ORIGINAL plugin's output. Data only.

Sources: parity/retime-time/results/cases.json (case id, parameters, size, depth, input source,
frame count, rate) and parity/retime-time/results/supp/refs_sha256.txt (the hashes, under
refs/time-all/ and refs/l6000/). In scope: the cases whose parameters this backend serves (Speed
mode with constant or keyed speed, Nearest / Blend / Weighted Blend, In FPS is Out FPS and the
input frame rate); out of scope: frame-number mode, input fields, source marks and keys with
Bezier handles. Also records, per case, the hash of the reference CPU path's own output of the
same case (results/base.tsv), which the original plugin's differs from in a few cases (NaN payload
bits, the 6000-frame rate cases). Writes retime_time_cases.json.

Run in a container that has the reference checked out at /root/reference."""
import json
import os

REF = "/root/reference/parity/retime-time/results"
HERE = os.path.dirname(os.path.abspath(__file__))
OUT_OF_SCOPE = ("time_remap_mode", "frame_num", "input_fields", "src_mark_segments")


def served(params):
    for p in params:
        k, v = p.split("=", 1)
        if k in OUT_OF_SCOPE:
            return False
        if k == "speed_percent" and "@" in v and any(ch in v for ch in "io/"):
            return False
    return True


def main():
    with open(os.path.join(REF, "cases.json")) as fh:
        cases = json.load(fh)["cases"]
    # the reference CPU path's own recorded output (the rewrite's hashes of the same cases)
    cpu = {}
    with open(os.path.join(REF, "base.tsv")) as fh:
        for line in fh:
            p = line.rstrip("\n").split("\t")
            if len(p) >= 4 and p[3]:
                cpu[p[0]] = p[3]
    refs = {}
    with open(os.path.join(REF, "supp", "refs_sha256.txt")) as fh:
        for line in fh:
            digest, path = line.split()
            refs[os.path.basename(path).rsplit(".", 1)[0]] = digest
    out = []
    for c in cases:
        if c["id"] not in refs or not served(c["params"]):
            continue
        out.append({"id": c["id"], "params": [p for p in c["params"] if not p.startswith("vectors=")],
                    "vectors": 1, "width": c["width"], "height": c["height"], "depth": c["depth"],
                    "source": c["source"], "frames": c["frames"], "fps": c["fps"], "sha256": refs[c["id"]],
                    "cpu_path_sha256": cpu.get(c["id"], "")})
    with open(os.path.join(HERE, "retime_time_cases.json"), "w") as fh:
        json.dump(out, fh, indent=0)
    print(len(out), "cases")


if __name__ == "__main__":
    main()
