"""Writes tests/golden/vectors_in_cases.json from the reference's Vectors In parity data files.

Reads (as data, nothing is imported or executed) the case table of the reference's vectors-in
parity grid (parity/vectors-in/testsets/grid/cases.json and case_meta.json), the reference CPU
path's recorded output per case (parity/vectors-in/results/base.grid.tsv: exit code and SHA-256 of
`vfx apply` on the case) and the grade of each case against the original plugin
(parity/vectors-in/results/grade.tsv). Each case keeps its id, group, size, depth, rate, frame
count, parameters, input and layer file names (the inputs are regenerated by vectors_in_inputs.py),
the recorded exit code and hash, and the grade.

Run in a container that has the reference checked out at /root/reference:
    python tests/golden/make_vectors_in_fixture.py
"""
import json
import os

REF = "/root/reference/parity/vectors-in"
HERE = os.path.dirname(os.path.abspath(__file__))


def _tsv(path):
    with open(path) as fh:
        head = fh.readline().rstrip("\n").split("\t")
        return {row[0]: dict(zip(head, row)) for row in (line.rstrip("\n").split("\t") for line in fh) if row[0]}


def main():
    with open(os.path.join(REF, "testsets/grid/cases.json")) as fh:
        cases = json.load(fh)["cases"]
    with open(os.path.join(REF, "testsets/grid/case_meta.json")) as fh:
        meta = json.load(fh)
    base = _tsv(os.path.join(REF, "results/base.grid.tsv"))
    grade = _tsv(os.path.join(REF, "results/grade.tsv"))
    out = []
    for c in cases:
        b = base.get(c["id"], {})
        out.append({
            "id": c["id"], "group": meta[c["id"]]["group"], "width": c["width"], "height": c["height"],
            "depth": c["depth"], "fps": c["fps"], "frames": c["frames"], "params": c["params"], "input": c["input"],
            "layers": c["layers"], "flags": meta[c["id"]]["flags"], "stand_in": meta[c["id"]]["stand_in"],
            "exit": int(b["exit"]) if b.get("exit", "") != "" else None, "sha256": b.get("sha256") or None,
            "grade": grade.get(c["id"], {}).get("status"),
        })
    with open(os.path.join(HERE, "vectors_in_cases.json"), "w") as fh:
        json.dump(out, fh, indent=0)
    print(f"{len(out)} cases")


if __name__ == "__main__":
    main()
