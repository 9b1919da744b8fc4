"""Writes tests/golden/retime_flow_cases.json from the reference's parity data files.

Reads (as data, nothing is imported or executed) the case tables of the reference's retime-flow
parity grid (parity/retime-flow/cases/{grid,draft}/cases.json) and the SHA-256 of the ORIGINAL
plugin's output per case (parity/retime-time/results/supp/refs_sha256.txt, `refs/flow-all/...`
lines), and keeps every case with a hash. Each case gets a `scope` label: "core" for the cases the
frame-synthesis path serves today (warp code 0, no 360 video, no prep, no blur compensation, no
special float values, no Weighted Blend), otherwise the reason it is out of scope.

Run in a container that has the reference checked out at /root/reference:
    python tests/golden/make_retime_flow_fixture.py
"""
import json
import os

REF = "/root/reference/parity"
HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    hashes = {}
    with open(os.path.join(REF, "retime-time/results/supp/refs_sha256.txt")) as fh:
        for line in fh:
            h, path = line.split()
            if path.startswith("refs/flow-all/"):
                hashes[os.path.basename(path)] = h
    out = []
    for setn in ("grid", "draft"):
        with open(os.path.join(REF, "retime-flow/cases", setn, "cases.json")) as fh:
            doc = json.load(fh)
        for c in doc["cases"]:
            ref = c["reference"]
            if ref not in hashes:
                continue
            p = dict(x.split("=", 1) for x in c["params"])
            why = []
            if p.get("smart_blend", "0") != "0":
                why.append("smart_blend")
            if p.get("video_360", "0") != "0":
                why.append("video_360")
            if float(p.get("blur_compensation", "0")) != 0:
                why.append("blur")
            if p.get("prep", "1") != "1":
                why.append("prep")
            if c["source"].startswith("spec"):
                why.append("special_values")
            if p.get("frame_interp", "2") == "3":
                why.append("weighted_blend")
            out.append(dict(id=c["id"], params=c["params"], width=c["width"], height=c["height"], depth=c["depth"],
                            source=c["source"], frames=c["frames"], fps=c["fps"], quality=c.get("quality", "high"),
                            sha256=hashes[ref], scope="core" if not why else "+".join(why)))
    with open(os.path.join(HERE, "retime_flow_cases.json"), "w") as fh:
        json.dump(out, fh, indent=0)
        fh.write("\n")
    print(len(out), "cases,", sum(1 for c in out if c["scope"] == "core"), "core")


if __name__ == "__main__":
    main()
