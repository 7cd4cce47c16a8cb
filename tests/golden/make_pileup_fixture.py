"""Writes tests/golden/pileup_fp64.json: (N, kc, spectrum, pileUp) of archived FP64 reference runs.

The reference's cutoff pile-up (pocket-wind-tunnel/src/nslab.js:557-564; the GPU runner's identical rule,
pocket-wind-tunnel/gpu/nslab_gpu.py:357-372) depends only on the shell spectrum and kc = floor(N/3). The FP64 GPU
runs of the antiparallel-tube study archive both beside the value the reference computed from them
(research/nslab/tubes-Re2000-N{288,320,384}-gpu/final.json, `snapshots[]`). Every snapshot whose value is neither
1 (the band maximum at k8 itself) nor absent is kept, with the null (below-floor) snapshots and a few value-1 ones,
so a fixture of about 20 spectra tells apart a wrong k8, a wrong band end and a wrong floor.

Usage (only where the reference checkout is present; the fixture it writes is committed):
    python tests/golden/make_pileup_fixture.py /root/reference/research/nslab
"""
import json
import os
import sys

RUNS = ("tubes-Re2000-N288-gpu", "tubes-Re2000-N320-gpu", "tubes-Re2000-N384-gpu")


def main(root):
    out = []
    for run in RUNS:
        with open(os.path.join(root, run, "final.json")) as f:
            d = json.load(f)
        assert d["precision"] == "float64", run
        N = int(d["case"]["N"])
        ones = 0
        for s in d["snapshots"]:
            if "pileUp" not in s:
                continue
            v = s["pileUp"]
            keep = v is None or v != 1.0 or ones < 2
            if v == 1.0:
                ones += 1
            if keep:
                out.append({"run": run, "N": N, "kc": N // 3, "t": s["t"], "spectrum": s["spectrum"], "pileUp": v})
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "pileup_fp64.json")
    with open(dst, "w") as f:
        json.dump({"source": "research/nslab/<run>/final.json snapshots (FP64 GPU runner)", "cases": out}, f)
    print("wrote", dst, len(out), "spectra;", sum(1 for c in out if c["pileUp"] not in (None, 1.0)), "non-trivial")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/research/nslab")
