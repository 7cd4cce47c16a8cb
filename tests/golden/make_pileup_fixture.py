"""Writes tests/golden/pileup_fp64.json: (N, kc, spectrum, pileUp) of archived FP64 reference runs.

The reference's cutoff pile-up (pocket-wind-tunnel/src/nslab.js:557-564; the GPU runner's identical rule,
pocket-wind-tunnel/gpu/nslab_gpu.py:357-372) depends only on the shell spectrum and kc = floor(N/3). The FP64 GPU
runs of the antiparallel-tube study archive both beside the value the reference computed from them
(research/nslab/tubes-Re2000-N{288,320,384}-gpu/final.json, `snapshots[]`). Every snapshot whose value is neither
1 (the band maximum at k8 itself) nor absent is kept, with the null (below-floor) snapshots and a few value-1 ones,
so a fixture of about 20 spectra tells apart a wrong k8, a wrong band end and a wrong floor.

Also writes tests/golden/pileup_fp32.json from the FP32 640^3 run (research/nslab/expl-tubes-Re2000-N640-fp32-gpu,
runner 0.1.2, whose guard was a fixed 1e-20 of the peak): per snapshot t, the archived pileUp, the peak E(k) and the
band E(k8 .. kc). The runner's 0.1.3 notes (pocket-wind-tunnel/gpu/README.md:44-58) state what the precision-scaled
guard (100 eps^2 of the working precision) does to exactly this run: ungraded for t <= 5.0, graded from t = 5.5 on.
The spectra are kept as peak + band, the only bins the rule reads (the peak sits at k = 1 in every snapshot).

Usage (only where the reference checkout is present; the fixtures it writes are committed):
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

    run = "expl-tubes-Re2000-N640-fp32-gpu"
    with open(os.path.join(root, run, "final.json")) as f:
        d = json.load(f)
    assert d["precision"] == "float32", run
    N = int(d["case"]["N"])
    kc = N // 3
    k8 = int(kc * 0.8 + 0.5)
    snaps = []
    for s in d["snapshots"]:
        sp = s["spectrum"]
        peak = max(sp[1:])
        assert sp.index(peak) == 1 and max(sp[k8:kc + 1]) < peak, s["t"]
        snaps.append({"t": s["t"], "pileUp": s["pileUp"], "nb": len(sp), "peak": peak, "band": sp[k8:kc + 1]})
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "pileup_fp32.json")
    with open(dst, "w") as f:
        json.dump({"source": "research/nslab/%s/final.json snapshots (FP32 GPU runner 0.1.2)" % run, "N": N, "kc": kc,
                   "k8": k8, "snapshots": snaps}, f)
    print("wrote", dst, len(snaps), "snapshots")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/research/nslab")
