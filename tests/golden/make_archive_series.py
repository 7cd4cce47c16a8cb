"""Writes tests/golden/archive_series.json: the opening steps of archived tube runs of the reference's GPU runner.

Sources (research/nslab/<run>/final.json of the reference checkout, written by
pocket-wind-tunnel/gpu/nslab_gpu.py:529-533): the tube-pair IC with the default parameters (nslab.js:262-272), CFL
0.4, adaptive dt, at the grids the HIP library has size-specialised kernels for:
  * FP32 runs (the runner's --fp32 mode, complex64 throughout): expl-tubes-Re707-N128-fp32-gpu,
    expl-tubes-Re1000-N192-fp32-gpu, expl-tubes-Re707-N256-fp32-gpu;
  * FP64 runs: tubes-Re2000-N256-gpu, tubes-Re1000-N384-gpu.
Kept per run: the case, the first STEPS + 1 entries of every series (t, E, Z, eps, omMax, uMax, dt, ebal, zbal,
Pspec, Pal), the first snapshot's scalars, and the divergence and nonlinear-transfer rows of the archived health
report (the roundoff levels of the run's precision). Data only: numbers and strings read with json.

Usage (only where the reference checkout is present; the fixture it writes is committed):
    python tests/golden/make_archive_series.py /root/reference/research/nslab
"""
import json
import os
import sys

RUNS = ("expl-tubes-Re707-N128-fp32-gpu", "expl-tubes-Re1000-N192-fp32-gpu", "expl-tubes-Re707-N256-fp32-gpu",
        "tubes-Re2000-N256-gpu", "tubes-Re1000-N384-gpu")
STEPS = 40
SNAP_KEYS = ("t", "step", "E", "Z", "P", "eps", "omMax", "uMax", "Pspec", "Tnl", "divMax", "kmaxEta", "pileUp")


def main(root):
    out = []
    for run in RUNS:
        with open(os.path.join(root, run, "final.json")) as f:
            d = json.load(f)
        s0 = d["snapshots"][0]
        rows = [r for r in d["health"]["rows"] if r[0].startswith("divergence") or r[0].startswith("nonlinear")]
        out.append({"run": run, "precision": d["precision"], "instrument": d["instrument"], "case": d["case"],
                    "series": {k: v[:STEPS + 1] for k, v in d["series"].items()},
                    "snapshot0": {k: s0.get(k) for k in SNAP_KEYS}, "roundoff_rows": rows})
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "archive_series.json")
    with open(dst, "w") as f:
        json.dump({"steps": STEPS, "runs": out}, f)
    print("wrote", dst, [r["run"] for r in out])


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/research/nslab")
