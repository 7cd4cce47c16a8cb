"""Writes tests/golden/formats.json: the key schema of the reference's long-run outputs, as data.

Sources (read where the reference checkout is present; the fixture is committed):
  * final_cpu: research/nslab/tgv-Re1600-N192/final.json, written by pocket-wind-tunnel/test/run-ns-long.js:33-35;
  * final_gpu: research/nslab/tubes-Re2000-N256-gpu/final.json, written by pocket-wind-tunnel/gpu/nslab_gpu.py:529-533;
  * their run.log first lines (the log line format, run-ns-long.js:11 and nslab_gpu.py's log());
  * checkpoint_cpu / checkpoint_gpu: the keys run-ns-long.js:46 and nslab_gpu.py:541 write into checkpoint.json (no
    checkpoint is archived, so these two lists are transcribed from the writers);
  * analyse_reads: the fields research/nslab/analyse.js:9-34 reads from a final.json;
  * fp32_grading: what the GPU runner's FP32 mode changes in a health report, read from its source with Python's
    ast module (parsed, never run): the machine epsilons of the two precisions (pocket-wind-tunnel/gpu/
    nslab_gpu.py:64), the suffix of the two roundoff-limited row labels (:423) and the text appended to the note
    (:455-459).
Only key names, types and those output strings are kept.

Usage: python tests/golden/make_format_fixture.py /root/reference
"""
import ast
import json
import os
import sys


def keys(d):
    return sorted(d.keys())


def schema(path):
    with open(path) as f:
        d = json.load(f)
    return {"top": keys(d), "case": keys(d["case"]), "snapshot": keys(d["snapshots"][0]), "health": keys(d["health"]),
            "series": keys(d["series"]), "types": {k: type(v).__name__ for k, v in d.items()}}


def fp32_grading(path):
    """The GPU runner's FP32 grading strings and epsilons, from the syntax tree of its source."""
    with open(path) as f:
        tree = ast.parse(f.read())
    out = {}
    for node in ast.walk(tree):
        if isinstance(node, ast.Assign) and isinstance(node.value, ast.IfExp):
            for tgt in node.targets:
                name = tgt.id if isinstance(tgt, ast.Name) else tgt.attr if isinstance(tgt, ast.Attribute) else None
                if name == "eps":
                    out["eps32"], out["eps64"] = node.value.body.value, node.value.orelse.value
                if name == "rndNote":
                    out["label_suffix"] = node.value.body.value
        if isinstance(node, ast.AugAssign) and isinstance(node.target, ast.Name) and node.target.id == "note":
            out["note_fp32"] = node.value.value
    assert set(out) == {"eps32", "eps64", "label_suffix", "note_fp32"}, out
    out["precision"] = {"fp32": "float32", "fp64": "float64"}
    return out


def main(ref):
    nslab = os.path.join(ref, "research", "nslab")
    out = {
        "final_cpu": schema(os.path.join(nslab, "tgv-Re1600-N192", "final.json")),
        "final_gpu": schema(os.path.join(nslab, "tubes-Re2000-N256-gpu", "final.json")),
        "log_cpu": open(os.path.join(nslab, "tgv-Re1600-N192", "run.log")).readline().rstrip("\n"),
        "log_gpu": open(os.path.join(nslab, "tubes-Re2000-N256-gpu", "run.log")).readline().rstrip("\n"),
        "checkpoint_cpu": sorted(["N", "Re", "ic", "t", "step", "series", "outs", "nextSnap", "nextCkpt", "maxEbal",
                                  "maxZbal", "maxTnl", "divMax", "E0", "Z0", "elapsed"]),
        "checkpoint_gpu": sorted(["N", "Re", "ic", "t", "step", "series", "outs", "peakTrack", "peakGrid", "nextSnap",
                                  "nextCkpt", "maxEbal", "maxZbal", "maxTnl", "divMax", "E0", "Z0", "elapsed"]),
        "analyse_reads": {"top": ["build", "case", "elapsed_s", "final", "health", "instrument", "series", "snapshots",
                                  "steps"],
                          "case": ["N", "Re", "ic", "icParams", "nu", "tEnd"], "health": ["worst", "worstEnd"],
                          "series": ["E", "eps", "omMax", "t"], "build": ["version"]},
        "fp32_grading": fp32_grading(os.path.join(ref, "pocket-wind-tunnel", "gpu", "nslab_gpu.py")),
    }
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "formats.json")
    with open(dst, "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", dst)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
