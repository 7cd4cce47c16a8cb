"""Host logic of the request loop (js/serve.js) that needs no GPU: the request stream and the Reynolds-sweep summary.

- Without sweeps the request stream is the ladder stream of the serving configuration: every request is a 3-rung
  ladder from LADDERS at one Re in [707, 4000] and 200..1000 steps; asking for sweeps does not change the Poisson
  arrival of request 0.
- A sweep request is one grid of SWEEP_GRIDS at three consecutive Re of the sqrt(2) series 707 .. 4000.
- sweepSummary's slopes are the least-squares slopes of log peak against log Re: exact for a power law (peak = c Re^p
  gives p), and its verdict is the worst member health grade.
"""
import json
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SERVE = os.path.join(ROOT, "build-me-anything.github.io_amd", "js", "serve.js")
node = shutil.which("node")
pytestmark = pytest.mark.skipif(node is None, reason="node is not installed")


def run_js(body):
    # serve.js loads the addon, which needs the built library (no device call is made here)
    if not os.path.exists(os.path.join(ROOT, "build-me-anything.github.io_amd", "lib", "nslab_napi.node")):
        pytest.skip("addon not built")
    src = "const SV = require(%s);\n%s" % (json.dumps(SERVE), body)
    r = subprocess.run([node, "-e", src], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return json.loads(r.stdout.strip().splitlines()[-1])


def test_request_streams():
    out = run_js("""
const a = SV.requestSpec({ requests: 40, seed: 3 }), b = SV.requestSpec({ requests: 40, seed: 3, sweepFraction: 0.5 });
console.log(JSON.stringify({ a, b, ladders: SV.LADDERS, grids: SV.SWEEP_GRIDS, res: SV.SWEEP_RE }));
""")
    ladders = [tuple(x) for x in out["ladders"]]
    for q in out["a"]:
        assert q["kind"] == "ladder" and tuple(q["ladder"]) in ladders
        assert [m["N"] for m in q["members"]] == q["ladder"] and all(m["Re"] == q["Re"] for m in q["members"])
        assert 707 <= q["Re"] <= 4000 and 200 <= q["steps"] <= 1000
    assert out["a"][0]["arrival"] == out["b"][0]["arrival"]
    sweeps = [q for q in out["b"] if q["kind"] == "sweep"]
    assert 8 <= len(sweeps) <= 32
    res = out["res"]
    assert all(abs(res[i + 1] / res[i] - 2 ** 0.5) < 0.01 for i in range(len(res) - 1))
    for q in sweeps:
        k = res.index(q["Res"][0])
        assert q["Res"] == res[k:k + 3] and q["N"] in out["grids"]
        assert [m["N"] for m in q["members"]] == [q["N"]] * 3 and [m["Re"] for m in q["members"]] == q["Res"]


def test_sweep_summary_power_law():
    out = run_js("""
const lv = [707, 1000, 1414].map((Re, i) => ({ label: '64³', Re, peak: { omMax: 2 * Math.pow(Re, 0.5) },
  epsPeak: { v: 0.1 * Math.pow(Re, -0.25) }, Eend: 0.4, health: ['PASS', 'WARN', 'PASS'][i] }));
console.log(JSON.stringify(SV.sweepSummary(lv)));
""")
    assert abs(out["slope_omMax"] - 0.5) < 1e-12 and abs(out["slope_eps"] + 0.25) < 1e-12
    assert out["verdict"] == "WARN" and out["Re"] == [707, 1000, 1414]


def test_tuner_table_keeps_the_sizes_not_tuned(tmp_path):
    """js/tune.js rewrites only the rows of the sizes it tuned: a run over a subset of sizes keeps every other row of
    the table as it was, and the rows come out in order of N."""
    table = os.path.join(ROOT, "build-me-anything.github.io_amd", "csrc", "nsl_tune_table.inc")
    f = tmp_path / "table.inc"
    shutil.copy(table, f)
    before = [l for l in open(table) if l.startswith("{")]
    tune = os.path.join(ROOT, "build-me-anything.github.io_amd", "js", "tune.js")
    res = [{"N": 192, "pick": [4, 0, 8, 1, 0, 2, 0], "pick_what": "staged wave 8/CU (test)"},
           {"N": 100, "pick": [2, 0, 0, 1], "pick_what": "rows (test)"}]
    run_js("require(%s).writeTable(%s, %s); console.log('{}');" % (json.dumps(tune), json.dumps(str(f)), json.dumps(res)))
    after = [l for l in open(f) if l.startswith("{")]
    assert len(after) == len(before) + 1
    Ns = [int(l[1:].split(",")[0]) for l in after]
    assert Ns == sorted(Ns)
    for l in before:
        N = int(l[1:].split(",")[0])
        if N != 192:
            assert l in after, l
    assert "{192, 4, 0, 8, 1, 0, 2, 0},   // staged wave 8/CU (test)\n" in after
    assert "{100, 2, 0, 0, 1, 0, 0, 0},   // rows (test)\n" in after
