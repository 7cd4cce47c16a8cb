"""The Node host module (build-me-anything.github.io_amd/js/nslab.js) over the N-API addon."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SCRIPT = os.path.join(ROOT, "tests", "js", "validate_nslab.js")
MODULE = os.path.join(ROOT, "build-me-anything.github.io_amd", "js", "nslab.js")

node = shutil.which("node")
needs_node = pytest.mark.skipif(node is None, reason="node is not installed")


@needs_node
def test_module_loads_and_exports_reference_surface():
    """No GPU needed: the addon loads, the export list is the reference's, factorize throws its message."""
    js = ("const NS = require(%r);"
          "const keys = ['VERSION','IC_INFO','createSolver','studySummary','Spectral','fftPlan','fft1','factorize'];"
          "for (const k of keys) if (!(k in NS)) throw new Error('missing ' + k);"
          "if (JSON.stringify(NS.factorize(96)) !== '[4,4,2,3]') throw new Error('factorize');"
          "let m = ''; try { NS.factorize(7); } catch (e) { m = e.message; }"
          "if (m !== 'grid size must be 2^a·3^b ≥ 4 (got 7)') throw new Error('message ' + m);"
          "const s = NS.studySummary([{label:'a',N:16,peak:{omMax:1},epsPeak:{v:1,t:1},Eend:1},"
          "{label:'b',N:32,peak:{omMax:1.001},epsPeak:{v:1.001,t:1},Eend:1}], 'grid');"
          "if (s.verdict !== 'PASS') throw new Error('verdict');"
          "console.log('OK');") % MODULE
    r = subprocess.run([node, "-e", js], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and "OK" in r.stdout, r.stdout + r.stderr


@needs_node
def test_report_module_reproduces_reference_reports():
    """No GPU needed: js/report.js rebuilds health reports, verdicts and a ladder summary the reference wrote, and the
    host's pile-up rule reproduces the archived FP64 values (tests/js/validate_report.js)."""
    r = subprocess.run([node, os.path.join(ROOT, "tests", "js", "validate_report.js")], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0 and "ALL PASS" in r.stdout, r.stdout + r.stderr


@needs_node
@pytest.mark.gpu
@pytest.mark.parametrize("group", ["fft", "exact", "tgv", "misc", "config1"])
def test_validate_nslab_on_gpu(group):
    r = subprocess.run([node, SCRIPT, group], capture_output=True, text=True, timeout=600)
    print(r.stdout)
    assert r.returncode == 0 and "ALL PASS" in r.stdout, r.stdout + r.stderr


@needs_node
@pytest.mark.gpu
def test_pools_and_request_loop_on_gpu():
    script = os.path.join(ROOT, "tests", "js", "validate_pools.js")
    r = subprocess.run([node, script], capture_output=True, text=True, timeout=600)
    print(r.stdout)
    assert r.returncode == 0 and "ALL PASS" in r.stdout, r.stdout + r.stderr


@needs_node
@pytest.mark.gpu
def test_long_run_harness_on_gpu():
    script = os.path.join(ROOT, "tests", "js", "validate_run_long.js")
    r = subprocess.run([node, script], capture_output=True, text=True, timeout=600)
    print(r.stdout)
    assert r.returncode == 0 and "ALL PASS" in r.stdout, r.stdout + r.stderr


@needs_node
def test_parity_gate_grades_series():
    """No GPU needed: the gate of cloud/parity-check.py over synthetic final.json pairs."""
    gate = os.path.join(ROOT, "build-me-anything.github.io_amd", "js", "parity_check.js")
    js = ("const {parityCheck} = require(%r);"
          "const mk = (f, dev) => ({case: {N: 64, Re: 1600, ic: 'tgv', cfl: 0.4}, precision: 'float64', device: dev,"
          " series: Object.fromEntries(['t','E','Z','eps','omMax','Pspec'].map(k => [k, Array.from({length: 50},"
          " (_, i) => (k === 't' ? 0.01 * i : 1 + 0.01 * i) * (k === 't' ? 1 : f(i)))]))});"
          "const q = () => {}; const a = mk(() => 1, 'x');"
          "const r1 = parityCheck(mk(() => 1, 'y'), a, q).verdict, r2 = parityCheck(mk(i => 1 + 1e-8 * i / 50, 'y'), a, q).verdict,"
          " r3 = parityCheck(mk(i => 1 + 1e-4, 'y'), a, q).verdict;"
          "const c = mk(() => 1, 'y'); c.case.N = 32; const r4 = parityCheck(c, a, q).verdict;"
          "console.log([r1, r2, r3, r4].join(' '));") % gate
    r = subprocess.run([node, "-e", js], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and r.stdout.strip() == "PASS WARN FAIL FAIL", r.stdout + r.stderr


@needs_node
@pytest.mark.gpu
def test_autotuner_runs_and_checks_every_form(tmp_path):
    """js/tune.js times every compiled form at 32^3 and 256^3 (briefly), alone and in a pool, checks each against the
    current form's energy after 3 steps, and reports a pick among the candidates; --dry leaves the compiled-in table
    alone."""
    import json
    script = os.path.join(ROOT, "build-me-anything.github.io_amd", "js", "tune.js")
    rep = str(tmp_path / "tune.json")
    r = subprocess.run([node, script, "--sizes", "32,256", "--budget", "0.05", "--dry", "--report", rep],
                       capture_output=True, text=True, timeout=600)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    import capi
    res = json.load(open(rep))["results"]
    assert [x["N"] for x in res] == [32, 256]
    for x in res:
        N = x["N"]
        forms = [c["form"] for c in x["candidates"]]
        # search 1: every x-pass form and fusion at the default tile and schedules (the current form first)
        base = [list(c) for c in capi.tune_candidates(N) if not any(c[4:])]
        assert len(forms) == len(base) + (1 if any(forms[0][4:]) else 0)
        assert all(f in forms for f in base)
        # search 2: the tile and schedule variants next to the pick (32: tile, both schedules; 256: schedules)
        vforms = [c["form"] for c in x["variants"]]
        assert vforms and all(sum(1 for i in (4, 5, 6) if f[i] != r[i]) == 1 for f, r in
                              ((c["form"], next(p for p in forms + vforms if p[:4] == c["form"][:4] and
                                                sum(1 for i in (4, 5, 6) if p[i] != c["form"][i]) == 1)) for c in x["variants"]))
        assert x["pick"] in forms + vforms
        rows = x["candidates"] + x["variants"]
        assert all(c["rel_diff_E"] < 1e-11 and c["steps_per_s"] > 0 for c in rows)
        # both sizes are timed in pools too (up to --pool-max 256), and the pick has the best score within 2 %
        assert all(c["pool_solve_steps_per_s"] > 0 and c["pool_slots"] >= 2 for c in rows)
        best = max(c["score"] for c in x["candidates"])
        assert x["gain"] * 1.02 >= best
