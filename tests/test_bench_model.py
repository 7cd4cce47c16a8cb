"""bench.py's traffic and work model, and the committed rocprofv3 summaries it reads (CPU only, no GPU calls).

- B_alg(N) = 1152 h (b^2 + b N + N^2) bytes per RK4 step prices every update at 15 D complex values (D retained
  modes per component); the per-launch bytes of the fused kernels (each pass reads and writes every operand once,
  16 B per complex) sum to B_alg less 9 D: the stage-1 update reads no T or ACC (9 D instead of 15 D) and the stage-4
  update writes only S (12 D). The unfused form moves the nonlinear term through HBM, 6 D more per rhs. The per-N
  values are the quoted ones (13.73 GB at 256^3, 108.9 GB at 512^3).
- The measured HBM bytes per step (FETCH_SIZE + WRITE_SIZE, calibrated) are within a few per cent of B_alg: every
  pass moves its operands once.
- The FP64 flop counts are positive for every kernel class a step launches, the flop per step is four rhs of them,
  and the x pass does the most arithmetic.
"""
import json
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402

PER_STEP = {"k_zinv": 4, "k_yinv": 4, "k_xnl": 4, "k_yfwd": 4, "k_update<1>": 1, "k_update<2|3>": 2,
            "k_update<4>": 1}


@pytest.mark.parametrize("N", [32, 48, 64, 96, 128, 192, 256, 384, 512, 1024])
def test_pass_bytes_sum_to_b_alg_less_the_stage_savings(N):
    per = bench.alg_bytes_per_launch(N, fused=True)
    step = sum(n * per[k] for k, n in PER_STEP.items())
    h, b, D = bench.geometry(N)
    assert abs(step - (bench.b_alg_step(N) - 9 * D * 16)) <= 1e-9 * step
    # the unfused form moves the nonlinear term through HBM: 6 D complex values more per rhs
    unf = bench.alg_bytes_per_launch(N, fused=False)
    step_unf = sum(n * unf[k] for k, n in PER_STEP.items()) + 4 * unf["k_zfwd"]
    assert abs(step_unf - step - 4 * 6 * D * 16) <= 1e-9 * step


def test_b_alg_values():
    assert abs(bench.b_alg_step(256) / 1e9 - 13.73) < 0.01
    assert abs(bench.b_alg_step(512) / 1e9 - 108.9) < 0.1
    assert abs(bench.b_alg_step(1024) / 1e9 - 872.5) < 0.5


@pytest.mark.parametrize("N", [256, 512])
def test_measured_traffic_matches_the_model(N):
    hb = bench.hbm_bytes_per_step(N)
    assert hb is not None, "profiles/pmc_traffic.json lacks %d" % N
    assert 0.97 < hb / bench.b_alg_step(N) < 1.06, hb / bench.b_alg_step(N)


@pytest.mark.parametrize("N", [256, 512])
def test_fp64_work_from_the_profile(N):
    d = json.load(open(os.path.join(ROOT, "profiles", "pmc_valu.json")))
    assert d["peak_tflops"] == 78.6
    c = d[str(N)]
    assert all(c[k]["flop_per_launch"] > 0 and 0 < c[k]["valu_busy_pct"] < 100 for k in PER_STEP)
    assert max(PER_STEP, key=lambda k: c[k]["flop_per_launch"]) == "k_xnl"
    # fp64_vector over a synthetic run: flop per step is four rhs of the per-launch counts
    kern = {k: {"avg_us": 100.0} for k in PER_STEP}
    r = bench.fp64_vector({"N": N, "zfwd_fused": True, "kernels": kern, "steps_per_s": 10.0})
    flop = sum(n * c[k]["flop_per_launch"] for k, n in PER_STEP.items())
    assert abs(r["flop_per_step"] - flop) <= 1e-9 * flop
    assert abs(r["achieved"] - flop * 10.0 / 1e12) <= 1e-9 * r["achieved"] and r["unit"] == "TFLOP/s"
    assert bench.fp64_vector({"N": N, "zfwd_fused": False, "kernels": kern, "steps_per_s": 10.0}) is None
