#!/usr/bin/env python
"""Benchmark: FP64 RK4 steps/s of the NSLab pseudo-spectral Navier-Stokes solver on one MI355X.

Workload (the headline configuration): Taylor-Green vortex on [0, 2pi]^3 at 256^3, nu = 1/1600, fixed dt = 0.004,
FP64, 2/3 dealiasing, rotational form, classical RK4. A "step" is one RK4 step (4 right-hand sides). The state is
resident in HBM when the timed region starts; the timed region is K steps between two device synchronisations, with
nothing but the solver's kernels in the stream. The same K steps then run again in a second timed region with a HIP
event pair around every launch, for the per-kernel averages (the events add gaps between kernels, not kernel time).

Printed: one JSON line with the throughput, the roofline of the dominant kernel (algorithmic bytes per launch over
its mean launch time, from HIP events recorded on the solver's stream inside the second timed region), the whole-step
roofline fraction of the truncated-storage traffic model, a 512^3 line, FP32-storage lines at the headline size
and 512^3 (the GPU runner's fp32 mode: complex64 in HBM, single-precision transforms), the CPU baseline (the oracle, a
single-threaded C restatement of the reference algorithm, timed on this host on a bounded sample), and three lines
from the Node host (js/serve.js): the batched ensemble (128 solves of uneven N in slot pools, one hipGraph per step),
the ladder request loop (Poisson arrivals, admission by HBM footprint, p50/p99 latency, ladders per hour, against
sequential single-solve execution of the same work), and the same loop with half the requests Reynolds sweeps.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--N 256] [--no-512] [--no-cpu] [--no-ensemble]
                  [--no-serve] [--no-fp32] [--serve-requests R]
"""
import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

HBM_PEAK_GBS = 8000.0
CLASSES = ["k_zinv", "k_yinv", "k_xnl", "k_yfwd", "k_zfwd", "k_update<1>", "k_update<2|3>", "k_update<4>", "k_finalize"]


def geometry(N):
    kc = N // 3
    h, b = kc + 1, 2 * kc + 1
    return h, b, b * b * h


def alg_bytes_per_launch(N, fused=False, c=16.0):
    """Algorithmic HBM bytes per launch of each kernel class: every operand read once and written once
    (c bytes per stored complex value: 16 for FP64 solves, 8 for FP32 solves). h = floor(N/3)+1 retained kx,
    b = 2 floor(N/3)+1 retained ky/kz, D = b*b*h. fused: the z-forward runs inside the update (one kernel, timed as
    k_update<s>): its 3 pencils come in and the nonlinear term never goes to HBM."""
    h, b, D = geometry(N)
    z = 3 * h * b * N   # z-forward input pencils
    return {
        "k_zinv": (3 * D + 6 * h * b * N) * c,           # 3 velocity fields in, u and omega (6) out
        "k_yinv": (6 * h * b * N + 6 * h * N * N) * c,   # 6 in, 6 out
        "k_xnl": (6 * h * N * N + 3 * h * N * N) * c,    # 6 in, u x omega (3) out
        "k_yfwd": (3 * h * N * N + 3 * h * b * N) * c,   # 3 in, 3 out
        "k_zfwd": (z + 3 * D) * c,                       # 3 in, NL (3) out
        "k_update<1>": (z + 9 * D if fused else 12 * D) * c,       # (NL,) S in; ACC, T out
        "k_update<2|3>": (z + 15 * D if fused else 18 * D) * c,    # (NL,) T, S, ACC in; T, ACC out
        "k_update<4>": (z + 12 * D if fused else 15 * D) * c,      # (NL,) T, S, ACC in; S out
        "k_finalize": 0.0,
    }


def b_alg_step(N):
    """Minimum HBM bytes of one RK4 step for a 5-pass-per-rhs schedule over 2/3-truncated storage:
    B_alg(N) = 1152 h (b^2 + b N + N^2)."""
    h, b, _ = geometry(N)
    return 1152.0 * h * (b * b + b * N + N * N)


def measure(N, steps, warmup, dt, label, fp32=False):
    """fp32: an FP32 solve (complex64 storage, single-precision transforms): 8 bytes per stored complex value."""
    import capi
    import torch
    s = capi.Solver(N=N, Re=1600, ic="tgv", dt=dt, fp32=fp32)
    esz = 8.0 if fp32 else 16.0
    s.run(warmup)
    torch.cuda.synchronize()
    # timed region 1: the throughput, no instrumentation in the stream
    t0 = time.perf_counter()
    done = s.run(steps)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    # timed region 2: the same number of steps with a HIP event pair around every launch (nsl_profile) for the
    # per-kernel averages; the events add gaps between kernels, not kernel time
    L = capi.lib()
    capi.check(L.nsl_profile(s.h, 1))
    torch.cuda.synchronize()
    tp0 = time.perf_counter()
    done_p = s.run(steps)
    torch.cuda.synchronize()
    tp1 = time.perf_counter()
    ms = (ctypes.c_double * len(CLASSES))()
    cnt = (ctypes.c_int64 * len(CLASSES))()
    capi.check(L.nsl_profile_read(s.h, ms, cnt))
    capi.check(L.nsl_profile(s.h, 0))
    assert done == steps and done_p == steps, (done, done_p, steps)
    elapsed = t1 - t0
    fused = cnt[CLASSES.index("k_zfwd")] == 0   # no separate z-forward launches: the fused pass ran
    per = alg_bytes_per_launch(N, fused, esz)
    kern = {}
    for i, name in enumerate(CLASSES):
        if cnt[i]:
            avg_s = ms[i] / cnt[i] / 1e3
            kern[name] = {"launches": int(cnt[i]), "total_ms": ms[i], "avg_us": avg_s * 1e6,
                          "alg_bytes": per[name], "GBs": per[name] / avg_s / 1e9 if per[name] else 0.0}
    dom = max((k for k in kern if per[k] > 0), key=lambda k: kern[k]["total_ms"])
    E = s.last["E"]
    dev_bytes = s.device_bytes()
    s.close()
    del s
    return {"label": label, "N": N, "steps": steps, "elapsed_s": elapsed, "steps_per_s": steps / elapsed,
            "ms_per_step": 1e3 * elapsed / steps, "ms_per_step_profiled": 1e3 * (tp1 - tp0) / steps,
            "kernels": kern, "dominant": dom, "E_end": E, "zfwd_fused": fused, "device_bytes": dev_bytes,
            "step_roofline_frac": b_alg_step(N) * esz / 16.0 * steps / elapsed / (HBM_PEAK_GBS * 1e9)}


def measure_tubes(N, steps):
    """The memory-heavy configuration: antiparallel tubes along z (Gamma = 1, sigma = 0.3, centres pi +- (0.8 +
    0.1 cos z), seeded perturbation rms 1e-4), Re = Gamma/nu = 2000, CFL 0.5; `steps` RK4 steps timed after one."""
    import capi
    import torch
    t0 = time.perf_counter()
    s = capi.Solver(N=N, Re=2000, ic="tubesz", cfl=0.5)
    setup = time.perf_counter() - t0
    s.run(1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    done = s.run(steps)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    ser = s.series
    r = {"workload": "tube pair along z %d^3, Re 2000, CFL 0.5, FP64" % N, "steps": done, "steps_per_s": done / el,
         "step_roofline_frac": b_alg_step(N) * done / el / (HBM_PEAK_GBS * 1e9), "setup_s": setup, "run_s": el,
         "device_GB": s.device_bytes() / 1e9, "t": s.clock["t"], "omMax": s.last["omMax"], "E": s.last["E"],
         "max_ebal": max(ser["ebal"]), "max_zbal": max(ser["zbal"])}
    s.close()
    return r


def cpu_baseline(N, dt, budget_s=20.0):
    """The oracle (single-threaded C restatement of the reference algorithm, 5 rhs per step as nslab.js) timed on
    this host: one untimed set-up, then RK4 steps until about budget_s seconds of CPU work (at most 3)."""
    from oracle import oracle as O
    t0 = time.perf_counter()
    o = O.OracleSolver(N=N, Re=1600, ic="tgv", dt=dt)
    t_setup = time.perf_counter() - t0
    n, t0 = 0, time.perf_counter()
    while n < 3:
        o.run(1)
        n += 1
        if time.perf_counter() - t0 > budget_s:
            break
    el = time.perf_counter() - t0
    del o
    return {"value": n / el, "unit": "steps/s", "cores": 1, "kind": "port",
            "sample": "TGV %d^3 FP64 dt %g: %d RK4 steps timed after an untimed set-up (%.1f s), %.1f s on 1 host "
                      "thread; host %s, %d cores visible" % (N, dt, n, t_setup, el, _cpu_model(), os.cpu_count())}


def _oracle_steps(args):
    """Worker: one oracle solve (random IC, CFL 0.5) of size N, `steps` RK4 steps timed after its set-up."""
    N, steps, seed = args
    from oracle import oracle as O
    o = O.OracleSolver(N=N, Re=1000, ic="random", cfl=0.5, icParams={"seed": seed})
    t0 = time.perf_counter()
    o.run(steps)
    return (time.perf_counter() - t0) / steps


def cpu_ensemble(cores=None, budget_s=3.0):
    """The ensemble mix (64x32^3, 32x48^3, 24x64^3, 8x96^3, 300 steps) on all host cores: one oracle solve per core
    at a time, as independent processes. Per N, `cores` solves run at once for a few steps (memory contention
    included); the mix's time is then sum(count * 300 * t_step(N)) / cores."""
    import multiprocessing as mp
    cores = cores or max(1, min(os.cpu_count() or 1, 16))
    mix = ((32, 64), (48, 32), (64, 24), (96, 8))
    per = {}
    ctx = mp.get_context("spawn")
    with ctx.Pool(cores) as pool:
        for N, _ in mix:
            steps = max(1, int(budget_s / (2e-7 * N ** 3 * 3)))   # about budget_s per worker (oracle ~0.6 us/point)
            t = pool.map(_oracle_steps, [(N, steps, 1 + i) for i in range(cores)])
            per[N] = sum(t) / len(t)
    total = sum(n * 300 * per[N] for N, n in mix) / cores
    return {"solve_steps_per_s": 128 * 300 / total, "mix_s": total, "cores": cores, "kind": "port",
            "sec_per_step": {str(N): per[N] for N in per},
            "sample": "the C oracle, one solve per core (%d processes at once) per N, a few RK4 steps each; the "
                      "128-solve x 300-step mix time extrapolated from those per-step times" % cores}


def cpu_sizes(budget_s=12.0):
    """The oracle on one core at 32^3 (config 1: TGV dt 0.01, 200 steps) and a few steps at 64^3 and 128^3."""
    from oracle import oracle as O
    out = {}
    for N, steps, dt in ((32, 200, 0.01), (64, 10, 0.004), (128, 2, 0.004)):
        o = O.OracleSolver(N=N, Re=1600, ic="tgv", dt=dt)
        t0 = time.perf_counter()
        o.run(steps)
        el = time.perf_counter() - t0
        out[str(N)] = {"steps": steps, "steps_per_s": steps / el}
        del o
        if sum(1.0 / v["steps_per_s"] * v["steps"] for v in out.values()) > budget_s:
            break
    return out


def node_leg(args, timeout):
    """Runs js/serve.js under Node with `args`; returns its JSON object, or {"error": ...}."""
    import shutil
    import subprocess
    node = shutil.which("node")
    if node is None:
        return {"error": "node is not installed"}
    script = os.path.join(ROOT, "build-me-anything.github.io_amd", "js", "serve.js")
    try:
        r = subprocess.run([node, script] + args, capture_output=True, text=True, timeout=timeout)
    except subprocess.TimeoutExpired:
        return {"error": "timed out after %d s" % timeout}
    if r.returncode != 0:
        return {"error": (r.stderr or r.stdout)[-800:]}
    return json.loads(r.stdout.strip().splitlines()[-1])


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def hbm_bytes_per_step(N):
    """rocprofv3-measured HBM bytes of one RK4 step (FETCH_SIZE + WRITE_SIZE, calibrated) from the committed
    summary: 4 launches of each pass per step, update stages 1, 2|3, 4 once, twice, once."""
    p = os.path.join(ROOT, "profiles", "pmc_traffic.json")
    if not os.path.exists(p):
        return None
    try:
        d = json.load(open(p)).get(str(N), {})
    except (ValueError, OSError):
        return None
    per_step = {"k_zinv": 4, "k_yinv": 4, "k_xnl": 4, "k_yfwd": 4, "k_update<1>": 1, "k_update<2|3>": 2,
                "k_update<4>": 1}
    if not all(k in d for k in per_step):
        return None
    return sum(n * d[k]["bytes_per_launch"] for k, n in per_step.items())


def fp64_vector(r):
    """FP64 vector work of a `measure` result: flop per launch of each kernel class from the committed rocprofv3
    SQ-counter summary (64 x (ADD + MUL + 2 FMA + TRANS) FP64 VALU instructions per dispatch), over this run's mean
    launch times (HIP events), and the whole step's flop x steps/s, against the FP64 vector peak. The summary is of
    the fused z-forward form; None if this run used another form or no summary exists for N."""
    p = os.path.join(ROOT, "profiles", "pmc_valu.json")
    if not os.path.exists(p) or not r["zfwd_fused"]:
        return None
    try:
        d = json.load(open(p))
    except (ValueError, OSError):
        return None
    c = d.get(str(r["N"]), {})
    per_step = {"k_zinv": 4, "k_yinv": 4, "k_xnl": 4, "k_yfwd": 4, "k_update<1>": 1, "k_update<2|3>": 2,
                "k_update<4>": 1}
    if not all(k in c and k in r["kernels"] for k in per_step):
        return None
    peak = d["peak_tflops"]
    flop = sum(n * c[k]["flop_per_launch"] for k, n in per_step.items())
    tf = flop * r["steps_per_s"] / 1e12
    ker = {k: {"tflops": round(c[k]["flop_per_launch"] / (r["kernels"][k]["avg_us"] * 1e-6) / 1e12, 2),
               "valu_busy_pct": c[k]["valu_busy_pct"]} for k in per_step}
    return {"flop_per_step": flop, "achieved": tf, "peak": peak, "unit": "TFLOP/s", "frac": tf / peak,
            "kernels": ker}


def copy_ceiling_GBs(mib=2048, reps=10):
    """This GPU's practical HBM ceiling: the library's plain copy kernel over `mib` MiB, bytes read + written per
    second over `reps` copies (HIP events)."""
    import capi
    v = ctypes.c_double(0.0)
    capi.check(capi.lib().nsl_copy_probe(mib << 20, reps, ctypes.byref(v)))
    return v.value


def traffic_from_profile(kernel, N):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 PMC summary, if one exists."""
    p = os.path.join(ROOT, "profiles", "pmc_traffic.json")
    if not os.path.exists(p):
        return None
    try:
        d = json.load(open(p))
        return d.get(str(N), {}).get(kernel, {}).get("bytes_per_launch")
    except (ValueError, OSError):
        return None


def serving_bound(sv, copy_GBs, rate):
    """How far batching can take the served mix: its RK4 steps priced at B_alg per step, run at this GPU's copy
    ceiling and at the 8 TB/s peak, against the measured sequential time; and the service rate against the arrival
    rate (above it the queue, and with it the latency, grows with the number of requests)."""
    spg = {int(N): n for N, n in sv["steps_per_grid"].items()}
    alg = sum(n * b_alg_step(N) for N, n in spg.items())
    t_copy, t_peak = alg / (copy_GBs * 1e9), alg / (HBM_PEAK_GBS * 1e9)
    mu = sv["requests"] / sv["wall_s"]
    return {"steps_per_grid": spg, "b_alg_GB": alg / 1e9,
            "bound_s_at_copy_ceiling": t_copy, "bound_vs_sequential": sv["sequential_estimate_s"] / t_copy,
            "bound_vs_sequential_at_peak": sv["sequential_estimate_s"] / t_peak,
            "served_b_alg_GBs": alg / sv["wall_s"] / 1e9,
            "arrival_rate_per_s": rate, "service_rate_per_s": mu,
            "queue": "stable" if mu > rate else "arrivals outpace service: latency grows with the request count"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--N", type=int, default=256)
    ap.add_argument("--dt", type=float, default=0.004)
    ap.add_argument("--no-512", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-tubes", action="store_true")
    ap.add_argument("--no-ensemble", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-fp32", action="store_true")
    ap.add_argument("--serve-requests", type=int, default=30)
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="launch one device-to-device copy of this many MiB first (PMC counter calibration)")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the solver has no multi-GPU path (each solve runs on one device; "
              "ensembles shard by solve). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)
    import torch
    assert torch.cuda.is_available(), "no GPU visible"
    torch.cuda.set_device(0)
    if a.copy_probe_mib:
        n = a.copy_probe_mib * (1 << 20) // 8
        x = torch.ones(n, dtype=torch.float64, device="cuda")
        y = torch.empty_like(x)
        torch.mul(x, 1.0, out=y)   # one elementwise kernel: reads n*8 bytes, writes n*8 bytes
        torch.cuda.synchronize()
        del x, y
    main_r = measure(a.N, a.steps, a.warmup, a.dt, "headline")
    extra = None
    if not a.no_512 and a.N != 512:
        extra = measure(512, max(4, a.steps // 4), 1, 0.002, "512^3")
    cpu = None if a.no_cpu else cpu_baseline(min(a.N, 256), a.dt)
    dom = main_r["dominant"]
    k = main_r["kernels"][dom]
    traffic = traffic_from_profile(dom, a.N)
    out = {
        "metric": "FP64 RK4 steps/s at %d^3 (TGV, %% of HBM roofline in roofline)" % a.N,
        "value": main_r["steps_per_s"],
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": main_r["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "Taylor-Green vortex %d^3, nu=1/1600, dt=%g fixed, FP64, 2/3 dealiased, RK4" % (a.N, a.dt),
                   "N": a.N, "nu": 1 / 1600, "dt": a.dt, "ic": "tgv"},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": k["GBs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": k["GBs"] / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_source": "profiles/pmc_traffic.json (rocprofv3 FETCH_SIZE and WRITE_SIZE of this kernel, "
                                       "committed; not measured in this run)" if traffic else None,
                     "alg_bytes_per_launch": k["alg_bytes"], "avg_launch_us": k["avg_us"]},
        "step_roofline_frac": main_r["step_roofline_frac"],
        "zfwd_fused_with_update": main_r["zfwd_fused"],
        "ms_per_step_profiled": main_r["ms_per_step_profiled"],
        "b_alg_per_step_GB": b_alg_step(a.N) / 1e9,
        "kernels": {n: {"avg_us": round(v["avg_us"], 2), "GBs": round(v["GBs"], 1), "launches": v["launches"]}
                    for n, v in main_r["kernels"].items()},
        "cpu_baseline": cpu,
    }
    out["fp64_vector"] = fp64_vector(main_r)
    hb = hbm_bytes_per_step(a.N)
    if hb:
        out["hbm_bytes_per_step_measured"] = hb
        out["hbm_bytes_per_step_vs_b_alg"] = hb / b_alg_step(a.N)
        out["hbm_bytes_per_step_source"] = "profiles/pmc_traffic.json (rocprofv3, committed; not measured in this run)"
    copy = copy_ceiling_GBs()
    out["hbm_copy_GBs"] = copy
    if not a.no_cpu:
        out["cpu_sizes"] = cpu_sizes()
    if extra:
        hb512 = hbm_bytes_per_step(512)
        out["n512"] = {"workload": "Taylor-Green 512^3 FP64, dt=0.002 fixed", "steps_per_s": extra["steps_per_s"],
                       "hbm_bytes_per_step_measured": hb512,
                       "hbm_bytes_per_step_vs_b_alg": hb512 / b_alg_step(512) if hb512 else None,
                       "ms_per_step": extra["ms_per_step"], "step_roofline_frac": extra["step_roofline_frac"],
                       "fp64_vector": fp64_vector(extra),
                       "dominant": extra["dominant"],
                       "dominant_GBs": extra["kernels"][extra["dominant"]]["GBs"],
                       "kernels": {n: {"avg_us": round(v["avg_us"], 2), "GBs": round(v["GBs"], 1)}
                                   for n, v in extra["kernels"].items()}}
    if not a.no_fp32:
        # the GPU runner's FP32 mode: complex64 in HBM (half the bytes of every pass), single-precision transforms
        f32 = [measure(a.N, a.steps, a.warmup, a.dt, "fp32", fp32=True)]
        if not a.no_512 and a.N != 512:
            f32.append(measure(512, max(4, a.steps // 4), 1, 0.002, "fp32 512^3", fp32=True))
        fp64_by_n = {a.N: main_r, 512: extra}
        out["fp32"] = [{
            "workload": "Taylor-Green %d^3, FP32 (complex64 storage, single-precision transforms, FP64 update), dt=%g fixed"
                        % (r["N"], a.dt if r["N"] == a.N else 0.002),
            "steps_per_s": r["steps_per_s"], "ms_per_step": r["ms_per_step"],
            "vs_fp64": r["steps_per_s"] / fp64_by_n[r["N"]]["steps_per_s"] if fp64_by_n.get(r["N"]) else None,
            "step_roofline_frac": r["step_roofline_frac"], "b_alg_per_step_GB": b_alg_step(r["N"]) / 2e9,
            "device_bytes": r["device_bytes"],
            "device_bytes_vs_fp64": r["device_bytes"] / fp64_by_n[r["N"]]["device_bytes"] if fp64_by_n.get(r["N"]) else None,
            "dominant": r["dominant"], "dominant_GBs": r["kernels"][r["dominant"]]["GBs"],
            "kernels": {n: {"avg_us": round(v["avg_us"], 2), "GBs": round(v["GBs"], 1)} for n, v in r["kernels"].items()}}
            for r in f32]
    if not a.no_tubes:
        # config 3: 500 RK4 steps at 512^3 (the full run), 3 at 1024^3
        out["tubes"] = [measure_tubes(512, 500), measure_tubes(1024, 3)]
    if not a.no_ensemble:
        e = node_leg(["ensemble", "--steps", "300", "--seq-per-n", "4"], 600)
        out["ensemble"] = e if "error" in e else {
            "workload": "128 solves in slot pools: 64x32^3, 32x48^3, 24x64^3, 8x96^3, random IC k0=4 E0=0.5, "
                        "Re log-uniform [707,4000], CFL 0.5, 300 RK4 steps, one hipGraph per pool step",
            "solve_steps_per_s": e["solve_steps_per_s"], "run_s": e["run_s"], "admit_s": e["admit_s"],
            "sequential_s": e.get("sequential_s"), "speedup_vs_sequential": e.get("speedup_vs_sequential"),
            "sequential_with_setup_s": e.get("sequential_with_setup_s"),
            "speedup_with_setup": e.get("speedup_with_setup"),
            "pool_equals_single": e.get("pool_equals_single"), "graphed": e["graphed"],
            "b_alg_GBs": sum(n * b_alg_step(N) for N, n in ((32, 64), (48, 32), (64, 24), (96, 8))) * 300
                         / e["run_s"] / 1e9}
        # the same ensemble with every pool's step in one captured graph (one launch per ensemble step)
        e1 = node_leg(["ensemble", "--steps", "300", "--seq-per-n", "0", "--one-graph"], 600)
        if "error" not in e and "error" not in e1:
            out["ensemble"]["one_graph"] = {"run_s": e1["run_s"], "solve_steps_per_s": e1["solve_steps_per_s"],
                                            "speedup_vs_sequential": e["sequential_s"] / e1["run_s"]
                                            if e.get("sequential_s") else None}
        if not a.no_cpu and "error" not in e:
            ce = cpu_ensemble()
            out["ensemble"]["cpu_all_cores"] = ce
            out["ensemble"]["vs_cpu_all_cores"] = e["solve_steps_per_s"] / ce["solve_steps_per_s"]
    if not a.no_serve:
        sv = node_leg(["serve", "--requests", str(a.serve_requests), "--rate", "2", "--seed", "1",
                       "--seq-sample", "2"], 900)
        out["serving"] = sv if "error" in sv else {
            "workload": "%d ladder requests, Poisson 2/s, ladders (48,64,96)/(64,96,128)/(96,128,192)/(128,192,256), "
                        "tubes or random IC, Re log-uniform [707,4000], 200-1000 steps on the finest rung, CFL 0.5"
                        % sv["requests"],
            "ladders_per_hour": sv["ladders_per_hour"], "rung_solves_per_hour": sv["rung_solves_per_hour"],
            "p50_latency_s": sv["p50_latency_s"], "p99_latency_s": sv["p99_latency_s"], "wall_s": sv["wall_s"],
            "admit_host_s": sv.get("admit_host_s"), "warm_start_s": sv.get("warm_start_s"),
            "sequential_estimate_s": sv["sequential_estimate_s"],
            "throughput_vs_sequential": sv["throughput_vs_sequential"],
            "sequential_sample_s": sv.get("sequential_sample_s"),
            "sequential_sample_estimate_s": sv.get("sequential_sample_estimate_s"),
            "served_equals_sequential": sv.get("served_equals_sequential")}
        if "error" not in sv:
            out["serving"].update(serving_bound(sv, copy, 2.0))
        # the same loop with half the requests Reynolds sweeps (one grid of 64..192, three Re of the sqrt(2) series)
        sm = node_leg(["serve", "--requests", str(max(4, a.serve_requests // 2)), "--rate", "2", "--seed", "2",
                       "--sweeps", "0.5", "--seq-sample", "2"], 900)
        out["serving_mixed"] = sm if "error" in sm else {
            "workload": "%d requests, Poisson 2/s, half refinement ladders as above, half Reynolds sweeps (one grid "
                        "from 64/96/128/192, three consecutive Re of 707..4000 in steps of sqrt(2)), tubes or random "
                        "IC, 200-1000 steps, CFL 0.5" % sm["requests"],
            "requests_per_hour": sm["requests_per_hour"], "sweeps": sm["sweeps"],
            "rung_solves_per_hour": sm["rung_solves_per_hour"], "p50_latency_s": sm["p50_latency_s"],
            "p99_latency_s": sm["p99_latency_s"], "wall_s": sm["wall_s"],
            "throughput_vs_sequential": sm["throughput_vs_sequential"],
            "served_equals_sequential": sm.get("served_equals_sequential")}
        if "error" not in sm:
            out["serving_mixed"].update(serving_bound(sm, copy, 2.0))
    print(json.dumps(out))


if __name__ == "__main__":
    main()
