"""Parity of the HIP solver (through the C ABI) with the reference: golden vectors written by the reference
(tests/golden/) and the oracle on the same inputs. Tolerances: fields 1e-13 relative (L-infinity over the state),
scalar series 1e-13 relative, except quantities that are roundoff by construction (Tnl, Pspec at t = 0, ebal, zbal),
which are compared in absolute terms."""
import numpy as np
import pytest

import capi
import fixtures as F
from conftest import rel_inf
from oracle import oracle as O

pytestmark = pytest.mark.gpu

SERIES_REL = ["E", "Z", "eps", "omMax", "uMax", "Pal", "t", "dt"]


def test_factorize_and_fft1_match_reference():
    g = F.load_json("fft1.json")
    for N, case in g.items():
        N = int(N)
        assert capi.factorize(N) == case["factors"]
        fr, fi = capi.fft1(case["xr"], case["xi"], -1)
        assert rel_inf(np.concatenate([fr, fi]), np.concatenate(case["fwd"])) < 1e-14
        br, bi = capi.fft1(case["xr"], case["xi"], 1)
        assert rel_inf(np.concatenate([br, bi]), np.concatenate(case["inv"])) < 1e-14


def test_fft1_many_lines_against_numpy():
    rng = np.random.default_rng(5)
    for N in (4, 6, 8, 12, 18, 27 * 2, 96, 128, 192, 256, 384, 512, 1024):
        xr = rng.standard_normal((37, N)); xi = rng.standard_normal((37, N))
        fr, fi = capi.fft1(xr, xi, -1)
        ref = np.fft.fft(xr + 1j * xi, axis=1)
        assert np.max(np.abs(fr + 1j * fi - ref)) < 1e-12 * np.max(np.abs(ref)), N
        br, bi = capi.fft1(xr, xi, 1)
        ref = np.fft.ifft(xr + 1j * xi, axis=1) * N
        assert np.max(np.abs(br + 1j * bi - ref)) < 1e-12 * np.max(np.abs(ref)), N


def test_spectral_forward_inverse_match_reference():
    N = 24
    h = 2 * np.pi / N
    l, j, i = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")
    f = np.sin(3 * i * h) * np.cos(2 * j * h) + 0.5 * np.cos(5 * l * h) * np.sin(i * h) + 0.1
    re, im = capi.spectral_forward(f)
    g = F.load_f64("spectral24_fwd.f64.gz")
    NS = N * N * (N // 2 + 1)
    assert np.max(np.abs(re - g[:NS])) < 1e-15 and np.max(np.abs(im - g[NS:])) < 1e-15
    back = capi.spectral_inverse(N, re, im)
    assert np.max(np.abs(back - f.ravel())) < 1e-12
    assert abs(re[0] - 0.1) < 1e-14


def test_tgv32_fixed_dt_100_and_200_steps_match_reference():
    g = F.load_json("tgv32_dt01.json")
    s = capi.Solver(N=32, Re=1600, ic="tgv", dt=0.01)
    S0 = F.load_state("tgv32_dt01_s000.f64.gz", 32)
    assert rel_inf(s.state(), S0) < 1e-14
    for k, v in g["last0"].items():
        if k in ("Tnl", "Pspec"):
            assert abs(s.last[k]) < 1e-15
        else:
            assert abs(s.last[k] - v) <= 1e-14 * abs(v), k
    assert s.run(100) == 100
    assert rel_inf(s.state(), F.load_state("tgv32_dt01_s100.f64.gz", 32)) < 1e-13
    assert s.run(100) == 100
    assert rel_inf(s.state(), F.load_state("tgv32_dt01_s200.f64.gz", 32)) < 1e-13
    for k in SERIES_REL + ["Pspec"]:
        assert rel_inf(s.series[k], g["series"][k]) < 1e-13, k
    assert np.max(np.abs(np.array(s.series["ebal"]) - g["series"]["ebal"])) < 1e-13
    assert np.max(np.abs(np.array(s.series["zbal"]) - g["series"]["zbal"])) < 1e-12
    st = g["st200"]
    c = s.clock
    assert c["step"] == st["step"] and c["t"] == st["t"] and c["dt"] == st["dt"]


def test_random_ic_and_steps_match_reference():
    g = F.load_json("random.json")
    for seed in (1, 7, 42):
        s = capi.Solver(N=24, Re=500, ic="random", icParams={"seed": seed, "E0": 0.5})
        assert rel_inf(s.state(), F.load_state("random24_seed%d_s000.f64.gz" % seed, 24)) < 1e-13
        assert abs(s.E0 - 0.5) < 1e-13
        if seed == 7:
            assert s.run(10) == 10
            assert rel_inf(s.state(), F.load_state("random24_seed7_s010.f64.gz", 24)) < 1e-13
            for k in SERIES_REL:
                assert rel_inf(s.series[k], g["7"]["series"][k]) < 1e-13, k
    s = capi.Solver(N=32, Re=2000, ic="random", icParams={"seed": 3, "k0": 6, "E0": 0.25})
    assert rel_inf(s.state(), F.load_state("random32_seed3_k06_s000.f64.gz", 32)) < 1e-13


def test_tubes_ic_matches_reference():
    s = capi.Solver(N=24, Re=4000, ic="tubes")
    assert rel_inf(s.state(), F.load_state("tubes24_s000.f64.gz", 24)) < 1e-13
    p = capi.Solver(N=32, Re=2000, ic="tubes", icParams={"amp": 5, "sigma": 0.3, "sep": 0.8, "pert": 0.1})
    assert rel_inf(p.state(), F.load_state("tubes32_params_s000.f64.gz", 32)) < 1e-13
    u = capi.Solver(N=24, Re=500, ic="tubes")
    u.run(20)
    assert rel_inf(u.state(), F.load_state("tubes24_re500_s020.f64.gz", 24)) < 1e-13
    for k in SERIES_REL:
        assert rel_inf(u.series[k], F.load_json("tubes.json")["re500"]["series"][k]) < 1e-13, k


def test_hip_matches_oracle_adaptive_cfl():
    for (N, ic, Re) in ((16, "abc", 10), (48, "tgv", 1600), (36, "random", 800), (32, "tubesz", 2000)):
        s = capi.Solver(N=N, Re=Re, ic=ic, cfl=0.4)
        o = O.OracleSolver(N=N, Re=Re, ic=ic, cfl=0.4)
        assert s.run(12) == 12 and o.run(12) == 12
        assert rel_inf(s.state(), o.state()) < 1e-13, (N, ic)
        for k in SERIES_REL:
            assert rel_inf(s.series[k], o.series[k]) < 1e-13, (N, ic, k)


def test_diagnose_matches_reference_tgv32():
    g = F.load_json("tgv32_dt01.json")
    s = capi.Solver(N=32, Re=1600, ic="tgv", dt=0.01)
    s.run(200)
    d = s.diagnose()
    gd = g["diag"]
    for k in ("E", "Z", "P", "eps", "omMax", "uMax", "Pspec", "Pphys", "skew", "kmaxEta", "skewIso", "uL3", "Rel",
              "lambda", "eta"):
        assert abs(d[k] - gd[k]) <= 1e-12 * abs(gd[k]) + 1e-15, k
    assert d["divMax"] < 1e-12
    assert np.allclose(d["align"], gd["align"], rtol=1e-12, atol=0)
    assert np.allclose(d["spectrum"], gd["spectrum"], rtol=1e-12, atol=1e-30)
    assert np.max(np.abs(np.array(d["alignHist"]) - np.array(gd["alignHist"]))) <= 2.0 / 32768
    assert abs(d["omMaxI"] - gd["omMaxI"]) < 1e-10
    assert abs(d["omMaxNode"] - gd["omMaxNode"]) < 1e-12


def test_exact_solutions():
    g = F.load_json("exact.json")
    s = capi.Solver(N=16, Re=10, ic="abc", cfl=0.4)
    s.run(10 ** 6, 1.0)
    ex = s.exactError()
    assert ex["linf"] < 1e-10
    assert abs(ex["linf"] - g["abc"]["exact"]["linf"]) < 1e-13
    assert len(s.series["t"]) == len(g["abc"]["series"]["t"])
    b = capi.Solver(N=16, Re=10, ic="tgv2d", cfl=0.4)
    b.run(10 ** 6, 1.0)
    assert b.exactError()["linf"] < 1e-8
