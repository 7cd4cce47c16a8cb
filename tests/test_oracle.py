"""The oracle (oracle/nslab_oracle.c) pinned against the reference's golden vectors and closed forms, on the CPU.
Tolerances: states 1e-15 relative (the oracle follows the reference's arithmetic order; only libm last-ulp
differences remain), scalars 1e-14 relative."""
import numpy as np
import pytest

import fixtures as F
from conftest import rel_inf
from oracle import oracle as O


def test_factorize_follows_reference():
    assert O.factorize(96) == [4, 4, 2, 3]
    assert O.factorize(512) == [4, 4, 4, 4, 2]
    assert O.factorize(1024) == [4, 4, 4, 4, 4]
    for bad in (7, 2, 10, 1):
        with pytest.raises(ValueError):
            O.factorize(bad)


def test_fft1_golden_and_direct_dft():
    g = F.load_json("fft1.json")
    for N, case in g.items():
        N = int(N)
        fr, fi = O.fft1(case["xr"], case["xi"], -1)
        assert rel_inf(np.concatenate([fr, fi]), np.concatenate(case["fwd"])) < 1e-15
        x = np.array(case["xr"]) + 1j * np.array(case["xi"])
        n = np.arange(N)
        dft = np.exp(-2j * np.pi * np.outer(n, n) / N) @ x            # brute force O(N^2)
        assert np.max(np.abs(fr + 1j * fi - dft)) < 1e-9
        br, bi = O.fft1(case["xr"], case["xi"], 1)
        assert np.max(np.abs(br + 1j * bi - np.conj(np.exp(-2j * np.pi * np.outer(n, n) / N)) @ x)) < 1e-9


def test_spectral_transforms_golden_numpy_and_round_trip():
    N = 24
    h = 2 * np.pi / N
    l, j, i = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")
    f = np.sin(3 * i * h) * np.cos(2 * j * h) + 0.5 * np.cos(5 * l * h) * np.sin(i * h) + 0.1
    re, im = O.spectral_forward(f)
    g = F.load_f64("spectral24_fwd.f64.gz")
    NS = N * N * (N // 2 + 1)
    assert np.max(np.abs(re - g[:NS])) < 1e-16 and np.max(np.abs(im - g[NS:])) < 1e-16
    ref = (np.fft.rfftn(f) / N ** 3).ravel()                         # the convention u_k = rfftn(f)/N^3
    assert np.max(np.abs(re + 1j * im - ref)) < 1e-15
    assert np.max(np.abs(O.spectral_inverse(N, re, im) - f.ravel())) < 1e-12


def test_tgv_closed_forms_and_100_steps():
    g = F.load_json("tgv32_dt01.json")
    s = O.OracleSolver(N=32, Re=1600, ic="tgv", dt=0.01)
    st = s.st
    assert abs(st["E0"] - 0.125) < 1e-15 and abs(st["Z0"] - 0.375) < 1e-15          # E = 1/8, Z = 3/8
    assert abs(s.last["P"] - 1.125) < 1e-14                                          # all modes |k|^2 = 3
    assert abs(s.last["Pspec"]) < 1e-15 and abs(s.last["Tnl"]) < 1e-15
    assert rel_inf(s.state(), F.load_state("tgv32_dt01_s000.f64.gz", 32)) < 1e-15
    s.run(100)
    assert rel_inf(s.state(), F.load_state("tgv32_dt01_s100.f64.gz", 32)) < 1e-15
    for k in ("E", "Z", "eps", "omMax", "uMax", "Pal", "t", "dt"):
        assert rel_inf(s.series[k], g["series"][k][:101]) < 1e-15, k
    assert s.st["step"] == 100 and s.st["t"] == g["st100"]["t"]


def test_random_and_tubes_ics_golden():
    for seed in (1, 7, 42):
        s = O.OracleSolver(N=24, Re=500, ic="random", icParams={"seed": seed, "E0": 0.5})
        assert rel_inf(s.state(), F.load_state("random24_seed%d_s000.f64.gz" % seed, 24)) < 1e-14
    t = O.OracleSolver(N=24, Re=4000, ic="tubes")
    assert rel_inf(t.state(), F.load_state("tubes24_s000.f64.gz", 24)) < 1e-14
    g = F.load_json("tubes.json")["re4000"]
    assert abs(t.st["E0"] - g["E0"]) < 1e-14 * g["E0"]


def test_exact_abc_decay_and_rk4_order():
    s = O.OracleSolver(N=16, Re=10, ic="abc", cfl=0.4)
    s.run(10 ** 6, 1.0)
    ex = s.exactError()
    assert ex["linf"] < 1e-10
    E = s.last["E"]
    assert abs(E / (1.5 * np.exp(-2 * 0.1 * s.st["t"])) - 1) < 1e-10                 # E = 1.5 e^{-2 nu t}
    g = F.load_json("exact.json")
    assert s.st["step"] == g["abc"]["st"]["step"] and abs(ex["linf"] - g["abc"]["exact"]["linf"]) < 1e-14
    errs = []
    for dt in (0.1, 0.05, 0.025):
        q = O.OracleSolver(N=16, Re=10, ic="abc", dt=dt)
        q.run(10 ** 6, 1.0)
        errs.append(q.exactError()["linf"])
    assert 3.7 < np.log2(errs[0] / errs[1]) < 4.3 and 3.7 < np.log2(errs[1] / errs[2]) < 4.3


def test_diagnose_golden_tgv_and_tubes():
    s = O.OracleSolver(N=32, Re=1600, ic="tgv", dt=0.01)
    s.set_state(F.load_state("tgv32_dt01_s200.f64.gz", 32))
    g = F.load_json("tgv32_dt01.json")["diag"]
    d = s.diagnose()
    for k in ("E", "Pspec", "Pphys", "skew", "skewIso", "uL3", "kmaxEta", "omMaxI", "omMaxNode"):
        assert abs(d[k] - g[k]) <= 1e-13 * abs(g[k]), k
    assert np.allclose(d["align"], g["align"], rtol=1e-13)
    assert np.allclose(d["spectrum"], g["spectrum"], rtol=1e-13, atol=1e-30)
    t = O.OracleSolver(N=24, Re=4000, ic="tubes")
    dt = t.diagnose()
    assert abs(dt["omMaxI"] - 7.73221) < 2e-4 and abs(dt["omMax"] - 7.69712) < 2e-4   # GPU-runner constants
    gt = F.load_json("tubes.json")["re4000"]["diag0"]
    assert abs(dt["imageGap"] - gt["imageGap"]) < 1e-12 and abs(dt["zCentroid"] - np.pi) < 1e-6


def test_planted_mistake_is_caught():
    """A wrong sign in the viscous term changes the 100-step state far beyond the parity tolerance: the state
    comparison used by every parity test detects it."""
    s = O.OracleSolver(N=16, Re=10, ic="abc", dt=0.05)
    s.run(20)
    good = s.state()
    bad = O.OracleSolver(N=16, nu=-0.1, ic="abc", dt=0.05)
    bad.run(20)
    assert rel_inf(bad.state(), good) > 1e-3
