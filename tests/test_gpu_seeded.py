"""The size-specialised kernels of every grid the ensemble and the ladders run below 192^3 (32, 48, 64, 96, 128)
against the reference itself on a field with no symmetry: a seeded random solenoidal field (k0 4, E0 0.5), Re 1000,
adaptive CFL 0.5, stepped 100 times by nslab.js (tests/golden/random{N}_s{seed}*, written by make_golden.js seeded).
The initial realisation and the state after 100 steps must agree to 1e-13 relative (L-infinity over the stored
modes: all retained modes at 32^3 and 48^3, every `stride`-th one, about 2000, at 64^3 and up), and so must the
per-step series and the step clock."""
import numpy as np
import pytest

import capi
import fixtures as F
from conftest import rel_inf

pytestmark = pytest.mark.gpu

SIZES = [(32, 132), (48, 148), (64, 164), (96, 196), (128, 228)]
SERIES = ["t", "dt", "E", "Z", "eps", "omMax", "uMax", "Pal"]


def sampled(N, S, stride):
    """Every stride-th retained mode of a reference-layout state, in the fixture's order: (6, m)."""
    idx = F.mask_index(N)[::stride]
    return S[:, idx]


def fixture_modes(name, N, stride):
    raw = F.load_f64(name)
    return raw.reshape(6, -1)


@pytest.mark.parametrize("N,seed", SIZES)
def test_seeded_random_field_100_steps_matches_reference(N, seed):
    g = F.load_json("random%d_s%d.json" % (N, seed))
    stride = g["stride"]
    print("seeded random field N=%d seed=%d stride=%d" % (N, seed, stride), flush=True)
    s = capi.Solver(N=N, Re=1000, cfl=0.5, ic="random", icParams={"seed": seed, "k0": 4, "E0": 0.5})
    a0 = sampled(N, s.state(), stride)
    b0 = fixture_modes("random%d_s%d_s000.f64.gz" % (N, seed), N, stride)
    assert a0.shape == b0.shape
    assert rel_inf(a0, b0) < 1e-13, N
    assert s.run(100) == 100
    a1 = sampled(N, s.state(), stride)
    b1 = fixture_modes("random%d_s%d_s100.f64.gz" % (N, seed), N, stride)
    err = rel_inf(a1, b1)
    print("  state after 100 steps: rel err %.2e" % err, flush=True)
    assert err < 1e-13, N
    for k in SERIES:
        assert rel_inf(s.series[k], g["series"][k]) < 1e-13, (N, k)
    c = s.clock
    assert c["step"] == g["st"]["step"]
    assert abs(c["t"] - g["st"]["t"]) <= 1e-13 * g["st"]["t"]
    for k in ("E", "Z", "P", "eps", "omMax", "uMax"):
        assert abs(s.last[k] - g["last"][k]) <= 1e-13 * abs(g["last"][k]), (N, k)
    s.close()
