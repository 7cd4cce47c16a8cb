"""Cutoff pile-up (nslab.js:557-564) pinned by the reference's own archived FP64 spectra.

tests/golden/pileup_fp64.json holds (N, kc, spectrum, pileUp) from the FP64 GPU-runner snapshots of the Re 2000 tube
runs at 288^3, 320^3 and 384^3 (research/nslab/tubes-Re2000-N*-gpu/final.json): six values between 1.01 and 1.70,
value-1 snapshots and null ones. The rule has three parts, and each is pinned separately:
  * k8 = round(0.8 kc) and the band k8..kc: the non-trivial archived values, reproduced exactly;
  * the noise floor 100 eps^2 * max_{k>=1} E(k): closed-form spectra just below and just above it. (The archive was
    written by runner 0.1.2, whose guard was a fixed E(k8) > 1e-20; one archived snapshot, 384^3 at t = 0.5, has
    E(k8) = 9.2e-23 between the two floors: the current rule defines it.)
The oracle (C restatement) and the library (C ABI, host arithmetic, no GPU) are both checked; the Node host calls
the library's function (tests/js/validate_report.js). Mutants of the rule fail these checks (last test)."""
import json
import os

import numpy as np
import pytest

import capi
from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
EPS = np.finfo(np.float64).eps


def cases():
    with open(os.path.join(HERE, "golden", "pileup_fp64.json")) as f:
        return json.load(f)["cases"]


def expected(c):
    """The archived value, except where only the old runner's fixed 1e-20 guard made it null."""
    sp = c["spectrum"]
    k8 = int(np.floor(0.8 * c["kc"] + 0.5))
    if c["pileUp"] is None and sp[k8] <= 1e-20 and sp[k8] > 100 * EPS * EPS * max(sp[1:]):
        return "defined"
    return c["pileUp"]


IMPLS = {"oracle": O.cutoff_pile_up, "library": capi.cutoff_pile_up}


@pytest.mark.parametrize("impl", sorted(IMPLS))
def test_archived_spectra(impl):
    f = IMPLS[impl]
    nontrivial = 0
    for c in cases():
        got, want = f(c["kc"], c["spectrum"]), expected(c)
        if want == "defined":
            assert got is not None, (c["N"], c["t"])
        else:
            assert got == want, (c["N"], c["t"], got, want)   # exact: one division and a max per bin
            nontrivial += want not in (None, 1.0)
    assert nontrivial == 6


def synthetic(kc, base, band, peak):
    """A spectrum of ceil(sqrt 3 kc) + 2 bins: `base` at k8, band[k - k8 - 1] above it up to kc, `peak` at kc + 1."""
    nb = int(np.ceil(np.sqrt(3) * kc)) + 2
    k8 = int(np.floor(0.8 * kc + 0.5))
    sp = np.full(nb, 1e-3)
    sp[0] = 0.0
    sp[k8 - 1] = 1e-9            # below the band: must not count
    sp[k8] = base
    sp[k8 + 1:kc + 1] = band
    sp[kc + 1] = peak            # past the cutoff: must not count
    return sp


@pytest.mark.parametrize("impl", sorted(IMPLS))
def test_floor_and_band_closed_form(impl):
    f = IMPLS[impl]
    floor = 100 * EPS * EPS
    for kc in (16, 21, 32, 85):   # 0.8 kc = 12.8, 16.8, 25.6, 68: k8 rounds up three times
        k8 = int(np.floor(0.8 * kc + 0.5))
        n = kc - k8
        band = np.linspace(2.0, 3.0, n)
        band[-1] = 5.0            # the cutoff bin holds the maximum
        assert f(kc, synthetic(kc, 1.0, band, 1e3)) == 5.0
        # the floor scales with the largest bin, k >= 1
        assert f(kc, synthetic(kc, 1e3 * floor * 1.01, band, 1e3)) is not None
        assert f(kc, synthetic(kc, 1e3 * floor * 0.99, band, 1e3)) is None
        assert f(kc, synthetic(kc, 0.0, band, 1e3)) is None
        sp = synthetic(kc, 1e3 * floor * 1.01, band, 1e3)
        sp[0] = 1e9               # the mean-flow bin is not part of the peak
        assert f(kc, sp) is not None
        # a band that falls from k8 gives exactly 1
        assert f(kc, synthetic(kc, 1.0, np.full(n, 0.5), 1e3)) == 1.0


def test_library_refuses_a_short_spectrum():
    with pytest.raises(capi.NslError, match="does not reach kc"):
        capi.cutoff_pile_up(16, np.ones(16))


def test_mutants_of_the_rule_are_caught():
    """Plausible mistakes, restated here on the CPU, each fail one of the checks above: a truncated k8, a band
    ending one bin early or late, a fixed 1e-20 guard, a floor without the peak, a floor over bin 0 too."""
    def rule(kc, sp, k8f=lambda kc: int(np.floor(0.8 * kc + 0.5)), lo=0, hi=0, floor=None):
        sp = np.asarray(sp, float)
        k8 = k8f(kc)
        f = floor(sp) if floor else 100 * EPS * EPS * max(sp[1:])
        if not sp[k8] > f:
            return None
        return max(sp[k] / sp[k8] for k in range(k8 + lo, kc + 1 + hi))

    mutants = {
        "k8 truncated": dict(k8f=lambda kc: int(0.8 * kc)),
        "band ends at kc-1": dict(hi=-1),
        "band ends at kc+1": dict(hi=1),
        "fixed 1e-20 guard": dict(floor=lambda sp: 1e-20),
        "floor without the peak": dict(floor=lambda sp: 100 * EPS * EPS),
        "peak over bin 0 too": dict(floor=lambda sp: 100 * EPS * EPS * max(sp)),
    }

    def passes(f):
        try:
            for c in cases():
                got, want = f(c["kc"], c["spectrum"]), expected(c)
                if (got is None) if want == "defined" else got != want:
                    return False
            floor = 100 * EPS * EPS
            for kc in (16, 21, 32, 85):
                k8 = int(np.floor(0.8 * kc + 0.5))
                band = np.linspace(2.0, 3.0, kc - k8)
                band[-1] = 5.0
                if f(kc, synthetic(kc, 1.0, band, 1e3)) != 5.0:
                    return False
                if f(kc, synthetic(kc, 1e3 * floor * 0.99, band, 1e3)) is not None:
                    return False
                if f(kc, synthetic(kc, 1e3 * floor * 1.01, band, 1e3)) is None:
                    return False
                sp = synthetic(kc, 1e3 * floor * 1.01, band, 1e3)
                sp[0] = 1e9
                if f(kc, sp) is None:
                    return False
            return True
        except IndexError:
            return False

    assert passes(lambda kc, sp: rule(kc, sp))
    for name, m in mutants.items():
        assert not passes(lambda kc, sp, m=m: rule(kc, sp, **m)), name


# ------------------------------------------------------------------------------------------------ FP32 guard
# The GPU runner's guard in its FP32 mode is 100 eps^2 with eps = 1.1920929e-07 (nslab_gpu.py:61-65,369-372).
# tests/golden/pileup_fp32.json is the archived 640^3 FP32 run (runner 0.1.2, fixed 1e-20 guard), and the runner's
# 0.1.3 notes (gpu/README.md:44-58) say what the precision-scaled guard does to it: t = 0.5 (E(k8)/peak 8.76e-20,
# archived 15.38) is ungraded, every snapshot up to t = 5.0 is ungraded, and grading starts at t = 5.5, where
# E(k8)/peak = 1.9e-12.
EPS32 = 1.1920929e-07
IMPLS_EPS = {"oracle": O.cutoff_pile_up, "library": capi.cutoff_pile_up}


def fp32_run():
    with open(os.path.join(HERE, "golden", "pileup_fp32.json")) as f:
        d = json.load(f)
    for s in d["snapshots"]:
        sp = np.zeros(s["nb"])
        sp[1] = s["peak"]
        sp[d["k8"]:d["kc"] + 1] = s["band"]
        s["spectrum"] = sp
    return d


@pytest.mark.parametrize("impl", sorted(IMPLS_EPS))
def test_fp32_guard_on_the_archived_640_run(impl):
    f = IMPLS_EPS[impl]
    d = fp32_run()
    kc = d["kc"]
    assert d["N"] == 640 and kc == 213 and d["k8"] == 170
    graded = []
    for s in d["snapshots"]:
        # the FP64 guard (4.9e-30 of the peak) grades every snapshot, as the old 1e-20 guard did: the archived values
        assert f(kc, s["spectrum"]) == s["pileUp"], s["t"]
        assert f(kc, s["spectrum"], eps=np.finfo(np.float64).eps) == s["pileUp"], s["t"]
        v = f(kc, s["spectrum"], eps=EPS32)
        if v is not None:
            assert v == s["pileUp"], s["t"]
            graded.append(round(s["t"], 1))
    # the README's account: ungraded through t = 5.0 (15.38 at t = 0.5 among them), graded from t = 5.5 on
    assert graded[0] == 5.5 and len(graded) == len(d["snapshots"]) - 11, graded
    assert [round(s["t"], 1) for s in d["snapshots"]][11] == 5.5
    t05 = d["snapshots"][1]
    assert abs(t05["pileUp"] - 15.38) < 0.005 and f(kc, t05["spectrum"], eps=EPS32) is None


@pytest.mark.parametrize("impl", sorted(IMPLS_EPS))
def test_fp32_floor_closed_form(impl):
    f = IMPLS_EPS[impl]
    floor32 = 100 * EPS32 * EPS32
    for kc in (16, 85, 213):
        k8 = int(np.floor(0.8 * kc + 0.5))
        band = np.linspace(2.0, 3.0, kc - k8)
        assert f(kc, synthetic(kc, 1e3 * floor32 * 1.01, band, 1e3), eps=EPS32) is not None
        assert f(kc, synthetic(kc, 1e3 * floor32 * 0.99, band, 1e3), eps=EPS32) is None
        # between the two floors: graded in FP64, ungraded in FP32
        sp = synthetic(kc, 1e3 * 1e-20, band, 1e3)
        assert f(kc, sp) is not None and f(kc, sp, eps=EPS32) is None


def test_library_refuses_a_bad_epsilon():
    for eps in (0.0, -1e-7, 1.0, float("nan")):
        with pytest.raises(capi.NslError, match="machine epsilon"):
            capi.cutoff_pile_up(16, np.ones(40), eps=eps)


def test_mutants_of_the_fp32_guard_are_caught():
    """A floor of 100 eps (not squared), the FP64 epsilon in FP32 mode, or eps^2 without the margin of 100 each
    disagree with the archived run's account above."""
    d = fp32_run()
    kc, k8 = d["kc"], d["k8"]

    def graded_from(floor):
        out = []
        for s in d["snapshots"]:
            sp = s["spectrum"]
            if sp[k8] > floor * max(sp[1:]):
                out.append(round(s["t"], 1))
        return out

    assert graded_from(100 * EPS32 ** 2)[0] == 5.5
    for floor in (100 * EPS32, 100 * np.finfo(np.float64).eps ** 2, EPS32 ** 2):
        g = graded_from(floor)
        assert not (g and g[0] == 5.5 and len(g) == len(d["snapshots"]) - 11), floor
