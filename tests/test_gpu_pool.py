"""Slot pools (nsl_pool_*): solves of one N stepped together by one launch per pass over all slots.

The bar is bit-identity: a solve in a pool, alone or beside other solves (other ICs, viscosities, step counts,
slots that are released and re-admitted mid-run), gives exactly the state, clock and series of the same solve run by
itself through nsl_create/nsl_run. Each slot's launch shape and reduction order are those of a single solve, so
equality is exact (np.array_equal), not a tolerance."""
import numpy as np
import pytest

import capi

pytestmark = pytest.mark.gpu

KEYS = ["t", "E", "Z", "eps", "omMax", "uMax", "dt", "ebal", "zbal", "Pspec", "Pal"]


def same(a, b):
    assert np.array_equal(a.state(), b.state())
    assert a.clock == b.clock
    for k in KEYS:
        assert a.series[k] == b.series[k], k
    assert a.last == b.last


def single(N, steps, **kw):
    s = capi.Solver(N=N, **kw)
    s.run(steps)
    return s


CASES = [
    (40, dict(Re=900, ic="random", icParams={"seed": 5}, cfl=0.4)),
    (25, dict(Re=2500, ic="tubes", cfl=0.5)),
    (60, dict(Re=1600, ic="tgv", dt=0.01)),
]


@pytest.mark.parametrize("N", [32, 48])
def test_pool_slots_match_single_solves_bitwise(N):
    pool = capi.Pool(N, 4)
    assert pool.info()["graphed"] == 1
    members = []
    for steps, kw in CASES:
        m = pool.admit(**kw)
        pool.set_target(m, steps)
        members.append(m)
    # chunks of 16 steps: solves finish in different chunks and the pool keeps stepping the rest
    chunks = pool.run_all(16)
    assert chunks == 4
    for m, (steps, kw) in zip(members, CASES):
        assert m.clock["step"] == steps
        same(m, single(N, steps, **kw))
    pool.close()


def test_alone_and_in_the_mix_agree_with_slot_reuse():
    N = 64
    alone = capi.Pool(N, 1)
    a1 = alone.admit(Re=1200, ic="random", icParams={"seed": 9, "k0": 3}, cfl=0.5)
    alone.set_target(a1, 50)
    alone.run_all(10)

    mix = capi.Pool(N, 3)
    other = mix.admit(Re=700, ic="tubes", cfl=0.5)
    mix.set_target(other, 20)
    mix.launch(10); mix.collect(10)
    a2 = mix.admit(Re=1200, ic="random", icParams={"seed": 9, "k0": 3}, cfl=0.5)   # joins mid-run
    mix.set_target(a2, 50)
    third = mix.admit(Re=3000, ic="tgv", cfl=0.4)
    mix.set_target(third, 7)
    mix.launch(10); mix.collect(10)
    mix.release(other)                                                              # leaves; its slot is reused
    fourth = mix.admit(Re=2000, ic="abc", cfl=0.4)
    mix.set_target(fourth, 33)
    mix.run_all(10)
    same(a1, a2)
    same(a2, single(N, 50, Re=1200, ic="random", icParams={"seed": 9, "k0": 3}, cfl=0.5))
    same(fourth, single(N, 33, Re=2000, ic="abc", cfl=0.4))
    same(third, single(N, 7, Re=3000, ic="tgv", cfl=0.4))
    alone.close(); mix.close()


def test_pool_slot_handles_support_the_solver_calls():
    pool = capi.Pool(32, 2)
    m = pool.admit(Re=1600, ic="tgv", dt=0.01)
    pool.set_target(m, 10)
    pool.run_all()
    ref = single(32, 10, Re=1600, ic="tgv", dt=0.01)
    d, r = m.diagnose(), ref.diagnose()
    for k in ("E", "Z", "P", "eps", "omMax", "skew", "divMax"):
        assert d[k] == r[k], k
    assert np.array_equal(m.spectrum(), ref.spectrum())
    m.run(5); ref.run(5)            # nsl_run on the slot handle: that slot only
    same(m, ref)
    pool.set_target(m, 3, tEnd=m.clock["t"] + 0.015)   # tEnd stops it after 2 of the 3 steps (dt = 0.01)
    pool.run_all()
    assert m.clock["step"] == 17
    pool.close()


def test_pool_rejects_bad_requests():
    with pytest.raises(capi.NslError, match="compiled kernels"):
        capi.Pool(36, 2)
    pool = capi.Pool(32, 1)
    pool.admit(Re=100, ic="tgv")
    with pytest.raises(capi.NslError, match="full"):
        pool.admit(Re=100, ic="tgv")
    with pytest.raises(capi.NslError, match="does not match"):
        capi.Solver(N=48, Re=100, ic="tgv", pool=pool)
    pool.close()
