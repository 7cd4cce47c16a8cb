"""Runs split into pieces give the same bits as one run. The sums of each rhs are booked by the launch after it (the
next rhs's z-inverse, the next step's start, or a booking kernel after the last step enqueued), so the last step of
every piece is booked at a different place than the steps inside a piece; every split must still give the same
series, clock and state, for single solves (FP64 and FP32), for runs stopped by an end time inside a piece, and for
pool slots launched a few steps at a time."""
import numpy as np
import pytest

import capi

pytestmark = pytest.mark.gpu

KEYS = ["t", "dt", "E", "Z", "Pal", "eps", "omMax", "uMax", "Pspec", "ebal", "zbal"]


def same(a, b):
    assert a.clock["t"] == b.clock["t"] and a.clock["step"] == b.clock["step"], (a.clock, b.clock)
    for k in KEYS:
        x, y = np.array(a.series[k]), np.array(b.series[k])
        assert x.shape == y.shape and np.array_equal(x, y), k
    assert np.array_equal(a.state(), b.state())


@pytest.mark.parametrize("fp32", [False, True])
def test_pieces_equal_one_run(fp32):
    N = 48
    whole = capi.Solver(N=N, Re=1000, ic="random", cfl=0.5, fp32=fp32)
    assert whole.run(70) == 70   # the library enqueues at most 64 steps between host round trips
    parts = capi.Solver(N=N, Re=1000, ic="random", cfl=0.5, fp32=fp32)
    done = sum(parts.run(k) for k in (1, 2, 5, 62))
    assert done == 70
    same(whole, parts)
    by_ones = capi.Solver(N=N, Re=1000, ic="random", cfl=0.5, fp32=fp32)
    assert by_ones.run(70, chunk=1) == 70
    same(whole, by_ones)
    for s in (whole, parts, by_ones):
        s.close()


def test_end_time_inside_a_piece():
    a = capi.Solver(N=32, Re=1600, ic="tgv", cfl=0.4)
    b = capi.Solver(N=32, Re=1600, ic="tgv", cfl=0.4)
    na = a.run(1000, tEnd=1.0)
    nb = sum(b.run(7, tEnd=1.0) for _ in range(20))
    assert na == nb and 0 < na < 140, (na, nb)
    assert abs(a.clock["t"] - 1.0) < 0.05
    same(a, b)
    # after the stop, further steps of the same run call do nothing
    assert b.run(5, tEnd=1.0) == 0
    same(a, b)
    a.close(); b.close()


def test_pool_slots_launched_a_few_steps_at_a_time():
    N = 32
    singles = [capi.Solver(N=N, Re=700 + 300 * i, ic="random", cfl=0.5) for i in range(2)]
    for s, n in zip(singles, (11, 17)):
        assert s.run(n) == n
    for k in (1, 3, 64):
        pool = capi.Pool(N, 2)
        ms = [pool.admit(Re=700 + 300 * i, ic="random", cfl=0.5) for i in range(2)]
        for m, n in zip(ms, (11, 17)):
            pool.set_target(m, n)
        pool.run_all(k)
        for m, s in zip(ms, singles):
            same(m, s)
        pool.close()
    for s in singles:
        s.close()
