"""nsl_spectrum_noalloc, the entry point that allocates nothing, against nsl_spectrum through the C ABI: bit-identical
shell spectra on two calls in a row (the second on a smaller grid, so the shared device scratch holds the first
call's partials when it starts), again after the state moved on, no change in free device memory across the calls,
and a clear refusal for a grid whose partials exceed the fixed scratch."""
import ctypes

import numpy as np
import pytest

import capi

pytestmark = pytest.mark.gpu


def noalloc_spectrum(s):
    nb = capi.check(capi.lib().nsl_spectrum_len(s.h))
    out = np.full(nb, np.nan)
    capi.check(capi.lib().nsl_spectrum_noalloc(s.h, capi._p(out)))
    return out


def free_bytes():
    f, t = ctypes.c_int64(0), ctypes.c_int64(0)
    capi.check(capi.lib().nsl_device_memory(ctypes.byref(f), ctypes.byref(t)))
    return f.value


def test_noalloc_spectrum_matches_spectrum():
    big = capi.Solver(N=96, Re=1000, cfl=0.5, ic="random", icParams={"seed": 5})
    small = capi.Solver(N=32, Re=1600, ic="tgv", dt=0.01)
    ref_big, ref_small = big.spectrum(), small.spectrum()   # also loads the kernels
    f0 = free_bytes()
    a = noalloc_spectrum(big)
    b = noalloc_spectrum(small)      # the second call, smaller: fewer partials than the first left behind
    assert free_bytes() == f0, "the allocation-free entry point changed free device memory"
    assert np.array_equal(a, ref_big) and np.array_equal(b, ref_small)
    assert len(b) < len(a)
    big.run(3)
    assert np.array_equal(noalloc_spectrum(big), big.spectrum())
    big.close()
    small.close()


def test_noalloc_spectrum_refuses_a_grid_beyond_its_scratch():
    s = capi.Solver(N=1024, Re=1600, ic="tgv", dt=0.001)
    with pytest.raises(capi.NslError, match="needs .* bytes of scratch"):
        noalloc_spectrum(s)
    s.close()
