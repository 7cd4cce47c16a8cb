"""Every compiled kernel form of a size gives the tuned form's step: the autotuner may pick any of them.

For every size with a size-specialised path from 32 to 512 (row-pair tiles 1, 2, 4, 8, one row per workgroup, one
wave per row group, the staged wave form with and without persistent waves, at 32, 48 and 64 the plane pass (y-inverse,
x pass and y-forward in one workgroup per z-plane), each with the z-forward fused into the update or not, crossed with the pencil passes' alternative tile width and the alternative radix schedules of the
staged x pass and the fused update), a random-IC solve (CFL 0.5) run for 3 steps under each form matches the same
solve under the table's form to 1e-13 relative in the state and the series. The forms differ only in the radix factorisation and tiling of
the same transforms, so the expected difference is roundoff (a few 1e-16). Also: a pool created under an override
uses it for every slot, and the slot stays bit-identical to a single solve created under the same override."""
import numpy as np
import pytest

import capi

pytestmark = pytest.mark.gpu


def run3(N, cfg):
    capi.tune_set(N, cfg)
    try:
        s = capi.Solver(N=N, Re=1000, ic="random", cfl=0.5)
        s.run(3)
        st = s.state()[:, ::97].copy()
        E = np.array(s.series["E"])
        s.close()
        return st, E
    finally:
        capi.tune_set(N)


@pytest.mark.parametrize("N", [32, 48, 64, 96, 128, 192, 256, 384, 512])
def test_every_form_matches_the_tuned_form(N):
    ref_cfg = capi.tune_get(N)
    ref, Eref = run3(N, ref_cfg)
    # the persistent-wave counts share one kernel: check no persistence and 16 per CU besides the tuned form
    forms = [f for f in capi.tune_candidates(N) if f[0] != 4 or f[2] in (0, 16)]
    assert any(f[4] for f in forms) == (N in (32, 48, 64, 96, 128, 192)), N
    assert any(f[5] == 1 for f in forms) == (N != 48) and any(f[6] for f in forms), N
    assert any(f[5] == 2 for f in forms) == (N in (48, 96, 192, 384)), N
    assert any(f[0] == 5 for f in forms) == (N in (32, 48, 64)), N   # the plane pass
    for cfg in forms:
        if cfg == ref_cfg:
            continue
        st, E = run3(N, cfg)
        rel = float(np.max(np.abs(st - ref)) / np.max(np.abs(ref)))
        assert rel < 1e-13, (N, cfg, rel)
        assert np.max(np.abs(E - Eref) / Eref) < 1e-13, (N, cfg)


@pytest.mark.parametrize("N", [32, 48, 96, 128, 192, 384])
def test_fused_and_unfused_z_forward_agree(N):
    """The z-forward fused with the update (radix schedules ending in 8: 4*8, 2*3*8, 4*3*8, 16*8, 8*3*8, 16*3*8) and
    the two-kernel form, each against the table's form."""
    ref_cfg = capi.tune_get(N)
    flip = ref_cfg[:3] + (1 - ref_cfg[3],) + ref_cfg[4:6] + ((ref_cfg[6] if ref_cfg[3] == 0 else 0),)
    assert flip in capi.tune_candidates(N), (N, flip)
    ref, Eref = run3(N, ref_cfg)
    st, E = run3(N, flip)
    rel = float(np.max(np.abs(st - ref)) / np.max(np.abs(ref)))
    assert rel < 1e-13, (N, flip, rel)
    assert np.max(np.abs(E - Eref) / Eref) < 1e-13, (N, flip)


def test_pool_uses_the_override_and_matches_a_single_solve():
    N, cfg = 64, (1, 8, 0, 0)
    assert capi.tune_get(N) != cfg
    capi.tune_set(N, cfg)
    try:
        s = capi.Solver(N=N, Re=900, ic="random", cfl=0.5, icParams={"seed": 5})
        s.run(4)
        a = s.state()
        s.close()
        pool = capi.Pool(N, 2)
        m = pool.admit(Re=900, ic="random", cfl=0.5, icParams={"seed": 5})
        pool.set_target(m, 4)
        pool.run_all()
        b = m.state()
        pool.close()
        assert np.array_equal(a, b)
    finally:
        capi.tune_set(N)
    # the table's form gives the same step to roundoff (the row-pair tile changes no arithmetic at all)
    s = capi.Solver(N=N, Re=900, ic="random", cfl=0.5, icParams={"seed": 5})
    s.run(4)
    c = s.state()
    s.close()
    assert float(np.max(np.abs(a - c)) / np.max(np.abs(c))) < 1e-13


def test_1024_forms_match_the_tuned_form():
    """1024^3 (88.6 GB per solve): the row form and the staged form with and without the fused z-forward, and the
    alternative schedules of the staged x pass and the fused update, against the default form."""
    N = 1024
    ref_cfg = capi.tune_get(N)
    ref, Eref = run3(N, ref_cfg)
    for cfg in ((2, 0, 0, 0), (4, 0, 16, 0), (4, 0, 16, 1), (4, 0, 16, 1, 0, 1, 0), (4, 0, 16, 1, 0, 0, 1)):
        if cfg == ref_cfg:
            continue
        st, E = run3(N, cfg)
        rel = float(np.max(np.abs(st - ref)) / np.max(np.abs(ref)))
        assert rel < 1e-13, (cfg, rel)
        assert np.max(np.abs(E - Eref) / Eref) < 1e-13, cfg
