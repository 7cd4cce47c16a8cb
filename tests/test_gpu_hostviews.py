"""GPU paths behind the health report and the host views: the cutoff pile-up of a diagnosed state (closed form),
slice() and imageDiag() against the oracle, and the pool's refusal to touch a slot of the chunk in flight."""
import math

import numpy as np
import pytest

import capi
from conftest import rel_inf
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def shell_state(N, shells):
    """A state of single modes k = (kx, ky, 0), kx > 0, each carrying velocity perpendicular to k: shells maps
    (kx, ky) -> (a, b) with u = a e_z + b (-ky, kx, 0)/|k|. Each such mode holds energy hw/2 |u|^2 = a^2 + b^2
    (hw = 2 off the kx = 0 and kx = N/2 planes) in the shell round(|k|)."""
    NH = N // 2 + 1
    S = np.zeros((6, N * N * NH))
    for (kx, ky), (a, b) in shells.items():
        idx = (0 * N + (ky % N)) * NH + kx
        k = np.hypot(kx, ky)
        S[0, idx] = -b * ky / k
        S[2, idx] = b * kx / k
        S[4, idx] = a
    return S


def test_pile_up_of_a_diagnosed_state_closed_form():
    """N = 48: kc = 16, k8 = round(12.8) = 13. Energies (dyadic, so every shell sum is exact): E(12) tiny, E(13) =
    1/64, E(14) = 5/64, E(15) = 8/64, E(16) = 20/64 and E(17) = 1 from k = (16, 6, 0) (|k| = 17.09: past the cutoff
    although every component is retained). pileUp = E(16)/E(13) = 20. A truncated k8 (12), a band ending at 15 or at
    17 give other values; E(13) below 100 eps^2 E(17) gives null."""
    N = 48
    base = {(12, 0): (2.0 ** -10, 0.0), (14, 0): (0.125, 0.25), (15, 0): (0.25, 0.25), (16, 0): (0.5, 0.25),
            (16, 6): (1.0, 0.0)}
    s = capi.Solver(N=N, Re=1600, ic="tgv", dt=0.001)
    for e13, want in ((0.125, 20.0), (2.0 ** -45, 20.0 / 2.0 ** -84), (2.0 ** -50, None)):
        shells = dict(base)
        shells[(13, 0)] = (e13, 0.0)
        s.set_state(shell_state(N, shells))
        sp = s.spectrum()
        assert sp[13] == e13 * e13 and sp[16] == 0.3125 and sp[17] == 1.0 and sp[12] == 2.0 ** -20
        d = s.diagnose(interp=False)
        assert d["pileUp"] == want, (e13, d["pileUp"], want)
        assert capi.cutoff_pile_up(16, d["spectrum"]) == want
    s.close()


@pytest.fixture(scope="module")
def tubes24():
    s = capi.Solver(N=24, Re=500, ic="tubes")
    o = O.OracleSolver(N=24, Re=500, ic="tubes")
    assert s.run(6) == 6 and o.run(6) == 6
    d = s.diagnose()
    od = o.diagnose()
    return s, o, d, od


def test_image_diag_matches_oracle(tubes24):
    s, o, d, od = tubes24
    im = s.image_diag()
    for k in ("zCentroid", "zExtent", "imageGap"):
        assert abs(im[k] - od[k]) <= 1e-12 * max(abs(od[k]), 1.0), k
        assert im[k] == d[k], k   # the standalone call and diagnose() share the code
    assert im["zBands"] == od["zBands"]


def test_slice_matches_oracle_fields(tubes24):
    s, o, d, od = tubes24
    N = 24
    u = [o.physical(c).reshape(N, N, N) for c in range(3)]       # [z][y][x]
    w = [o.physical(3 + c).reshape(N, N, N) for c in range(3)]
    q, pst = o.physical(6).reshape(N, N, N), o.physical(7).reshape(N, N, N)
    mag = lambda f: np.sqrt(f[0] ** 2 + f[1] ** 2 + f[2] ** 2)
    ref = {"vort": mag(w), "speed": mag(u), "ke": 0.5 * (u[0] ** 2 + u[1] ** 2 + u[2] ** 2), "q": q, "stretch": pst,
           "u": u[0], "v": u[1], "w": u[2], "omx": w[0], "omy": w[1], "omz": w[2]}
    # float32 planes of FP64 fields that agree to roundoff: within a few float32 ulps of the field's scale (a
    # component's scale is its vector's magnitude: u_x of these tubes is itself roundoff)
    scale = {k: np.max(np.abs(ref["speed" if k in ("u", "v", "w", "ke") else "vort" if k.startswith("om") else k]))
             for k in ref}
    for kind, F in ref.items():
        for axis, ix in (("z", 5), ("y", 17), ("x", 0), ("x", 99)):
            i = min(max(ix, 0), N - 1)
            want = F[i, :, :] if axis == "z" else F[:, i, :] if axis == "y" else F[:, :, i]   # rows b, columns a
            got = s.slice(kind, axis, ix).reshape(N, N)
            err = float(np.max(np.abs(got.astype(np.float64) - want)))
            assert err <= 1e-6 * scale[kind], "%s %s: %.3e vs scale %.3e" % (kind, axis, err, scale[kind])
    assert not np.any(s.slice("nonsense", "z", 3))


def test_pool_refuses_a_slot_of_the_chunk_in_flight():
    pool = capi.Pool(32, 3)
    a = pool.admit(Re=1600, ic="tgv", dt=0.01)
    pool.set_target(a, 200)
    pool.launch(64)
    with pytest.raises(capi.NslError, match="in flight"):
        pool.set_target(a, 10)
    with pytest.raises(capi.NslError, match="in flight"):
        a.run(1)
    # a slot admitted after the launch was not part of the chunk: it may be set up meanwhile
    b = pool.admit(Re=800, ic="random", cfl=0.5)
    pool.set_target(b, 5)
    pool.collect()
    pool.set_target(a, 10)   # collected: allowed again
    pool.run_all()
    assert a.clock["step"] == 74 and b.clock["step"] == 5
    pool.close()


def test_tube_pair_along_z_at_512():
    """The memory-heavy configuration's initial field at its real size (512^3), checked as the oracle's is at 64^3:
    curl of the recovered velocity against the analytic tube vorticity (on 16 z-planes), the circulation -Gamma of
    the tube at x < pi, k . u = 0, and the perturbation energy 3/2 pert^2; then 5 RK4 steps at CFL 0.5 keep the
    energy budget closed."""
    N, G, sig, d0 = 512, 1.0, 0.3, 0.8
    print("tubes along z at 512^3", flush=True)
    s = capi.Solver(N=N, Re=2000, ic="tubesz", icParams={"pert": 0}, cfl=0.5)
    h = 2 * np.pi / N
    wz = s.field(5).reshape(N, N, N)
    wx = s.field(3).reshape(N, N, N)
    y, x = np.meshgrid(np.arange(N) * h, np.arange(N) * h, indexing="ij")
    peak = G / (np.pi * sig ** 2)
    for l in range(0, N, N // 16):
        z = l * h
        ez = np.zeros_like(x); ex = np.zeros_like(x)
        for sgn in (-1, 1):
            xc = np.pi + sgn * (d0 + 0.1 * np.cos(z))
            best = np.full_like(x, np.inf)
            for px in (-1, 0, 1):
                for py in (-1, 0, 1):
                    best = np.minimum(best, (x - xc + px * 2 * np.pi) ** 2 + (y - np.pi + py * 2 * np.pi) ** 2)
            om = sgn * G / (np.pi * sig ** 2) * np.exp(-best / sig ** 2)
            ez += om
            ex += -sgn * 0.1 * np.sin(z) * om
        assert np.max(np.abs(wz[l] - ez)) < 1e-6 * peak, l
        assert np.max(np.abs(wx[l] - ex)) < 1e-6 * peak, l
    # circulation through x < pi at z = 0: the tube there (centre pi - 0.9) less the share erfc(3)/2 of its
    # Gaussian beyond x = pi, plus the same share of the opposite tube (centre pi + 0.9): -Gamma (1 - erfc(3)). The
    # grid sum over the columns x < pi differs from that integral by the Euler-Maclaurin term h^2/12 F'(pi) (F the
    # y-integral of omega_z, whose two tube slopes at the cut add: F'(pi) = 2 Gamma (2a/sigma^2) e^{-a^2/sigma^2} /
    # (sigma sqrt(pi)), a = 0.9); the next term is below 1e-10 at 512^3.
    a = d0 + 0.1
    Fp = 2 * G * (2 * a / sig ** 2) * math.exp(-a * a / sig ** 2) / (sig * math.sqrt(math.pi))
    circ = np.sum(wz[0][:, : N // 2]) * h * h
    assert abs(circ - (-G * (1 - math.erfc(a / sig)) + h * h / 12 * Fp)) < 1e-9
    del wz, wx
    S = s.state()
    NH = N // 2 + 1
    kw = np.fft.fftfreq(N, 1.0 / N)
    scale = np.max(np.abs(S))
    for z0 in range(0, N, 64):   # k . u per kz slab
        sl = slice(z0 * N * NH, (z0 + 64) * N * NH)
        U = [(S[2 * c, sl] + 1j * S[2 * c + 1, sl]).reshape(64, N, NH) for c in range(3)]
        div = np.arange(NH)[None, None, :] * U[0] + kw[None, :, None] * U[1] + kw[z0:z0 + 64, None, None] * U[2]
        assert np.max(np.abs(div)) < 1e-14 * scale * N
    del U, div
    p = capi.Solver(N=N, Re=2000, ic="tubesz", icParams={"pert": 1e-4}, cfl=0.5)
    D = p.state() - S
    del S
    hw = np.where((np.arange(NH) == 0) | (np.arange(NH) == N // 2), 1.0, 2.0)
    Ep = 0.5 * np.sum(hw[None, None, :] * (D.reshape(6, N, N, NH) ** 2).sum(axis=0))
    del D
    assert abs(Ep - 1.5e-8) < 1e-13
    s.close()
    assert p.run(5) == 5
    ser = p.series
    assert max(ser["ebal"]) < 1e-9 and max(ser["zbal"]) < 1e-7
    assert all(b < a for a, b in zip(ser["E"], ser["E"][1:]))
    assert abs(ser["omMax"][0] - peak) < 2e-2 * peak
    p.close()
