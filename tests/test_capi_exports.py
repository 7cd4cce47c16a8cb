"""The C ABI library loads without a GPU and exports every function include/nslab_hip.h declares."""
import ctypes
import re
import subprocess

import capi


def declared_functions():
    text = open(capi.HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(nsl_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_solver_surface():
    names = declared_functions()
    for want in ("nsl_create", "nsl_run", "nsl_get_state", "nsl_set_state", "nsl_diagnose", "nsl_fft1",
                 "nsl_spectral_forward", "nsl_spectral_inverse", "nsl_last_error", "nsl_destroy"):
        assert want in names


def test_library_exports_every_declared_symbol():
    L = capi.lib()
    for name in declared_functions():
        assert hasattr(L, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", capi.LIB], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\bT (nsl_[a-z0-9_]+)", out))
    assert set(declared_functions()) <= exported


def test_host_side_errors_without_gpu():
    assert capi.lib().nsl_version().startswith(b"nslab-hip")
    assert capi.factorize(96) == [4, 4, 2, 3]
    try:
        capi.factorize(7)
        raise AssertionError("expected an error")
    except capi.NslError as e:
        assert "2^a·3^b ≥ 4 (got 7)" in str(e)


def test_pool_slot_bytes_cover_the_solver_buffers():
    """One slot holds S, T, ACC, NL (3 D each), X1 (6 N b hp), X2 (6 N^2 hp) complex FP64 plus small control data."""
    prev = 0
    for N in (32, 48, 64, 96, 128, 192, 256, 512):
        kc = N // 3
        hx, b = kc + 1, 2 * kc + 1
        hp = (hx + 7) & ~7
        D = b * b * hp
        big = 16 * (12 * D + 6 * N * b * hp + 6 * N * N * hp)
        got = capi.Pool.slot_bytes(N)
        assert big <= got <= big * 1.05 + (1 << 20), N
        assert got % 4096 == 0 and got > prev
        prev = got
