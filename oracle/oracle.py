"""ORACLE — test infrastructure only (the parity checker). ctypes wrapper over liboracle.so, the CPU restatement
of the NSLab solver in nslab_oracle.c. Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module; the product library never does.

Initial-condition parameter defaults follow the reference (pocket-wind-tunnel/src/nslab.js:255-274):
  abc A, B, C = 1;  tubes amp 8, sigma 0.4, sep 0.7, pert 0.2;  random seed (seed || 1) | 0, k0 || 4, E0 || 0.5.
"""
import ctypes
import math
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")

IC_KINDS = {"tgv": 0, "tgv2d": 1, "abc": 2, "tubes": 3, "random": 4, "tubesz": 5}
DIAG_KEYS = ["t", "step", "E", "Z", "P", "eps", "omMax", "uMax", "Pspec", "Pphys", "Tnl", "skew", "divMax", "eta",
             "kmaxEta", "Rel", "lambda", "skewIso"]
REC_KEYS = ["t", "E", "Z", "eps", "omMax", "uMax", "dt", "ebal", "zbal", "Pspec", "Pal"]

_lib = None


def build():
    subprocess.check_call(["make", "-s", "-C", HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        dp = ctypes.POINTER(ctypes.c_double)
        L.or_create.restype = ctypes.c_void_p
        L.or_create.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_int, dp]
        L.or_free.argtypes = [ctypes.c_void_p]
        L.or_step.argtypes = [ctypes.c_void_p, dp]
        L.or_run.restype = ctypes.c_long
        L.or_run.argtypes = [ctypes.c_void_p, ctypes.c_long, ctypes.c_double, dp]
        for f in ("or_get_state", "or_set_state", "or_get_last", "or_get_st", "or_set_st", "or_norms", "or_spectrum"):
            getattr(L, f).argtypes = [ctypes.c_void_p, dp]
        L.or_get_physical.argtypes = [ctypes.c_void_p, ctypes.c_int, dp]
        L.or_rhs_current.argtypes = [ctypes.c_void_p]
        L.or_diagnose.argtypes = [ctypes.c_void_p, ctypes.c_int, dp]
        L.or_interpMax.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, dp]
        L.or_exactError.restype = ctypes.c_int
        L.or_exactError.argtypes = [ctypes.c_void_p, dp]
        L.or_spectrum_len.restype = ctypes.c_int
        L.or_spectrum_len.argtypes = [ctypes.c_void_p]
        L.or_NS.restype = ctypes.c_long
        L.or_NS.argtypes = [ctypes.c_void_p]
        L.or_fft1.restype = ctypes.c_int
        L.or_fft1.argtypes = [ctypes.c_int, ctypes.c_int, dp, dp]
        L.or_factorize.restype = ctypes.c_int
        L.or_factorize.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
        L.or_spectral_forward.restype = ctypes.c_int
        L.or_spectral_forward.argtypes = [ctypes.c_int, dp, dp, dp]
        L.or_spectral_inverse.restype = ctypes.c_int
        L.or_spectral_inverse.argtypes = [ctypes.c_int, dp, dp, dp]
        L.or_cutoff_pile_up.restype = ctypes.c_int
        L.or_cutoff_pile_up.argtypes = [ctypes.c_int, dp, ctypes.c_int, dp]
        L.or_cutoff_pile_up_eps.restype = ctypes.c_int
        L.or_cutoff_pile_up_eps.argtypes = [ctypes.c_int, dp, ctypes.c_int, ctypes.c_double, dp]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def ic_params(kind, prm=None):
    """The four IC parameters with the reference's defaults (nslab.js:255-274)."""
    prm = dict(prm or {})
    g = lambda k, d: d if prm.get(k) is None else float(prm[k])
    truthy = lambda k, d: float(prm[k]) if prm.get(k) else d
    if kind == "abc":
        return [g("A", 1.0), g("B", 1.0), g("C", 1.0), 0.0]
    if kind == "tubes":
        return [g("amp", 8.0), g("sigma", 0.4), g("sep", 0.7), g("pert", 0.2)]
    if kind == "tubesz":   # the memory-heavy benchmark's tube pair along z (no reference counterpart)
        return [g("Gamma", 1.0), g("sigma", 0.3), g("sep", 0.8), g("pert", 1e-4)]
    if kind == "random":
        seed = int(prm["seed"]) if prm.get("seed") else 1
        seed = ((seed + 2 ** 31) % 2 ** 32) - 2 ** 31
        return [float(seed), truthy("k0", 4.0), truthy("E0", 0.5), 0.0]
    if kind in ("tgv", "tgv2d"):
        return [0.0, 0.0, 0.0, 0.0]
    raise ValueError("unknown initial condition " + str(kind))


def factorize(N):
    f = (ctypes.c_int * 40)()
    k = lib().or_factorize(int(N), f)
    if k < 0:
        raise ValueError("grid size must be 2^a·3^b ≥ 4 (got %d)" % N)
    return list(f[:k])


def fft1(xr, xi, sign):
    re = np.ascontiguousarray(xr, dtype=np.float64).copy()
    im = np.ascontiguousarray(xi, dtype=np.float64).copy()
    if lib().or_fft1(len(re), int(sign), _p(re), _p(im)) != 0:
        raise ValueError("bad N")
    return re, im


def cutoff_pile_up(kc, spec, eps=None):
    """cutoffPileUp (nslab.js:557-564) of a shell spectrum on a grid with kc = floor(N/3); None below the floor.
    eps: machine epsilon of the working precision (the GPU runner's specFloor = 100 eps^2, nslab_gpu.py:61-65);
    None is FP64."""
    sp = np.ascontiguousarray(spec, dtype=np.float64)
    out = np.zeros(1)
    if eps is None:
        return float(out[0]) if lib().or_cutoff_pile_up(int(kc), _p(sp), len(sp), _p(out)) else None
    return float(out[0]) if lib().or_cutoff_pile_up_eps(int(kc), _p(sp), len(sp), float(eps), _p(out)) else None


def spectral_forward(f):
    N = f.shape[0]
    NS = N * N * (N // 2 + 1)
    re = np.zeros(NS); im = np.zeros(NS)
    ff = np.ascontiguousarray(f, dtype=np.float64).ravel()
    if lib().or_spectral_forward(N, _p(ff), _p(re), _p(im)) != 0:
        raise ValueError("bad N")
    return re, im


def spectral_inverse(N, re, im):
    re = np.ascontiguousarray(re, dtype=np.float64); im = np.ascontiguousarray(im, dtype=np.float64)
    f = np.zeros(N ** 3)
    if lib().or_spectral_inverse(N, _p(re), _p(im), _p(f)) != 0:
        raise ValueError("bad N")
    return f


class OracleSolver:
    """createSolver(opts) restated on the CPU (nslab.js:171-649)."""

    def __init__(self, N=32, Re=None, nu=None, cfl=None, dt=None, ic="tgv", icParams=None):
        self.N = int(N)
        self.nu = float(nu) if nu is not None else 1.0 / float(Re or 100)
        self.cfl = float(cfl) if cfl else 0.4
        self.dtFixed = float(dt) if dt else 0.0
        self.ic = ic
        prm = (ctypes.c_double * 4)(*ic_params(ic, icParams))
        self.h = lib().or_create(self.N, self.nu, self.cfl, self.dtFixed, IC_KINDS[ic], prm)
        if not self.h:
            raise ValueError("grid size must be 2^a·3^b ≥ 4 (got %d)" % self.N)
        self.NS = lib().or_NS(self.h)
        self.series = {k: [] for k in REC_KEYS}
        self._record0()

    def _record0(self):
        d = self.last
        st = self.st
        for k, v in (("t", st["t"]), ("E", d["E"]), ("Z", d["Z"]), ("eps", d["eps"]), ("omMax", d["omMax"]),
                     ("uMax", d["uMax"]), ("dt", 0.0), ("Pspec", d["Pspec"]), ("Pal", d["P"])):
            self.series[k].append(v)

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_free(self.h)
            self.h = None

    def run(self, n, tEnd=None):
        rec = np.zeros(11 * max(int(min(n, 10 ** 7)), 1))
        if tEnd is not None and n > 10 ** 5:
            # long adaptive runs: step in chunks so the record buffer stays small
            k = 0
            while k < n:
                got = self.run(min(4096, n - k), tEnd)
                if got == 0:
                    break
                k += got
            return k
        k = lib().or_run(self.h, int(n), float("nan") if tEnd is None else float(tEnd), _p(rec))
        r = rec[:11 * k].reshape(k, 11)
        for i, key in enumerate(REC_KEYS):
            self.series[key].extend(r[:, i].tolist())
        return k

    def step(self):
        return self.run(1)

    @property
    def last(self):
        o = np.zeros(8); lib().or_get_last(self.h, _p(o))
        return dict(zip(["E", "Z", "P", "eps", "Tnl", "Pspec", "omMax", "uMax"], o.tolist()))

    @property
    def st(self):
        o = np.zeros(9); lib().or_get_st(self.h, _p(o))
        d = dict(zip(["t", "step", "dt", "E0", "Z0", "maxEbal", "maxZbal", "maxTnl", "divMax"], o.tolist()))
        d["step"] = int(d["step"])
        return d

    def set_st(self, d):
        o = np.array([d[k] for k in ["t", "step", "dt", "E0", "Z0", "maxEbal", "maxZbal", "maxTnl", "divMax"]], dtype=np.float64)
        lib().or_set_st(self.h, _p(o))

    def state(self):
        o = np.zeros(6 * self.NS); lib().or_get_state(self.h, _p(o))
        return o.reshape(6, self.NS)

    def set_state(self, S):
        a = np.ascontiguousarray(S, dtype=np.float64).reshape(6 * self.NS)
        lib().or_set_state(self.h, _p(a))

    def rhs_current(self):
        lib().or_rhs_current(self.h)

    def physical(self, which):
        o = np.zeros(self.N ** 3); lib().or_get_physical(self.h, int(which), _p(o))
        return o

    def norms(self):
        o = np.zeros(3); lib().or_norms(self.h, _p(o))
        return dict(E=o[0], Z=o[1], P=o[2])

    def spectrum(self):
        nb = lib().or_spectrum_len(self.h)
        o = np.zeros(nb); lib().or_spectrum(self.h, _p(o))
        return o

    def diagnose(self, interp=True):
        nb = lib().or_spectrum_len(self.h)
        o = np.zeros(80 + nb); lib().or_diagnose(self.h, 1 if interp else 0, _p(o))
        d = dict(zip(DIAG_KEYS, o[:18].tolist()))
        d["step"] = int(d["step"])
        d["align"] = o[18:21].tolist()
        d["uL3"] = o[21]
        d["pileUp"] = None if math.isnan(o[22]) else o[22]
        d["zCentroid"], d["zExtent"], d["imageGap"], d["zBands"] = o[23], o[24], o[25], int(o[26])
        if not math.isnan(o[27]):
            d["omMaxI"] = o[27]; d["omMaxIpos"] = o[28:31].tolist(); d["omMaxNode"] = o[31]
        d["alignHist"] = o[32:80].reshape(3, 16).tolist()
        d["spectrum"] = o[80:].tolist()
        return d

    def exactError(self):
        o = np.zeros(2)
        if not lib().or_exactError(self.h, _p(o)):
            return None
        return dict(linf=o[0], rel=o[1])
