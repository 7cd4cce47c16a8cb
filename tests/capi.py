"""ctypes binding of the C ABI (include/nslab_hip.h) for the tests and bench.py. It loads the in-tree
build-me-anything.github.io_amd/lib/libnslab_hip.so and raises if it is missing: there is no fallback path."""
import ctypes
import math
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "build-me-anything.github.io_amd")
LIB = os.path.join(PKG, "lib", "libnslab_hip.so")
HEADER = os.path.join(ROOT, "include", "nslab_hip.h")

IC = {"tgv": 0, "tgv2d": 1, "abc": 2, "tubes": 3, "random": 4, "tubesz": 5}
DIAG_KEYS = ["t", "step", "E", "Z", "P", "eps", "omMax", "uMax", "Pspec", "Pphys", "Tnl", "skew", "divMax", "eta",
             "kmaxEta", "Rel", "lambda", "skewIso"]
LAST_KEYS = ["E", "Z", "P", "eps", "Tnl", "Pspec", "omMax", "uMax"]


class Opts(ctypes.Structure):
    _fields_ = [("N", ctypes.c_int32), ("ic", ctypes.c_int32), ("nu", ctypes.c_double), ("cfl", ctypes.c_double),
                ("dt", ctypes.c_double), ("icp", ctypes.c_double * 4)]


class RhsDiag(ctypes.Structure):
    _fields_ = [(k, ctypes.c_double) for k in LAST_KEYS]

    def asdict(self):
        return {k: getattr(self, k) for k in LAST_KEYS}


class Tune(ctypes.Structure):
    _fields_ = [("xpass", ctypes.c_int32), ("lp", ctypes.c_int32), ("xwaves", ctypes.c_int32), ("zfused", ctypes.c_int32),
                ("ptile", ctypes.c_int32), ("xsched", ctypes.c_int32), ("zsched", ctypes.c_int32)]

    def astuple(self):
        return (self.xpass, self.lp, self.xwaves, self.zfused, self.ptile, self.xsched, self.zsched)


class StepRec(ctypes.Structure):
    _fields_ = [("t", ctypes.c_double), ("dt", ctypes.c_double), ("step", ctypes.c_int64), ("ebal", ctypes.c_double),
                ("zbal", ctypes.c_double), ("last", RhsDiag), ("d0", RhsDiag)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError("libnslab_hip.so is not built (run build-me-anything.github.io_amd/build.py): " + LIB)
        L = ctypes.CDLL(LIB)
        dp = ctypes.POINTER(ctypes.c_double)
        vp = ctypes.c_void_p
        L.nsl_create.argtypes = [ctypes.POINTER(Opts), ctypes.POINTER(vp)]
        L.nsl_create_fp32.argtypes = [ctypes.POINTER(Opts), ctypes.POINTER(vp)]
        L.nsl_precision.argtypes = [vp]
        L.nsl_set_ic.argtypes = [vp, ctypes.c_int32, dp]
        L.nsl_destroy.argtypes = [vp]
        L.nsl_destroy.restype = None
        L.nsl_run.argtypes = [vp, ctypes.c_int64, ctypes.c_int32, ctypes.c_double, ctypes.POINTER(StepRec),
                              ctypes.POINTER(ctypes.c_int64)]
        L.nsl_set_clock.argtypes = [vp, ctypes.c_double, ctypes.c_int64, ctypes.c_double]
        L.nsl_get_clock.argtypes = [vp, dp, ctypes.POINTER(ctypes.c_int64), dp]
        L.nsl_last.argtypes = [vp, ctypes.POINTER(RhsDiag)]
        L.nsl_norms.argtypes = [vp, dp]
        L.nsl_get_state.argtypes = [vp, ctypes.POINTER(dp)]
        L.nsl_set_state.argtypes = [vp, ctypes.POINTER(dp)]
        L.nsl_diagnose.argtypes = [vp, ctypes.c_int32, dp]
        L.nsl_get_divmax.argtypes = [vp, dp]
        L.nsl_set_divmax.argtypes = [vp, ctypes.c_double]
        L.nsl_spectrum_len.argtypes = [vp]
        L.nsl_spectrum.argtypes = [vp, dp]
        L.nsl_interp_max.argtypes = [vp, ctypes.c_int32, ctypes.c_int32, dp]
        L.nsl_field.argtypes = [vp, ctypes.c_int32, dp]
        L.nsl_exact_error.argtypes = [vp, dp]
        L.nsl_device_bytes.argtypes = [vp]
        L.nsl_device_bytes.restype = ctypes.c_int64
        L.nsl_fft1.argtypes = [ctypes.c_int32, ctypes.c_int32, ctypes.c_int64, dp, dp]
        L.nsl_spectral_forward.argtypes = [ctypes.c_int32, dp, dp, dp]
        L.nsl_spectral_inverse.argtypes = [ctypes.c_int32, dp, dp, dp]
        L.nsl_factorize.argtypes = [ctypes.c_int32, ctypes.POINTER(ctypes.c_int32), ctypes.c_int32]
        pp = ctypes.POINTER(ctypes.c_void_p)
        i32p = ctypes.POINTER(ctypes.c_int32)
        i64p = ctypes.POINTER(ctypes.c_int64)
        L.nsl_pool_slot_bytes.argtypes = [ctypes.c_int32]
        L.nsl_pool_slot_bytes.restype = ctypes.c_int64
        L.nsl_pool_create.argtypes = [ctypes.c_int32, ctypes.c_int32, pp]
        L.nsl_pool_destroy.argtypes = [vp]
        L.nsl_pool_destroy.restype = None
        L.nsl_pool_info.argtypes = [vp, i32p, i32p, i32p, i32p, i64p]
        L.nsl_pool_admit.argtypes = [vp, ctypes.POINTER(Opts), i32p]
        L.nsl_pool_slot.argtypes = [vp, ctypes.c_int32, pp]
        L.nsl_pool_set_target.argtypes = [vp, ctypes.c_int32, ctypes.c_int64, ctypes.c_int32, ctypes.c_double]
        L.nsl_pool_launch.argtypes = [vp, ctypes.c_int32]
        L.nsl_pool_ready.argtypes = [vp]
        L.nsl_pool_collect.argtypes = [vp, ctypes.POINTER(StepRec), ctypes.c_int32, i64p, i32p]
        L.nsl_pool_release.argtypes = [vp, ctypes.c_int32]
        L.nsl_tune_get.argtypes = [ctypes.c_int32, ctypes.POINTER(Tune)]
        L.nsl_tune_set.argtypes = [ctypes.c_int32, ctypes.POINTER(Tune)]
        L.nsl_tune_candidates.argtypes = [ctypes.c_int32, ctypes.POINTER(Tune), ctypes.c_int32]
        L.nsl_spectrum_noalloc.argtypes = [vp, dp]
        L.nsl_copy_probe.argtypes = [ctypes.c_int64, ctypes.c_int32, dp]
        L.nsl_device_memory.argtypes = [i64p, i64p]
        L.nsl_slice.argtypes = [vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(ctypes.c_float)]
        L.nsl_image_diag.argtypes = [vp, dp]
        L.nsl_cutoff_pile_up.argtypes = [ctypes.c_int32, dp, ctypes.c_int32, dp]
        L.nsl_cutoff_pile_up_eps.argtypes = [ctypes.c_int32, dp, ctypes.c_int32, ctypes.c_double, dp]
        L.nsl_grid_size.argtypes = [vp]
        L.nsl_xpass_plan.argtypes = [ctypes.c_int32, ctypes.POINTER(Tune), ctypes.c_int32, ctypes.c_int64, i64p]
        L.nsl_last_error.restype = ctypes.c_char_p
        L.nsl_version.restype = ctypes.c_char_p
        _lib = L
    return _lib


class NslError(RuntimeError):
    pass


def check(rc):
    if rc < 0:
        raise NslError(lib().nsl_last_error().decode("utf-8"))
    return rc


def _p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def make_opts(N, Re=None, nu=None, cfl=None, dt=None, ic="tgv", icParams=None):
    o = Opts()
    o.N = int(N)
    o.ic = IC[ic] if ic in IC else 99
    o.nu = float(nu) if nu is not None else 1.0 / float(Re or 100)
    o.cfl = float(cfl) if cfl else 0.4
    o.dt = float(dt) if dt else 0.0
    for i, v in enumerate(ic_params(ic, icParams) if ic in IC else [0, 0, 0, 0]):
        o.icp[i] = v
    return o


def ic_params(kind, prm=None):
    """The reference's IC parameter defaults (pocket-wind-tunnel/src/nslab.js:255-274)."""
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


class Solver:
    """createSolver(opts) through the C ABI, with the reference's st/series bookkeeping (nslab.js:188-190, 300-302)."""

    def __init__(self, N=32, Re=None, nu=None, cfl=None, dt=None, ic="tgv", icParams=None, pool=None, fp32=False):
        o = make_opts(N, Re, nu, cfl, dt, ic, icParams)
        h = ctypes.c_void_p()
        self.pool, self.slot = pool, None
        if fp32:   # the GPU runner's fp32 mode (nslab_gpu.py:56-65): complex64 storage
            if pool is not None:
                raise ValueError("FP32 solves are never pool slots")
            check(lib().nsl_create_fp32(ctypes.byref(o), ctypes.byref(h)))
        elif pool is None:
            check(lib().nsl_create(ctypes.byref(o), ctypes.byref(h)))
        else:   # a slot of a pool: createSolver into the slot, then the slot's solver handle
            sl = ctypes.c_int32(-1)
            check(lib().nsl_pool_admit(pool.h, ctypes.byref(o), ctypes.byref(sl)))
            self.slot = sl.value
            check(lib().nsl_pool_slot(pool.h, self.slot, ctypes.byref(h)))
        self.h = h
        self.N = o.N
        self.nu = o.nu
        self.NS = self.N * self.N * (self.N // 2 + 1)
        self.series = {k: [] for k in ["t", "E", "Z", "eps", "omMax", "uMax", "dt", "ebal", "zbal", "Pspec", "Pal"]}
        self.maxEbal = self.maxZbal = self.maxTnl = 0.0
        d = self.last
        self.E0, self.Z0 = d["E"], d["Z"]
        self._record(0.0, 0.0, d)

    def _record(self, t, dt, d):
        s = self.series
        s["t"].append(t); s["E"].append(d["E"]); s["Z"].append(d["Z"]); s["eps"].append(d["eps"])
        s["omMax"].append(d["omMax"]); s["uMax"].append(d["uMax"]); s["dt"].append(dt)
        s["Pspec"].append(d["Pspec"]); s["Pal"].append(d["P"])

    def close(self):
        if getattr(self, "h", None):
            if self.pool is None:
                lib().nsl_destroy(self.h)
            elif self.pool.h:
                check(lib().nsl_pool_release(self.pool.h, self.slot))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _book(self, recs, got):
        for i in range(got):
            r = recs[i]
            d0 = r.d0.asdict()
            self.maxTnl = max(self.maxTnl, abs(d0["Tnl"]) / max(d0["eps"], 1e-300))
            self._record(r.t, r.dt, r.last.asdict())
            self.series["ebal"].append(r.ebal); self.series["zbal"].append(r.zbal)
            self.maxEbal = max(self.maxEbal, r.ebal); self.maxZbal = max(self.maxZbal, r.zbal)

    def run(self, n, tEnd=None, chunk=1 << 20):
        total = 0
        while total < n:
            k = int(min(n - total, chunk))
            recs = (StepRec * k)()
            got = ctypes.c_int64(0)
            check(lib().nsl_run(self.h, k, 0 if tEnd is None else 1, 0.0 if tEnd is None else float(tEnd), recs,
                                ctypes.byref(got)))
            self._book(recs, got.value)
            total += got.value
            if got.value < k:
                break
        return total

    @property
    def precision(self):
        return check(lib().nsl_precision(self.h))

    @property
    def last(self):
        d = RhsDiag()
        check(lib().nsl_last(self.h, ctypes.byref(d)))
        return d.asdict()

    @property
    def clock(self):
        t = ctypes.c_double(); st = ctypes.c_int64(); dt = ctypes.c_double()
        check(lib().nsl_get_clock(self.h, ctypes.byref(t), ctypes.byref(st), ctypes.byref(dt)))
        return dict(t=t.value, step=st.value, dt=dt.value)

    def set_clock(self, t, step, E0):
        check(lib().nsl_set_clock(self.h, float(t), int(step), float(E0)))

    def state(self):
        S = np.zeros((6, self.NS))
        arr = (ctypes.POINTER(ctypes.c_double) * 6)(*[_p(S[i]) for i in range(6)])
        check(lib().nsl_get_state(self.h, arr))
        return S

    def set_state(self, S):
        S = np.ascontiguousarray(S, dtype=np.float64).reshape(6, self.NS)
        arr = (ctypes.POINTER(ctypes.c_double) * 6)(*[_p(S[i]) for i in range(6)])
        check(lib().nsl_set_state(self.h, arr))

    def norms(self):
        o = np.zeros(3)
        check(lib().nsl_norms(self.h, _p(o)))
        return dict(E=o[0], Z=o[1], P=o[2])

    def spectrum(self):
        nb = check(lib().nsl_spectrum_len(self.h))
        o = np.zeros(nb)
        check(lib().nsl_spectrum(self.h, _p(o)))
        return o

    def spectrum_noalloc(self):
        nb = check(lib().nsl_spectrum_len(self.h))
        o = np.zeros(nb)
        check(lib().nsl_spectrum_noalloc(self.h, _p(o)))
        return o

    def diagnose(self, interp=True):
        nb = check(lib().nsl_spectrum_len(self.h))
        o = np.zeros(80 + nb)
        check(lib().nsl_diagnose(self.h, 1 if interp else 0, _p(o)))
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

    def interp_max(self, top=24, iters=12):
        o = np.zeros(5)
        check(lib().nsl_interp_max(self.h, top, iters, _p(o)))
        return dict(omMaxI=o[0], omMaxIpos=o[1:4].tolist(), omMaxNode=o[4])

    def field(self, which):
        o = np.zeros(self.N ** 3)
        check(lib().nsl_field(self.h, int(which), _p(o)))
        return o

    def exactError(self):
        o = np.zeros(2)
        if check(lib().nsl_exact_error(self.h, _p(o))) == 0:
            return None
        return dict(linf=o[0], rel=o[1])

    SLICE = {"vort": 0, "speed": 1, "ke": 2, "q": 3, "stretch": 4, "u": 5, "v": 6, "w": 7, "omx": 8, "omy": 9,
             "omz": 10}

    def slice(self, kind, axis, index):
        out = np.zeros(self.N * self.N, dtype=np.float32)
        k = self.SLICE.get(kind, -1)
        a = 2 if axis == "z" else 1 if axis == "y" else 0
        check(lib().nsl_slice(self.h, k, a, int(index), out.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
        return out

    def image_diag(self):
        o = np.zeros(4)
        check(lib().nsl_image_diag(self.h, _p(o)))
        return {"zCentroid": o[0], "zExtent": o[1], "imageGap": o[2], "zBands": o[3]}

    def device_bytes(self):
        return lib().nsl_device_bytes(self.h)


def device_memory():
    """(free, total) bytes of the current device."""
    f = ctypes.c_int64(0)
    t = ctypes.c_int64(0)
    check(lib().nsl_device_memory(ctypes.byref(f), ctypes.byref(t)))
    return f.value, t.value


def fft1(xr, xi, sign):
    re = np.ascontiguousarray(xr, dtype=np.float64).copy()
    im = np.ascontiguousarray(xi, dtype=np.float64).copy()
    lines = re.size // re.shape[-1]
    check(lib().nsl_fft1(re.shape[-1], int(sign), lines, _p(re), _p(im)))
    return re, im


def spectral_forward(f):
    N = f.shape[0]
    NS = N * N * (N // 2 + 1)
    re = np.zeros(NS); im = np.zeros(NS)
    ff = np.ascontiguousarray(f, dtype=np.float64).ravel()
    check(lib().nsl_spectral_forward(N, _p(ff), _p(re), _p(im)))
    return re, im


def spectral_inverse(N, re, im):
    re = np.ascontiguousarray(re, dtype=np.float64); im = np.ascontiguousarray(im, dtype=np.float64)
    f = np.zeros(N ** 3)
    check(lib().nsl_spectral_inverse(N, _p(re), _p(im), _p(f)))
    return f


def cutoff_pile_up(kc, spec, eps=None):
    """The library's cutoff pile-up of a shell spectrum (host arithmetic; no GPU needed). None below the floor
    100 eps^2 * peak; eps None: the FP64 guard of nsl_cutoff_pile_up, else nsl_cutoff_pile_up_eps."""
    sp = np.ascontiguousarray(spec, dtype=np.float64)
    out = np.zeros(1)
    if eps is None:
        rc = check(lib().nsl_cutoff_pile_up(int(kc), _p(sp), len(sp), _p(out)))
    else:
        rc = check(lib().nsl_cutoff_pile_up_eps(int(kc), _p(sp), len(sp), float(eps), _p(out)))
    return float(out[0]) if rc == 1 else None


def xpass_plan(N, cfg, cus=256, cap_pairs=None):
    """Workgroups the x pass launches for form cfg at grid N (its maxima pairs checked against cap_pairs)."""
    t = Tune(*cfg)
    g = ctypes.c_int64(0)
    check(lib().nsl_xpass_plan(int(N), ctypes.byref(t), int(cus), int(N * N if cap_pairs is None else cap_pairs),
                               ctypes.byref(g)))
    return g.value


def tune_get(N):
    t = Tune()
    check(lib().nsl_tune_get(N, ctypes.byref(t)))
    return t.astuple()


def tune_set(N, cfg=None):
    """Override the kernel forms of N for solvers created afterwards; None drops the override."""
    if cfg is None:
        check(lib().nsl_tune_set(N, None))
    else:
        t = Tune(*cfg)
        check(lib().nsl_tune_set(N, ctypes.byref(t)))


def tune_candidates(N):
    n = check(lib().nsl_tune_candidates(N, None, 0))
    arr = (Tune * max(n, 1))()
    check(lib().nsl_tune_candidates(N, arr, n))
    return [arr[i].astuple() for i in range(n)]


def factorize(N):
    f = (ctypes.c_int32 * 40)()
    k = check(lib().nsl_factorize(int(N), f, 40))
    return list(f[:k])


class Pool:
    """A slot pool (nsl_pool_*): solves of one N stepped together, one launch per pass over all slots."""

    CHUNK = 64

    def __init__(self, N, slots):
        h = ctypes.c_void_p()
        check(lib().nsl_pool_create(int(N), int(slots), ctypes.byref(h)))
        self.h, self.N, self.C = h, int(N), int(slots)
        self.members = {}

    @staticmethod
    def slot_bytes(N):
        return check(lib().nsl_pool_slot_bytes(int(N)))

    def info(self):
        v = [ctypes.c_int32() for _ in range(4)] + [ctypes.c_int64()]
        check(lib().nsl_pool_info(self.h, *[ctypes.byref(x) for x in v]))
        return dict(N=v[0].value, slots=v[1].value, used=v[2].value, graphed=v[3].value, bytes=v[4].value)

    def admit(self, **kw):
        s = Solver(N=self.N, pool=self, **kw)
        self.members[s.slot] = s
        return s

    def release(self, s):
        self.members.pop(s.slot, None)
        s.close()

    def set_target(self, s, n, tEnd=None):
        check(lib().nsl_pool_set_target(self.h, s.slot, int(n), 0 if tEnd is None else 1,
                                        0.0 if tEnd is None else float(tEnd)))

    def launch(self, k=None):
        check(lib().nsl_pool_launch(self.h, int(k or self.CHUNK)))

    def collect(self, k=None):
        """Waits for the launched chunk; books each member's records; returns (steps done per slot, running)."""
        k = int(k or self.CHUNK)
        recs = (StepRec * (self.C * k))()
        done = (ctypes.c_int64 * self.C)()
        running = ctypes.c_int32(0)
        check(lib().nsl_pool_collect(self.h, recs, k, done, ctypes.byref(running)))
        for slot, s in self.members.items():
            if done[slot]:
                base = ctypes.addressof(recs) + slot * k * ctypes.sizeof(StepRec)
                s._book((StepRec * k).from_address(base), done[slot])
        return list(done), running.value

    def run_all(self, k=None):
        """Launch/collect until no slot has steps left; returns the number of chunks."""
        n = 0
        while True:
            self.launch(k)
            _, running = self.collect(k)
            n += 1
            if running == 0:
                return n

    def close(self):
        if getattr(self, "h", None):
            for s in list(self.members.values()):
                s.h = None
            self.members = {}
            lib().nsl_pool_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
