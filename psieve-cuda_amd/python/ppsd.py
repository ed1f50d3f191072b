"""ctypes binding of libppsd.so (include/ppsd.h): the MI355X PPSieve hot path.

This module is a thin wrapper over the C ABI; all compute runs in the HIP
library.  Importing it fails loudly when libppsd.so has not been built.

Reference semantics (Ken-g6/PSieve-CUDA): the factor set of a run is every
``p | k*2^n+c`` with p prime in [pmin, pmax), k == koffset (mod kstep) in
[kmin, kmax], nmin <= n < nmax, and no prime 3 <= q < 32768 dividing
k*2^n+c (pps/app.c:150-186, pps/factor_proth.c:246-253); lines are formatted
as pps/app.c:134 does.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "libppsd.so")

PPSD_OK = 0
PPSD_ERR_FAILURE = -1
PPSD_ERR_INVALID_PARAM = -178
PPSD_ERR_INSUFFICIENT_RESOURCE = -198
PPSD_ERR_NOT_IMPLEMENTED = -150
PPSD_ERR_STATE = -117


class Params(C.Structure):
    _fields_ = [
        ("pmin", C.c_uint64), ("pmax", C.c_uint64),
        ("kmin", C.c_uint64), ("kmax", C.c_uint64),
        ("qmax", C.c_uint64),
        ("nmin", C.c_uint32), ("nmax", C.c_uint32),
        ("kstep", C.c_uint32), ("koffset", C.c_uint32),
        ("sign", C.c_int32),
        ("nstep_ref", C.c_uint32), ("nstep", C.c_uint32),
        ("status", C.c_int32), ("msg", C.c_char * 160),
    ]


class Stats(C.Structure):
    _fields_ = [
        ("primes", C.c_uint64), ("prime_sum", C.c_uint64),
        ("window_steps", C.c_uint64), ("raw_hits", C.c_uint64),
        ("factors", C.c_uint64), ("walk_launches", C.c_uint64),
        ("walk_ms", C.c_double), ("sieve_ms", C.c_double), ("filter_ms", C.c_double),
        ("retries", C.c_uint64),
    ]


FACTOR_DTYPE = np.dtype([("p", "<u8"), ("k", "<u8"), ("n", "<u4"), ("c", "<i4")])

class ServeConfig(C.Structure):
    _fields_ = [("mem_budget", C.c_uint64), ("slice_ms", C.c_double), ("max_active", C.c_uint32),
                ("raw", C.c_int32)]


class ServeResult(C.Structure):
    _fields_ = [("id", C.c_uint64), ("status", C.c_int32), ("slices", C.c_uint32), ("primes", C.c_uint64),
                ("prime_sum", C.c_uint64), ("factors", C.c_uint64), ("window_steps", C.c_uint64),
                ("t_arrival", C.c_double), ("t_admit", C.c_double), ("t_start", C.c_double), ("t_done", C.c_double)]


class ServeStats(C.Structure):
    _fields_ = [(f, C.c_uint64) for f in ("submitted", "admitted", "completed", "queued", "active", "slices",
                                          "primes", "window_steps", "mem_in_use", "mem_peak", "mem_budget")] + \
               [("busy_s", C.c_double), ("now", C.c_double)]


_P = C.POINTER
_u64p = _P(C.c_uint64)


def _load():
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"libppsd.so not built at {LIB_PATH}: run __graft_entry__.build()")
    lib = C.CDLL(LIB_PATH)
    sig = {
        "ppsd_derive_params": (C.c_int, [_P(Params)]),
        "ppsd_device_count": (C.c_int, []),
        "ppsd_device_clock": (C.c_int, [C.c_int, _P(C.c_uint32)]),
        "ppsd_device_info": (C.c_int, [C.c_int, C.c_char_p, C.c_uint32, _P(C.c_uint32)]),
        "ppsd_open": (C.c_int, [C.c_int, _P(Params), C.c_uint32, _P(C.c_void_p)]),
        "ppsd_batch_size": (C.c_uint32, [C.c_void_p]),
        "ppsd_submit_primes": (C.c_int, [C.c_void_p, _u64p, C.c_uint32]),
        "ppsd_submit_range": (C.c_int, [C.c_void_p, C.c_uint64, C.c_uint64, _u64p]),
        "ppsd_collect": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64, _u64p, _u64p, _u64p]),
        "ppsd_collect_raw": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64, _u64p, _u64p, _u64p]),
        "ppsd_set_tuning": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int]),
        "ppsd_get_tuning": (C.c_int, [C.c_void_p, _P(C.c_int), _P(C.c_int), _P(C.c_int)]),
        "ppsd_set_hit_capacity": (C.c_int, [C.c_void_p, C.c_uint64]),
        "ppsd_set_pretest": (C.c_int, [C.c_void_p, C.c_int]),
        "ppsd_get_pretest": (C.c_int, [C.c_void_p, _P(C.c_int), _P(C.c_int), _P(C.c_int)]),
        "ppsd_walk_form": (C.c_int, [_P(Params), C.c_int, C.c_int, C.c_int, _P(C.c_int), _P(C.c_int)]),
        "ppsd_set_candidates": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint64, C.c_uint64]),
        "ppsd_get_stats": (C.c_int, [C.c_void_p, _P(Stats)]),
        "ppsd_reset_stats": (C.c_int, [C.c_void_p]),
        "ppsd_close": (C.c_int, [C.c_void_p]),
        "ppsd_last_error": (C.c_char_p, []),
        "ppsd_debug_state": (C.c_int, [C.c_int, _u64p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, _u64p]),
        "ppsd_batch_run": (C.c_int, [C.c_void_p, _P(Params), C.c_uint32, C.c_int]),
        "ppsd_batch_prepare": (C.c_int, [C.c_void_p, C.c_uint64]),
        "ppsd_batch_results": (C.c_int, [C.c_void_p, C.c_void_p, _P(C.c_uint32), C.c_uint64, _u64p]),
        "ppsd_batch_unit_primes": (C.c_int, [C.c_void_p, _u64p, _u64p, C.c_uint32]),
        "ppsd_batch_set_tuning": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_uint32, C.c_uint32]),
        "ppsd_batch_get_tuning": (C.c_int, [C.c_void_p, _P(C.c_int), _P(C.c_int), _P(C.c_int),
                                            _P(C.c_uint32), _P(C.c_uint32)]),
        "ppsd_batch_set_hit_capacity": (C.c_int, [C.c_void_p, C.c_uint64]),
        "ppsd_serve_open": (C.c_int, [C.c_int, _P(ServeConfig), _P(C.c_void_p)]),
        "ppsd_serve_submit": (C.c_int, [C.c_void_p, _P(Params), _u64p]),
        "ppsd_serve_next": (C.c_int, [C.c_void_p, C.c_int, _P(ServeResult)]),
        "ppsd_serve_factors": (C.c_int, [C.c_void_p, C.c_uint64, C.c_void_p, C.c_uint64, _u64p]),
        "ppsd_serve_release": (C.c_int, [C.c_void_p, C.c_uint64]),
        "ppsd_serve_get_stats": (C.c_int, [C.c_void_p, _P(ServeStats)]),
        "ppsd_serve_close": (C.c_int, [C.c_void_p]),
        "ppsd_walk_primes_noalloc": (C.c_int, [C.c_int, _P(Params), C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint64,
                                               C.c_int, C.c_void_p, _u64p]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


_lib = None


def lib():
    global _lib
    if _lib is None:
        _lib = _load()
    return _lib


class PPSDError(RuntimeError):
    def __init__(self, code: int, where: str):
        msg = lib().ppsd_last_error().decode(errors="replace")
        super().__init__(f"{where} failed ({code}): {msg}")
        self.code = code


def _check(rc: int, where: str):
    if rc != 0:
        raise PPSDError(rc, where)


@dataclass
class SearchParams:
    """Options with the reference's meanings; 0 = not given."""
    pmin: int = 0
    pmax: int = 0
    kmin: int = 0
    kmax: int = 0
    nmin: int = 0
    nmax: int = 0
    kstep: int = 0
    koffset: int = 0
    sign: int = 1  # +1 Proth, -1 Riesel (-R), 0 both
    qmax: int = 0  # -Q; 0 = every odd prime <= sqrt(pmax - 1)

    def to_c(self) -> Params:
        p = Params()
        for f in ("pmin", "pmax", "kmin", "kmax", "nmin", "nmax", "kstep", "koffset", "sign", "qmax"):
            setattr(p, f, getattr(self, f))
        return p


def derive(sp: SearchParams) -> Params:
    """Apply the reference's parameter rules; raises PPSDError when invalid."""
    p = sp.to_c()
    _check(lib().ppsd_derive_params(C.byref(p)), "ppsd_derive_params")
    return p


def manual_params(pmin, pmax, kmin, kmax, nmin, nmax, sign=1, nstep=32, kstep=2, koffset=1) -> Params:
    """Parameters set directly (no reference validation), for work units the
    walk supports but the reference rejects; ppsd_open checks the domain."""
    return Params(pmin=pmin, pmax=pmax, kmin=kmin, kmax=kmax, nmin=nmin, nmax=nmax, kstep=kstep,
                  koffset=koffset, sign=sign, nstep_ref=nstep, nstep=nstep, status=0)


def device_count() -> int:
    return lib().ppsd_device_count()


def device_info(device: int = 0):
    """(architecture name, compute units) of a device."""
    buf = C.create_string_buffer(128)
    cus = C.c_uint32()
    _check(lib().ppsd_device_info(device, buf, len(buf), C.byref(cus)), "ppsd_device_info")
    return buf.value.decode(), cus.value


def device_clock_mhz(device: int = 0) -> int:
    """Peak shader clock of a device (its reported clock rate), MHz."""
    mhz = C.c_uint32()
    _check(lib().ppsd_device_clock(device, C.byref(mhz)), "ppsd_device_clock")
    return mhz.value


def walk_form(params: Params, r: int, u: int, batched: bool = False):
    """(form: 1 exact / 2 range / 3 odd-low, chunk length) the automatic
    choice gives a unit at kernel shape (r, u)."""
    f, c = C.c_int(), C.c_int()
    _check(lib().ppsd_walk_form(C.byref(params), r, u, 1 if batched else 0, C.byref(f), C.byref(c)), "ppsd_walk_form")
    return f.value, c.value


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(_u64p)


class Context:
    """One device context (the reference's per-thread GPU state)."""

    def __init__(self, params: Params, device: int = 0, batch: int = 0):
        self._h = C.c_void_p()
        self.params = params
        _check(lib().ppsd_open(device, C.byref(params), batch, C.byref(self._h)), "ppsd_open")
        self.batch = lib().ppsd_batch_size(self._h)

    def close(self):
        if self._h:
            lib().ppsd_close(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_tuning(self, r: int = 0, u: int = 0, wg: int = 0):
        _check(lib().ppsd_set_tuning(self._h, r, u, wg), "ppsd_set_tuning")

    def tuning(self):
        r, u, wg = C.c_int(), C.c_int(), C.c_int()
        _check(lib().ppsd_get_tuning(self._h, C.byref(r), C.byref(u), C.byref(wg)), "ppsd_get_tuning")
        return r.value, u.value, wg.value

    def submit_primes(self, P: np.ndarray):
        P = np.ascontiguousarray(P, dtype=np.uint64)
        _check(lib().ppsd_submit_primes(self._h, _ptr(P), len(P)), "ppsd_submit_primes")

    def submit_range(self, p0: int, p1: int) -> int:
        nxt = C.c_uint64()
        _check(lib().ppsd_submit_range(self._h, p0, p1, C.byref(nxt)), "ppsd_submit_range")
        return nxt.value

    def _collect(self, fn, where):
        n, np_, ps = C.c_uint64(), C.c_uint64(), C.c_uint64()
        _check(fn(self._h, None, 0, C.byref(n), C.byref(np_), C.byref(ps)), where)
        out = np.zeros(n.value, dtype=FACTOR_DTYPE)
        if n.value:
            _check(fn(self._h, out.ctypes.data_as(C.c_void_p), n.value, C.byref(n), C.byref(np_), C.byref(ps)), where)
        return out, np_.value, ps.value

    def collect(self):
        """(factors, prime count, prime sum) of the outstanding submission."""
        return self._collect(lib().ppsd_collect, "ppsd_collect")

    def collect_raw(self):
        return self._collect(lib().ppsd_collect_raw, "ppsd_collect_raw")

    def set_pretest(self, mode: int):
        """0 automatic, 1 exact odd-part test, 2 range test."""
        _check(lib().ppsd_set_pretest(self._h, mode), "ppsd_set_pretest")

    def pretest(self):
        """(mode set, form the plain walk uses: 1 exact / 2 range / 3 odd low word, its chunk u)."""
        m, f, u = C.c_int(), C.c_int(), C.c_int()
        _check(lib().ppsd_get_pretest(self._h, C.byref(m), C.byref(f), C.byref(u)), "ppsd_get_pretest")
        return m.value, f.value, u.value

    def set_hit_capacity(self, cap: int):
        _check(lib().ppsd_set_hit_capacity(self._h, cap), "ppsd_set_hit_capacity")

    def set_candidates(self, bitmap, n0: int = 0, b0: int = 0):
        """Candidate list (a 2-D uint8 array: rows n0.., bit b-b0 for k =
        kstep*b + koffset), or None to clear."""
        if bitmap is None:
            _check(lib().ppsd_set_candidates(self._h, None, 0, 0, 0, 0), "ppsd_set_candidates")
            return
        bm = np.ascontiguousarray(bitmap, dtype=np.uint8)
        _check(lib().ppsd_set_candidates(self._h, bm.ctypes.data_as(C.c_void_p), n0, bm.shape[0], b0, bm.shape[1]),
               "ppsd_set_candidates")

    def stats(self) -> Stats:
        st = Stats()
        _check(lib().ppsd_get_stats(self._h, C.byref(st)), "ppsd_get_stats")
        return st

    def reset_stats(self):
        _check(lib().ppsd_reset_stats(self._h), "ppsd_reset_stats")

    def run_range(self, p0: int, p1: int, raw: bool = False):
        """Every factor of [p0, p1) with on-device primes; returns
        (factors, prime count, prime sum)."""
        parts, cnt, psum = [], 0, 0
        p = p0
        while p < p1:
            nxt = self.submit_range(p, p1)
            f, c, s = self.collect_raw() if raw else self.collect()
            parts.append(f)
            cnt += c
            psum = (psum + s) & (2**64 - 1)
            p = nxt
        out = np.concatenate(parts) if parts else np.zeros(0, dtype=FACTOR_DTYPE)
        return out, cnt, psum

    def run_primes(self, P: np.ndarray, raw: bool = False):
        """Every factor for host-supplied primes, batch by batch."""
        P = np.ascontiguousarray(P, dtype=np.uint64)
        parts = []
        for i in range(0, len(P), self.batch):
            self.submit_primes(P[i:i + self.batch])
            f, _, _ = self.collect_raw() if raw else self.collect()
            parts.append(f)
        return np.concatenate(parts) if parts else np.zeros(0, dtype=FACTOR_DTYPE)

    # ---- batched search: many independent units in one persistent launch
    def batch_prepare(self, pmax_all: int):
        """The batched path's one-time set-up for units with pmax <= pmax_all."""
        _check(lib().ppsd_batch_prepare(self._h, pmax_all), "ppsd_batch_prepare")

    def batch_run(self, units, raw: bool = False):
        """Search every unit (a sequence of Params) completely; returns
        (hits, unit index per hit, primes per unit, prime sum per unit)."""
        arr = (Params * max(1, len(units)))(*units)
        _check(lib().ppsd_batch_run(self._h, arr, len(units), 1 if raw else 0), "ppsd_batch_run")
        n = C.c_uint64()
        _check(lib().ppsd_batch_results(self._h, None, None, 0, C.byref(n)), "ppsd_batch_results")
        out = np.zeros(n.value, dtype=FACTOR_DTYPE)
        unit = np.zeros(n.value, dtype=np.uint32)
        if n.value:
            _check(lib().ppsd_batch_results(self._h, out.ctypes.data_as(C.c_void_p),
                                            unit.ctypes.data_as(_P(C.c_uint32)), n.value, C.byref(n)),
                   "ppsd_batch_results")
        cnt = np.zeros(len(units), dtype=np.uint64)
        sums = np.zeros(len(units), dtype=np.uint64)
        if len(units):
            _check(lib().ppsd_batch_unit_primes(self._h, _ptr(cnt), _ptr(sums), len(units)),
                   "ppsd_batch_unit_primes")
        return out, unit, cnt, sums

    def set_batch_tuning(self, r=0, u=0, waves=0, seg_min=0, seg_max=0):
        _check(lib().ppsd_batch_set_tuning(self._h, r, u, waves, seg_min, seg_max), "ppsd_batch_set_tuning")

    def batch_tuning(self):
        r, u, w = C.c_int(), C.c_int(), C.c_int()
        a, b = C.c_uint32(), C.c_uint32()
        _check(lib().ppsd_batch_get_tuning(self._h, C.byref(r), C.byref(u), C.byref(w), C.byref(a), C.byref(b)),
               "ppsd_batch_get_tuning")
        return r.value, u.value, w.value, a.value, b.value

    def set_batch_hit_capacity(self, cap: int):
        _check(lib().ppsd_batch_set_hit_capacity(self._h, cap), "ppsd_batch_set_hit_capacity")


class Server:
    """Serving loop on one device (ppsd_serve_*): submit searches at any
    time, take completed ones with their timings and factors."""

    def __init__(self, device: int = 0, mem_budget: int = 0, slice_ms: float = 0.0, max_active: int = 0,
                 raw: bool = False):
        self._h = C.c_void_p()
        cfg = ServeConfig(mem_budget=mem_budget, slice_ms=slice_ms, max_active=max_active, raw=1 if raw else 0)
        _check(lib().ppsd_serve_open(device, C.byref(cfg), C.byref(self._h)), "ppsd_serve_open")

    def submit(self, params: Params) -> int:
        rid = C.c_uint64()
        _check(lib().ppsd_serve_submit(self._h, C.byref(params), C.byref(rid)), "ppsd_serve_submit")
        return rid.value

    def next(self, timeout_ms: int = 100):
        """(ServeResult, factors) of the next completed request, or None."""
        res = ServeResult()
        rc = lib().ppsd_serve_next(self._h, timeout_ms, C.byref(res))
        if rc == 1:
            return None
        _check(rc, "ppsd_serve_next")
        n = C.c_uint64()
        out = np.zeros(res.factors, dtype=FACTOR_DTYPE)
        _check(lib().ppsd_serve_factors(self._h, res.id, out.ctypes.data_as(C.c_void_p), len(out), C.byref(n)),
               "ppsd_serve_factors")
        _check(lib().ppsd_serve_release(self._h, res.id), "ppsd_serve_release")
        return res, out

    def stats(self) -> ServeStats:
        st = ServeStats()
        _check(lib().ppsd_serve_get_stats(self._h, C.byref(st)), "ppsd_serve_get_stats")
        return st

    def close(self):
        if self._h:
            lib().ppsd_serve_close(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def debug_state(P, e: int, nwin: int, s: int, device: int = 0) -> np.ndarray:
    """2^-(e + nwin*s) mod P[i] by the device arithmetic."""
    P = np.ascontiguousarray(P, dtype=np.uint64)
    out = np.zeros_like(P)
    _check(lib().ppsd_debug_state(device, _ptr(P), len(P), e, nwin, s, _ptr(out)), "ppsd_debug_state")
    return out


def format_factor(f) -> str:
    """One ppfactors.txt line, "%llu | %llu*2^%u%+d" (pps/app.c:134)."""
    return f"{int(f['p'])} | {int(f['k'])}*2^{int(f['n'])}{int(f['c']):+d}"


def factor_set(arr) -> set:
    return {(int(x["p"]), int(x["k"]), int(x["n"]), int(x["c"])) for x in arr}
