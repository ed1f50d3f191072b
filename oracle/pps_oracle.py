"""ctypes binding of the CPU oracle (liboracle.so, built from pps_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  The product never uses it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")
REF_FACTOR_PROTH = os.path.join(_HERE, "_ref", "factor_proth")


class OrcParams(C.Structure):
    _fields_ = [
        ("pmin", C.c_uint64), ("pmax", C.c_uint64),
        ("kmin", C.c_uint64), ("kmax", C.c_uint64),
        ("nmin", C.c_uint32), ("nmax", C.c_uint32),
        ("kstep", C.c_uint32), ("koffset", C.c_uint32),
        ("sign", C.c_int32),
        ("nstep_ref", C.c_uint32), ("nstep", C.c_uint32),
        ("status", C.c_int32), ("msg", C.c_char * 160),
    ]


HIT_DTYPE = np.dtype([("p", "<u8"), ("k", "<u8"), ("n", "<u4"), ("c", "<i4")])
_u64p = C.POINTER(C.c_uint64)


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        P = C.POINTER(OrcParams)
        L.orc_derive.argtypes = [P]
        L.orc_derive.restype = C.c_int
        L.orc_isqrt.argtypes = [C.c_uint64]
        L.orc_isqrt.restype = C.c_uint64
        L.orc_is_prime.argtypes = [C.c_uint64]
        L.orc_is_prime.restype = C.c_int
        L.orc_primes.argtypes = [C.c_uint64, C.c_uint64, _u64p, C.c_uint64]
        L.orc_primes.restype = C.c_int64
        L.orc_count_primes.argtypes = [C.c_uint64, C.c_uint64, C.c_int, _u64p, _u64p]
        L.orc_count_primes.restype = C.c_int
        L.orc_count_primes_many.argtypes = [C.c_uint32, _u64p, _u64p, C.c_int, _u64p, _u64p]
        L.orc_count_primes_many.restype = C.c_int
        L.orc_neg_inv64.argtypes = [C.c_uint64]
        L.orc_neg_inv64.restype = C.c_uint64
        L.orc_inv2pow.argtypes = [C.c_uint64, C.c_uint32]
        L.orc_inv2pow.restype = C.c_uint64
        L.orc_step.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint32]
        L.orc_step.restype = C.c_uint64
        L.orc_try_all_factors.argtypes = [C.c_uint64, C.c_uint32, C.c_int]
        L.orc_try_all_factors.restype = C.c_int
        L.orc_test_factor.argtypes = [P, C.c_uint64, C.c_uint32, C.c_int]
        L.orc_test_factor.restype = C.c_int
        L.orc_check_hits.argtypes = [P, C.c_void_p, C.c_uint64, C.c_int, C.c_void_p]
        L.orc_check_hits.restype = C.c_int64
        L.orc_walk.argtypes = [P, _u64p, C.c_uint64, C.c_void_p, C.c_uint64, C.c_int, C.c_int]
        L.orc_walk.restype = C.c_int64
        L.orc_run_range.argtypes = [P, C.c_uint64, C.c_uint64, C.c_void_p, C.c_uint64, C.c_int, C.c_int,
                                    _u64p, _u64p]
        L.orc_run_range.restype = C.c_int64
        _lib = L
    return _lib


def derive(pmin=0, pmax=0, kmin=0, kmax=0, nmin=0, nmax=0, kstep=0, koffset=0, sign=1) -> OrcParams:
    p = OrcParams(pmin=pmin, pmax=pmax, kmin=kmin, kmax=kmax, nmin=nmin, nmax=nmax, kstep=kstep,
                  koffset=koffset, sign=sign)
    lib().orc_derive(C.byref(p))
    return p


def primes(p0: int, p1: int) -> np.ndarray:
    n = lib().orc_primes(p0, p1, None, 0)
    out = np.zeros(n, dtype=np.uint64)
    if n:
        lib().orc_primes(p0, p1, out.ctypes.data_as(_u64p), n)
    return out


def count_primes(p0: int, p1: int, threads: int = 8):
    """(count, sum mod 2^64) of the primes in [p0, p1), for large ranges."""
    c, s = C.c_uint64(), C.c_uint64()
    rc = lib().orc_count_primes(p0, p1, threads, C.byref(c), C.byref(s))
    if rc:
        raise ValueError("orc_count_primes: range outside its domain")
    return c.value, s.value


def count_primes_many(ranges, threads: int = 8):
    """[(count, sum mod 2^64)] of the primes in each [p0, p1) of ranges,
    sieved together (one table of sieving primes for all)."""
    n = len(ranges)
    p0 = np.array([r[0] for r in ranges], dtype=np.uint64)
    p1 = np.array([r[1] for r in ranges], dtype=np.uint64)
    c = np.zeros(n, dtype=np.uint64)
    s = np.zeros(n, dtype=np.uint64)
    if n and lib().orc_count_primes_many(n, p0.ctypes.data_as(_u64p), p1.ctypes.data_as(_u64p), threads,
                                         c.ctypes.data_as(_u64p), s.ctypes.data_as(_u64p)):
        raise ValueError("orc_count_primes_many: a range outside its domain")
    return [(int(a), int(b)) for a, b in zip(c, s)]


def test_factor_mask(pr: OrcParams, hits) -> np.ndarray:
    """orc_test_factor (pps/app.c:150-186) of every (k, n, c) of hits."""
    f = lib().orc_test_factor
    return np.array([f(C.byref(pr), int(h["k"]), int(h["n"]), int(h["c"])) != 0 for h in hits], dtype=bool)


BAD_PRIME, BAD_DIVIDES, BAD_K, BAD_N, BAD_SIGN, KEPT = 1, 2, 4, 8, 16, 32


def check_hits(pr: OrcParams, hits: np.ndarray, threads: int = 8) -> np.ndarray:
    """Per hit (HIT_DTYPE): the BAD_* bits of what is wrong with it as a
    divisor of the search pr (p not prime, not a divisor, k, n or sign out of
    range), or KEPT when it is valid and test_factor keeps it."""
    hits = np.ascontiguousarray(hits, dtype=HIT_DTYPE)
    flags = np.zeros(len(hits), dtype=np.uint8)
    if len(hits):
        lib().orc_check_hits(C.byref(pr), hits.ctypes.data_as(C.c_void_p), len(hits), threads,
                             flags.ctypes.data_as(C.c_void_p))
    return flags


def walk(pr: OrcParams, P: np.ndarray, filter: bool = True, threads: int = 1) -> np.ndarray:
    P = np.ascontiguousarray(P, dtype=np.uint64)
    n = lib().orc_walk(C.byref(pr), P.ctypes.data_as(_u64p), len(P), None, 0, int(filter), threads)
    out = np.zeros(n, dtype=HIT_DTYPE)
    if n:
        lib().orc_walk(C.byref(pr), P.ctypes.data_as(_u64p), len(P), out.ctypes.data_as(C.c_void_p), n,
                       int(filter), threads)
    return out


def run_range(pr: OrcParams, p0: int, p1: int, filter: bool = True, threads: int = 1):
    """(hits, prime count, prime sum) for primes in [p0, p1)."""
    P = primes(p0, p1)
    return walk(pr, P, filter, threads), len(P), int(P.sum(dtype=np.uint64)) if len(P) else 0


def hit_set(arr) -> set:
    return {(int(x["p"]), int(x["k"]), int(x["n"]), int(x["c"])) for x in arr}


def parse_line(line: str):
    """'p | k*2^n+c' -> (p, k, n, c)"""
    p, rest = line.strip().split(" | ")
    k, e = rest.split("*2^")
    return int(p), int(k), int(e[:-2]), int(e[-2:])
