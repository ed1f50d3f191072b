"""The allocation-free entry point (ppsd_walk_primes_noalloc), through the
C ABI, against the allocating path it mirrors (ppsd_submit_primes +
ppsd_collect / ppsd_collect_raw on a context): the same raw divisors and
filtered factors, on two calls in a row with the second smaller (the device
array keeps the first call's contents), over a raw-hit ring overflow, with a
short output buffer, and its error returns.  Device buffers for the primes
and the output come from the HIP runtime the library is linked to."""
import ctypes as C

import numpy as np
import pytest

import pps_oracle as orc
import ppsd

pytestmark = pytest.mark.gpu

MANY_N = dict(pmin=42070000000000, pmax=42070050000000, kmin=1201, kmax=9999, nmax=2000000)


class Hip:
    """hipMalloc / hipMemcpy / hipFree of libamdhip64, the runtime of libppsd."""

    def __init__(self):
        ppsd.lib()  # loads libamdhip64 as a dependency
        self.h = C.CDLL("libamdhip64.so.7")
        self.h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.h.hipFree.argtypes = [C.c_void_p]
        self.h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.h.hipDeviceSynchronize.argtypes = []

    def alloc(self, nbytes):
        p = C.c_void_p()
        assert self.h.hipMalloc(C.byref(p), max(nbytes, 8)) == 0
        return p

    def upload(self, arr):
        arr = np.ascontiguousarray(arr)
        d = self.alloc(arr.nbytes)
        assert self.h.hipMemcpy(d, arr.ctypes.data, arr.nbytes, 1) == 0  # hipMemcpyHostToDevice
        return d

    def download(self, d, n, dtype):
        out = np.zeros(n, dtype=dtype)
        if n:
            assert self.h.hipMemcpy(out.ctypes.data, d, out.nbytes, 2) == 0  # hipMemcpyDeviceToHost
        return out

    def free(self, d):
        assert self.h.hipFree(d) == 0


@pytest.fixture(scope="module")
def hip():
    return Hip()


def noalloc(hip, params, P_dev, n, cap, raw, out_dev=None):
    """(return code, result count, results) of one call."""
    own = out_dev is None
    if own:
        out_dev = hip.alloc(cap * ppsd.FACTOR_DTYPE.itemsize)
    n_out = C.c_uint64(12345)
    rc = ppsd.lib().ppsd_walk_primes_noalloc(0, C.byref(params), P_dev, n, out_dev, cap, 1 if raw else 0, None,
                                             C.byref(n_out))
    res = hip.download(out_dev, min(n_out.value, cap), ppsd.FACTOR_DTYPE) if rc == 0 else None
    if own:
        hip.free(out_dev)
    return rc, n_out.value, res


def test_noalloc_equals_submit_and_collect(hip):
    """The pps/many_n_test.sh search on 7e5 primes, then again on its first
    3000 primes: raw divisors and factors equal the context path's."""
    params = ppsd.derive(ppsd.SearchParams(**MANY_N))
    P = orc.primes(MANY_N["pmin"], MANY_N["pmin"] + 2 * 10**7)
    with ppsd.Context(params) as ctx:
        want_raw = ppsd.factor_set(ctx.run_primes(P, raw=True))
        want = ppsd.factor_set(ctx.run_primes(P))
        want_raw_small = ppsd.factor_set(ctx.run_primes(P[:3000], raw=True))
        want_small = ppsd.factor_set(ctx.run_primes(P[:3000]))
    assert len(want_raw) > 50 and want and len(want_raw_small) < len(want_raw)
    d = hip.upload(P)
    try:
        for raw, big, small in ((True, want_raw, want_raw_small), (False, want, want_small)):
            rc, n, res = noalloc(hip, params, d, len(P), 4096, raw)
            assert rc == 0, ppsd.lib().ppsd_last_error()
            assert n == len(res) == len(big) and ppsd.factor_set(res) == big
            rc, n, res = noalloc(hip, params, d, 3000, 4096, raw)  # second, smaller call
            assert rc == 0 and n == len(small) and ppsd.factor_set(res) == small
    finally:
        hip.free(d)


def test_noalloc_ring_overflow_and_short_output(hip):
    """A dense unit (s = 1, kmax 2^30 at p near 2^33: about 600 divisors per
    prime) whose 2000 primes hold more raw divisors than the allocation-free
    ring: it is walked in pieces, and results equal the context path's;
    with an output buffer shorter than the result the count is still the
    total and the written records are results."""
    kw = dict(pmin=2**33 + 10**6, pmax=2**33 + 10**6 + 10**5, kmin=1, kmax=2**30 + 1, nmin=64, nmax=10064, sign=0,
              nstep=1)
    params = ppsd.manual_params(**kw)
    P = orc.primes(kw["pmin"], kw["pmax"])[:2000]
    with ppsd.Context(params) as ctx:
        want_raw = ppsd.factor_set(ctx.run_primes(P, raw=True))
        want = ppsd.factor_set(ctx.run_primes(P))
    assert len(want_raw) > 2**19  # PPSD_NOALLOC_RING_HITS
    d = hip.upload(P)
    try:
        rc, n, res = noalloc(hip, params, d, len(P), len(want_raw) + 10, True)
        assert rc == 0, ppsd.lib().ppsd_last_error()
        assert n == len(want_raw) and ppsd.factor_set(res) == want_raw
        rc, n, res = noalloc(hip, params, d, len(P), len(want) + 10, False)
        assert rc == 0 and n == len(want) and ppsd.factor_set(res) == want
        rc, n, res = noalloc(hip, params, d, 700, 1000, True)  # smaller again, short buffer
        assert rc == 0 and n > 1000 and len(res) == 1000
        assert ppsd.factor_set(res) <= {t for t in want_raw if t[0] in set(P[:700].tolist())}
    finally:
        hip.free(d)


def test_noalloc_errors(hip):
    params = ppsd.derive(ppsd.SearchParams(**MANY_N))
    P = orc.primes(MANY_N["pmin"], MANY_N["pmin"] + 10**5)
    bad = P.copy()
    bad[7] += 1  # even
    d = hip.upload(bad)
    try:
        rc, _, _ = noalloc(hip, params, d, len(bad), 16, False)
        assert rc == ppsd.PPSD_ERR_INVALID_PARAM
        assert b"not odd" in ppsd.lib().ppsd_last_error()
        rc, _, _ = noalloc(hip, params, d, (1 << 24) + 1, 16, False)
        assert rc == ppsd.PPSD_ERR_INVALID_PARAM
        assert b"allocation-free limit" in ppsd.lib().ppsd_last_error()
        rc, n, _ = noalloc(hip, params, d, 0, 16, False)
        assert rc == 0 and n == 0
    finally:
        hip.free(d)
