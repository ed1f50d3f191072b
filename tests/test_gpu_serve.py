"""Serving loop (ppsd_serve_*): every request's result equals searching it
alone, whatever else is in flight, under any admission budget and slice
length."""
import random
import threading
import time

import pytest

import ppsd
import ppsd_serve
from conftest import GOLDEN  # noqa: F401
from unitgen import _dense_unit, _oracle, _ref_unit

pytestmark = pytest.mark.gpu


def _drain(srv, n, timeout=120):
    out = {}
    t_end = time.time() + timeout
    while len(out) < n and time.time() < t_end:
        got = srv.next(200)
        if got is not None:
            out[got[0].id] = got
    assert len(out) == n, f"{len(out)} of {n} requests completed"
    return out


def _check(units, results, ids, raw):
    for i, (pr, _) in enumerate(units):
        res, fac = results[ids[i]]
        c, s, want = _oracle(pr, raw)
        assert res.status == 0
        assert (res.primes, res.prime_sum) == (c, s), f"request {i}: prime count/sum"
        assert ppsd.factor_set(fac) == want, f"request {i}"
        assert res.t_arrival <= res.t_admit <= res.t_start <= res.t_done


@pytest.mark.parametrize("slice_ms,budget,raw", [(1.0, 0, True), (0.2, 0, False), (5.0, 1, True)])
def test_serve_burst_vs_oracle(slice_ms, budget, raw):
    """A burst of mixed requests; tiny slices spread each over many slices,
    budget 1 admits one request at a time."""
    rng = random.Random(8000 + int(slice_ms * 10) + budget)
    units = [_dense_unit(rng) if i % 2 else _ref_unit(rng) for i in range(16)]
    with ppsd.Server(slice_ms=slice_ms, mem_budget=budget, raw=raw) as srv:
        ids = [srv.submit(u[1]) for u in units]
        results = _drain(srv, len(units))
        st = srv.stats()
    _check(units, results, ids, raw)
    if budget == 1:
        assert st.mem_peak <= max(1, max(r[0].primes for r in results.values()) * 8 * 1.3 + (1 << 20))
    assert st.completed == len(units) and st.queued == 0 and st.active == 0


def test_serve_small_requests_pass_a_large_one():
    """Continuous batching: requests arriving while a large one is in
    progress join the next slices and finish first, with exact results; the
    large one's result equals searching it alone."""
    # about 1e12 window-steps: a few hundred ms of walking
    big = ppsd.derive(ppsd.SearchParams(pmin=42070000000000, pmax=42070500000000, kmin=1201, kmax=9999,
                                        nmax=2000000))
    rng = random.Random(8100)
    small = [_ref_unit(rng) for _ in range(8)]
    with ppsd.Server(slice_ms=5.0) as srv:
        big_id = srv.submit(big)
        time.sleep(0.02)
        ids = [srv.submit(u[1]) for u in small]
        results = _drain(srv, len(small) + 1)
    order = sorted(results, key=lambda i: results[i][0].t_done)
    assert order[-1] == big_id
    _check(small, results, ids, raw=False)
    with ppsd.Context(big) as ctx:
        alone, cnt, psum = ctx.run_range(big.pmin, big.pmax)
    res, fac = results[big_id]
    assert ppsd.factor_set(fac) == ppsd.factor_set(alone)
    assert (res.primes, res.prime_sum) == (cnt, psum)
    assert res.slices > 1


def test_serve_concurrent_submitters():
    rng = random.Random(8200)
    units = [_dense_unit(rng) for _ in range(12)]
    ids = [None] * len(units)
    with ppsd.Server(slice_ms=1.0, raw=True) as srv:
        def feed(lo, hi):
            for i in range(lo, hi):
                ids[i] = srv.submit(units[i][1])
                time.sleep(0.002)
        ths = [threading.Thread(target=feed, args=(j * 4, j * 4 + 4)) for j in range(3)]
        for t in ths:
            t.start()
        for t in ths:
            t.join()
        results = _drain(srv, len(units))
    _check(units, results, ids, raw=True)


def test_serve_errors_and_close_with_work_pending():
    rng = random.Random(8300)
    pr, params = _ref_unit(rng)
    with ppsd.Server() as srv:
        bad = ppsd.manual_params(pr.pmin, pr.pmax, pr.kmin, pr.kmax, 100, 200, nstep=8)
        bad.kmin = bad.kmax + 1
        with pytest.raises(ppsd.PPSDError) as e:
            srv.submit(bad)
        assert e.value.code == ppsd.PPSD_ERR_INVALID_PARAM
        assert srv.next(0) is None
        big = ppsd.derive(ppsd.SearchParams(pmin=2**50, pmax=2**50 + 5 * 10**10, kmin=1201, kmax=9999,
                                            nmin=50000, nmax=6000000))
        srv.submit(big)  # about 50 s of work: dropped by close
        srv.submit(params)
        got = None
        t_end = time.time() + 30
        while got is None and time.time() < t_end:
            got = srv.next(200)
        assert got is not None and got[0].id == 2
    # close() returned with the large request unfinished


def test_poisson_driver_sub_saturating():
    """The Poisson driver at a low rate: every request completes, results
    spot-checked against searching alone."""
    rep, rec = ppsd_serve.run(rate=8.0, duration=2.0, drain=60.0, slice_ms=20.0, keep_factors=True)
    assert rep["backlog"] == 0 and rep["failed"] == 0 and rep["completed"] == rep["requests"] > 0
    assert rep["latency_s"]["p50"] is not None
    import ppsd_units
    import numpy as np
    rng = np.random.Generator(np.random.PCG64(ppsd_units.MIX_SEED))
    units = [ppsd_units.draw(rng) for _ in range(rep["requests"])]
    checked = 0
    for i, u in enumerate(units):
        if u.work_est > 2e10:
            continue
        with ppsd.Context(u.params) as ctx:
            alone, cnt, psum = ctx.run_range(u.params.pmin, u.params.pmax)
        assert rec[i]["factor_set"] == ppsd.factor_set(alone), i
        assert (rec[i]["primes"], rec[i]["prime_sum"]) == (cnt, psum), i
        checked += 1
    assert checked >= 3
