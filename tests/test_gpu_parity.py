"""Parity of the HIP path (libppsd.so through its C ABI) with the oracle and
the golden fixtures.  Integer work: every comparison is exact set equality."""
import ctypes
import json
import os
import random

import numpy as np
import pytest
import sympy

import pps_oracle as orc
import ppsd
from conftest import GOLDEN, read_lines, windowed_run

pytestmark = pytest.mark.gpu

RUNS = json.load(open(os.path.join(GOLDEN, "runs.json")))
MANY_N = dict(pmin=42070000000000, pmax=42070050000000, kmin=1201, kmax=9999, nmax=2000000)


def golden_set(name):
    return {orc.parse_line(l) for l in read_lines(name)}


def context(batch=0, **kw):
    return ppsd.Context(ppsd.derive(ppsd.SearchParams(**kw)), batch=batch)


def test_device_present():
    assert ppsd.device_count() >= 1


def test_state_vectors_against_pow():
    """2^-(e + w*s) mod p from the device arithmetic, p across 2^32..2^62."""
    rng = random.Random(5)
    P = [sympy.nextprime(rng.randrange(2**32, 2**62)) for _ in range(300)]
    P += [sympy.nextprime(2**32), sympy.prevprime(2**62), sympy.nextprime(2**50), sympy.nextprime(2**61),
          sympy.nextprime(10**12)]
    P = np.array(P, dtype=np.uint64)
    for e, nwin, s in [(64, 0, 32), (76, 3, 32), (100, 7404, 27), (50000, 1000, 32), (100000, 517, 32),
                       (2**31 - 1, 5, 1), (6000000, 31, 17), (1000, 2000, 31)]:
        got = ppsd.debug_state(P, e, nwin, s)
        want = [pow(2, -(e + nwin * s), int(p)) for p in P]
        assert got.tolist() == want, (e, nwin, s)


def test_every_state_stays_canonical():
    """Each of the first 160 states of a walk equals 2^-n mod p exactly (not
    merely mod p): the walk's pre-test relies on canonical states.  Primes
    whose low word is extreme stress the ceiling term of the step.  At s = 32
    the device takes the three-multiply step for primes with
    lo32(p) + hi32(p) < 2^32 and the four-multiply one for the rest: the
    primes on either side of that bound cover both."""
    rng = random.Random(6)
    P = [sympy.nextprime(rng.randrange(2**32, 2**62)) for _ in range(100)]
    for hi in (1, 2**20, 2**29, 2**30 - 1):
        P += [sympy.prevprime(hi * 2**32 + 2**32 - 1), sympy.nextprime(hi * 2**32)]
        edge = hi * 2**32 + 2**32 - 1 - hi  # lo32 + hi32 == 2^32 - 1
        P += [sympy.prevprime(edge + 1), sympy.nextprime(edge)]
    P += [sympy.prevprime(2**62)]
    fused = [(p & 0xFFFFFFFF) + (p >> 32) < 2**32 for p in P]
    assert 20 < sum(fused) < len(P) - 5
    P = np.array(P, dtype=np.uint64)
    for s in (32, 31, 7, 1):
        for nwin in range(160):
            got = ppsd.debug_state(P, 100, nwin, s)
            want = [pow(2, -(100 + nwin * s), int(p)) for p in P]
            assert got.tolist() == want, (s, nwin)


def _fused_edge(h: int, s: int) -> int:
    """Largest p = h*2^32 + lo with ((p-1) >> s) + lo <= 2^32 - 1."""
    lo, hi = 1, 2**32 - 1
    ok = lambda x: ((h * 2**32 + x - 1) >> s) + x <= 2**32 - 1  # noqa: E731
    while lo < hi:
        mid = (lo + hi + 1) // 2
        lo, hi = (mid, hi) if ok(mid) else (lo, mid - 1)
    return h * 2**32 + lo


@pytest.mark.parametrize("s", [31, 24, 13, 7, 1])
def test_fused_step_bound_each_window(s):
    """The three-multiply step at s < 32 is exact while
    ((p-1) >> s) + lo32(p) < 2^32, and the device state kernel takes it for
    exactly those primes.  Primes on both sides of that bound, for several
    high words, plus random primes small enough for it, walk their first
    states to 2^-n mod p exactly."""
    rng = random.Random(100 + s)
    P = []
    for h in sorted({1, min(2 ** (s - 1), 2**30 - 1), min(2**s - 1, 2**30 - 1)}):
        edge = _fused_edge(h, s)
        P += [sympy.prevprime(edge + 1), sympy.nextprime(edge)]
    top = min(2 ** (32 + s), 2**62)
    P += [sympy.nextprime(rng.randrange(2**32, top)) for _ in range(40)]
    fused = [((p - 1) >> s) + (p & 0xFFFFFFFF) < 2**32 for p in P]
    assert any(fused) and not all(fused)
    P = np.array(P, dtype=np.uint64)
    for nwin in list(range(0, 40)) + [97, 160]:
        got = ppsd.debug_state(P, 100, nwin, s)
        want = [pow(2, -(100 + nwin * s), int(p)) for p in P]
        assert got.tolist() == want, (s, nwin)


@pytest.mark.parametrize("sign", [1, -1, 0])
def test_walk_across_fused_step_bound(sign):
    """A p-range across the run of primes that fail the three-multiply
    step's bound at the top of a 2^32 block (here s = 25, so the run is the
    top 2^19 of the block, about 18,000 primes), with primes that pass it on
    both sides: the plain walk (device and host primes) and the batched walk
    equal the oracle's divisor and factor sets."""
    B = 2**44 + 2**32
    kw = dict(pmin=B - 2**20, pmax=B + 2**19, kmin=3, kmax=1000001, nmin=100000, nmax=200000, sign=sign)
    params = ppsd.derive(ppsd.SearchParams(**kw))
    s = params.nstep
    pr = orc.derive(**kw)
    P = orc.primes(kw["pmin"], kw["pmax"])
    slow = sum(1 for p in P.tolist() if ((p - 1) >> s) + (p & 0xFFFFFFFF) >= 2**32)
    assert s == 25 and 10000 < slow < len(P) - 10000, (s, slow, len(P))
    want_raw = orc.hit_set(orc.walk(pr, P, filter=False, threads=8))
    want = orc.hit_set(orc.walk(pr, P, filter=True, threads=8))
    assert len(want_raw) > 100
    with ppsd.Context(params) as ctx:
        raw, cnt, _ = ctx.run_range(kw["pmin"], kw["pmax"], raw=True)
        got, _, _ = ctx.run_range(kw["pmin"], kw["pmax"])
        host = ctx.run_primes(P, raw=True)
        braw, bunit, bcnt, _ = ctx.batch_run([params], raw=True)
    assert cnt == len(P) and int(bcnt[0]) == len(P)
    assert ppsd.factor_set(raw) == want_raw
    assert ppsd.factor_set(host) == want_raw
    assert ppsd.factor_set(braw) == want_raw
    assert ppsd.factor_set(got) == want


def test_walk_across_fused_step_bound_s32():
    """As above at s = 32, where the bound is lo32(p) + hi32(p) < 2^32: near
    2^56 the failing run is the top 2^24 of a 2^32 block, starting at
    R = 2^56 + 2^32 - 2^24.  The range covers 2^21 of primes below R and
    2^22 above it; kmax 2^25 - 1 keeps s at 32."""
    R = 2**56 + 2**32 - 2**24
    kw = dict(pmin=R - 2**21, pmax=R + 2**22, kmin=3, kmax=2**25 - 1, nmin=100000, nmax=1100000, sign=1)
    params = ppsd.derive(ppsd.SearchParams(**kw))
    pr = orc.derive(**kw)
    P = orc.primes(kw["pmin"], kw["pmax"])
    slow = sum(1 for p in P.tolist() if (p & 0xFFFFFFFF) + (p >> 32) >= 2**32)
    assert params.nstep == 32 and 50000 < slow < len(P) - 30000, (params.nstep, slow, len(P))
    want_raw = orc.hit_set(orc.walk(pr, P, filter=False, threads=16))
    assert len(want_raw) > 20
    with ppsd.Context(params) as ctx:
        raw, cnt, _ = ctx.run_range(kw["pmin"], kw["pmax"], raw=True)
        braw, _, bcnt, _ = ctx.batch_run([params], raw=True)
    assert cnt == len(P) and int(bcnt[0]) == len(P)
    assert ppsd.factor_set(raw) == want_raw
    assert ppsd.factor_set(braw) == want_raw


_S32_RUN_CACHE = {}


def _s32_run_near_2p61(sign):
    """Parameters and oracle sets for a p-range across the start of the run
    of primes that fail the three-multiply bound in the 2^32 block right
    above 2^61 (lo32(p) > 2^32 - 1 - 2^29: the top eighth of the block).
    kmax = 2^29 - 1 keeps s at 32 (33 before the reference's clamp) and gives
    each prime about 2.3e-4 divisors per sign over the 2e6 exponents."""
    if sign not in _S32_RUN_CACHE:
        R = 2**61 + 2**32 - 2**29
        kw = dict(pmin=R - 2**21, pmax=R + 2**22, kmin=3, kmax=2**29 - 1, nmin=1000000, nmax=3000000, sign=sign)
        pr = orc.derive(**kw)
        P = orc.primes(kw["pmin"], kw["pmax"])
        want_raw = orc.hit_set(orc.walk(pr, P, filter=False, threads=16))
        want = {t for t in want_raw if orc.lib().orc_test_factor(ctypes.byref(pr), t[1], t[2], t[3])}
        _S32_RUN_CACHE[sign] = (kw, P, want_raw, want)
    return _S32_RUN_CACHE[sign]


@pytest.mark.parametrize("form", [1, 2, 3])
@pytest.mark.parametrize("sign", [1, -1, 0])
def test_four_multiply_loop_s32_near_2p61(sign, form):
    """Every s = 32 walk loop on its four-multiply branch, with hits: each
    sign (+1, -1, both) and each pre-test form forced (1 exact, 2 range, 3
    odd-low) over a range where about a third of the primes fail the
    three-multiply bound, so whole waves take the four-multiply loop and
    others the fused one.  Raw divisors and kept factors as the oracle's."""
    kw, P, want_raw, want = _s32_run_near_2p61(sign)
    params = ppsd.derive(ppsd.SearchParams(**kw))
    slow = sum(1 for p in P.tolist() if (p & 0xFFFFFFFF) + (p >> 32) >= 2**32)
    assert params.nstep == 32 and params.nstep_ref == 32 and 30000 < slow < len(P) - 30000
    assert len(want_raw) >= 20 * (2 if sign == 0 else 1) and want
    if sign == 0:
        assert {t[3] for t in want_raw} == {1, -1}
    with ppsd.Context(params) as ctx:
        ctx.set_pretest(form)
        assert ctx.pretest()[1] == form
        raw, cnt, psum = ctx.run_range(kw["pmin"], kw["pmax"], raw=True)
        got, _, _ = ctx.run_range(kw["pmin"], kw["pmax"])
    assert (cnt, psum) == (len(P), int(P.sum(dtype=np.uint64)))
    assert ppsd.factor_set(raw) == want_raw
    assert ppsd.factor_set(got) == want


def test_many_n_raw_and_filtered():
    """pps/many_n_test.sh range: all 296 divisors (whose 3/5/7-free part is
    pps/oldsort.txt) and the 39 factors today's filter keeps; prime count and
    sum as the oracle's."""
    with context(**MANY_N) as ctx:
        raw, cnt, psum = ctx.run_range(MANY_N["pmin"], MANY_N["pmax"], raw=True)
        assert ppsd.factor_set(raw) == golden_set("many_n_raw.txt")
        assert len(raw) == 296
        assert (cnt, psum) == (RUNS["many_n_raw"]["count"], RUNS["many_n_raw"]["sum"])
        got, _, _ = ctx.run_range(MANY_N["pmin"], MANY_N["pmax"])
        assert ppsd.factor_set(got) == golden_set("many_n_filtered.txt")


@pytest.mark.parametrize("r,u,wg", [(1, 4, 256), (1, 8, 128), (1, 16, 256), (2, 4, 64), (2, 8, 256),
                                    (2, 16, 192)])
def test_many_n_every_kernel_shape(r, u, wg):
    with context(**MANY_N) as ctx:
        ctx.set_tuning(r, u, wg)
        raw, _, _ = ctx.run_range(MANY_N["pmin"], MANY_N["pmin"] + 10**7, raw=True)
        want = {t for t in golden_set("many_n_raw.txt") if t[0] < MANY_N["pmin"] + 10**7}
        assert ppsd.factor_set(raw) == want


def test_many_n_host_primes():
    """check_ns-style path: primes supplied by the host, several batches."""
    lo, hi = MANY_N["pmin"], MANY_N["pmin"] + 5 * 10**6
    P = orc.primes(lo, hi)
    with context(batch=50000, **MANY_N) as ctx:
        raw = ctx.run_primes(P, raw=True)
    want = {t for t in golden_set("many_n_raw.txt") if t[0] < hi}
    assert ppsd.factor_set(raw) == want


def test_hit_ring_overflow_rewalks():
    with context(**MANY_N) as ctx:
        ctx.set_hit_capacity(1)
        raw, _, _ = ctx.run_range(MANY_N["pmin"], MANY_N["pmax"], raw=True)
        assert ppsd.factor_set(raw) == golden_set("many_n_raw.txt")
        assert ctx.stats().retries > 0


def _golden_run(tag, name):
    r = RUNS[tag]
    kw = dict(r["params"])
    with context(**kw) as ctx:
        got, cnt, psum = ctx.run_range(r["params"]["pmin"], r["params"]["pmax"])
        st = ctx.stats()
    assert ppsd.factor_set(got) == golden_set(name)
    assert (cnt, psum) == (r["count"], r["sum"])
    return st


def test_cfg1_full_range():
    """-p1e12 -P1000100M -k3 -K9999 -n100 -N200000 (s = 27), every factor."""
    st = _golden_run("cfg1", "cfg1_factors.txt")
    assert st.primes == RUNS["cfg1"]["count"]


@pytest.mark.parametrize("tag", ["cfg2_windows", "cfg3_windows"])
def test_windows_of_the_large_searches(tag):
    """The 2^50 search (k 1201..9999, n to 6e6) and the 2^61 both-signs
    search (k 3..100001, n to 1e7) over p-windows around divisors: every
    raw divisor and every kept factor of each window, and the windows' prime
    count and sum, as the oracle's complete walk of them."""
    kw, wins = windowed_run(tag)
    raw, got, cnt, psum = set(), set(), 0, 0
    with context(**kw) as ctx:
        for a, b in wins:
            r, c, s = ctx.run_range(a, b, raw=True)
            f, _, _ = ctx.run_range(a, b)
            raw |= ppsd.factor_set(r)
            got |= ppsd.factor_set(f)
            cnt += c
            psum = (psum + s) % 2**64
    assert raw == golden_set(tag + "_raw.txt") and len(raw) >= 20
    assert got == golden_set(tag + ".txt") and got
    assert (cnt, psum) == (RUNS[tag]["count"], RUNS[tag]["sum"])


def _random_config(rng):
    """A config inside the reference's validation domain."""
    pmin = rng.choice([rng.randrange(2**40, 2**44), rng.randrange(2**44, 2**50)])
    kmax = rng.randrange(3, min(2**22, pmin * pmin >> 63))
    kmin = rng.randrange(1, kmax + 1)
    nmin = rng.randrange(64, 5000)
    nmax = nmin + rng.randrange(0, 400000)
    sign = rng.choice((1, -1, 0))
    ks = rng.choice([(0, 0), (0, 0), (6, 3)])
    return dict(pmin=pmin, pmax=pmin + rng.randrange(1000, 40000), kmin=kmin, kmax=kmax, nmin=nmin, nmax=nmax,
                sign=sign, kstep=ks[0], koffset=ks[1])


def _compare(pr_orc, params, seed, batch=0):
    P = orc.primes(pr_orc.pmin, pr_orc.pmax)
    want_raw = orc.hit_set(orc.walk(pr_orc, P, filter=False, threads=8))
    want = orc.hit_set(orc.walk(pr_orc, P, filter=True, threads=8))
    with ppsd.Context(params, batch=batch) as ctx:
        ctx.set_tuning(r=1 + seed % 2, u=(4, 8, 16)[seed % 3])
        raw, cnt, _ = ctx.run_range(pr_orc.pmin, pr_orc.pmax, raw=True)
        got, _, _ = ctx.run_range(pr_orc.pmin, pr_orc.pmax)
        host = ctx.run_primes(P, raw=True)
    assert cnt == len(P)
    assert ppsd.factor_set(raw) == want_raw
    assert ppsd.factor_set(host) == want_raw
    assert ppsd.factor_set(got) == want
    return len(want_raw)


@pytest.mark.parametrize("seed", range(32))
def test_random_configs_vs_oracle(seed):
    rng = random.Random(1000 + seed)
    kw = _random_config(rng)
    pr = orc.derive(**kw)
    if pr.status != 0:
        with pytest.raises(ppsd.PPSDError):
            ppsd.derive(ppsd.SearchParams(**kw))
        return
    _compare(pr, ppsd.derive(ppsd.SearchParams(**kw)), seed)


@pytest.mark.parametrize("seed", range(32))
def test_dense_hit_configs_vs_oracle(seed):
    """Units outside the reference's validation (small p for their kmax) where
    most lanes of a wave hit: compaction, ring and filter under load."""
    rng = random.Random(2000 + seed)
    pmin = rng.randrange(2**32 + 1, 2**36)
    # small windows leave room for a large kmax: hits per prime ~ n-range * kmax / 2p
    s = rng.randrange(1, 15) if seed < 14 or (seed >= 16 and seed % 8 < 6) else 32 - seed % 2
    kmax = max(1, min((pmin - 1) >> (s - 1), rng.randrange(2**12, 2**26)))
    kmin = rng.randrange(1, kmax + 1)
    nmin = rng.randrange(64, 3000)
    nmax = nmin + rng.randrange(1, 20000)
    sign = rng.choice((1, -1, 0))
    pmax = pmin + rng.randrange(500, 20000)
    kw = dict(pmin=pmin, pmax=pmax, kmin=kmin, kmax=kmax, nmin=nmin, nmax=nmax, sign=sign, nstep=s)
    pr = orc.OrcParams(kstep=2, koffset=1, nstep_ref=s, **kw)
    params = ppsd.manual_params(**kw)
    n = _compare(pr, params, seed, batch=rng.choice([0, 257, 4096]))
    if s < 15:
        assert n > 0


@pytest.mark.parametrize("sign", (1, -1))
def test_filter_high_k_vs_oracle(sign):
    """k up to 2^62 (s = 1, p just below 2^62): the khigh guards of
    test_factor and its wrapping 64-bit pre-tests (pps/app.c:171-179) on the
    device, against the oracle's restatement of them."""
    pmin, pmax = 2**62 - 20000, 2**62
    kw = dict(pmin=pmin, pmax=pmax, kmin=2**46, kmax=pmin - 1, nmin=64, nmax=1000, sign=sign, nstep=1)
    pr = orc.OrcParams(kstep=2, koffset=1, nstep_ref=1, **kw)
    P = orc.primes(pmin, pmax)
    want = orc.hit_set(orc.walk(pr, P, filter=True, threads=8))
    with ppsd.Context(ppsd.manual_params(**kw)) as ctx:
        got, cnt, _ = ctx.run_range(pmin, pmax)
    assert cnt == len(P)
    assert ppsd.factor_set(got) == want
    assert sum(1 for t in want if t[1] >= 2**61) > 100


def test_empty_inputs():
    # nmin == nmax: no exponent in range (the reference rejects it: nstep 1)
    params = ppsd.manual_params(MANY_N["pmin"], MANY_N["pmax"], 1201, 9999, 76, 76, nstep=32)
    with ppsd.Context(params) as ctx:
        got, cnt, _ = ctx.run_range(MANY_N["pmin"], MANY_N["pmin"] + 10**6)
        assert len(got) == 0 and cnt > 0
    with context(**MANY_N) as ctx:
        assert ctx.submit_range(MANY_N["pmin"], MANY_N["pmin"]) == MANY_N["pmin"]
        got, cnt, _ = ctx.collect()
        assert len(got) == 0 and cnt == 0
        ctx.submit_primes(np.zeros(0, dtype=np.uint64))
        got, cnt, _ = ctx.collect()
        assert len(got) == 0 and cnt == 0


def test_errors():
    with context(**MANY_N) as ctx:
        with pytest.raises(ppsd.PPSDError) as e:
            ctx.submit_primes(np.array([MANY_N["pmin"] + 2], dtype=np.uint64))  # even
        assert e.value.code == ppsd.PPSD_ERR_INVALID_PARAM
        with pytest.raises(ppsd.PPSDError) as e:
            ctx.submit_range(MANY_N["pmin"] - 10, MANY_N["pmin"] + 10)
        assert e.value.code == ppsd.PPSD_ERR_INVALID_PARAM
        ctx.submit_range(MANY_N["pmin"], MANY_N["pmin"] + 1000)
        with pytest.raises(ppsd.PPSDError) as e:
            ctx.submit_range(MANY_N["pmin"], MANY_N["pmin"] + 1000)
        assert e.value.code == ppsd.PPSD_ERR_STATE
        ctx.collect()
        with pytest.raises(ppsd.PPSDError) as e:
            ctx.set_tuning(r=3)
        assert e.value.code == ppsd.PPSD_ERR_INVALID_PARAM
