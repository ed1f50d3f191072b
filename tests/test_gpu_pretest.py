"""The walk's three pre-test forms (ppsd_set_pretest): the exact odd-part
test, the range test K <= kmax*2^(s-1) (or >= p minus that) and the odd-low
test (the low 32 bits of the odd part, minimised over a chunk, <= kmax).
Each is only a filter in front of the same exact re-scan, so every result
must be the same under all three and equal the oracle's; the range and
odd-low forms are forced onto dense units where they fire in most chunks,
which drives the re-scan path hard."""
import json
import os
import random

import pytest

import pps_oracle as orc
import ppsd
import ppsd_units
from conftest import GOLDEN, read_lines

pytestmark = pytest.mark.gpu

RUNS = json.load(open(os.path.join(GOLDEN, "runs.json")))
EXACT, RANGE, ODDLO = 1, 2, 3
FORMS = (EXACT, RANGE, ODDLO)


def golden_set(name):
    return {orc.parse_line(l) for l in read_lines(name)}


@pytest.mark.parametrize("mode", FORMS)
@pytest.mark.parametrize("tag,name", [("cfg2_windows", "cfg2_windows_raw.txt"), ("cfg3_windows", "cfg3_windows_raw.txt"),
                                      ("many_n", "many_n_filtered.txt")])
def test_golden_ranges_each_form(mode, tag, name):
    r = RUNS[tag]
    kw = dict(r["params"])
    wins = [tuple(w) for w in r.get("windows", [(kw["pmin"], kw["pmax"])])]
    raw = name.endswith("_raw.txt")
    got, cnt, psum = set(), 0, 0
    with ppsd.Context(ppsd.derive(ppsd.SearchParams(**kw))) as ctx:
        ctx.set_pretest(mode)
        assert ctx.pretest()[:2] == (mode, mode)
        for a, b in wins:
            f, c, s = ctx.run_range(a, b, raw=raw)
            got |= ppsd.factor_set(f)
            cnt += c
            psum = (psum + s) % 2**64
    assert got == golden_set(name) and got
    assert (cnt, psum) == (r["count"], r["sum"])


@pytest.mark.parametrize("seed", range(12))
def test_dense_units_each_form(seed):
    """Small p for its kmax (many hits per prime), every sign, s from 1 to 32,
    each kernel shape: raw divisors and filtered factors as the oracle's."""
    rng = random.Random(7000 + seed)
    pmin = rng.randrange(2**32 + 1, 2**38)
    s = (1, 3, 9, 17, 31, 32)[seed % 6]
    kmax = max(1, min((pmin - 1) >> (s - 1), rng.randrange(2**8, 2**24)))
    kmin = rng.randrange(1, kmax + 1)
    nmin = rng.randrange(64, 3000)
    nmax = nmin + rng.randrange(1, 30000)
    sign = (1, -1, 0)[seed % 3]
    pmax = pmin + rng.randrange(500, 15000)
    kw = dict(pmin=pmin, pmax=pmax, kmin=kmin, kmax=kmax, nmin=nmin, nmax=nmax, sign=sign, nstep=s)
    pr = orc.OrcParams(kstep=2, koffset=1, nstep_ref=s, **kw)
    P = orc.primes(pmin, pmax)
    want_raw = orc.hit_set(orc.walk(pr, P, filter=False, threads=8))
    want = orc.hit_set(orc.walk(pr, P, filter=True, threads=8))
    with ppsd.Context(ppsd.manual_params(**kw)) as ctx:
        ctx.set_tuning(r=1 + seed % 2, u=(4, 8, 16)[seed % 3])
        for mode in FORMS:
            ctx.set_pretest(mode)
            assert ctx.pretest()[1] == mode
            raw, cnt, _ = ctx.run_range(pmin, pmax, raw=True)
            got, _, _ = ctx.run_range(pmin, pmax)
            assert cnt == len(P)
            assert ppsd.factor_set(raw) == want_raw, mode
            assert ppsd.factor_set(got) == want, mode


def test_automatic_choice():
    """Automatic: the range form where kmax*2^(s-1)/p is tiny (2^61, both
    signs, kmax 1e5: 0.1 false alarms per chunk of 8 windows of a wave,
    each re-checked from the chunk's kept states); the odd-low form where
    kmax/2^32 is (2^50 and
    4.2e13 with kmax 9999: 1.2e-3 false alarms per chunk); the exact form
    where neither is (2^46, kmax 1e6); a forced form keeps the tuned chunk."""
    big_k = dict(pmin=2**46, pmax=2**46 + 10**6, kmin=3, kmax=999999, nmin=100000, nmax=2000000)
    for tag, kw, want in (("cfg3", RUNS["cfg3_windows"]["params"], (0, RANGE, 8)),
                          ("cfg2", RUNS["cfg2_windows"]["params"], (0, ODDLO, 8)),
                          ("many_n", RUNS["many_n"]["params"], (0, ODDLO, 8)),
                          ("2^46 kmax 1e6", big_k, (0, EXACT, 8))):
        with ppsd.Context(ppsd.derive(ppsd.SearchParams(**kw))) as ctx:
            ctx.set_tuning(r=1, u=8)
            assert ctx.pretest() == want, tag
            for other in FORMS:
                ctx.set_pretest(other)
                assert ctx.pretest() == (other, other, 8)
    with ppsd.Context(ppsd.derive(ppsd.SearchParams(**RUNS["cfg2_windows"]["params"]))) as ctx:
        with pytest.raises(ppsd.PPSDError):
            ctx.set_pretest(4)
        with pytest.raises(ppsd.PPSDError):
            ctx.set_pretest(-1)


def _manual(kw):
    """Parameters with the largest window that keeps one candidate per
    window (kmax*2^(s-1) < pmin), outside the reference's rules."""
    s = 32
    while (kw["kmax"] << (s - 1)) >= kw["pmin"]:
        s -= 1
    params = ppsd.manual_params(nstep=s, **kw)
    pr = orc.OrcParams(kstep=2, koffset=1, nstep_ref=s, nstep=s, **kw)
    return params, pr


def test_oddlo_needs_kmax_below_2p32():
    """kmax >= 2^32: a forced odd-low form falls back to the exact form, and
    the results are the oracle's."""
    pmin = 2**45 + 1
    kw = dict(pmin=pmin, pmax=pmin + 4000, kmin=2**33 - 2**26 + 1, kmax=2**33 + 1, nmin=64, nmax=30000, sign=0)
    params, pr = _manual(kw)
    P = orc.primes(kw["pmin"], kw["pmax"])
    want_raw = orc.hit_set(orc.walk(pr, P, filter=False, threads=8))
    assert want_raw
    with ppsd.Context(params) as ctx:
        ctx.set_pretest(ODDLO)
        assert ctx.pretest()[:2] == (ODDLO, EXACT)
        raw, cnt, _ = ctx.run_range(kw["pmin"], kw["pmax"], raw=True)
    assert cnt == len(P)
    assert ppsd.factor_set(raw) == want_raw


@pytest.mark.parametrize("seed", range(6))
def test_oddlo_extremes(seed):
    """The odd-low form where its compare is tightest and loosest: kmax of 32
    bits (odd parts compared in the full low word), and p just above 2^32
    with kmax of 10 bits, where about half of all hits are states below 2^32
    and many states have a zero low word; every sign; raw divisors as the
    oracle's."""
    rng = random.Random(7700 + seed)
    if seed % 2 == 0:
        kmax = rng.randrange(2**31, 2**32) | 1
        pmin = rng.randrange(2**40, 2**42)
        kmin = kmax - rng.randrange(2**28, 2**31)
    else:
        kmax = rng.randrange(2**9, 2**10) | 1
        pmin = rng.randrange(2**32 + 1, 2**33)
        kmin = rng.randrange(1, kmax)
    sign = (1, -1, 0)[seed % 3]
    nmin = rng.randrange(64, 2000)
    span = rng.randrange(4000, 20000) if seed % 2 == 0 else rng.randrange(10**6, 2 * 10**6)
    kw = dict(pmin=pmin, pmax=pmin + span, kmin=kmin, kmax=kmax, nmin=nmin,
              nmax=nmin + rng.randrange(20000, 60000), sign=sign)
    params, pr = _manual(kw)
    P = orc.primes(kw["pmin"], kw["pmax"])
    want_raw = orc.hit_set(orc.walk(pr, P, filter=False, threads=8))
    assert want_raw
    with ppsd.Context(params) as ctx:
        ctx.set_tuning(r=1 + seed % 2, u=(4, 16)[seed % 2])
        ctx.set_pretest(ODDLO)
        assert ctx.pretest()[1] == ODDLO
        raw, cnt, _ = ctx.run_range(kw["pmin"], kw["pmax"], raw=True)
    assert cnt == len(P)
    assert ppsd.factor_set(raw) == want_raw


def test_batch_mix_each_form():
    """The persistent batched walk over 24 units of the seeded mix, pieces
    forced to each form: per-unit results equal, and equal the plain walk's."""
    units = ppsd_units.mix_units(24)
    params = [u.params for u in units]
    # a small slice of each unit keeps the oracle-free comparison quick
    for p in params:
        p.pmax = min(p.pmax, p.pmin + 2000000)
    per_mode = {}
    with ppsd.Context(params[0]) as ctx:
        for mode in FORMS:
            ctx.set_pretest(mode)
            out, unit, cnt, sums = ctx.batch_run(params, raw=True)
            hits = {(int(u), (int(f["p"]), int(f["k"]), int(f["n"]), int(f["c"]))) for f, u in zip(out, unit)}
            assert len(hits) == len(out)
            per_mode[mode] = (hits, [int(x) for x in cnt], [int(x) for x in sums])
    assert per_mode[EXACT] == per_mode[RANGE] == per_mode[ODDLO]
    for i in (0, 5, 11, 23):
        with ppsd.Context(params[i]) as ctx:
            ctx.set_pretest(RANGE)
            raw, cnt, psum = ctx.run_range(params[i].pmin, params[i].pmax, raw=True)
        assert {t for u, t in per_mode[RANGE][0] if u == i} == ppsd.factor_set(raw)
        assert (cnt, psum) == (per_mode[RANGE][1][i], per_mode[RANGE][2][i])
