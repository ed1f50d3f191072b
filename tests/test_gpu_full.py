"""A full-size run on the GPU: the whole 2^61 both-signs range (4.73e7
primes, 309,375 windows each).  Too large for the oracle's walk, so it is
checked by size-independent properties: the prime count and sum equal the
oracle's own sieve over the range, every reported divisor divides
k*2^n+c with p prime and k, n in range, and the device filter keeps exactly
the divisors the oracle's test_factor keeps.  (The full 2^50 range is run
by bench.py, which checks its count and sum the same way.)"""
import json
import os

import pytest

import pps_oracle as orc
import ppsd
from conftest import GOLDEN

pytestmark = pytest.mark.gpu

RUNS = json.load(open(os.path.join(GOLDEN, "runs.json")))


def test_cfg3_full_range():
    r = RUNS["cfg3_full"]
    kw = r["params"]
    params = ppsd.derive(ppsd.SearchParams(**kw))
    pr = orc.derive(**kw)
    print("2^61 range, both signs:", kw, flush=True)
    with ppsd.Context(params) as ctx:
        raw, cnt, psum = ctx.run_range(kw["pmin"], kw["pmax"], raw=True)
        got, cnt2, psum2 = ctx.run_range(kw["pmin"], kw["pmax"])
    assert (cnt, psum) == (r["count"], r["sum"]) == (cnt2, psum2)
    raw_set = ppsd.factor_set(raw)
    assert len(raw_set) == len(raw) > 0
    for p, k, n, c in raw_set:
        assert (k * pow(2, n, p) + c) % p == 0 and orc.lib().orc_is_prime(p)
        assert k % 2 == 1 and pr.kmin <= k <= pr.kmax and pr.nmin <= n < pr.nmax and c in (1, -1)
    keep = {t for t in raw_set if orc.lib().orc_test_factor(pr, t[1], t[2], t[3])}
    assert ppsd.factor_set(got) == keep
