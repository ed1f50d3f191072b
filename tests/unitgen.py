"""Generators of small search units for the batched and serving GPU tests,
with their oracle answers."""
import numpy as np

import pps_oracle as orc
import ppsd


def _dense_unit(rng):
    """A unit with many hits per prime (small windows, large kmax)."""
    pmin = rng.randrange(2**32 + 1, 2**36)
    s = rng.randrange(1, 15) if rng.random() < 0.85 else rng.choice((31, 32))
    kmax = max(1, min((pmin - 1) >> (s - 1), rng.randrange(2**12, 2**26)))
    kw = dict(pmin=pmin, pmax=pmin + rng.randrange(500, 20000), kmin=rng.randrange(1, kmax + 1), kmax=kmax,
              nmin=rng.randrange(64, 3000), sign=rng.choice((1, -1, 0)), nstep=s)
    kw["nmax"] = kw["nmin"] + rng.randrange(0, 20000)
    ks = rng.choice([(2, 1), (2, 1), (6, 3), (4, 1)])
    pr = orc.OrcParams(kstep=ks[0], koffset=ks[1], nstep_ref=s, **kw)
    return pr, ppsd.manual_params(kstep=ks[0], koffset=ks[1], **kw)


def _ref_unit(rng):
    """A unit inside the reference's validation domain (derived window)."""
    while True:
        pmin = rng.randrange(2**40, 2**50)
        kmax = rng.randrange(3, min(2**22, pmin * pmin >> 63))
        kw = dict(pmin=pmin, pmax=pmin + rng.randrange(1000, 40000), kmin=rng.randrange(1, kmax + 1), kmax=kmax,
                  nmin=rng.randrange(64, 5000), sign=rng.choice((1, -1, 0)))
        kw["nmax"] = kw["nmin"] + rng.randrange(0, 400000)
        pr = orc.derive(**kw)
        if pr.status == 0:
            return pr, ppsd.derive(ppsd.SearchParams(**kw))


def _oracle(pr, raw):
    P = orc.primes(pr.pmin, pr.pmax)
    hits = orc.hit_set(orc.walk(pr, P, filter=not raw, threads=8))
    return len(P), int(P.sum(dtype=np.uint64)), hits
