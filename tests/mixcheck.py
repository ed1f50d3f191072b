"""Oracle checks of work units searched on the device at their full sizes
(the seeded mixed workload, the serving driver), where walking every prime
on the CPU would take hours.  For each unit:

  * its prime count and sum equal the oracle's sieve of [pmin, pmax)
    (orc_count_primes_many);
  * every reported divisor is valid: p prime, p | k*2^n + c, k odd in
    [kmin, kmax], n in [nmin, nmax), c the unit's sign (orc_check_hits:
    Miller-Rabin and 128-bit powmod, independent of the walk);
  * raw divisors: the filtered set equals test_factor applied to them;
    filtered factors: every one passes test_factor;
  * completeness on p-slices: the hits among the unit's first and last
    `edge` primes equal the oracle walk of those primes, and for units of at
    most `full_work` window-steps the hits of the whole unit equal the
    oracle walk of all of its primes.
"""
import math
import time

import numpy as np

import pps_oracle as orc
import ppsd

THREADS = 16


def orc_params(p) -> "orc.OrcParams":
    """The oracle's parameters for a device unit (ppsd.Params), window
    included: units outside the reference's validation carry the smallest
    window with one candidate each, which the oracle walks as given."""
    return orc.OrcParams(pmin=p.pmin, pmax=p.pmax, kmin=p.kmin, kmax=p.kmax, nmin=p.nmin, nmax=p.nmax,
                         kstep=p.kstep, koffset=p.koffset, sign=p.sign, nstep_ref=p.nstep_ref, nstep=p.nstep,
                         status=0)


def to_hits(s) -> np.ndarray:
    """A set of (p, k, n, c) as an oracle hit array."""
    a = np.zeros(len(s), dtype=orc.HIT_DTYPE)
    for i, (p, k, n, c) in enumerate(sorted(s)):
        a[i] = (p, k, n, c)
    return a


def work(p) -> float:
    """Expected window-steps of a unit (both signs count twice)."""
    primes = (p.pmax - p.pmin) / math.log(p.pmin)
    wins = (p.nmax - p.nmin + p.nstep - 1) // p.nstep if p.nmax > p.nmin else 0
    return primes * wins * (2 if p.sign == 0 else 1)


def edge_primes(p, m):
    """The first and the last m primes of [pmin, pmax) (fewer if the range
    holds fewer), by the oracle's listing sieve on windows sized from the
    prime density."""
    span = int(m * math.log(p.pmax) * 1.6) + 1000
    lo = orc.primes(p.pmin, min(p.pmax, p.pmin + span))
    if p.pmin + span >= p.pmax:
        return lo
    hi = orc.primes(max(p.pmin, p.pmax - span), p.pmax)
    while len(lo) < m or len(hi) < m:  # a prime gap far above the mean: widen
        span *= 2
        lo = orc.primes(p.pmin, min(p.pmax, p.pmin + span))
        hi = orc.primes(max(p.pmin, p.pmax - span), p.pmax)
        if p.pmin + span >= p.pmax:
            return lo
    return np.concatenate([lo[:m], hi[-m:]])


def check_units(params, counts, sums, hit_sets, raw, kept_sets=None, edge=1000, full_work=2e9):
    """Runs the checks above over units params[i] with device results
    counts[i], sums[i] and hit_sets[i] (raw divisors when raw, else filtered
    factors); kept_sets[i], when given with raw, is the device's filtered set
    of the same unit.  Returns (units walked fully, hits checked)."""
    want_cs = orc.count_primes_many([(p.pmin, p.pmax) for p in params], threads=THREADS)
    print(f"oracle counts of {len(params)} units done", flush=True)
    full, checked = 0, 0
    last = time.time()
    for i, p in enumerate(params):
        if time.time() - last > 20:
            print(f"oracle checks: unit {i} of {len(params)}", flush=True)
            last = time.time()
        tag = f"unit {i} (p {p.pmin}..{p.pmax}, k {p.kmin}..{p.kmax}, n {p.nmin}..{p.nmax}, sign {p.sign}, s {p.nstep})"
        assert (int(counts[i]), int(sums[i])) == want_cs[i], tag + ": prime count/sum"
        pr = orc_params(p)
        hits = to_hits(hit_sets[i])
        flags = orc.check_hits(pr, hits, threads=THREADS)
        bad = np.nonzero(flags & ~np.uint8(orc.KEPT))[0]
        assert len(bad) == 0, f"{tag}: invalid divisor {hits[bad[0]]} flags {flags[bad[0]]}"
        kept_mask = (flags & orc.KEPT) != 0
        if raw:
            if kept_sets is not None:
                assert orc.hit_set(hits[kept_mask]) == kept_sets[i], tag + ": filtered set"
        else:
            assert kept_mask.all(), tag + ": a factor test_factor drops"
        checked += len(hits)
        if work(p) <= full_work:
            P = orc.primes(p.pmin, p.pmax)
            full += 1
        else:
            P = edge_primes(p, edge)
        want = orc.hit_set(orc.walk(pr, P, filter=not raw, threads=THREADS))
        in_slice = set(P.tolist())
        got = {t for t in hit_sets[i] if t[0] in in_slice}
        assert got == want, f"{tag}: hits of {len(P)} primes: {len(got)} against the oracle's {len(want)}"
    return full, checked


def factor_set(arr) -> set:
    return ppsd.factor_set(arr)
