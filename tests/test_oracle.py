"""Pins the CPU oracle (oracle/pps_oracle.c) before anything is compared with
it: reference fixtures, the compiled reference small-prime driver, closed
forms (Python pow, sympy) and brute force on tiny inputs."""
import json
import os
import random
import subprocess

import numpy as np
import pytest
import sympy

import pps_oracle as orc
from conftest import GOLDEN, read_lines

SMALL = list(sympy.primerange(3, 32768))


def smallest_factor(k, n, c):
    for q in SMALL:
        if (k * pow(2, n, q) + c) % q == 0:
            return q
    return 0


# --- parameter rules (main.c:916-932, pps/app.c:655-712, pps/appcu.cu:155-239)

def test_derive_many_n_defaults():
    pr = orc.derive(pmin=42070 * 10**9, pmax=42070050 * 10**6, kmin=1201, kmax=9999, nmax=2000000)
    assert pr.status == 0
    # nmin = 2*lg2(pmin) - lg2(kmax) - 1 = 2*45 - 13 - 1
    assert pr.nmin == 76
    assert (pr.kmin, pr.kmax) == (1201, 9999)
    assert pr.nstep_ref == 32 and pr.nstep == 32


@pytest.mark.parametrize("kw,nstep_ref,nstep", [
    # cfg1: 9999 * 2^27 >= 1e12 > 9999 * 2^26
    (dict(pmin=10**12, pmax=10**12 + 10**8, kmin=3, kmax=9999, nmin=100, nmax=200000), 27, 27),
    # cfg2: s = 37 clamped to 32
    (dict(pmin=2**50, pmax=2**50 + 5 * 10**10, kmin=1201, kmax=9999, nmin=50000, nmax=6000000), 32, 32),
    # cfg3: s = 45 clamped to 32
    (dict(pmin=2**61, pmax=2**61 + 2 * 10**9, kmin=3, kmax=100001, nmin=100000, nmax=10000000), 32, 32),
    # s >= 48 is not clamped by the reference; the walk uses 32
    (dict(pmin=2**61, pmax=2**61 + 10**6, kmin=3, kmax=3, nmin=1000, nmax=100000), 60, 32),
])
def test_derive_nstep(kw, nstep_ref, nstep):
    pr = orc.derive(**kw)
    assert pr.status == 0, pr.msg
    assert (pr.nstep_ref, pr.nstep) == (nstep_ref, nstep)
    # uniqueness of a hit per window: kmax * 2^(s-1) < pmin
    assert pr.kmax << (pr.nstep - 1) < pr.pmin


def test_derive_k_alignment():
    pr = orc.derive(pmin=10**12, kmin=1200, kmax=10000, nmin=100, nmax=1000)
    assert (pr.kmin, pr.kmax) == (1201, 10001)   # even kmax rounds up
    assert pr.pmax == 10**12 + 10**9            # default pmax
    pr = orc.derive(pmin=10**12, kmin=5, kmax=100, nmin=100, nmax=1000, kstep=6, koffset=3)
    assert (pr.kmin, pr.kmax) == (3, 99)


@pytest.mark.parametrize("kw,msg", [
    (dict(pmin=10**12, kmin=0, kmax=99, nmin=100, nmax=200), "specify"),
    (dict(pmin=10**12, kmin=101, kmax=99, nmin=100, nmax=200), "kmin <= kmax"),
    (dict(pmin=1000, kmin=3, kmax=1000, nmin=100, nmax=200), "kmax < pmin"),
    (dict(pmin=2**62 - 10, pmax=2**62, kmin=1, kmax=3 * 2**36 + 1, nmin=100, nmax=200), "3*2^36"),
    (dict(pmin=10**12, kmin=3, kmax=99, nmin=300, nmax=200), "nmin <= nmax"),
    (dict(pmin=10**12, kmin=3, kmax=99, nmin=32, nmax=200), "nmin too small"),
    (dict(pmin=2**32, pmax=2**33, kmin=1, kmax=1, nmin=100, nmax=200), "PMin is too small"),
    # nstep limited by nmax-nmin+1 = 2 -> 2^62 > pmin
    (dict(pmin=10**12, kmin=3, kmax=99, nmin=100, nmax=101), "not large enough"),
    (dict(pmin=10**12, pmax=10**12, kmin=3, kmax=99, nmin=100, nmax=200), "pmax must be greater"),
])
def test_derive_errors(kw, msg):
    pr = orc.derive(**kw)
    assert pr.status != 0
    assert msg in pr.msg.decode()


# --- prime supply (sieve.c, main.c:750-777)

@pytest.mark.parametrize("p0,p1", [(3, 5000), (10**9, 10**9 + 20000), (2**32 - 5000, 2**32 + 5000),
                                   (10**12, 10**12 + 30000)])
def test_primes_vs_sympy(p0, p1):
    got = orc.primes(p0, p1)
    want = list(sympy.primerange(p0, p1))
    assert got.tolist() == want


def test_primes_high_window_miller_rabin():
    p0 = 2**61
    got = orc.primes(p0, p0 + 4000)
    assert got.tolist() == [x for x in range(p0 | 1, p0 + 4000, 2) if sympy.isprime(x)]
    p0 = 2**62 - 3000
    got = orc.primes(p0, 2**62)
    assert got.tolist() == [x for x in range(p0 | 1, 2**62, 2) if sympy.isprime(x)]


def test_primes_sieve_equals_miller_rabin():
    lo = 10**12
    got = orc.primes(lo, lo + 2 * 10**5)        # segmented-sieve path (range >= isqrt/8)
    mr = [x for x in range(lo + 1, lo + 2 * 10**5, 2) if orc.lib().orc_is_prime(x)]
    assert got.tolist() == mr


# --- Montgomery arithmetic (pps/app.c:895-1035)

PRIMES = [2**32 + 15, 1000000000039, 2**50 + 55, 2**61 - 1, 2**61 + 15, 2**62 - 57]


@pytest.mark.parametrize("p", PRIMES)
def test_neg_inv64(p):
    assert sympy.isprime(p)
    ps = orc.lib().orc_neg_inv64(p)
    assert (p * ps) % 2**64 == 2**64 - 1


@pytest.mark.parametrize("p", PRIMES)
def test_inv2pow_and_step_closed_form(p):
    rng = random.Random(p)
    ps = orc.lib().orc_neg_inv64(p)
    for n in [64, 65, 127, 1000, 2**20 + 3, 6000000, 2**31 - 1] + [rng.randrange(64, 2**31) for _ in range(20)]:
        K = orc.lib().orc_inv2pow(p, n)
        assert K == pow(2, -n, p)
        for s in (1, 17, 27, 31, 32, 45):
            assert orc.lib().orc_step(K, p, ps, s) == pow(2, -(n + s), p)


def test_pps4test_halving_table():
    """pps4test.ods: 2^(1-n) mod 10007 by repeated halving."""
    p = 10007
    rows = [tuple(map(int, l.split())) for l in read_lines("pps4test_halving.txt")]
    assert len(rows) >= 40
    for n, v in rows:
        assert v == pow(2, 1 - n, p)
        if n >= 1:
            assert orc.lib().orc_inv2pow(p, n - 1) == v


# --- walk: brute force on tiny inputs

def brute(p_list, kmin, kmax, nmin, nmax, signs):
    """every (p, k, n, c) with p | k*2^n + c, k odd in [kmin, kmax]: since
    kmax < p, k must be the residue -c*2^-n mod p itself"""
    out = set()
    for p in p_list:
        inv = pow(2, -1, p)
        x = pow(inv, nmin, p)
        for n in range(nmin, nmax):
            for c in signs:
                k = (-c * x) % p
                if kmin <= k <= kmax and k & 1:
                    out.add((p, k, n, c))
            x = x * inv % p
    return out


def manual_params(pmin, pmax, kmin, kmax, nmin, nmax, sign, s):
    """Walk parameters outside the reference's validation domain (which needs
    2^(64-s) <= pmin, leaving almost no hits at testable sizes); the walk only
    needs kmax * 2^(s-1) < pmin for one candidate per window."""
    assert kmax << (s - 1) < pmin
    return orc.OrcParams(pmin=pmin, pmax=pmax, kmin=kmin, kmax=kmax, nmin=nmin, nmax=nmax, kstep=2, koffset=1,
                         sign=sign, nstep=s, nstep_ref=s)


@pytest.mark.parametrize("pmin,sign,kmax,nmin,nmax,s", [(10**6, 1, 1001, 64, 3000, 9),
                                                         (10**6 + 3, -1, 1001, 100, 2500, 7),
                                                         (3 * 10**6, 0, 3001, 64, 1500, 9),
                                                         (2**33, 1, 2**20 + 1, 64, 700, 13),
                                                         (2**40, -1, 2**21 + 1, 64, 900, 19)])
def test_walk_matches_brute_force(pmin, sign, kmax, nmin, nmax, s):
    pr = manual_params(pmin, pmin + 6000, 1, kmax, nmin, nmax, sign, s)
    P = orc.primes(pr.pmin, pr.pmax)
    got = orc.hit_set(orc.walk(pr, P, filter=False))
    want = brute(P.tolist(), pr.kmin, pr.kmax, pr.nmin, pr.nmax, (1, -1) if sign == 0 else (sign,))
    assert got == want
    if pmin < 2**32:
        assert len(want) > 50


def test_walk_window_invariance():
    """The factor set does not depend on the window: s = 32 vs the derived 27."""
    pr = orc.derive(pmin=10**12, pmax=10**12 + 200000, kmin=3, kmax=9999, nmin=100, nmax=60000)
    assert pr.nstep == 27
    P = orc.primes(pr.pmin, pr.pmax)
    a = orc.hit_set(orc.walk(pr, P, filter=False, threads=4))
    pr.nstep = 13
    b = orc.hit_set(orc.walk(pr, P, filter=False, threads=4))
    assert a == b and len(a) > 0


# --- filter (pps/app.c:150-186, pps/factor_proth.c)

def test_try_all_factors_vs_trial_division():
    rng = random.Random(7)
    for _ in range(300):
        k = rng.randrange(1, 2**40) | 1
        n = rng.randrange(64, 2**24)
        c = rng.choice((1, -1))
        q = orc.lib().orc_try_all_factors(k, n, c)
        want = next((q for q in SMALL if q >= 11 and (k * pow(2, n, q) + c) % q == 0), 0)
        assert q == want, (k, n, c)


def test_try_all_factors_vs_reference_driver():
    """tests/golden/factor_proth_ref.txt holds the compiled reference
    pps/factor_proth.c driver's output for 400 "+1" lines (its table starts
    at 3: sieve_small_primes(0))."""
    rows = [l.split("\t") for l in read_lines("factor_proth_ref.txt")]
    assert len(rows) == 400
    for inp, out in rows:
        _, k, n, c = orc.parse_line(inp)
        q3 = next((q for q in (3, 5, 7) if (k * pow(2, n, q) + c) % q == 0), 0)
        q = q3 or orc.lib().orc_try_all_factors(k, n, c)
        if q == 0:
            assert out == inp
        else:
            assert out == f"sm {q} | {k}*2^{n}+1"


@pytest.mark.skipif(not os.path.exists(orc.REF_FACTOR_PROTH), reason="the reference's factor_proth driver is built only where its sources are")
def test_golden_factors_reverify_in_reference_factor_proth():
    """Every +1 line of the golden factor sets -- which the device's output
    must equal, line for line -- passes the reference's own factor_proth.c
    check: its stdin driver echoes a line with no odd factor below 2^15
    unchanged (pps/factor_proth.c:255-282; the driver handles +1 only)."""
    lines = []
    for name in ("cfg1_factors.txt", "cfg2_windows.txt", "cfg3_windows.txt", "many_n_filtered.txt", "OKsort.txt",
                 "pps_test80t.txt"):
        lines += [l for l in read_lines(name) if l.endswith("+1")]
    out = subprocess.run([orc.REF_FACTOR_PROTH], input="\n".join(lines) + "\n", capture_output=True, text=True,
                         check=True, timeout=120).stdout.strip().split("\n")
    assert len(lines) > 450
    assert out == lines


def test_test_factor_rules():
    pr = orc.derive(pmin=10**12, kmin=3, kmax=9999, nmin=100, nmax=1000)
    rng = random.Random(3)
    for _ in range(400):
        k = rng.randrange(3, 10000)
        n = rng.randrange(50, 1100)
        c = rng.choice((1, -1))
        want = (k & 1) and 100 <= n < 1000 and smallest_factor(k, n, c) == 0
        assert orc.lib().orc_test_factor(pr, k, n, c) == int(bool(want)), (k, n, c)
    # large k: the khigh guards (pps/app.c:171-179) skip pre-tests from
    # k >= 2^46 on, which try_all_factors then covers; below 2^61 no shift
    # of the pre-tests wraps, so the kept set is exactly "no factor < 32768"
    pr = orc.OrcParams(pmin=2**62 - 2**40, pmax=2**62, kmin=1, kmax=2**62 - 1, nmin=64, nmax=2**31, kstep=2,
                       koffset=1, sign=1, nstep=1, nstep_ref=1)
    edges = [2**46, 2**52, 2**53, 2**54, 2**56, 2**61]
    for _ in range(600):
        e = rng.choice(edges)
        k = (rng.randrange(e - 2**20, e + 2**20) if rng.random() < 0.5 else rng.randrange(2**46, 2**61)) | 1
        k = min(k, 2**61 - 1)
        n = rng.randrange(64, 2**24)
        c = rng.choice((1, -1))
        want = smallest_factor(k, n, c) == 0
        assert orc.lib().orc_test_factor(pr, k, n, c) == int(want), (k, n, c)
    # -M 6 -s 3: only k == 3 (mod 6)
    pr = orc.derive(pmin=10**12, kmin=3, kmax=9999, nmin=100, nmax=1000, kstep=6, koffset=3)
    for k in range(3, 400, 2):
        for n in (101, 500):
            want = k % 6 == 3 and smallest_factor(k, n, 1) == 0
            assert orc.lib().orc_test_factor(pr, k, n, 1) == int(want)


# --- reference fixtures

def test_many_n_raw_reproduces_oldsort():
    """The oracle's divisor set of the many_n_test range, cut to the lines
    with no factor 3, 5 or 7 (the pre-test of the build that wrote it), is
    exactly pps/oldsort.txt."""
    raw = read_lines("many_n_raw.txt")
    old = read_lines("oldsort.txt")
    cut = [l for l in raw if all((lambda t: (t[1] * pow(2, t[2], q) + t[3]) % q)(orc.parse_line(l))
                                 for q in (3, 5, 7))]
    assert sorted(cut) == sorted(old)
    assert len(old) == 144 and len(raw) == 296


def test_many_n_filtered_golden():
    old = read_lines("oldsort.txt")
    filt = read_lines("many_n_filtered.txt")
    assert filt == [l for l in old if smallest_factor(*orc.parse_line(l)[1:]) == 0]
    assert len(filt) == 39


def test_many_n_oracle_slice():
    """Re-run the oracle on the first 2e6 of the many_n range."""
    lo, hi = 42070000000000, 42070002000000
    pr = orc.derive(pmin=lo, pmax=hi, kmin=1201, kmax=9999, nmin=76, nmax=2000000)
    hits, cnt, _ = orc.run_range(pr, lo, hi, filter=False, threads=os.cpu_count() or 4)
    got = {orc.parse_line(l) for l in read_lines("many_n_raw.txt") if int(l.split()[0]) < hi}
    assert orc.hit_set(hits) == got
    assert cnt == 63668


def test_ok_fixture_is_subset_of_filtered():
    """pps/OKsort.txt (= pps/pps_test_ok.txt): what a run over the many_n
    range with a sieve file wrote; every line is in the oracle's filtered
    divisor set of that range."""
    ok = read_lines("OKsort.txt")
    assert len(ok) == 15
    assert set(ok) <= set(read_lines("many_n_filtered.txt"))


def test_big_p_fixture_lines_kept():
    """pps/pps_test80t.txt (the golden of pps/big_p_test.sh, p in [80T,
    80000050M)): each line is a prime divisor that the oracle's test_factor
    keeps."""
    lines = read_lines("pps_test80t.txt")
    assert len(lines) == 8
    pr = orc.derive(pmin=80 * 10**12, pmax=80000050 * 10**6, kmin=1241, kmax=9209, nmin=50848, nmax=1487243)
    for l in lines:
        p, k, n, c = orc.parse_line(l)
        assert 80 * 10**12 <= p < 80000050 * 10**6 and sympy.isprime(p)
        assert (k * pow(2, n, p) + c) % p == 0
        assert orc.lib().orc_test_factor(pr, k, n, c) == 1


def test_negative_fixtures_rejected():
    """Lines of pps/newsort.txt and pps/test.txt that are not divisors."""
    lines = read_lines("negative_lines.txt")
    assert len(lines) == 64
    for l in lines:
        p, k, n, c = orc.parse_line(l)
        assert (k * pow(2, n, p) + c) % p != 0


def test_goldens_are_true_factors():
    for name in ("cfg1_factors.txt", "cfg2_windows.txt", "cfg3_windows.txt", "cfg2_windows_raw.txt",
                 "cfg3_windows_raw.txt", "many_n_filtered.txt", "many_n_raw.txt"):
        for l in read_lines(name):
            p, k, n, c = orc.parse_line(l)
            assert (k * pow(2, n, p) + c) % p == 0
            assert sympy.isprime(p)
    runs = json.load(open(os.path.join(GOLDEN, "runs.json")))
    assert runs["many_n_raw"]["count"] == 1592285
    assert runs["cfg1"]["nstep"] == 27


@pytest.mark.parametrize("p0,p1", [(3, 1000), (10**12, 10**12 + 10**7), (2**50, 2**50 + 3 * 10**7 + 7),
                                   (2**40 + 1, 2**40 + 2**28)])
def test_count_primes_equals_listing(p0, p1):
    """The threaded counting sieve behind the full-range goldens agrees with
    the listing sieve (itself pinned by sympy above) on count and sum."""
    P = orc.primes(p0, p1)
    assert orc.count_primes(p0, p1, threads=4) == (len(P), int(P.sum(dtype=np.uint64)))


def test_no_golden_is_empty():
    """Every fixture under tests/golden holds data: a golden that is empty
    would pin a run's factor set vacuously."""
    names = [n for n in os.listdir(GOLDEN) if n.endswith(".txt")]
    assert len(names) >= 12
    for n in names:
        assert read_lines(n), n


def test_count_primes_many_equals_listing():
    """Several ranges counted together (one table of sieving primes) give
    each range's listing count and sum; the listing sieve's presieved
    Miller-Rabin path (narrow windows high up) agrees with sympy."""
    ranges = [(10**12, 10**12 + 10**6), (3, 10**5), (2**50 + 12345, 2**50 + 3 * 10**6), (2**61, 2**61 + 10**5),
              (2**62 - 10**5, 2**62)]
    got = orc.count_primes_many(ranges, threads=4)
    for (a, b), (c, s) in zip(ranges, got):
        P = orc.primes(a, b)
        assert (c, s) == (len(P), int(P.sum(dtype=np.uint64))), (a, b)
    P = orc.primes(2**61, 2**61 + 3000)
    assert P.tolist() == list(sympy.primerange(2**61, 2**61 + 3000))


def test_check_hits_flags():
    """orc_check_hits against Python's pow and sympy: golden divisors are
    valid (and those test_factor keeps are flagged so); a non-divisor, a
    composite p, k even or out of range, n out of range and a sign the
    search does not walk are each flagged."""
    pr = orc.derive(pmin=42070000000000, pmax=42070050000000, kmin=1201, kmax=9999, nmax=2000000, sign=1)
    raw = [orc.parse_line(l) for l in read_lines("many_n_raw.txt")]
    kept = {orc.parse_line(l) for l in read_lines("many_n_filtered.txt")}
    flags = orc.check_hits(pr, _hits(raw), threads=3)
    for t, f in zip(raw, flags):
        assert int(f) & ~orc.KEPT == 0, t
        assert bool(int(f) & orc.KEPT) == (t in kept), t
    p, k, n, c = raw[0]
    kr = pow(2, -n, p)  # kr * 2^n - 1 is a multiple of p: a divisor of the -1 form
    q = int(sympy.nextprime(2**31) * sympy.nextprime(2**31 + 10**6))  # composite, about 2^62
    kq = (q - pow(2, -n, q)) % q  # kq * 2^n + 1 == 0 (mod q)

    def k_bad(kk):
        return 0 if kk % 2 and 1201 <= kk <= 9999 else orc.BAD_K

    cases = [((p, k, n + 1, c), orc.BAD_DIVIDES),
             ((p, k + 1, n, c), orc.BAD_DIVIDES | orc.BAD_K),
             ((p, k, 2000000, c), orc.BAD_DIVIDES | orc.BAD_N),
             ((p, k, n, -c), orc.BAD_DIVIDES | orc.BAD_SIGN),
             ((p, kr, n, -1), orc.BAD_SIGN | k_bad(kr)),
             ((q, kq, n, 1), orc.BAD_PRIME | k_bad(kq))]
    got = orc.check_hits(pr, _hits([t for t, _ in cases]))
    for (t, want), f in zip(cases, got):
        assert int(f) == want, (t, int(f), want)
        p_, k_, n_, c_ = t
        assert ((k_ * pow(2, n_, p_) + c_) % p_ == 0) == (not int(f) & orc.BAD_DIVIDES)
        assert sympy.isprime(p_) == (not int(f) & orc.BAD_PRIME)


def _hits(tuples):
    a = np.zeros(len(tuples), dtype=orc.HIT_DTYPE)
    for i, t in enumerate(tuples):
        a[i] = t
    return a
