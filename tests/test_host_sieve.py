"""The host sieve of `ppsieve --host-primes` (host/pps_hostsieve.c), through
its check driver: prime count and sum of each range equal the oracle's
counting sieve, primes come out ascending, and threads change nothing."""
import os
import subprocess

import pytest

import pps_oracle as orc
from conftest import REPO

HOST = os.path.join(REPO, "psieve-cuda_amd", "host")


@pytest.fixture(scope="module")
def exe(tmp_path_factory):
    """The check driver (tests/hostsieve_check.c) built against the host sieve."""
    out = str(tmp_path_factory.mktemp("hostsieve") / "hostsieve_check")
    subprocess.run(["gcc", "-O3", "-Wall", "-std=gnu11", "-I", HOST, "-o", out,
                    os.path.join(REPO, "tests", "hostsieve_check.c"), os.path.join(HOST, "pps_hostsieve.c"), "-lm",
                    "-lpthread"], check=True)
    return out


def sieve(exe, ranges, threads=1):
    args = [exe, f"-t{threads}"] + [str(x) for r in ranges for x in r]
    out = subprocess.run(args, capture_output=True, text=True, timeout=300, check=True).stdout.split("\n")
    res = []
    for line in out:
        if line.strip():
            p0, p1, cnt, s, order = line.split()
            res.append((int(cnt), int(s, 16), order))
    return res


@pytest.mark.parametrize("p0,p1", [
    (3, 10**6),                                     # the presieved primes 3..13 themselves
    (2**32 + 1, 2**32 + 10**8),
    (1125901156842624, 1125902406842624),           # 33554467^2: a first multiple far beyond the bucket ring
    (2**56 - 10**7, 2**56 + 10**7),
    (10**12 + 1, 10**12 + 2),                       # one candidate
    (10**12 + 6, 10**12 + 7),                       # none
])
def test_ranges_vs_oracle(exe, p0, p1):
    (cnt, s, order), = sieve(exe, [(p0, p1)])
    want = orc.count_primes(p0, p1)
    assert (cnt, s) == (want[0], want[1] % 2**64)
    assert order == "ascending"


def test_threads_split_the_range(exe):
    """2^28 numbers at 2^40 in 4 parts (parts are at least 2^26 numbers)."""
    r = (2**40 + 12345, 2**40 + 2**28 + 12345)
    want = orc.count_primes(*r)
    for t in (1, 4, 7):
        (cnt, s, order), = sieve(exe, [r], threads=t)
        assert (cnt, s, order) == (want[0], want[1] % 2**64, "ascending"), t


def test_small_prime_table(exe):
    """Primes below 10^6 from 3: pi(10^6) - 1 of them."""
    (cnt, _, _), = sieve(exe, [(3, 10**6)])
    assert cnt == 78497
