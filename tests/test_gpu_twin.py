"""The twin-prime sieve (tpsieve: k = 6a+3, k*2^n+1 and k*2^n-1 in one
pass; the reference's SEARCH_TWIN build, pps/app.c:53-59, 1071-1110) on the
GPU against the oracle: every factor of either form, with and without a
twin sieve file."""
import os
import subprocess

import pytest

import pps_oracle as orc
from conftest import REPO
from sievefiles import candidates, write_abcd_twin, write_newpgen_twin

pytestmark = pytest.mark.gpu

TP = os.path.join(REPO, "psieve-cuda_amd", "host", "tpsieve")
REF = os.path.join(REPO, "oracle", "_ref", "tpsieve_ref_on_ppsd")
LO, HI = 10**12, 10**12 + 6 * 10**6


def run(args, cwd, exe=TP, env=None):
    print(os.path.basename(exe), " ".join(args), flush=True)
    r = subprocess.run([exe] + args, cwd=cwd, capture_output=True, text=True, timeout=600,
                       env=dict(os.environ, **(env or {})))
    assert r.returncode == 0, r.stderr
    with open(os.path.join(cwd, "tpfactors.txt")) as f:
        return {orc.parse_line(l.rstrip("\n")) for l in f if "|" in l}


def oracle(filter, kmax=9999, nmax=200000):
    pr = orc.derive(pmin=LO, pmax=HI, kmin=3, kmax=kmax, nmin=100, nmax=nmax, kstep=6, koffset=3, sign=0)
    return orc.hit_set(orc.walk(pr, orc.primes(LO, HI), filter=filter, threads=8))


def test_twin_factors(tmp_path):
    want = oracle(filter=True)
    got = run([f"-p{LO}", f"-P{HI}", "-k3", "-K9999", "-n100", "-N200000", "-q"], str(tmp_path))
    assert got == want
    assert {c for _, _, _, c in want} == {1, -1} and all(k % 6 == 3 for _, k, _, _ in want)


@pytest.mark.parametrize("fmt", ["abcd", "newpgen"])
def test_twin_sieve_file(tmp_path, fmt):
    raw = {h for h in oracle(filter=False) if h[1] % 6 == 3}
    terms, chosen = candidates(raw, 3, 9999, 100, 200000, seed=21, decoys=0)
    terms = sorted(set(terms) | {(6 * a + 3, n) for a, n in zip(range(5, 2000, 7), range(150, 199000, 700))})
    want = {h for h in raw if (h[1], h[2]) in chosen}
    assert len(raw) > 20 and 0 < len(want) < len(raw)
    path = os.path.join(tmp_path, "t.txt")
    (write_abcd_twin if fmt == "abcd" else write_newpgen_twin)(path, terms)
    assert run([f"-p{LO}", f"-P{HI}", "-i", path, "-k3", "-K9999", "-n100", "-N200000", "-q"], str(tmp_path)) == want


@pytest.mark.skipif(not os.path.exists(REF), reason="reference drop-in build absent (no reference tree)")
def test_reference_twin_host(tmp_path):
    """The reference's own twin-prime host (main.c + pps/app.c built with
    SEARCH_TWIN) on libppsd: its CPU re-walk (test_one_p's twin branch) and
    filter consume the device's subsection flags; its factors are ours."""
    want = oracle(filter=True)
    got = run([f"-p{LO}", f"-P{HI}", "-k3", "-K9999", "-n100", "-N200000", "-q"], str(tmp_path), exe=REF,
              env={"PPSD_APPCU_TWIN": "1"})
    assert got == want
