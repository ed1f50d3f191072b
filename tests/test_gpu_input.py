"""Sieve-file input (-i) on the GPU: the factors written are exactly the
divisors whose (k, n) the file lists (pps/app.c:181-184), for both file
formats and both signs, and the same as the reference's own host program
(linked on libppsd) writes for the same file."""
import os
import subprocess

import pytest

import pps_oracle as orc
from conftest import REPO, read_lines
from sievefiles import candidates, write_abcd, write_newpgen

pytestmark = pytest.mark.gpu

EXE = os.path.join(REPO, "psieve-cuda_amd", "host", "ppsieve")
REF = os.path.join(REPO, "oracle", "_ref", "ppsieve_ref_on_ppsd")
LO, HI = 10**12, 10**12 + 4 * 10**6


def run(exe, args, cwd):
    print(os.path.basename(exe), " ".join(args), flush=True)
    r = subprocess.run([exe] + args, cwd=cwd, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr
    with open(os.path.join(cwd, "ppfactors.txt")) as f:
        return {orc.parse_line(l.rstrip("\n")) for l in f if "|" in l}


def expected(sign, nmax, seed):
    pr = orc.derive(pmin=LO, pmax=HI, kmin=3, kmax=9999, nmin=100, nmax=nmax, sign=sign)
    raw = orc.hit_set(orc.walk(pr, orc.primes(LO, HI), filter=False, threads=8))
    terms, chosen = candidates(raw, 3, 9999, 100, nmax, seed)
    want = {h for h in raw if (h[1], h[2]) in chosen}
    assert len(raw) > 50 and 0 < len(want) < len(raw)
    return terms, want


@pytest.mark.parametrize("fmt,sign", [("abcd", 1), ("newpgen", -1)])
def test_sieve_file(tmp_path, fmt, sign):
    terms, want = expected(sign, 200000, seed=11 + sign)
    path = os.path.join(tmp_path, "sieve.txt")
    (write_abcd if fmt == "abcd" else write_newpgen)(path, terms, sign)
    # ranges (and sign) from the file
    d1 = tmp_path / "scan"
    d1.mkdir()
    assert run(EXE, [f"-p{LO}", f"-P{HI}", "-i", path, "-q"], str(d1)) == want
    # every range on the command line (the sign then too)
    d2 = tmp_path / "given"
    d2.mkdir()
    args = [f"-p{LO}", f"-P{HI}", "-i", path, "-k3", "-K9999", "-n100", "-N200000", "-q"] + (["-R"] if sign < 0 else [])
    assert run(EXE, args, str(d2)) == want
    if os.path.exists(REF):
        d3 = tmp_path / "ref"
        d3.mkdir()
        assert run(REF, [f"-p{LO}", f"-P{HI}", "-i", path, "-q"], str(d3)) == want


def _fixture_run(tmp_path, name, p0, p1, fmt):
    """A sieve file holding exactly the (k, n) of a reference golden's lines,
    searched over the golden's p-range: every factor of those terms in the
    range is a line of the golden (the reference's run listed them all), so
    the factors written must be exactly the golden."""
    want = {orc.parse_line(l) for l in read_lines(name)}
    terms = sorted({(k, n) for _, k, n, _ in want})
    path = os.path.join(tmp_path, "sieve.txt")
    (write_abcd if fmt == "abcd" else write_newpgen)(path, terms, 1)
    for exe, sub in ((EXE, "ours"), (REF, "ref")):
        if not os.path.exists(exe):
            continue
        d = tmp_path / sub
        d.mkdir()
        assert run(exe, [f"-p{p0}", f"-P{p1}", "-i", path, "-q"], str(d)) == want


def test_big_p_reference_golden(tmp_path):
    """pps/big_p_test.sh: -p80T -P80000050M over a sieve file, golden
    pps/pps_test80t.txt (8 lines).  The script's sieve file
    (pps/ppse_37TE1.txt) is not in the tree, so the file here lists just
    the golden's terms."""
    _fixture_run(tmp_path, "pps_test80t.txt", "80T", "80000050M", "newpgen")


def test_ok_reference_golden(tmp_path):
    """pps/OKsort.txt (15 lines) over the many_n range, from an ABCD file of
    its terms."""
    _fixture_run(tmp_path, "OKsort.txt", "42070e9", "42070050e6", "abcd")
