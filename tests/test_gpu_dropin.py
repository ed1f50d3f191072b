"""The drop-in check: the reference's own host program (main.c + pps/app.c
+ sieve.c + putil.c + clock.c + pps/factor_proth.c, compiled from their
sources into oracle/_ref) linked against libppsd.so in place of
pps/appcu.cu.  Its factors file must equal the goldens: the reference's CPU
re-walk (test_one_p) and filter (test_factor) consume the flags that
libppsd's get_factors_found produces."""
import json
import os
import subprocess

import pytest

import pps_oracle as orc
from conftest import GOLDEN, REPO, read_lines

pytestmark = pytest.mark.gpu

EXE = os.path.join(REPO, "oracle", "_ref", "ppsieve_ref_on_ppsd")
RUNS = json.load(open(os.path.join(GOLDEN, "runs.json")))

needs_exe = pytest.mark.skipif(not os.path.exists(EXE), reason="reference drop-in build absent (no reference tree)")


def run(args, cwd):
    print("ppsieve_ref_on_ppsd", " ".join(args), flush=True)
    r = subprocess.run([EXE] + args, cwd=cwd, capture_output=True, text=True, timeout=600)
    return r.returncode, r.stdout, r.stderr


def factors(cwd):
    with open(os.path.join(cwd, "ppfactors.txt")) as f:
        return sorted(l.rstrip("\n") for l in f if "|" in l)


@needs_exe
def test_reference_host_many_n(tmp_path):
    rc, out, err = run(["-p42070e9", "-P42070050e6", "-k", "1201", "-K", "9999", "-N", "2000000", "-q"], tmp_path)
    assert rc == 0, err
    assert factors(tmp_path) == sorted(read_lines("many_n_filtered.txt"))
    r = RUNS["many_n"]
    assert f"count={r['count']},sum=0x{r['sum']:016x}" in err
    assert "Found 39 factors" in out
    assert "Computation Error" not in err


@needs_exe
def test_reference_host_cfg1_riesel(tmp_path):
    lo, hi = 10**12, 10**12 + 3 * 10**6
    rc, out, err = run([f"-p{lo}", f"-P{hi}", "-k3", "-K9999", "-n100", "-N40000", "-q", "-R", "-m", "8"], tmp_path)
    assert rc == 0, err
    pr = orc.derive(pmin=lo, pmax=hi, kmin=3, kmax=9999, nmin=100, nmax=40000, sign=-1)
    want = orc.walk(pr, orc.primes(lo, hi), filter=True, threads=8)
    assert factors(tmp_path) == sorted(f"{int(x['p'])} | {int(x['k'])}*2^{int(x['n'])}{int(x['c']):+d}" for x in want)
    assert "Computation Error" not in err
