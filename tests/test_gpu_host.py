"""End to end through the ppsieve program on the GPU: the factors file
(pps/app.c:134 lines) against the goldens and the oracle, the host-prime
batch path (app_thread_fun), and interruption + resume from a checkpoint."""
import json
import os
import signal
import subprocess
import time

import pytest

import pps_oracle as orc
from conftest import GOLDEN, REPO, read_lines

pytestmark = pytest.mark.gpu

EXE = os.path.join(REPO, "psieve-cuda_amd", "host", "ppsieve")
RUNS = json.load(open(os.path.join(GOLDEN, "runs.json")))
MANY_N = ["-p42070e9", "-P42070050e6", "-k", "1201", "-K", "9999", "-N", "2000000"]


def run(args, cwd, timeout=300):
    print("ppsieve", " ".join(args), flush=True)
    r = subprocess.run([EXE] + args, cwd=cwd, capture_output=True, text=True, timeout=timeout)
    return r.returncode, r.stdout, r.stderr


def factors(cwd, name="ppfactors.txt"):
    with open(os.path.join(cwd, name)) as f:
        return sorted(l.rstrip("\n") for l in f if "|" in l)


def test_many_n_factors_file(tmp_path):
    rc, out, err = run(MANY_N + ["-q"], tmp_path)
    assert rc == 0, err
    assert factors(tmp_path) == sorted(read_lines("many_n_filtered.txt"))
    r = RUNS["many_n"]
    assert f"count={r['count']},sum=0x{r['sum']:016x}" in err
    assert "Found 39 factors" in out
    assert "Sieve complete: 42070000000000 <= p < 42070050000000" in err
    assert not os.path.exists(os.path.join(tmp_path, "ppcheck42070e9.txt"))


def test_host_prime_batches(tmp_path):
    """--host-primes: host sieve, batches through app_thread_fun/app_thread_fun1."""
    hi = 42070000000000 + 10**7
    rc, out, err = run(["-p42070e9", f"-P{hi}", "-k1201", "-K9999", "-N2000000", "-q", "--host-primes", "-m", "64"],
                       tmp_path)
    assert rc == 0, err
    want = sorted(l for l in read_lines("many_n_filtered.txt") if int(l.split()[0]) < hi)
    assert factors(tmp_path) == want
    P = orc.primes(42070000000000, hi)
    assert f"count={len(P)},sum=0x{int(P.sum(dtype='uint64')):016x}" in err


@pytest.mark.parametrize("sign_args,sign", [(["-R"], -1), (["--both"], 0)])
def test_riesel_and_both_vs_oracle(tmp_path, sign_args, sign):
    lo, hi = 10**12, 10**12 + 4 * 10**6
    rc, out, err = run([f"-p{lo}", f"-P{hi}", "-k3", "-K9999", "-n100", "-N30000", "-q"] + sign_args, tmp_path)
    assert rc == 0, err
    pr = orc.derive(pmin=lo, pmax=hi, kmin=3, kmax=9999, nmin=100, nmax=30000, sign=sign)
    want = orc.walk(pr, orc.primes(lo, hi), filter=True, threads=8)
    lines = sorted(f"{int(x['p'])} | {int(x['k'])}*2^{int(x['n'])}{int(x['c']):+d}" for x in want)
    assert factors(tmp_path) == lines
    assert len(lines) > 3


def test_interrupt_and_resume(tmp_path):
    """SIGINT mid-run writes a checkpoint at the last finished chunk; the rerun
    resumes there (main.c:332-372) and the factor file and count/sum end up
    exactly as in one uninterrupted run."""
    args = ["-p1e12", "-P1012e9", "-k3", "-K9999", "-n100", "-N2000000", "-q", "-r", "1", "--chunk", "64M"]
    full = tmp_path / "full"
    part = tmp_path / "part"
    full.mkdir()
    part.mkdir()
    rc, _, err_full = run(args, full)
    assert rc == 0, err_full
    tail_full = [l for l in err_full.splitlines() if l.startswith("count=")]

    p = subprocess.Popen([EXE] + args, cwd=part, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True,
                         bufsize=1)
    t0 = time.time()
    for line in p.stdout:  # the first status report ("p=..., ... p/sec") comes after a second
        if line.startswith("p=") or "p/sec" in line or time.time() - t0 > 60:
            break
    p.send_signal(signal.SIGINT)
    _, err1 = p.communicate(timeout=120)
    assert "Sieve incomplete" in err1, err1
    assert os.path.exists(part / "ppcheck1e12.txt")
    rc, _, err2 = run(args, part)
    assert rc == 0, err2
    assert "Resuming from checkpoint" in err2
    assert [l for l in err2.splitlines() if l.startswith("count=")] == tail_full
    assert factors(part) == factors(full)
    assert len(factors(full)) > 1000


def test_periodic_checkpoint(tmp_path):
    args = ["-p1e12", "-P1001e9", "-k3", "-K9999", "-n100", "-N2000000", "-q", "-c", "1", "--chunk", "32M"]
    rc, out, err = run(args, tmp_path)
    assert rc == 0, err
    assert "Sieve complete" in err
    assert not os.path.exists(tmp_path / "ppcheck1e12.txt")
