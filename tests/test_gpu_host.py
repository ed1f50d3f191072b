"""End to end through the ppsieve program on the GPU: the factors file
(pps/app.c:134 lines) against the goldens and the oracle, the host-prime
batch path (app_thread_fun), and interruption + resume from a checkpoint."""
import json
import os
import re
import signal
import subprocess
import time

import numpy as np
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


RESUME_ARGS = ["-p1e12", "-P1012e9", "-k3", "-K9999", "-n100", "-N2000000", "-q", "-r", "1", "--chunk", "64M"]


def _interrupt_then_resume(args, cwd):
    """Starts ppsieve, sends SIGINT after its first status line, and runs it
    again to completion from the checkpoint it wrote.  Returns (checkpoint p,
    the status line, stderr of the resumed run)."""
    p = subprocess.Popen([EXE] + args, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True,
                         bufsize=1)
    t0 = time.time()
    status = ""
    for line in iter(p.stdout.readline, ""):  # the first report ("p=..., ... p/sec") comes after a second
        if "p/sec" in line or time.time() - t0 > 60:
            status = line
            break
    p.send_signal(signal.SIGINT)
    _, err1 = p.communicate(timeout=120)
    assert "Sieve incomplete" in err1, err1
    with open(os.path.join(cwd, "ppcheck1e12.txt")) as f:
        ck = int(f.readline().split(",")[1].split("=")[1])
    rc, _, err2 = run(args, cwd)
    assert rc == 0, err2
    assert f"Resuming from checkpoint p={ck}" in err2
    return ck, status, err2


def _check_resumed_run_vs_oracle(cwd, ck, err):
    """The resumed run's count=/sum= line is the oracle's sieve of the whole
    range; every factor line is a valid divisor that test_factor keeps,
    none appears twice, and around the resume point -- 5e6 on either side,
    one chunk finished before the interruption and one after -- the lines
    are exactly the oracle's walk of those primes."""
    lo, hi = 10**12, 1012 * 10**9
    c, s = orc.count_primes(lo, hi, threads=16)
    assert f"count={c},sum=0x{s:016x}" in err
    lines = factors(cwd)
    assert len(lines) == len(set(lines)) > 1000
    pr = orc.derive(pmin=lo, pmax=hi, kmin=3, kmax=9999, nmin=100, nmax=2000000)
    hits = np.array([orc.parse_line(l) for l in lines], dtype=orc.HIT_DTYPE)
    flags = orc.check_hits(pr, hits, threads=16)
    assert np.all(flags == orc.KEPT), "a factor line is not a kept divisor"
    assert lo < ck < hi
    a, b = max(lo, ck - 5 * 10**6), min(hi, ck + 5 * 10**6)
    want = orc.walk(pr, orc.primes(a, b), filter=True, threads=16)
    got = sorted(l for l in lines if a <= int(l.split()[0]) < b)
    assert got == sorted(f"{int(x['p'])} | {int(x['k'])}*2^{int(x['n'])}{int(x['c']):+d}" for x in want)
    assert got


def test_interrupt_and_resume(tmp_path):
    """SIGINT mid-run writes a checkpoint at the last finished chunk; the rerun
    resumes there (main.c:332-372).  The result is checked against the
    oracle (count, sum, every line, the lines around the resume point) and
    equals one uninterrupted run.  The status line carries the ETA field
    (main.c:276-291)."""
    full = tmp_path / "full"
    part = tmp_path / "part"
    full.mkdir()
    part.mkdir()
    rc, _, err_full = run(RESUME_ARGS, full)
    assert rc == 0, err_full
    ck, status, err2 = _interrupt_then_resume(RESUME_ARGS, str(part))
    assert re.search(r"p=\d+, [\d.]+[KM]? p/sec, [\d.]+ CPU cores, [\d.]+% done\. ETA \d\d \w{3} \d\d:\d\d", status), status
    _check_resumed_run_vs_oracle(str(part), ck, err2)
    assert [l for l in err2.splitlines() if l.startswith("count=")] == \
        [l for l in err_full.splitlines() if l.startswith("count=")]
    assert factors(part) == factors(full)


def test_two_workers_share_device_0(tmp_path):
    """-t 2 -d 0: two worker threads, each with its own contexts, take p-chunks
    of one range from the shared frontier (main.c:1026-1031, pps/app.c:801-805)
    on the one device: the pps/many_n_test.sh range in 2 Mi chunks gives the
    golden factors and count/sum, and an interrupted two-worker run resumes
    to the oracle's result."""
    rc, out, err = run(MANY_N + ["-q", "-t", "2", "-d", "0", "--chunk", "2m"], tmp_path)
    assert rc == 0, err
    assert "Thread 0 starting" in err and "Thread 1 starting" in err
    assert factors(tmp_path) == sorted(read_lines("many_n_filtered.txt"))
    r = RUNS["many_n"]
    assert f"count={r['count']},sum=0x{r['sum']:016x}" in err
    part = tmp_path / "part"
    part.mkdir()
    args = RESUME_ARGS[:-1] + ["16M", "-t", "2", "-d", "0"]
    ck, _, err2 = _interrupt_then_resume(args, str(part))
    _check_resumed_run_vs_oracle(str(part), ck, err2)


def test_periodic_checkpoint(tmp_path):
    args = ["-p1e12", "-P1001e9", "-k3", "-K9999", "-n100", "-N2000000", "-q", "-c", "1", "--chunk", "32M"]
    rc, out, err = run(args, tmp_path)
    assert rc == 0, err
    assert "Sieve complete" in err
    assert not os.path.exists(tmp_path / "ppcheck1e12.txt")
