"""Sieve-file input (-i) of the ppsieve program: format detection, the
ranges and sign a file supplies, term counts and errors.  All of it runs
before a device is opened, so it is checked on the CPU (the run then stops
for want of a GPU)."""
import os

import pytest

from sievefiles import write_abcd, write_newpgen
from test_host_cli import run

TERMS = [(1203, 120), (1203, 121), (1203, 130), (1205, 150), (3001, 100), (9999, 4000), (4321, 2000)]


@pytest.fixture
def wd(tmp_path):
    return str(tmp_path)


def test_abcd_scan_and_read(wd):
    write_abcd(os.path.join(wd, "s.txt"), TERMS, sign=+1)
    _, out, _ = run(["-p1e12", "-i", "s.txt", "-q"], wd)
    assert "Scanning ABCD file..." in out
    assert "Found K's from 1203 to 9999." in out
    assert "Found N's from 100 to 4001." in out
    assert "Reading ABCD file." in out
    assert "Read 7 terms from ABCD format input file `s.txt'" in out
    assert "ppsieve initialized: 1203 <= k <= 9999, 100 <= n < 4001" in out


def test_newpgen_scan_and_read(wd):
    write_newpgen(os.path.join(wd, "s.txt"), TERMS, sign=-1)
    _, out, _ = run(["-p1e12", "-i", "s.txt", "-q"], wd)
    assert "Scanning ABCD file..." not in out
    assert "Found K's from 1203 to 9999." in out and "Found N's from 100 to 4001." in out
    assert "Read 7 terms from NewPGen format input file `s.txt'" in out


def test_command_line_narrows_the_file(wd):
    write_newpgen(os.path.join(wd, "s.txt"), TERMS, sign=+1)
    _, out, _ = run(["-p1e12", "-i", "s.txt", "-k1205", "-n125", "-q"], wd)
    assert "ppsieve initialized: 1205 <= k <= 9999, 125 <= n < 4001" in out
    # (1203, *) is below kmin, (3001, 100) below nmin
    assert "Read 3 terms from NewPGen format input file" in out
    # every range given: the file is not scanned for them, its terms still read
    _, out, _ = run(["-p1e12", "-i", "s.txt", "-k3", "-K9999", "-n100", "-N130", "-q"], wd)
    assert "Found K's" not in out and "ppsieve initialized: 3 <= k <= 9999, 100 <= n < 130" in out
    assert "Read 4 terms" in out  # n <= nmax is kept in the bitmap (pps/app.c:394), as the reference


def test_abcd_counts_every_term(wd):
    write_abcd(os.path.join(wd, "s.txt"), TERMS, sign=+1)
    _, out, _ = run(["-p1e12", "-i", "s.txt", "-k3", "-K9999", "-n100", "-N130", "-q"], wd)
    assert "Read 7 terms from ABCD format input file" in out  # pps/app.c:459 counts all


@pytest.mark.parametrize("text,msg", [
    ("hello\n1 2\n", "Invalid header in input file `s.txt'"),
    ("1000:P:1:2:7\n3 100\n", "Invalid header in input file `s.txt'"),
    ("1000:P:1:2:1\n", "Invalid line 2 in input file `s.txt'"),
    ("1000:P:1:2:1\n1203 100\n1204 101\n", "Invalid line 2 in input file `s.txt'"),
])
def test_input_errors(wd, text, msg):
    with open(os.path.join(wd, "s.txt"), "w") as f:
        f.write(text)
    rc, out, err = run(["-p1e12", "-i", "s.txt", "-q"] + (["-k3", "-K9999", "-n100", "-N200"] if "1204" in text else []),
                       wd)
    assert rc == (-140) & 0xFF and msg in err


def test_missing_input_file(wd):
    rc, _, err = run(["-p1e12", "-i", "nope.txt", "-q"], wd)
    assert rc == (-140) & 0xFF and "nope.txt: cannot open" in err


# ---- the twin-prime build (tpsieve): k = 6a+3, both signs

TP = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "psieve-cuda_amd", "host", "tpsieve")
TWIN_TERMS = [(3, 120), (9, 120), (27, 120), (15, 130), (999, 400), (1005, 100)]


def run_tp(args, cwd):
    import subprocess
    env = dict(os.environ, HIP_VISIBLE_DEVICES="")
    r = subprocess.run([TP] + args, cwd=cwd, capture_output=True, text=True, timeout=120, env=env)
    return r.returncode, r.stdout, r.stderr


def test_twin_defaults(wd):
    _, out, _ = run_tp(["-p1e12", "-k1", "-K1000", "-n100", "-N1000", "-q"], wd)
    assert "tpsieve initialized: 3 <= k <= 999, 100 <= n < 1000" in out
    assert os.path.exists(os.path.join(wd, "tpfactors.txt"))


def test_twin_abcd(wd):
    from sievefiles import write_abcd_twin
    write_abcd_twin(os.path.join(wd, "t.txt"), TWIN_TERMS)
    _, out, _ = run_tp(["-p1e12", "-i", "t.txt", "-q"], wd)
    assert "Found K's from 3 to 1005." in out and "Found N's from 100 to 401." in out
    assert "Read 6 terms from ABCD format input file `t.txt'" in out
    assert "tpsieve initialized: 3 <= k <= 1005, 100 <= n < 401" in out


def test_twin_newpgen(wd):
    from sievefiles import write_newpgen_twin
    write_newpgen_twin(os.path.join(wd, "t.txt"), TWIN_TERMS)
    _, out, _ = run_tp(["-p1e12", "-i", "t.txt", "-q"], wd)
    assert "Read 6 terms from NewPGen format input file `t.txt'" in out
    # a Proth-form file is not a twin file
    write_newpgen(os.path.join(wd, "p.txt"), TWIN_TERMS, sign=+1)
    rc, _, err = run_tp(["-p1e12", "-i", "p.txt", "-q"], wd)
    assert rc == (-140) & 0xFF and "Invalid header" in err
