"""The ppsieve program (framework + application hooks over libppsd): option
grammar, ppconfig.txt, parameter errors and checkpoint validation.  These
paths all run before a device is opened, so they are checked on the CPU."""
import os
import subprocess

import pytest

from conftest import REPO

EXE = os.path.join(REPO, "psieve-cuda_amd", "host", "ppsieve")
MANY_N = ["-p42070e9", "-P42070050e6", "-k", "1201", "-K", "9999", "-N", "2000000"]


def run(args, cwd, env_extra=None):
    env = dict(os.environ)
    env["HIP_VISIBLE_DEVICES"] = env.get("PPSIEVE_TEST_DEVICES", "")
    if env_extra:
        env.update(env_extra)
    r = subprocess.run([EXE] + args, cwd=cwd, capture_output=True, text=True, timeout=120, env=env)
    return r.returncode, r.stdout, r.stderr


@pytest.fixture
def wd(tmp_path):
    return str(tmp_path)


def test_derived_parameters_printed(wd):
    _, out, err = run(MANY_N + ["-q"], wd)
    assert "nstart=76, nstep=32" in out
    assert "ppsieve initialized: 1201 <= k <= 9999, 76 <= n < 2000000" in out
    assert "Sieve started: 42070000000000 <= p < 42070050000000" in err


@pytest.mark.parametrize("args,line", [
    (["-p1p", "-k3", "-K10K", "-n64", "-N1e6"], "3 <= k <= 10001, 64 <= n < 1000000"),
    (["-p", "1125899906842624", "-k1201", "-K", "9999", "-n", "5e4", "-N6M"], "1201 <= k <= 9999, 50000 <= n < 6000000"),
    (["-p2b50", "-k3", "-K9999", "-n100", "-N200k"], "3 <= k <= 9999, 100 <= n < 204800"),
    (["-p1p", "-k5", "-K100", "-n100", "-N1000", "-M6", "-s3"], "3 <= k <= 99, 100 <= n < 1000"),
])
def test_value_suffixes(wd, args, line):
    _, out, _ = run(args, wd)
    assert f"ppsieve initialized: {line}" in out


def test_config_file(wd):
    with open(os.path.join(wd, "ppconfig.txt"), "w") as f:
        f.write("# comment\npmin=42070e9\npmax 42070050e6\nkmin=1201\nkmax=9999\nnmax=2M\nquiet\n")
    _, out, err = run([], wd)
    assert "ppsieve initialized: 1201 <= k <= 9999, 76 <= n < 2000000" in out
    # the command line overrides the file
    _, out, _ = run(["-K", "5001"], wd)
    assert "1201 <= k <= 5001" in out


def test_config_file_errors(wd):
    with open(os.path.join(wd, "ppconfig.txt"), "w") as f:
        f.write("bogus=3\n")
    rc, _, err = run(MANY_N, wd)
    assert rc == 1 and "unrecognised option `bogus'" in err
    with open(os.path.join(wd, "ppconfig.txt"), "w") as f:
        f.write("quiet=1\n")
    rc, _, err = run(MANY_N, wd)
    assert rc == 1 and "doesn't allow an argument" in err
    with open(os.path.join(wd, "ppconfig.txt"), "w") as f:
        f.write("kmin\n")
    rc, _, err = run(MANY_N, wd)
    assert rc == 1 and "requires an argument" in err


@pytest.mark.parametrize("args,code,msg", [
    (["-p1000", "-k3", "-K9999", "-n100", "-N200"], -178, "kmax < pmin is required"),
    (["-p1e12", "-k9", "-K3", "-n100", "-N200"], -178, "kmin <= kmax is required"),
    (["-p1e12", "-k3", "-K99", "-n300", "-N200"], -178, "nmin <= nmax is required"),
    (["-p1e12", "-k3", "-K99", "-n32", "-N200"], -178, "nmin too small"),
    (["-p1e12", "-n100", "-N200"], -178, "Please specify"),
    (["-p1e12", "-k3", "-K99", "-n100", "-N200", "-iabcd.txt"], -140, "abcd.txt: cannot open"),
    (["-p4e18", "-P3e18", "-k3", "-K99"], 1, "pmax must be greater than pmin"),
])
def test_parameter_errors(wd, args, code, msg):
    rc, _, err = run(args, wd)
    assert rc == code & 0xFF
    assert msg in err


@pytest.mark.parametrize("args,msg", [(["-p1x"], "invalid argument -p 1x"),
                                      (["-p", "5e30"], "out of range argument -p 5e30"),
                                      (["--kmax=1x"], "invalid argument --kmax 1x")])
def test_option_value_errors(wd, args, msg):
    rc, _, err = run(args, wd)
    assert rc == 1 and msg in err


def test_invalid_checkpoint_ignored(wd):
    with open(os.path.join(wd, "ppcheck42070e9.txt"), "w") as f:
        f.write("pmin=42070000000000,p=42070000500001,count=5,sum=0x0000000000000001,checksum=0x0000000000000000\n"
                "nmin=76,nmax=2000000,factor_count=0\n")
    _, _, err = run(MANY_N, wd)
    assert "Ignoring invalid checkpoint in ppcheck42070e9.txt" in err


def test_valid_checkpoint_needs_factors_file(wd):
    pmin, p, count, s = 42070000000000, 42070000500001, 5, 1
    with open(os.path.join(wd, "ppcheck42070e9.txt"), "w") as f:
        f.write(f"pmin={pmin},p={p},count={count},sum=0x{s:016x},checksum=0x{pmin + p + count + s:016x}\n"
                "nmin=76,nmax=2000000,factor_count=3\n")
    _, _, err = run(MANY_N, wd)
    assert "Factors file missing.  Restarting test." in err
    assert "Ignoring invalid checkpoint" in err
    with open(os.path.join(wd, "ppfactors.txt"), "w") as f:
        f.write("x\n")
    _, _, err = run(MANY_N, wd)
    assert f"Resuming from checkpoint p={p} in ppcheck42070e9.txt" in err


def test_no_device_fails(wd):
    rc, _, err = run(MANY_N, wd)
    assert rc == 1
    assert "not usable" in err


def _ckpt_line(pmin, p, count, s, fmt_sum="0x{:016x}"):
    return f"pmin={pmin},p={p},count={count},sum={fmt_sum.format(s)},checksum=0x{pmin + p + count + s:016x}\n"


@pytest.mark.parametrize("first,second,resumes", [
    (_ckpt_line(42070000000000, 42070000500001, 5, 10), "nmin=76,nmax=2000000,factor_count=0\n", True),
    (_ckpt_line(42070000000000, 42070000500001, 5, 10, "0x{:X}"), "nmin=76,nmax=2000000,factor_count=0\n", True),
    # fields out of order, missing, or trailing text: not a checkpoint record
    ("p=42070000500001,pmin=42070000000000,count=5,sum=0xa,checksum=0x0\n", "nmin=76,nmax=2000000,factor_count=0\n",
     False),
    (_ckpt_line(42070000000000, 42070000500001, 5, 10).replace(",count=5", ""),
     "nmin=76,nmax=2000000,factor_count=0\n", False),
    (_ckpt_line(42070000000000, 42070000500001, 5, 10)[:-1] + ",x=1\n", "nmin=76,nmax=2000000,factor_count=0\n",
     False),
    # the application's record names another exponent range, or is cut short
    (_ckpt_line(42070000000000, 42070000500001, 5, 10), "nmin=77,nmax=2000000,factor_count=0\n", False),
    (_ckpt_line(42070000000000, 42070000500001, 5, 10), "nmin=76,nmax=2000000\n", False),
    # resume point outside (pmin, pmax)
    (_ckpt_line(42070000000000, 42070000000000, 5, 10), "nmin=76,nmax=2000000,factor_count=0\n", False),
    (_ckpt_line(42070000000000, 42070050000000, 5, 10), "nmin=76,nmax=2000000,factor_count=0\n", False),
])
def test_checkpoint_records(wd, first, second, resumes):
    """The two checkpoint records are read field by field in the reference's
    format (main.c:305-372, pps/app.c:1301-1352): anything but the exact
    fields in order is ignored, as is a record for other parameters."""
    with open(os.path.join(wd, "ppcheck42070e9.txt"), "w") as f:
        f.write(first + second)
    _, _, err = run(MANY_N, wd)
    if resumes:
        assert "Resuming from checkpoint p=42070000500001 in ppcheck42070e9.txt" in err
    else:
        assert "Ignoring invalid checkpoint in ppcheck42070e9.txt" in err


def test_config_file_spacing(wd):
    """Blank lines, leading blanks, tabs and '=' with blanks around it."""
    with open(os.path.join(wd, "ppconfig.txt"), "w") as f:
        f.write("\n   # indented comment\n\tpmin = 42070e9\npmax\t42070050e6\n  kmin=1201\nkmax =9999\n\nnmax= 2M\n")
    _, out, err = run([], wd)
    assert "ppsieve initialized: 1201 <= k <= 9999, 76 <= n < 2000000" in out, out + err
    with open(os.path.join(wd, "ppconfig.txt"), "w") as f:
        f.write("pmin=42070e9\nkmax=9x9\n")
    rc, _, err = run([], wd)
    assert rc == 1 and "invalid argument kmax 9x9" in err
