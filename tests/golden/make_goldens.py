"""Writes the golden fixtures under tests/golden/ (run in the dev container,
where the reference tree and the oracle build exist).

Fixtures taken from the reference's own test data (data files, copied as-is
or reduced by closed form):
  oldsort.txt            pps/oldsort.txt, the golden of pps/many_n_test.sh
  many_n_filtered.txt    the oldsort lines with no prime factor 3 <= q < 32768
                         (trial division): what today's filter (pps/app.c:150-186)
                         keeps for the many_n_test range
  negative_lines.txt     non-divisor lines of pps/newsort.txt and pps/test.txt
  pps4test_halving.txt   column A/B of pps4test.ods: n, 2^(1-n) mod 10007
  factor_proth_ref.txt   inputs and outputs of the compiled reference
                         pps/factor_proth.c stdin driver (oracle/_ref)

Fixtures computed by the oracle and cross-checked here before writing:
  many_n_raw.txt         every divisor of the many_n_test range (no filter);
                         its 3/5/7-free subset must equal oldsort.txt exactly
  cfg1_factors.txt       factors of -p1e12 -P1000100M -k3 -K9999 -n100 -N200000
  cfg2_slice.txt / cfg3_slice.txt   factors of p-slices at the 2^50 / 2^61 sizes
  runs.json              prime count/sum and factor counts of each oracle run
"""
import json
import os
import random
import re
import subprocess
import sys
import zipfile

import sympy

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
REF = "/root/reference"
sys.path.insert(0, os.path.join(REPO, "oracle"))
import pps_oracle as orc  # noqa: E402

SMALL = list(sympy.primerange(3, 32768))


def parse(line):
    return orc.parse_line(line)


def fmt(t):
    p, k, n, c = t
    return f"{p} | {k}*2^{n}{c:+d}"


def smallest_factor(k, n, c):
    for q in SMALL:
        if (k * pow(2, n, q) + c) % q == 0:
            return q
    return 0


def write_lines(name, lines):
    with open(os.path.join(HERE, name), "w") as f:
        for line in lines:
            f.write(line + "\n")


def main():
    runs = {}
    # reference fixtures -----------------------------------------------------
    old = [l.strip() for l in open(os.path.join(REF, "pps/oldsort.txt")) if l.strip()]
    write_lines("oldsort.txt", old)
    filt = [l for l in old if smallest_factor(*parse(l)[1:]) == 0]
    write_lines("many_n_filtered.txt", filt)
    neg = []
    for fn in ("pps/newsort.txt", "pps/test.txt"):
        for l in open(os.path.join(REF, fn)):
            if "|" not in l:
                continue
            p, k, n, c = parse(l)
            if (k * pow(2, n, p) + c) % p != 0:
                neg.append(l.strip())
    write_lines("negative_lines.txt", neg)

    x = zipfile.ZipFile(os.path.join(REF, "pps4test.ods")).read("content.xml").decode()
    rows = re.findall(r"<table:table-row[^>]*>(.*?)</table:table-row>", x, re.S)
    halving = []
    for r in rows[1:]:
        vals = re.findall(r'office:value="([^"]*)"', r)
        if len(vals) >= 2:
            halving.append(f"{int(float(vals[0]))} {int(float(vals[1]))}")
    write_lines("pps4test_halving.txt", halving[:80])

    # reference factor_proth driver on random + lines ------------------------
    subprocess.run(["make", "-s", "-C", os.path.join(REPO, "oracle"), "ref"], check=True)
    rng = random.Random(20101)
    inp = []
    for _ in range(400):
        k = rng.randrange(1, 2**20) | 1
        n = rng.randrange(64, 3_000_000)
        inp.append(f"1 | {k}*2^{n}+1")
    out = subprocess.run([orc.REF_FACTOR_PROTH], input="\n".join(inp) + "\n", capture_output=True, text=True,
                         check=True).stdout.strip().split("\n")
    assert len(out) == len(inp)
    write_lines("factor_proth_ref.txt", [f"{a}\t{b}" for a, b in zip(inp, out)])

    # oracle runs -------------------------------------------------------------
    def run(tag, threads=8, raw=False, **kw):
        pr = orc.derive(**kw)
        assert pr.status == 0, pr.msg
        hits, cnt, psum = orc.run_range(pr, pr.pmin, pr.pmax, filter=not raw, threads=threads)
        lines = sorted(fmt(t) for t in orc.hit_set(hits))
        runs[tag] = {"params": kw, "nstep": pr.nstep, "nstep_ref": pr.nstep_ref, "nmin": pr.nmin,
                     "nmax": pr.nmax, "kmin": pr.kmin, "kmax": pr.kmax, "count": cnt,
                     "sum": psum & (2**64 - 1), "lines": len(lines), "raw": raw}
        return lines

    raw = run("many_n_raw", raw=True, pmin=42070000000000, pmax=42070050000000, kmin=1201, kmax=9999,
              nmax=2000000)
    free357 = [l for l in raw if all((lambda t: (t[1] * pow(2, t[2], q) + t[3]) % q)(parse(l)) for q in (3, 5, 7))]
    assert free357 == sorted(old), "oracle raw set does not reproduce pps/oldsort.txt"
    for l in raw:
        p, k, n, c = parse(l)
        assert (k * pow(2, n, p) + c) % p == 0 and sympy.isprime(p)
    write_lines("many_n_raw.txt", raw)
    filtered = run("many_n", pmin=42070000000000, pmax=42070050000000, kmin=1201, kmax=9999, nmax=2000000)
    assert filtered == sorted(filt), "oracle filter disagrees with trial division on oldsort"

    write_lines("cfg1_factors.txt", run("cfg1", pmin=10**12, pmax=10**12 + 10**8, kmin=3, kmax=9999, nmin=100,
                                        nmax=200000))
    write_lines("cfg2_slice.txt", run("cfg2_slice", pmin=2**50, pmax=2**50 + 3 * 10**6, kmin=1201, kmax=9999,
                                      nmin=50000, nmax=6000000))
    write_lines("cfg3_slice.txt", run("cfg3_slice", pmin=2**61, pmax=2**61 + 10**6, kmin=3, kmax=100001,
                                      nmin=100000, nmax=10000000, sign=0))
    with open(os.path.join(HERE, "runs.json"), "w") as f:
        json.dump(runs, f, indent=1, sort_keys=True)
    print(json.dumps(runs, indent=1))


FULL = {
    # the single large unit at 2^50 and the high-p both-signs unit at 2^61
    "cfg2_full": dict(pmin=2**50, pmax=2**50 + 5 * 10**10, kmin=1201, kmax=9999, nmin=50000, nmax=6000000),
    "cfg3_full": dict(pmin=2**61, pmax=2**61 + 2 * 10**9, kmin=3, kmax=100001, nmin=100000, nmax=10000000, sign=0),
}


def full_counts():
    """Prime count and sum (the reference's count=/sum=) of the full 2^50 and
    2^61 ranges, by the oracle's threaded sieve; merged into runs.json.  Too
    large for the oracle's walk: the GPU runs over these ranges are checked
    by these counts and by the validity of every factor they report."""
    path = os.path.join(HERE, "runs.json")
    with open(path) as f:
        runs = json.load(f)
    for tag, kw in FULL.items():
        cnt, psum = orc.count_primes(kw["pmin"], kw["pmax"], threads=8)
        runs[tag] = {"params": kw, "count": cnt, "sum": psum}
        print(tag, cnt, hex(psum), flush=True)
    with open(path, "w") as f:
        json.dump(runs, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    if sys.argv[1:] == ["--full-counts"]:
        full_counts()
    else:
        main()
