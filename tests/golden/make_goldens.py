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
  cfg2_windows[_raw].txt / cfg3_windows[_raw].txt
                         factors (and raw divisors) of p-windows of the 2^50
                         and 2^61 (both signs) searches, each window around a
                         divisor (see windows())
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
    with open(os.path.join(HERE, "runs.json"), "w") as f:
        json.dump(runs, f, indent=1, sort_keys=True)
    windows()


# Window centres: p-values of divisors that a device run over the whole
# 2^50 search (24 of its 3601 factors, evenly spaced) and over the whole 2^61
# both-signs search (all 22 of its raw divisors) reported.  They only place
# the windows, so that the goldens hold hits (a contiguous slice of either
# search the oracle can walk holds next to none); what each window contains
# is decided by the oracle's complete walk of every prime in it.
CFG2_CENTRES = [1125901037288939, 1125903237675401, 1125905246522519, 1125907220481349, 1125909255010373,
                1125911391247831, 1125913426274423, 1125915649379629, 1125917588337161, 1125919643038243,
                1125921460578587, 1125923800585219, 1125926030070847, 1125928229982109, 1125930444025969,
                1125932566212599, 1125934883685323, 1125936917007559, 1125939040888273, 1125940925008381,
                1125942897869017, 1125945114077401, 1125947009987603, 1125948965101183]
CFG3_CENTRES = [2305843009378763033, 2305843009432867523, 2305843009554438901, 2305843009571129923,
                2305843009741738669, 2305843009813309303, 2305843009835655683, 2305843010018116507,
                2305843010100330797, 2305843010303526457, 2305843010330208617, 2305843010377028831,
                2305843010400013697, 2305843010443865203, 2305843010487221651, 2305843010568480431,
                2305843010612587969, 2305843010809514747, 2305843010938955149, 2305843011046819063,
                2305843011047201137, 2305843011190992643]
WINDOWED = {
    "cfg2_windows": (dict(pmin=2**50, pmax=2**50 + 5 * 10**10, kmin=1201, kmax=9999, nmin=50000, nmax=6000000),
                     CFG2_CENTRES, 350000),
    "cfg3_windows": (dict(pmin=2**61, pmax=2**61 + 2 * 10**9, kmin=3, kmax=100001, nmin=100000, nmax=10000000,
                          sign=0), CFG3_CENTRES, 200000),
}


def merged_windows(centres, half, lo, hi):
    """[c - half, c + half) around each centre, clipped to [lo, hi), with
    overlapping windows merged."""
    out = []
    for c in sorted(centres):
        a, b = max(lo, c - half), min(hi, c + half)
        if out and a <= out[-1][1]:
            out[-1][1] = max(out[-1][1], b)
        else:
            out.append([a, b])
    return out


def windows():
    """Oracle runs over p-windows of the full 2^50 and 2^61 searches:
    <tag>_raw.txt every divisor, <tag>.txt the factors test_factor keeps,
    and in runs.json the windows, their prime count and sum."""
    path = os.path.join(HERE, "runs.json")
    with open(path) as f:
        runs = json.load(f)
    runs.pop("cfg2_slice", None)
    runs.pop("cfg3_slice", None)
    for tag, (kw, centres, half) in WINDOWED.items():
        pr = orc.derive(**kw)
        assert pr.status == 0, pr.msg
        wins = merged_windows(centres, half, kw["pmin"], kw["pmax"])
        raw, kept, cnt, psum = set(), set(), 0, 0
        for a, b in wins:
            P = orc.primes(a, b)
            raw |= orc.hit_set(orc.walk(pr, P, filter=False, threads=8))
            kept |= orc.hit_set(orc.walk(pr, P, filter=True, threads=8))
            cnt += len(P)
            psum += int(P.sum(dtype="uint64"))
        for p, k, n, c in raw:
            assert (k * pow(2, n, p) + c) % p == 0 and sympy.isprime(p)
        assert kept <= raw and len(kept) > 0 and len(raw) >= 20, (tag, len(kept), len(raw))
        write_lines(tag + "_raw.txt", sorted(fmt(t) for t in raw))
        write_lines(tag + ".txt", sorted(fmt(t) for t in kept))
        runs[tag] = {"params": kw, "windows": wins, "nstep": pr.nstep, "count": cnt, "sum": psum & (2**64 - 1),
                     "lines": len(kept), "raw_lines": len(raw)}
        print(tag, len(wins), "windows", cnt, "primes", len(raw), "divisors", len(kept), "factors", flush=True)
    with open(path, "w") as f:
        json.dump(runs, f, indent=1, sort_keys=True)


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
    elif sys.argv[1:] == ["--windows"]:
        windows()
    else:
        main()
