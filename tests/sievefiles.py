"""Writers of the two sieve-file formats the reference reads with -i
(pps/app.c:188-517), for tests: NewPGen ("<p>:P:<x>:2:<t>" then "k n"
lines) and ABCD (per k a header "ABCD <k>*2^$a<+|->1 [<n>]" and then one
line per further n holding its increment)."""
import random


def write_abcd(path, terms, sign, comment=" // Sieved to 1000000"):
    """terms: iterable of (k, n)."""
    byk = {}
    for k, n in terms:
        byk.setdefault(k, set()).add(n)
    with open(path, "w") as f:
        for k in sorted(byk):
            ns = sorted(byk[k])
            f.write(f"ABCD {k}*2^$a{'+' if sign > 0 else '-'}1 [{ns[0]}]{comment}\n")
            for a, b in zip(ns, ns[1:]):
                f.write(f"{b - a}\n")


def write_newpgen(path, terms, sign, p0=1000000, code=None):
    if code is None:
        code = 1 if sign > 0 else 255
    with open(path, "w") as f:
        f.write(f"{p0}:P:1:2:{code}\n")
        for k, n in sorted(terms, key=lambda t: (t[1], t[0])):
            f.write(f"{k} {n}\n")


def candidates(hits, kmin, kmax, nmin, nmax, seed, decoys=3000, keep=0.5):
    """A candidate list: about half of the divisors' (k, n) plus random odd-k
    decoys; returns (terms, kept hit pairs)."""
    rng = random.Random(seed)
    pairs = sorted({(k, n) for _, k, n, _ in hits})
    chosen = {t for t in pairs if rng.random() < keep}
    terms = set(chosen)
    while len(terms) < len(chosen) + decoys:
        k = rng.randrange(kmin | 1, kmax + 1, 2)
        terms.add((k, rng.randrange(nmin, nmax)))
    return sorted(terms), chosen


def write_abcd_twin(path, terms, comment=" // Sieved to 1000000"):
    """The twin-prime ABCD form: per n a header with the first a (k = 6a+3),
    then one line per further k holding the increment of a."""
    byn = {}
    for k, n in terms:
        assert k % 6 == 3
        byn.setdefault(n, set()).add((k - 3) // 6)
    with open(path, "w") as f:
        for n in sorted(byn):
            a = sorted(byn[n])
            f.write(f"ABCD (6*$a+3)*2^{n}+1 & (6*$a+3)*2^{n}-1 [{a[0]}]{comment}\n")
            for x, y in zip(a, a[1:]):
                f.write(f"{y - x}\n")


def write_newpgen_twin(path, terms, p0=1000000):
    with open(path, "w") as f:
        f.write(f"{p0}:T:0:2:3\n")
        for k, n in sorted(terms, key=lambda t: (t[1], t[0])):
            f.write(f"{k} {n}\n")
