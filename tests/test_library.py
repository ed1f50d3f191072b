"""CPU-side checks of libppsd.so: it loads without a GPU, exports every entry
point declared in include/*.h, and its parameter rules equal the oracle's."""
import ctypes
import os
import random
import re

import pytest

import pps_oracle as orc
import ppsd
from conftest import REPO


def declared_functions():
    names = set()
    inc = os.path.join(REPO, "include")
    for fn in os.listdir(inc):
        if fn.endswith(".h"):
            txt = open(os.path.join(inc, fn)).read()
            txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
            names |= set(re.findall(r"\b(ppsd_\w+)\s*\(", txt))
    return names


def test_exports_every_declared_symbol():
    lib = ctypes.CDLL(ppsd.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 16
    for name in sorted(names):
        assert hasattr(lib, name), name


def _same(kw):
    a = orc.derive(**kw)
    sp = ppsd.SearchParams(**kw)
    b = sp.to_c()
    rc = ppsd.lib().ppsd_derive_params(ctypes.byref(b))
    assert rc == b.status
    assert a.status == b.status, (kw, a.msg, b.msg)
    if a.status == 0:
        for f in ("pmin", "pmax", "kmin", "kmax", "nmin", "nmax", "nstep", "nstep_ref"):
            assert getattr(a, f) == getattr(b, f), (kw, f)
    else:
        assert a.msg == b.msg


def test_derive_matches_oracle_random():
    rng = random.Random(11)
    for _ in range(3000):
        pmin = rng.choice([rng.randrange(3, 2**33), rng.randrange(2**32, 2**62), 10**12, 2**50, 2**61])
        pmax = rng.choice([0, pmin + rng.randrange(0, 10**9), 2**62 + 5])
        kmin = rng.choice([0, rng.randrange(1, 10**6)])
        kmax = rng.choice([0, rng.randrange(1, 10**7), rng.randrange(1, 2**40)])
        nmin = rng.choice([0, rng.randrange(1, 200), rng.randrange(1, 10**7)])
        nmax = rng.choice([0, nmin + rng.randrange(0, 10**7), rng.randrange(1, 10**7)])
        sign = rng.choice((1, -1, 0))
        ks = rng.choice([(0, 0), (2, 1), (6, 3), (4, 1)])
        _same(dict(pmin=pmin, pmax=pmax, kmin=kmin, kmax=kmax, nmin=nmin, nmax=nmax, sign=sign,
                   kstep=ks[0], koffset=ks[1]))


def test_derive_rejects_bad_sign():
    with pytest.raises(ppsd.PPSDError) as e:
        ppsd.derive(ppsd.SearchParams(pmin=10**12, kmin=3, kmax=99, nmin=100, nmax=200, sign=2))
    assert e.value.code == ppsd.PPSD_ERR_INVALID_PARAM


def test_open_needs_derived_params():
    p = ppsd.SearchParams(pmin=10**12, kmin=3, kmax=99, nmin=100, nmax=200).to_c()
    p.status = -178
    h = ctypes.c_void_p()
    rc = ppsd.lib().ppsd_open(0, ctypes.byref(p), 0, ctypes.byref(h))
    assert rc == ppsd.PPSD_ERR_INVALID_PARAM and not h.value


def test_collect_without_submit_is_an_error():
    rc = ppsd.lib().ppsd_collect(None, None, 0, None, None, None)
    assert rc == ppsd.PPSD_ERR_INVALID_PARAM
    assert ppsd.lib().ppsd_close(None) == 0


def test_derive_rejects_exponents_past_31_bits():
    """-n/-N parse at most 2^31-1 (pps/app.c:579-580); the library rejects
    more instead of letting a window start or hit exponent wrap 32 bits."""
    base = dict(pmin=2**50, pmax=2**50 + 10**6, kmin=3, kmax=9999, nmin=100)
    _same(dict(base, nmax=2**31 - 1))
    for nmax in (2**31, 2**32 - 1):
        _same(dict(base, nmax=nmax))
        with pytest.raises(ppsd.PPSDError) as e:
            ppsd.derive(ppsd.SearchParams(**dict(base, nmax=nmax)))
        assert e.value.code == ppsd.PPSD_ERR_INVALID_PARAM
    p = ppsd.manual_params(2**50, 2**50 + 10**6, 3, 9999, 100, 2**31, nstep=32)
    h = ctypes.c_void_p()
    assert ppsd.lib().ppsd_open(0, ctypes.byref(p), 0, ctypes.byref(h)) == ppsd.PPSD_ERR_INVALID_PARAM


def test_walk_form_choice_without_a_device():
    """The automatic pre-test choice (ppsd_walk_form): the range form for the
    2^61 both-signs search with kmax 1e5 on the plain walk's longest chunk,
    the odd-low form for the 2^50 search with kmax 9999, the exact form at
    2^46 with kmax 1e6; the batched walk keeps the chunk it is given."""
    cfg3 = ppsd.derive(ppsd.SearchParams(pmin=2**61, pmax=2**61 + 2 * 10**9, kmin=3, kmax=100001, nmin=100000,
                                         nmax=10000000, sign=0))
    cfg2 = ppsd.derive(ppsd.SearchParams(pmin=2**50, pmax=2**50 + 5 * 10**10, kmin=1201, kmax=9999, nmin=50000,
                                         nmax=6000000))
    big_k = ppsd.derive(ppsd.SearchParams(pmin=2**46, pmax=2**46 + 10**6, kmin=3, kmax=999999, nmin=100000,
                                          nmax=2000000))
    assert ppsd.walk_form(cfg3, 1, 16) == (2, 16)
    assert ppsd.walk_form(cfg3, 1, 8) == (2, 8)
    assert ppsd.walk_form(cfg2, 1, 16) == (3, 16)
    assert ppsd.walk_form(big_k, 1, 16)[0] == 1
    assert ppsd.walk_form(cfg3, 2, 8, batched=True)[1] == 8
    with pytest.raises(ppsd.PPSDError):
        ppsd.walk_form(cfg3, 3, 16)
