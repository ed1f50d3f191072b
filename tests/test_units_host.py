"""Host logic of the batched and serving workloads (no GPU): the seeded unit
generator and the Poisson arrival process."""
import math

import numpy as np

import ppsd
import ppsd_serve
import ppsd_units


def test_units_deterministic_and_prefix():
    a = ppsd_units.mix_units(64)
    b = ppsd_units.mix_units(128)
    assert [u.search for u in a] == [u.search for u in b[:64]]


def test_units_distribution_and_domain():
    us = ppsd_units.mix_units()
    assert len(us) == 4096
    for u in us:
        s, p = u.search, u.params
        assert s["sign"] in (1, -1) and s["kmin"] == 3
        assert 1e3 <= s["kmax"] <= 1e6 + 1 and 1000 <= s["nmin"] <= 10**6
        assert 1e4 - 1 <= s["nmax"] - s["nmin"] <= 5e6 + 1
        assert 1e11 - 1 <= s["pmin"] <= 2**60 + 1 and 1e5 - 1 <= s["pmax"] - s["pmin"] <= 1e9 + 1
        # the walk's own domain: one candidate per window, s <= 32
        assert 1 <= p.nstep <= 32 and (p.kmax << (p.nstep - 1)) < p.pmin and p.status == 0
    rej = sum(not u.ref_domain for u in us) / len(us)
    s48 = sum(u.ref_s48 for u in us) / len(us)
    assert 0.09 < rej < 0.14 and 0.01 < s48 < 0.03
    # log-uniform p-width: median near sqrt(1e5 * 1e9)
    w = np.array([u.search["pmax"] - u.search["pmin"] for u in us], dtype=float)
    assert abs(math.log10(np.median(w)) - 7) < 0.2


def test_extended_params_match_derive_where_reference_accepts():
    for u in ppsd_units.mix_units(512):
        if not u.ref_domain or u.ref_s48:
            continue
        e = ppsd_units.extended_params(**u.search)
        if u.params.nstep_ref < 32:
            assert (e.nstep, e.kmin, e.kmax) == (u.params.nstep, u.params.kmin, u.params.kmax)


def test_poisson_arrivals():
    t = ppsd_serve.poisson_arrivals(200.0, 60.0, seed=1)
    assert np.all(np.diff(t) > 0) and t[0] >= 0 and t[-1] < 60
    assert abs(len(t) - 12000) < 4 * math.sqrt(12000)
    gaps = np.diff(t)
    assert abs(gaps.mean() - 1 / 200) < 0.05 / 200
    assert abs(np.median(gaps) - math.log(2) / 200) < 0.05 / 200  # exponential gaps
    assert np.array_equal(t, ppsd_serve.poisson_arrivals(200.0, 60.0, seed=1))


def test_percentiles():
    p = ppsd_serve.percentiles(list(range(101)))
    assert p == {"p50": 50.0, "p99": 99.0}
    assert ppsd_serve.percentiles([]) == {"p50": None, "p99": None}


def test_tuning_cache_roundtrip(tmp_path, monkeypatch):
    import ppsd_tune
    monkeypatch.setattr(ppsd, "device_info", lambda device=0: ("gfx950:sramecc+:xnack-", 256))
    monkeypatch.setattr(ppsd_tune, "CACHE_DIR", str(tmp_path))
    key = ppsd_tune.device_key(0)
    assert key.startswith("gfx950-256cu-") and len(key.split("-")[-1]) == 12
    assert ppsd_tune.load() is None
    entry = {"device": key, "plain": {"best": {"r": 1, "u": 8, "wg": 64}},
             "batch": {"best": {"r": 2, "u": 4, "waves": 8, "seg_min": 2048, "seg_max": 65536}}}
    path = ppsd_tune.save(entry, out_dir=str(tmp_path))
    assert path.startswith(str(tmp_path)) and ppsd_tune.load() == entry

    class Ctx:
        def set_tuning(self, *a):
            self.plain = a

        def set_batch_tuning(self, *a):
            self.batch = a
    c = Ctx()
    assert ppsd_tune.apply_plain(c) and ppsd_tune.apply_batch(c)
    assert c.plain == (1, 8, 64) and c.batch == (2, 4, 8, 2048, 65536)
