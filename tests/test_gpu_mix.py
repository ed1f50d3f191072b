"""The seeded mixed workload (ppsd_units: 4096 units, sign +-1, kmax to 1e6,
nmin to 1e6, n-widths to 5e6, pmin from 1e11 to 2^60, p-widths to 1e9) and
the serving driver's Poisson arrivals of it, at full unit sizes on the
device, checked unit by unit against the oracle (tests/mixcheck.py): prime
count and sum of the whole range, every divisor valid, the filtered set as
test_factor makes it of the raw one, and complete hit sets on p-slices and
on the smaller units (pps/app.c:1037-1121, pps/app.c:150-186)."""
import numpy as np
import pytest

import mixcheck
import ppsd
import ppsd_serve
import ppsd_units

pytestmark = pytest.mark.gpu

MIX_PREFIX = 32


def _per_unit(out, unit, n):
    sets = [set() for _ in range(n)]
    for f, u in zip(out, unit):
        sets[int(u)].add((int(f["p"]), int(f["k"]), int(f["n"]), int(f["c"])))
    return sets


def test_seeded_mix_prefix_batched_vs_oracle():
    """The first 32 units of the mix at their full ranges, as one batched run
    (raw divisors) and a second one (filtered factors)."""
    units = ppsd_units.mix_units(MIX_PREFIX)
    params = [u.params for u in units]
    print(f"mix prefix: {MIX_PREFIX} units, {sum(mixcheck.work(p) for p in params):.3g} window-steps", flush=True)
    with ppsd.Context(params[0]) as ctx:
        raw, raw_unit, cnt, sums = ctx.batch_run(params, raw=True)
        kept, kept_unit, cnt2, sums2 = ctx.batch_run(params)
    assert np.array_equal(cnt, cnt2) and np.array_equal(sums, sums2)
    raw_sets = _per_unit(raw, raw_unit, len(params))
    kept_sets = _per_unit(kept, kept_unit, len(params))
    assert len(raw) == sum(len(s) for s in raw_sets) and len(kept) == sum(len(s) for s in kept_sets), \
        "a hit reported twice"
    full, checked = mixcheck.check_units(params, cnt, sums, raw_sets, raw=True, kept_sets=kept_sets, edge=2000)
    print(f"checked {checked} divisors, {len(kept)} factors; {full} units walked whole by the oracle", flush=True)
    assert checked > 10000 and len(kept) > 100 and full >= 10
    # units outside the reference's validation (smallest one-candidate window) are among them
    assert any(not u.ref_domain for u in units)


def test_poisson_serving_200_per_s_vs_oracle():
    """Config-5 arrivals: Poisson at 200 requests/s for 1 s (the mix's
    stream, about 3.4x the device's capacity), admitted by the memory budget and
    searched in 50 ms slices while others arrive; after a drain every
    request has completed and each result passes the oracle checks."""
    rep, rec = ppsd_serve.run(rate=200.0, duration=1.0, drain=240.0, slice_ms=50.0, keep_factors=True)
    print({k: rep[k] for k in ("requests", "completed", "backlog", "wall_s", "latency_s")}, flush=True)
    assert rep["failed"] == 0 and rep["backlog"] == 0 and rep["completed"] == rep["requests"] > 150
    rng = np.random.Generator(np.random.PCG64(ppsd_units.MIX_SEED))
    units = [ppsd_units.draw(rng) for _ in range(rep["requests"])]
    params = [u.params for u in units]
    full, checked = mixcheck.check_units(params, [rec[i]["primes"] for i in range(len(params))],
                                         [rec[i]["prime_sum"] for i in range(len(params))],
                                         [rec[i]["factor_set"] for i in range(len(params))], raw=False, edge=300,
                                         full_work=2e8)
    print(f"checked {checked} factors; {full} requests walked whole by the oracle", flush=True)
    assert checked > 1000 and full > 30
    # overlap: requests were in flight together (admitted before an earlier one finished)
    order = sorted(rec.values(), key=lambda r: r["admit"])
    assert any(b["admit"] < a["done"] for a, b in zip(order, order[1:]))
