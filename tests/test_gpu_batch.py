"""Batched search (ppsd_batch_run): many independent units in one persistent
launch.  Every unit's result must equal what the oracle (and the plain
per-unit path, ppsd_submit_range) gives for that unit alone."""
import json
import os
import random

import numpy as np
import pytest

import pps_oracle as orc
import ppsd
import ppsd_units
from conftest import GOLDEN, read_lines, windowed_run
from unitgen import _dense_unit, _oracle, _ref_unit

pytestmark = pytest.mark.gpu

RUNS = json.load(open(os.path.join(GOLDEN, "runs.json")))
U64 = 2**64 - 1


def golden_set(name):
    return {orc.parse_line(l) for l in read_lines(name)}


def _per_unit(out, unit, n):
    sets = [set() for _ in range(n)]
    for f, u in zip(out, unit):
        sets[int(u)].add((int(f["p"]), int(f["k"]), int(f["n"]), int(f["c"])))
    return sets


def _check_mix(units, ctx, raw):
    out, unit, cnt, sums = ctx.batch_run([u[1] for u in units], raw=raw)
    got = _per_unit(out, unit, len(units))
    assert len(out) == sum(len(g) for g in got), "a hit was reported twice"
    nhits = 0
    for i, (pr, _) in enumerate(units):
        c, s, want = _oracle(pr, raw)
        assert (int(cnt[i]), int(sums[i])) == (c, s), f"unit {i}: prime count/sum"
        assert got[i] == want, f"unit {i}: {len(got[i])} hits, oracle {len(want)}"
        nhits += len(want)
    return nhits


# (r, u, waves per SIMD, seg_min, seg_max): the default shape, the other
# kernel shapes, and tiny n-segments so that most tasks start mid-range
SHAPES = [(1, 8, 8, 0, 0), (2, 4, 4, 16, 64), (1, 16, 2, 16, 16), (2, 16, 8, 64, 4096), (1, 4, 16, 32, 256),
          (2, 8, 1, 2048, 65536)]


@pytest.mark.parametrize("seed", range(len(SHAPES)))
def test_batch_mix_vs_oracle(seed):
    rng = random.Random(3000 + seed)
    units = [_dense_unit(rng) if i % 3 else _ref_unit(rng) for i in range(12)]
    with ppsd.Context(units[0][1]) as ctx:
        ctx.set_batch_tuning(*SHAPES[seed])
        assert _check_mix(units, ctx, raw=True) > 0
        _check_mix(units, ctx, raw=False)


def test_batch_golden_ranges_together():
    """pps/many_n_test.sh's range, the 1e12 search (s = 27) and the p-windows
    of the 2^50 and 2^61 (both signs) searches, each window a unit of one
    batch: the factors (and, raw, the divisors) of each golden run, and its
    prime count and sum."""
    units, owner = [], []
    for tag in ("many_n", "cfg1"):
        units.append(RUNS[tag]["params"])
        owner.append(tag)
    for tag in ("cfg2_windows", "cfg3_windows"):
        kw, wins = windowed_run(tag)
        for a, b in wins:
            units.append(dict(kw, pmin=a, pmax=b))
            owner.append(tag)
    params = [ppsd.derive(ppsd.SearchParams(**u)) for u in units]
    with ppsd.Context(params[0]) as ctx:
        out, unit, cnt, sums = ctx.batch_run(params)
        raw, raw_unit, rcnt, _ = ctx.batch_run(params, raw=True)
    got = _per_unit(out, unit, len(units))
    got_raw = _per_unit(raw, raw_unit, len(units))
    for tag, name, raw_name in (("many_n", "many_n_filtered.txt", "many_n_raw.txt"), ("cfg1", "cfg1_factors.txt", None),
                                ("cfg2_windows", "cfg2_windows.txt", "cfg2_windows_raw.txt"),
                                ("cfg3_windows", "cfg3_windows.txt", "cfg3_windows_raw.txt")):
        idx = [i for i, o in enumerate(owner) if o == tag]
        assert set().union(*(got[i] for i in idx)) == golden_set(name), tag
        if raw_name:
            assert set().union(*(got_raw[i] for i in idx)) == golden_set(raw_name), tag
        assert sum(int(cnt[i]) for i in idx) == RUNS[tag]["count"], tag
        assert sum(int(sums[i]) for i in idx) % 2**64 == RUNS[tag]["sum"], tag
    assert [int(x) for x in rcnt] == [int(x) for x in cnt]


def test_round_sieve_in_several_launches():
    """9000 narrow units fit one round of the prime pool, but each takes at
    least one sieve segment (2^18 candidates) of the round's bitmap, which
    holds 2^13 segments per launch: the round is sieved in two launches.
    Every unit's prime count and sum equal the oracle's sieve, and the
    divisors of every 61st unit (and of the units on either side of the
    launch boundary) equal the oracle walk."""
    rng = random.Random(8000)
    units = []
    for _ in range(9000):
        pmin = rng.randrange(10**12, 10**14)
        kw = dict(pmin=pmin, pmax=pmin + rng.randrange(1000, 300000), kmin=3, kmax=rng.randrange(101, 3000),
                  nmin=rng.randrange(100, 5000), sign=rng.choice((1, -1)))
        kw["nmax"] = kw["nmin"] + rng.randrange(1, 2000)
        units.append(ppsd_units.extended_params(**kw))
    # the first unit whose segments start the second launch
    segs, boundary = 0, None
    for i, p in enumerate(units):
        segs += (((p.pmax - (p.pmin | 1) + 1) // 2) + 2**18 - 1) // 2**18
        if segs > 2**13 and boundary is None:
            boundary = i
    assert boundary is not None
    with ppsd.Context(units[0]) as ctx:
        out, unit, cnt, sums = ctx.batch_run(units, raw=True)
    want_cs = orc.count_primes_many([(p.pmin, p.pmax) for p in units], threads=16)
    for i in range(len(units)):
        assert (int(cnt[i]), int(sums[i])) == want_cs[i], f"unit {i}: prime count/sum"
    got = _per_unit(out, unit, len(units))
    check = sorted(set(range(0, len(units), 61)) | {boundary - 1, boundary, boundary + 1})
    for i in check:
        p = units[i]
        pr = orc.OrcParams(pmin=p.pmin, pmax=p.pmax, kmin=p.kmin, kmax=p.kmax, nmin=p.nmin, nmax=p.nmax,
                           kstep=p.kstep, koffset=p.koffset, sign=p.sign, nstep_ref=p.nstep_ref, nstep=p.nstep,
                           status=0)
        want = orc.hit_set(orc.walk(pr, orc.primes(p.pmin, p.pmax), filter=False, threads=8))
        assert got[i] == want, f"unit {i}: {len(got[i])} divisors, oracle {len(want)}"


def test_batch_equals_plain_path():
    """A unit's hits in a mix are the hits of searching it alone."""
    rng = random.Random(4000)
    units = [_ref_unit(rng) for _ in range(6)] + [_dense_unit(rng) for _ in range(6)]
    with ppsd.Context(units[0][1]) as ctx:
        out, unit, _, _ = ctx.batch_run([u[1] for u in units], raw=True)
    got = _per_unit(out, unit, len(units))
    for i, (pr, params) in enumerate(units):
        with ppsd.Context(params) as ctx:
            alone, _, _ = ctx.run_range(pr.pmin, pr.pmax, raw=True)
            single, su, _, _ = ctx.batch_run([params], raw=True)
        assert got[i] == ppsd.factor_set(alone), i
        assert ppsd.factor_set(single) == got[i] and set(su.tolist()) <= {0}


def test_batch_ring_overflow_reruns():
    rng = random.Random(5000)
    units = [_dense_unit(rng) for _ in range(8)]
    with ppsd.Context(units[0][1]) as ctx:
        ctx.set_batch_tuning(seg_min=16, seg_max=64)
        full, fu, _, _ = ctx.batch_run([u[1] for u in units], raw=True)
        launches = ctx.stats().walk_launches
        ctx.set_batch_hit_capacity(512)
        small, su, _, _ = ctx.batch_run([u[1] for u in units], raw=True)
        # the task set ran in many groups sized for the ring, re-run in
        # halves where a group still overflowed
        st = ctx.stats()
        assert st.walk_launches - launches > 2 * launches
    # more results than the device output holds with this ring (twice its
    # capacity): the output was moved to host memory mid-batch
    assert len(full) > 2 * 512
    assert sorted(zip(map(tuple, full.tolist()), fu.tolist())) == sorted(zip(map(tuple, small.tolist()), su.tolist()))


def test_batch_repeat_is_stable():
    """Replays of the captured round give the same result every time."""
    rng = random.Random(6000)
    units = [_dense_unit(rng) for _ in range(5)]
    with ppsd.Context(units[0][1]) as ctx:
        first = ctx.batch_run([u[1] for u in units], raw=True)
        for _ in range(3):
            again = ctx.batch_run([u[1] for u in units], raw=True)
            assert sorted(zip(map(tuple, again[0].tolist()), again[1].tolist())) == \
                sorted(zip(map(tuple, first[0].tolist()), first[1].tolist()))
            assert again[2].tolist() == first[2].tolist() and again[3].tolist() == first[3].tolist()


def test_batch_empty_and_errors():
    rng = random.Random(7000)
    pr, params = _ref_unit(rng)
    with ppsd.Context(params) as ctx:
        out, unit, cnt, _ = ctx.batch_run([])
        assert len(out) == 0 and len(cnt) == 0
        with pytest.raises(ppsd.PPSDError) as e:
            ctx.batch_prepare(2**62 + 1)
        assert e.value.code == ppsd.PPSD_ERR_INVALID_PARAM
        ctx.batch_prepare(2**62)  # the set-up for any unit; the runs below then do none of it
        # an empty n-range and an empty p-range
        e1 = ppsd.manual_params(pr.pmin, pr.pmax, pr.kmin, pr.kmax, 100, 100, nstep=8)
        e2 = ppsd.manual_params(pr.pmin, pr.pmin + 1, pr.kmin, pr.kmax, 100, 200, nstep=8)
        out, unit, cnt, _ = ctx.batch_run([e1, e2])
        assert len(out) == 0 and int(cnt[0]) > 0 and int(cnt[1]) <= 1
        bad = ppsd.manual_params(pr.pmin, pr.pmax, pr.kmin, pr.kmax, 100, 200, nstep=8)
        bad.kmin = bad.kmax + 1
        with pytest.raises(ppsd.PPSDError) as e:
            ctx.batch_run([params, bad])
        assert e.value.code == ppsd.PPSD_ERR_INVALID_PARAM
        with pytest.raises(ppsd.PPSDError) as e:
            ctx.set_batch_tuning(r=3)
        assert e.value.code == ppsd.PPSD_ERR_INVALID_PARAM
        ctx.submit_range(pr.pmin, pr.pmax)
        with pytest.raises(ppsd.PPSDError) as e:
            ctx.batch_run([params])
        assert e.value.code == ppsd.PPSD_ERR_STATE
        ctx.collect()
        ctx.batch_run([params])
