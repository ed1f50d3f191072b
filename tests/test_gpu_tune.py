"""Autotuner (ppsd_tune): every timed shape is a valid kernel shape, the
cache round-trips, and the tuned shape gives the same result as the
default."""
import pytest

import ppsd
import ppsd_tune
import ppsd_units

pytestmark = pytest.mark.gpu


def test_tune_and_apply(tmp_path, monkeypatch):
    monkeypatch.setattr(ppsd_tune, "CACHE_DIR", str(tmp_path))
    monkeypatch.setattr(ppsd_tune, "BATCH_GRID", dict(r=(1, 2), u=(4, 16), waves=(4, 8), seg=((2048, 65536),)))
    entry = {"device": ppsd_tune.device_key(0), "plain": ppsd_tune.tune_plain(reps=1),
             "batch": ppsd_tune.tune_batch(units=8)}
    assert len(entry["plain"]["trials"]) == 18 and len(entry["batch"]["trials"]) == 8
    ppsd_tune.save(entry, out_dir=str(tmp_path))
    assert ppsd_tune.load()["plain"]["best"] == entry["plain"]["best"]
    units = [u.params for u in ppsd_units.mix_units(6)]
    with ppsd.Context(units[0]) as ctx:
        base = ctx.batch_run(units, raw=True)
        assert ppsd_tune.apply_batch(ctx) and ppsd_tune.apply_plain(ctx)
        b = entry["batch"]["best"]
        assert ctx.batch_tuning() == (b["r"], b["u"], b["waves"], b["seg_min"], b["seg_max"])
        tuned = ctx.batch_run(units, raw=True)
    key = lambda r: sorted(zip(map(tuple, r[0].tolist()), r[1].tolist()))  # noqa: E731
    assert key(base) == key(tuned)
