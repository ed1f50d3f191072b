"""Autotuner for the walk kernels' shape: primes per lane (r), windows per
pre-test chunk (u), workgroup size or resident waves per SIMD, and (batched
walk) the n-segment lengths.  Every shape gives identical results; only the
time differs, so the tuner times a fixed sample and keeps the fastest.

Results are cached as JSON per device (name, CU count) and library build
(a hash of its sources) in psieve-cuda_amd/tuning/, and applied with
apply_plain / apply_batch.

    python psieve-cuda_amd/python/ppsd_tune.py          # tune both, write the cache
"""
import hashlib
import itertools
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
if HERE not in sys.path:
    sys.path.insert(0, HERE)

import ppsd  # noqa: E402
import ppsd_units  # noqa: E402

CACHE_DIR = os.path.join(os.path.dirname(HERE), "tuning")

PLAIN_GRID = dict(r=(1, 2), u=(4, 8, 16), wg=(64, 128, 256))
# waves per SIMD below the 8 a SIMD holds leave room for the next rounds'
# sieves, which run beside the walk on a second stream
BATCH_GRID = dict(r=(1, 2), u=(4, 8, 16), waves=(5, 6, 7, 8), seg=((2048, 65536),))

# the plain sample: one batch of the 2^50 workload (k <= 1e4, n <= 6e6)
PLAIN_SAMPLE = dict(pmin=2**50, pmax=2**50 + 5 * 10**10, kmin=1201, kmax=9999, nmin=50000, nmax=6000000)


def device_key(device: int = 0) -> str:
    """Device architecture, CU count and a hash of the library's sources: a
    cache entry is valid only for the kernels that were timed."""
    name, cus = ppsd.device_info(device)
    h = hashlib.sha1()
    csrc = os.path.join(os.path.dirname(HERE), "csrc")
    for fn in sorted(os.listdir(csrc)):
        with open(os.path.join(csrc, fn), "rb") as f:
            h.update(fn.encode() + b"\0" + f.read())
    return f"{name.split(':')[0]}-{cus}cu-{h.hexdigest()[:12]}"


def _cache_path(device: int = 0) -> str:
    return os.path.join(CACHE_DIR, device_key(device) + ".json")


def load(device: int = 0):
    try:
        with open(_cache_path(device)) as f:
            return json.load(f)
    except (OSError, ValueError):
        return None


def save(entry, device: int = 0, out_dir: str = CACHE_DIR):
    os.makedirs(out_dir, exist_ok=True)
    path = os.path.join(out_dir, os.path.basename(_cache_path(device)))
    with open(path, "w") as f:
        json.dump(entry, f, indent=1, sort_keys=True)
    return path


def tune_plain(device: int = 0, reps: int = 2, log=None):
    """Fastest (r, u, wg) for one 2^21-prime batch of the plain walk."""
    params = ppsd.derive(ppsd.SearchParams(**PLAIN_SAMPLE))
    rows = []
    with ppsd.Context(params, device=device) as ctx:
        ctx.submit_range(params.pmin, params.pmax)
        ctx.collect()
        for r, u, wg in itertools.product(*PLAIN_GRID.values()):
            ctx.set_tuning(r, u, wg)
            best = None
            for _ in range(reps):
                ctx.reset_stats()
                ctx.submit_range(params.pmin, params.pmax)
                ctx.collect()
                st = ctx.stats()
                rate = st.window_steps / (st.walk_ms * 1e-3)
                best = rate if best is None else max(best, rate)
            rows.append({"r": r, "u": u, "wg": wg, "window_steps_per_s": best})
            if log:
                log(f"plain r={r} u={u} wg={wg}: {best:.4g} window-steps/s")
    top = max(rows, key=lambda x: x["window_steps_per_s"])
    return {"best": {k: top[k] for k in ("r", "u", "wg")}, "trials": rows}


def tune_batch(device: int = 0, units: int = 256, log=None):
    """Fastest (r, u, waves per SIMD, segment lengths) for ppsd_batch_run on
    the first `units` units of the seeded mixed workload."""
    params = [u.params for u in ppsd_units.mix_units(units)]
    rows = []
    with ppsd.Context(params[0], device=device) as ctx:
        ppsd_units.warm_up(ctx, params)
        for r, u, w, (smin, smax) in itertools.product(*BATCH_GRID.values()):
            ctx.set_batch_tuning(r, u, w, smin, smax)
            t0 = time.perf_counter()
            ctx.batch_run(params)
            dt = time.perf_counter() - t0
            rows.append({"r": r, "u": u, "waves": w, "seg_min": smin, "seg_max": smax, "seconds": dt})
            if log:
                log(f"batch r={r} u={u} waves={w} seg={smin}..{smax}: {dt:.3f} s")
    top = min(rows, key=lambda x: x["seconds"])
    return {"best": {k: top[k] for k in ("r", "u", "waves", "seg_min", "seg_max")}, "units": units, "trials": rows}


def apply_plain(ctx, device: int = 0) -> bool:
    e = load(device)
    if not e or "plain" not in e:
        return False
    b = e["plain"]["best"]
    ctx.set_tuning(b["r"], b["u"], b["wg"])
    return True


def apply_batch(ctx, device: int = 0) -> bool:
    e = load(device)
    if not e or "batch" not in e:
        return False
    b = e["batch"]["best"]
    ctx.set_batch_tuning(b["r"], b["u"], b["waves"], b["seg_min"], b["seg_max"])
    return True


def main():
    import argparse
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=CACHE_DIR, help="directory of the cache file")
    ap.add_argument("--units", type=int, default=256, help="mixed-workload units timed per batched shape")
    a = ap.parse_args()
    log = lambda m: print(m, file=sys.stderr, flush=True)  # noqa: E731
    entry = {"device": device_key(0), "plain": tune_plain(log=log), "batch": tune_batch(units=a.units, log=log)}
    path = save(entry, out_dir=a.out)
    print(json.dumps({"cache": path, **{k: entry[k]["best"] for k in ("plain", "batch")}}))


if __name__ == "__main__":
    main()
