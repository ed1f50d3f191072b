"""Serving workload: Poisson arrivals of PPS searches drawn from the batched
mixed-unit distribution (ppsd_units), fed in real time to a ppsd.Server, with
latency split into queueing (arrival -> first slice) and service (first slice
-> done).

    python psieve-cuda_amd/python/ppsd_serve.py --rate 200 --duration 60

prints one JSON report.  Requests still unfinished when the run stops are
counted as backlog; their latencies are censored (at least the time they
waited), so the percentiles over completed requests are reported together
with the backlog and with percentiles over all requests where an unfinished
one counts as its censored lower bound.
"""
import argparse
import json
import os
import sys
import threading
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
if HERE not in sys.path:
    sys.path.insert(0, HERE)

import ppsd  # noqa: E402
import ppsd_units  # noqa: E402

ARRIVAL_SEED = 0xA7712026


def poisson_arrivals(rate: float, duration: float, seed: int = ARRIVAL_SEED) -> np.ndarray:
    """Arrival times in [0, duration) of a Poisson process of the given rate."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out, t = [], 0.0
    while True:
        t += rng.exponential(1.0 / rate)
        if t >= duration:
            return np.array(out)
        out.append(t)


def percentiles(x, qs=(50, 99)):
    if len(x) == 0:
        return {f"p{q}": None for q in qs}
    return {f"p{q}": float(np.percentile(np.asarray(x, dtype=float), q)) for q in qs}


def run(rate: float, duration: float, drain: float = 0.0, slice_ms: float = 50.0, mem_budget: int = 0,
        max_active: int = 0, seed: int = ppsd_units.MIX_SEED, arrival_seed: int = ARRIVAL_SEED,
        device: int = 0, keep_factors: bool = False, progress=None):
    """Offer Poisson(rate) requests for `duration` seconds, then wait up to
    `drain` seconds for the backlog.  Returns (report, records)."""
    times = poisson_arrivals(rate, duration, arrival_seed)
    rng = np.random.Generator(np.random.PCG64(seed))
    units = [ppsd_units.draw(rng) for _ in range(len(times))]
    records = {}
    srv = ppsd.Server(device=device, mem_budget=mem_budget, slice_ms=slice_ms, max_active=max_active)
    t0 = time.perf_counter()
    submitted = [0]

    def feeder():
        for i, (t, u) in enumerate(zip(times, units)):
            d = t0 + t - time.perf_counter()
            if d > 0:
                time.sleep(d)
            rid = srv.submit(u.params)
            assert rid == i + 1  # the server numbers requests from 1 in submission order
            submitted[0] += 1

    th = threading.Thread(target=feeder, daemon=True)
    th.start()
    stop_at = t0 + duration + drain
    last = time.perf_counter()
    while True:
        got = srv.next(100)
        now = time.perf_counter()
        if got is not None:
            res, fac = got
            i = res.id - 1
            rec = dict(unit=i, status=res.status, primes=res.primes, prime_sum=res.prime_sum,
                       factors=res.factors, window_steps=res.window_steps, slices=res.slices,
                       arrival=res.t_arrival, admit=res.t_admit, start=res.t_start, done=res.t_done)
            if keep_factors:
                rec["factor_set"] = ppsd.factor_set(fac)
            records[i] = rec
        if progress and now - last > 5:
            st = srv.stats()
            progress(f"t={now - t0:.1f}s submitted={st.submitted} done={st.completed} active={st.active} "
                     f"queued={st.queued} mem={st.mem_in_use / 2**20:.0f}MiB")
            last = now
        if not th.is_alive() and len(records) == len(times):
            break
        if now >= stop_at and not th.is_alive():
            break
    th.join()
    st = srv.stats()
    wall = time.perf_counter() - t0
    srv.close()

    done = [r for r in records.values() if r["status"] == 0]
    lat_all = []
    for i in range(len(times)):
        r = records.get(i)
        lat_all.append(r["done"] - r["arrival"] if r else max(0.0, wall - times[i]))
    report = {
        "offered_rate_per_s": rate,
        "duration_s": duration,
        "drain_s": drain,
        "slice_ms": slice_ms,
        "requests": len(times),
        "completed": len(done),
        "failed": len([r for r in records.values() if r["status"] != 0]),
        "backlog": len(times) - len(records),
        "offered_window_steps": float(sum(u.work_est for u in units)),
        "served_window_steps": float(st.window_steps),
        "primes_per_s": st.primes / wall,
        "window_steps_per_s": st.window_steps / wall,
        "busy_fraction": st.busy_s / wall,
        "latency_s": percentiles([r["done"] - r["arrival"] for r in done]),
        "latency_s_all_censored": percentiles(lat_all),
        "queue_s": percentiles([r["start"] - r["arrival"] for r in done]),
        "service_s": percentiles([r["done"] - r["start"] for r in done]),
        "mem_budget_bytes": st.mem_budget,
        "mem_peak_bytes": st.mem_peak,
        "slices": st.slices,
        "wall_s": wall,
    }
    return report, records


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--rate", type=float, default=200.0)
    ap.add_argument("--duration", type=float, default=60.0)
    ap.add_argument("--drain", type=float, default=0.0)
    ap.add_argument("--slice-ms", type=float, default=50.0)
    ap.add_argument("--mem-budget", type=int, default=0)
    ap.add_argument("--max-active", type=int, default=0)
    a = ap.parse_args()
    rep, _ = run(a.rate, a.duration, a.drain, a.slice_ms, a.mem_budget, a.max_active,
                 progress=lambda m: print(m, file=sys.stderr, flush=True))
    print(json.dumps(rep))


if __name__ == "__main__":
    main()
