"""Throughput of the MI355X PPSieve hot path on its headline workload.

Workload (single large work unit, plain kernel): Proth sign (+1), odd k in
[1201, 9999], n in [50000, 6000000), p over the primes of [2^50, 2^50+5e10),
exact uint64 arithmetic.  One step = one batch: the primes of the next
p-chunk are sieved on the device (2^21 primes by default), walked over the
whole n-range, hits filtered, and the batch's factors brought to the host.
Two library contexts are submitted to in turn, as the host program drives
them, so the next batch is queued while the current one walks.  By default
the steps cover the whole p-range (1.44e9 primes, about 690 steps), whose
prime count and sum are checked against the oracle's independent sieve;
--steps K times a prefix of K steps.

Prints one JSON line:  value = primes/s over the timed steps (inputs
generated on the device, nothing crosses PCIe but the factors),
roofline = the walk's single-rate VALU issue (its algorithmic instructions
per launch over the time per launch in the timed region) against the
chip's (CUs x 64 lanes x the peak shader clock; frac_at_measured_clock at
the clock the firmware reports meanwhile), cpu_baseline = the reference's
own CPU walk (its host program re-walking every prime with test_one_p) on
the host's cores over a bounded sample of the same workload, with the
oracle's C port of it beside it.
"""
import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "psieve-cuda_amd", "python"))

METRIC = "primes/sec sieved (k≤1e4, n≤6e6) batched vs plain; % of INT32 VALU peak"

WORKLOAD = dict(pmin=2**50, pmax=2**50 + 5 * 10**10, kmin=1201, kmax=9999, nmin=50000, nmax=6000000, sign=1)

# The walk is bound by VALU issue.  Every non-move instruction of its loop
# (v_mul_lo_u32, v_mad_u64_u32, v_ffbl_b32, v_alignbit_b32, v_min3_u32,
# v_cmp) is single-rate on gfx950: a wave64 instruction holds its SIMD's 16
# lanes for 4 clocks, so a CU issues at most 64 single-rate lane-instructions
# per clock.  The peak is CUs x 64 x the device's peak clock (2400 MHz on
# MI355X), which bounds the fraction by 1; under this load the chip runs
# below its peak clock, so the line also gives the fraction at the clock
# the firmware reports during the timed steps (ClockSampler).  The
# one v_mov_b32 per window retires beside the counted instructions and is
# not counted.
LANES_PER_CU_CLOCK = 64


def nonmove_per_window(r, u):
    """Non-move VALU instructions per prime-window of the headline loop (s =
    32, one sign, odd-low pre-test, three-multiply step), as compiled: per
    prime-window one v_mul_lo_u32, two v_mad_u64_u32, v_ffbl_b32,
    v_alignbit_b32 and half a v_min3_u32; per chunk of u windows of r
    primes one compare (89 per 16 windows at r = 1, u = 16)."""
    return 5.5 + 1.0 / (r * u)


# The 15-op model of one window-step (4 MUL32 + 11 ALU32) against the FP32
# vector peak's 78.6 T 32-bit ops/s: a ratio of models, reported beside the
# roofline, not a fraction of anything the hardware does.
MODEL_OPS_PER_STEP = 4 + 11
MODEL_PEAK_TOPS = 256 * 128 * 2.4e9 / 1e12
# HBM bytes one walk launch moves per prime walked, from the FETCH_SIZE and
# WRITE_SIZE counters (scaled by a 256 MiB copy kernel's readings) on a
# 2,088,811-prime launch: 16.0 MiB fetched + 4.6 KiB written.
TRAFFIC_BYTES_PER_PRIME = (16.0 * 2**20 + 4.6 * 2**10) / 2088811


def host_cpu():
    """(threads for CPU work, nproc, CPU model): the process's CPU share --
    OMP_NUM_THREADS where the machine sets it, else the CPUs this process may
    run on -- at most the reference's 32 workers (main.h:24)."""
    share = int(os.environ.get("OMP_NUM_THREADS") or 0) or len(os.sched_getaffinity(0))
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            model = next((l.split(":", 1)[1].strip() for l in f if l.startswith("model name")), "")
    except OSError:
        pass
    return max(1, min(share, 32)), os.cpu_count(), model


class ClockSampler:
    """The shader clock the device's firmware reports while the timed steps
    run: amdsmi's GPU metrics (current_gfxclks, one reading per XCD), read
    every 50 ms on a thread.  The firmware refreshes them about once a
    second, so what is read during the first settle_s seconds can still
    describe the idle GPU before the steps began: those readings are
    dropped, and a run too short to leave min_span_s seconds of readings
    after them reports no clock rather than a stale one.  The median over
    the rest is taken.  The GPU is the first one amdsmi lists (the only one
    on a one-GPU host).  Without amdsmi the readings stay empty."""

    def __init__(self, period_s=0.05, settle_s=1.5, min_span_s=2.0):
        self.period_s = period_s
        self.settle_s = settle_s
        self.min_span_s = min_span_s
        self.readings = []
        self._times = []
        self._stop = None
        self._thread = None
        self._smi = None

    def __enter__(self):
        import threading

        try:
            import amdsmi

            amdsmi.amdsmi_init()
            handle = amdsmi.amdsmi_get_processor_handles()[0]
        except Exception as e:  # no library or no device access: no measured clock
            print(f"no clock readings: {e}", file=sys.stderr)
            return self
        self._smi = amdsmi
        self._stop = threading.Event()

        def run():
            while not self._stop.is_set():
                try:
                    clks = amdsmi.amdsmi_get_gpu_metrics_info(handle)["current_gfxclks"]
                    vals = [c for c in clks if isinstance(c, int) and c > 0]
                    if vals:
                        self._times.append(time.perf_counter())
                        self.readings.append(sum(vals) / len(vals))
                except Exception:
                    pass
                self._stop.wait(self.period_s)

        self._thread = threading.Thread(target=run, daemon=True)
        self._thread.start()
        return self

    def __exit__(self, *exc):
        if self._thread:
            self._stop.set()
            self._thread.join()
        if self._smi:
            self._smi.amdsmi_shut_down()
        return False

    def settled(self):
        """Readings taken after the first settle_s seconds, or [] when they
        span less than min_span_s seconds."""
        if not self._times:
            return []
        t0 = self._times[0] + self.settle_s
        kept = [(t, r) for t, r in zip(self._times, self.readings) if t >= t0]
        if not kept or kept[-1][0] - kept[0][0] < self.min_span_s:
            return []
        return [r for _, r in kept]

    def median_mhz(self):
        v = sorted(self.settled())
        return v[len(v) // 2] if v else None


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=0, help="batches to time (0 = the whole p-range, about 690)")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batch", type=int, default=0, help="primes per step (0 = library default 2^21)")
    ap.add_argument("--tune", default="", help="r,u,wg kernel shape")
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="target length of the CPU baseline sample")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--mix-units", type=int, default=256,
                    help="units of the seeded mixed workload run batched and plain (0 = skip)")
    ap.add_argument("--serve-seconds", type=float, default=10.0,
                    help="seconds of Poisson arrivals at 200 requests/s through the serving loop (0 = skip)")
    return ap.parse_args()


def cpu_baseline_reference(threads):
    """The reference's own CPU walk, timed on the host's cores: its unmodified
    host (main.c, sieve.c, pps/app.c, pps/factor_proth.c, compiled into
    oracle/_ref by oracle/Makefile) on libppsd in the appcu layer's flag-all
    mode, where every n-subsection of every prime is flagged and the host
    re-walks each prime's whole n-range with its test_one_p
    (pps/app.c:1037-1121, 1197-1219).  `threads` workers (-t, -d 0) over the
    first 32768 * threads primes of the workload, batches of 32768 (-m 1).
    Rate = primes / the program's own sieve time, which covers prime supply
    and excludes start-up (main.c:1237-1244)."""
    import math
    import re
    import subprocess
    import tempfile

    exe = os.path.join(HERE, "oracle", "_ref", "ppsieve_ref_on_ppsd")
    if not os.path.exists(exe):
        return None
    lo = WORKLOAD["pmin"]
    hi = lo + int(32768 * threads * math.log(lo))
    with tempfile.TemporaryDirectory() as wd:
        env = dict(os.environ, PPSD_APPCU_FLAG_ALL="1")
        args = [exe, f"-p{lo}", f"-P{hi}", "-k1201", "-K9999", "-n50000", "-N6000000", "-q", "-d", "0", "-t",
                str(threads), "-m", "1", "-r", "20"]
        r = subprocess.run(args, cwd=wd, env=env, capture_output=True, text=True, timeout=900)
    if r.returncode != 0:
        print(f"reference CPU baseline failed ({r.returncode}): {r.stderr[-500:]}", file=sys.stderr)
        return None
    cnt = re.search(r"count=(\d+),sum=0x([0-9a-f]+)", r.stderr)
    el = re.search(r"Elapsed time: ([\d.]+) sec\. \(([\d.]+) init \+ ([\d.]+) sieve\)", r.stderr)
    if not cnt or not el:
        return None
    primes, sieve_s = int(cnt.group(1)), float(el.group(3))
    return {"value": primes / sieve_s, "unit": "primes/s", "cores": threads, "kind": "reference",
            "sample": f"{primes} primes from 2^50 ([{lo},{hi})), full n-range [50000,6000000), k 1201..9999: "
                      f"the reference's main.c + pps/app.c test_one_p re-walk of every prime ({threads} workers, "
                      f"libppsd flag-all), {sieve_s:.1f} s of the program's sieve time"}


def cpu_baseline_port(threads, target_s):
    """The oracle (pps_oracle.c, a C port of the reference's prime supply,
    test_one_p and test_factor) on the first primes of the workload, full
    n-range, `threads` threads; the clock covers the prime sieve as well."""
    sys.path.insert(0, os.path.join(HERE, "oracle"))
    import pps_oracle as orc

    pr = orc.derive(**WORKLOAD)
    P = orc.primes(WORKLOAD["pmin"], WORKLOAD["pmin"] + 40 * 10**6)
    orc.walk(pr, P[:threads], filter=True, threads=threads)  # first call's one-time costs
    pilot = 32  # primes per thread: long enough that thread start-up is noise
    t0 = time.perf_counter()
    orc.walk(pr, P[:threads * pilot], filter=True, threads=threads)
    per = (time.perf_counter() - t0) / pilot  # seconds per prime per thread
    m = int(max(threads * 4, min(len(P), target_s / max(per, 1e-9) * threads)))
    hi = int(P[m - 1]) + 1
    t0 = time.perf_counter()
    Q = orc.primes(WORKLOAD["pmin"], hi)
    orc.walk(pr, Q, filter=True, threads=threads)
    dt = time.perf_counter() - t0
    return {"value": len(Q) / dt, "unit": "primes/s", "cores": threads, "kind": "port",
            "sample": f"the {len(Q)} primes of [2^50, {hi}), full n-range [50000,6000000), sieved and walked "
                      f"in {dt:.1f} s on {threads} threads"}


# Non-move VALU instructions per prime-window of the walk loops' pre-tests,
# as compiled (one sign, both signs): odd-low v_ffbl_b32 + v_alignbit_b32 +
# half a v_min3_u32; exact v_ffbl_b32 + v_lshrrev_b64 + v_cmp_ge_u64; range
# half a v_min3_u32 (and half a v_max3_u32); both signs add p - K.
PRETEST_NONMOVE = {1: (3.0, 8.0), 2: (0.5, 1.0), 3: (2.5, 7.0)}


def walk_issue_seconds(params, primes, r, u, batched, cus, clock_hz):
    """A lower bound on the time any path needs to walk these units: per
    unit, its prime-windows times the non-move VALU instructions per
    prime-window of the loop it is walked with (the step's v_mul_lo_u32 and
    two v_mad_u64_u32, plus v_lshlrev_b64 at s < 32 and, for the primes past
    the three-multiply bound at the top of each 2^32 block, one more mad;
    the pre-test of its form; one compare per chunk), over 64 lanes, against
    one single-rate wave-instruction per CU per clock.  Prime supply and the
    filter are left out."""
    import ppsd

    total = 0.0
    for p, n in zip(params, primes):
        form, chunk = ppsd.walk_form(p, r, u, batched)
        both = p.sign == 0
        wins = (p.nmax - p.nmin + p.nstep - 1) // p.nstep if p.nmax > p.nmin else 0
        slow = min(1.0, ((p.pmin + p.pmax) // 2 >> p.nstep) / 2.0**32)
        step = 3.0 + (p.nstep < 32) + slow
        per = step + PRETEST_NONMOVE[form][1 if both else 0] + 1.0 / (r * chunk)
        total += float(n) * wins * per / 64.0
    return total / (cus * clock_hz)


def mixed_units(n):
    """The first n units of the seeded mixed workload (4096 units, one
    persistent launch) searched twice: batched (ppsd_batch_run) and one unit
    at a time on the plain path (contexts opened outside the clock).  Both
    must give the same factors, prime counts and sums per unit."""
    import numpy as np

    import ppsd
    import ppsd_units

    units = ppsd_units.mix_units(n)
    params = [u.params for u in units]
    import ppsd_tune

    ctx = ppsd.Context(params[0])
    ppsd_tune.apply_batch(ctx)  # the autotuner's cached shape for this device, when present
    shape = dict(zip(("r", "u", "waves", "seg_min", "seg_max"), ctx.batch_tuning()))
    ppsd_units.warm_up(ctx, params)  # first use, outside the clock
    ctx.reset_stats()
    t0 = time.perf_counter()
    out, unit, cnt, sums = ctx.batch_run(params)
    tb = time.perf_counter() - t0
    st = ctx.stats()
    ctx.close()
    tp, parts, same_counts = 0.0, [], True
    last = time.perf_counter()
    for i, p in enumerate(params):
        with ppsd.Context(p) as c:
            ppsd_tune.apply_plain(c)  # each path with its own tuned shape
            t0 = time.perf_counter()
            f, cn, ps = c.run_range(p.pmin, p.pmax)
            tp += time.perf_counter() - t0
        same_counts &= (cn, ps) == (int(cnt[i]), int(sums[i]))
        parts.append((np.full(len(f), i, np.uint32), f))
        if time.perf_counter() - last > 20:
            print(f"mixed units, plain path: {i + 1}/{n} searched", file=sys.stderr, flush=True)
            last = time.perf_counter()

    def canon(u, f):  # rows (unit, p, k, n, c) in lexicographic order
        order = np.lexsort((f["c"], f["n"], f["k"], f["p"], u))
        return u[order], f[order]

    pu, pf = canon(np.concatenate([u for u, _ in parts]), np.concatenate([f for _, f in parts]))
    bu, bf = canon(unit, out)
    identical = same_counts and len(pu) == len(bu) and np.array_equal(pu, bu) and all(
        np.array_equal(pf[name], bf[name]) for name in ("p", "k", "n", "c"))
    _, cus = ppsd.device_info(0)
    limit = walk_issue_seconds(params, cnt, shape["r"], shape["u"], True, cus, ppsd.device_clock_mhz(0) * 1e6)
    return {"units": n, "seed": ppsd_units.MIX_SEED, "primes": st.primes, "window_steps": st.window_steps,
            "factors": len(out), "batched_s": tb, "plain_s": tp, "batched_primes_per_s": st.primes / tb,
            "plain_primes_per_s": st.primes / tp, "batched_over_plain": tp / tb, "identical_per_unit": bool(identical),
            "batch_shape": shape, "walk_issue_limit_s": limit, "batched_frac_of_limit": limit / tb,
            "plain_frac_of_limit": limit / tp, "batched_over_plain_bound": tp / limit}


def serving(seconds):
    """The serving loop under the mixed workload's Poisson arrivals at 200
    requests/s (about 3.4x what one MI355X serves) for `seconds`, no drain: the
    requests completed, their latency percentiles (arrival to done, and its
    queueing and service parts) and the device's throughput meanwhile."""
    import ppsd_serve

    rep, _ = ppsd_serve.run(rate=200.0, duration=seconds, drain=0.0)
    keep = ("requests", "completed", "backlog", "latency_s", "queue_s", "service_s", "primes_per_s",
            "window_steps_per_s", "busy_fraction", "slices", "mem_peak_bytes")
    return dict({k: rep[k] for k in keep}, offered_rate_per_s=200.0, duration_s=seconds)


def _is_prime64(n):
    """Deterministic Miller-Rabin for n < 3.3e24 (the first 12 prime bases)."""
    if n < 2:
        return False
    small = (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37)
    for q in small:
        if n % q == 0:
            return n == q
    d, r = n - 1, 0
    while d % 2 == 0:
        d, r = d // 2, r + 1
    for a in small:
        x = pow(a, d, n)
        if x in (1, n - 1):
            continue
        for _ in range(r - 1):
            x = x * x % n
            if x == n - 1:
                break
        else:
            return False
    return True


def verify_factors(f, kmin, kmax, nmin, nmax):
    """Every factor line is valid: p prime, p | k*2^n+1, k odd in range, n in
    range, and no odd prime q < 2^15 divides k*2^n+1 (the rule the filter
    applies), the last by vectorised trial division."""
    import numpy as np

    p = [int(x) for x in f["p"]]
    k = [int(x) for x in f["k"]]
    n = [int(x) for x in f["n"]]
    ok = all(_is_prime64(pi) and (ki * pow(2, ni, pi) + 1) % pi == 0 and ki % 2 == 1 and kmin <= ki <= kmax
             and nmin <= ni < nmax for pi, ki, ni in zip(p, k, n))
    if not ok or not p:
        return ok
    sieve = np.ones(32768, dtype=bool)
    sieve[:2] = False
    for q in range(2, 182):
        if sieve[q]:
            sieve[q * q::q] = False
    ka = np.array(k, dtype=np.uint64)
    na = np.array(n, dtype=np.int64)
    for q in np.nonzero(sieve)[0][1:]:  # odd primes
        q = int(q)
        e = na % (q - 1)  # Fermat
        t = np.ones(len(na), dtype=np.int64)
        b = np.full(len(na), 2, dtype=np.int64)
        while e.any():
            t = np.where(e & 1, t * b % q, t)
            b = b * b % q
            e >>= 1
        if np.any(((ka % np.uint64(q)).astype(np.int64) * t + 1) % q == 0):
            return False
    return True


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {a.gpus} is not supported (primes shard by p-range "
              f"with no exchange; see the sharding note in the design document)", file=sys.stderr)
        sys.exit(2)
    import ppsd

    import ppsd_tune

    params = ppsd.derive(ppsd.SearchParams(**WORKLOAD))
    # Two library contexts, submitted to in turn, as the host program drives
    # them: the next batch is queued while the current one walks, so the
    # GPU goes from one walk to the next without waiting on the host.
    ctxs = [ppsd.Context(params, batch=a.batch) for _ in range(2)]
    for ctx in ctxs:
        if a.tune:
            r, u, wg = (int(x) for x in a.tune.split(","))
            ctx.set_tuning(r, u, wg)
        else:
            ppsd_tune.apply_plain(ctx)  # the autotuner's cached shape for this device, when present
    shape = dict(zip(("r", "u", "wg"), ctxs[0].tuning()))
    p = WORKLOAD["pmin"]
    for i in range(max(a.warmup, 2) if a.warmup else 0):
        p = ctxs[i % 2].submit_range(p, WORKLOAD["pmax"])
        ctxs[i % 2].collect()
    for ctx in ctxs:
        ctx.reset_stats()
    full = a.steps == 0
    if full:
        p = WORKLOAD["pmin"]  # the whole range, warm-up batches included
    p_start = p
    factors, nsteps, count, psum, found = 0, 0, 0, 0, []
    pending = []

    def take():
        nonlocal factors, count, psum, nsteps
        f, c, s = pending.pop(0).collect()
        factors += len(f)
        found.append(f)
        count += c
        psum = (psum + s) & (2**64 - 1)
        nsteps += 1
        if nsteps % 50 == 0 or not full:
            print(f"step {nsteps}: p < {p}, {time.perf_counter() - t0:.2f} s", file=sys.stderr, flush=True)

    submitted = 0
    with ClockSampler() as clock:
        t0 = time.perf_counter()
        while p < WORKLOAD["pmax"] and (full or submitted < a.steps):
            ctx = ctxs[submitted % 2]
            p = ctx.submit_range(p, WORKLOAD["pmax"])
            pending.append(ctx)
            submitted += 1
            if len(pending) == 2:
                take()
        while pending:
            take()
        dt = time.perf_counter() - t0
    sts = [ctx.stats() for ctx in ctxs]
    for ctx in ctxs:
        ctx.close()

    class Sum:  # the two contexts' counters together
        pass

    st = Sum()
    for name in ("primes", "window_steps", "walk_launches", "walk_ms", "sieve_ms", "filter_ms"):
        setattr(st, name, sum(getattr(x, name) for x in sts))
    checksum = None
    if full:  # the reference's count= and sum= of the whole range, from the oracle's own sieve
        with open(os.path.join(HERE, "tests", "golden", "runs.json")) as fh:
            g = json.load(fh)["cfg2_full"]
        checksum = {"count": count, "sum": f"0x{psum:016x}",
                    "matches_independent_sieve": (count, psum) == (g["count"], g["sum"])}

    import numpy as np

    valid = verify_factors(np.concatenate(found) if found else np.zeros(0, dtype=ppsd.FACTOR_DTYPE),
                           params.kmin, params.kmax, params.nmin, params.nmax)
    value = st.primes / dt
    steps_per_launch = st.window_steps / max(st.walk_launches, 1)
    # time per launch in the steady state of the timed region: the launches
    # follow one another with no gap, and the region's sieve, filter and host
    # work (which overlap them) count against the walk, so this is an upper
    # bound on the kernel's own time per launch.  HIP events around each
    # launch (walk_ms) overlap between the two streams: their mean is given
    # as launch_ms_events.
    t_launch = dt / max(st.walk_launches, 1)
    _, cus = ppsd.device_info(0)
    nonmove = nonmove_per_window(shape["r"], shape["u"])
    achieved = steps_per_launch * nonmove / t_launch / 1e12  # T lane-instructions/s
    peak = cus * LANES_PER_CU_CLOCK * ppsd.device_clock_mhz(0) * 1e6 / 1e12
    run_mhz = clock.median_mhz()
    model_tops = steps_per_launch * MODEL_OPS_PER_STEP / t_launch / 1e12
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "primes/s",
        "n_gpus": 1,
        "steps": nsteps,
        "warmup": a.warmup,
        "ms_per_step": dt / max(nsteps, 1) * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic",
        "config": {
            "workload": "plain kernel, single work unit: +1, k in [1201,9999], n in [50000,6000000), "
                        "p prime in [2^50, 2^50+5e10)" + ("" if full else " (timed prefix)"),
            "checksum": checksum,
            "p_range_timed": [p_start, p],
            "primes_timed": st.primes,
            "batch_primes": ctx.batch,
            "nstep": params.nstep,
            "window_steps": st.window_steps,
            "factors": factors,
            "factors_valid": valid,
            "walk_ms": st.walk_ms,
            "sieve_ms": st.sieve_ms,
            "filter_ms": st.filter_ms,
            "parallelism": "1 GPU",
            "contexts_in_flight": 2,
            "walk_shape": shape,
        },
        "roofline": {
            "bound": "valu",
            "achieved": achieved,
            "peak": peak,
            "unit": "T single-rate lane-instructions/s",
            "frac": achieved / peak if peak else None,
            "launch_ms": t_launch * 1e3,
            "launch_ms_events": st.walk_ms / max(st.walk_launches, 1),
            "traffic": TRAFFIC_BYTES_PER_PRIME * st.primes / max(st.walk_launches, 1),
            "peak_clock_mhz": ppsd.device_clock_mhz(0),
            "measured_clock_mhz": run_mhz,
            "clock_readings": len(clock.settled()),
            "frac_at_measured_clock": achieved / (cus * LANES_PER_CU_CLOCK * run_mhz * 1e-6) if run_mhz else None,
            "cus": cus,
            "nonmove_valu_per_window": nonmove,
            "model_ratio_15op": model_tops / MODEL_PEAK_TOPS,
        },
    }
    if a.mix_units:
        line["config"]["batched"] = mixed_units(a.mix_units)
    if a.serve_seconds > 0:
        line["config"]["serving"] = serving(a.serve_seconds)
    if not a.no_cpu_baseline:
        threads, nproc, model = host_cpu()
        host = {"nproc": nproc, "cpu_model": model, "threads_used": threads}
        port = cpu_baseline_port(threads, a.cpu_seconds)
        ref = cpu_baseline_reference(threads)
        line["cpu_baseline"] = dict(ref or port, **host)
        if ref:
            line["config"]["cpu_port"] = port
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
