"""Work units of the batched mixed workload.

A seeded generator (numpy PCG64) of independent PPS searches, each a full
ppsd.Params: sign uniform in {+1, -1}; kmax log-uniform in [1e3, 1e6] with
kmin = 3; nmin uniform in [1e3, 1e6]; n-width log-uniform in [1e4, 5e6];
pmin log-uniform in [1e11, 2^60]; p-width log-uniform in [1e5, 1e9].

Units the reference's parameter rules accept (pps/app.c:649-724,
pps/appcu.cu:155-209) get exactly the reference's parameters
(ppsd.derive).  The rest -- about one in nine, rejected by the reference
because 2^(64-nstep) > pmin (pps/appcu.cu:163-176) -- are searched in the
walk's wider domain: the same k rounding and the smallest window s with
kmax*2^s >= pmin, capped at 32, which keeps one candidate per window.  Each
unit records which case it is, and whether the reference's own window was
>= 48 (its CPU re-walk is wrong there: pps/app.c:871-874).
"""
import math
from dataclasses import dataclass

import numpy as np

import ppsd

MIX_SEED = 0x5EED2026
MIX_UNITS = 4096


@dataclass
class Unit:
    search: dict           # pmin, pmax, kmin, kmax, nmin, nmax, sign as drawn
    params: "ppsd.Params"  # what the device runs
    ref_domain: bool       # the reference accepts these parameters
    ref_s48: bool          # the reference's window would be >= 48

    @property
    def primes_est(self) -> float:
        p = self.params
        return (p.pmax - p.pmin) / math.log(p.pmin)

    @property
    def windows(self) -> int:
        p = self.params
        return (p.nmax - p.nmin + p.nstep - 1) // p.nstep if p.nmax > p.nmin else 0

    @property
    def work_est(self) -> float:
        """Expected window-steps (both signs count twice)."""
        return self.primes_est * self.windows * (2 if self.params.sign == 0 else 1)


def _log_uniform(rng, lo, hi):
    return math.exp(rng.uniform(math.log(lo), math.log(hi)))


def extended_params(pmin, pmax, kmin, kmax, nmin, nmax, sign) -> "ppsd.Params":
    """Parameters for a unit outside the reference's validation: its k
    rounding (pps/app.c:682-690), the smallest s with kmax*2^s >= pmin
    (pps/app.c:710-712), capped by nmax-nmin+1 and by 32."""
    kmin = 2 * (kmin // 2) + 1
    kmax = 2 * (kmax // 2) + 1
    s = 1
    while (kmax << s) < pmin:
        s += 1
    s = max(1, min(s, nmax - nmin + 1, 32))
    return ppsd.manual_params(pmin, pmax, kmin, kmax, nmin, nmax, sign=sign, nstep=s, kstep=2, koffset=1)


def draw(rng) -> Unit:
    sign = 1 if rng.integers(0, 2) == 0 else -1
    kmax = int(round(_log_uniform(rng, 1e3, 1e6)))
    nmin = int(rng.integers(1000, 1000001))
    nmax = nmin + int(round(_log_uniform(rng, 1e4, 5e6)))
    pmin = int(round(_log_uniform(rng, 1e11, 2.0**60)))
    pmax = min(pmin + int(round(_log_uniform(rng, 1e5, 1e9))), 2**62)
    search = dict(pmin=pmin, pmax=pmax, kmin=3, kmax=kmax, nmin=nmin, nmax=nmax, sign=sign)
    try:
        params = ppsd.derive(ppsd.SearchParams(**search))
        return Unit(search, params, True, params.nstep_ref >= 48)
    except ppsd.PPSDError:
        return Unit(search, extended_params(**search), False, False)


def warm_up(ctx, params):
    """First use of a context's batched path for these units, outside any
    timing: ppsd_batch_prepare (buffers, the round graphs, the sieving
    primes up to isqrt of the highest pmax -- context state, as ppsd_open
    builds it for the plain path), then a run of the first four units."""
    ctx.batch_prepare(max(p.pmax for p in params))
    ctx.batch_run(list(params[:4]))


def mix_units(n: int = MIX_UNITS, seed: int = MIX_SEED):
    """The first n units of the seeded stream (a prefix of a longer stream is
    the shorter stream)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    return [draw(rng) for _ in range(n)]


def main():
    """Search the first N units batched on device 0 and print one JSON line:
    python psieve-cuda_amd/python/ppsd_units.py --run N"""
    import argparse
    import json
    import time

    ap = argparse.ArgumentParser()
    ap.add_argument("--run", type=int, default=256, help="units of the seeded mix to search")
    ap.add_argument("--raw", action="store_true", help="raw divisors instead of filtered factors")
    ap.add_argument("--tuned", action="store_true", help="the tuning cache's batch shape, and a warm-up run")
    a = ap.parse_args()
    units = mix_units(a.run)
    params = [u.params for u in units]
    with ppsd.Context(params[0]) as ctx:
        if a.tuned:
            import ppsd_tune

            ppsd_tune.apply_batch(ctx)
            warm_up(ctx, params)
            ctx.reset_stats()
        t0 = time.perf_counter()
        out, unit, cnt, sums = ctx.batch_run(params, raw=a.raw)
        dt = time.perf_counter() - t0
        st = ctx.stats()
        shape = ctx.batch_tuning()
    print(json.dumps({"units": a.run, "seed": MIX_SEED, "batch_shape": shape, "seconds": dt, "primes": st.primes,
                      "window_steps": st.window_steps, "raw_hits": st.raw_hits, "factors": len(out),
                      "walk_ms": st.walk_ms, "sieve_ms": st.sieve_ms, "walk_launches": st.walk_launches,
                      "primes_per_s": st.primes / dt, "ref_rejected_units": sum(not u.ref_domain for u in units),
                      "ref_s48_units": sum(u.ref_s48 for u in units)}))


if __name__ == "__main__":
    main()
