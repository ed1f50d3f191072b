// ppsd_serve.cpp -- serving loop on one device: requests (full PPS searches)
// arrive at any time, are admitted under a device-memory budget, and are
// served by continuous batching: the device runs slices of bounded length,
// and every slice is one round of the batched persistent walk over a piece
// (p-subrange) of each admitted request.  A request that arrives while a
// slice runs joins the next one; a small request completes in its first
// slice even while huge ones are in progress.
//
// Resident state: the context's sieving primes (every prime below 2^31, so
// any p < 2^62), the prime pool the slices sieve into and the hit rings,
// allocated once at open.  Results equal searching each request alone.
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <thread>

#include "ppsd_ctx.h"

using namespace ppsd;

namespace {

using Clock = std::chrono::steady_clock;

struct Request {
  uint64_t id = 0;
  ppsd_params prm{};
  double t_arrival = 0, t_admit = -1, t_start = -1, t_done = -1;
  uint64_t cur = 0;      // next p not yet handed to a slice
  uint64_t need = 0;     // bytes reserved at admission
  uint64_t primes = 0, sum = 0, steps = 0;
  uint32_t slices = 0;
  uint32_t inflight = 0;  // planned slices not yet walked
  int status = 0;
  std::vector<ppsd_factor> hits;

  double per_prime() const {  // window-steps per prime
    const double w = windows(prm);
    return w * (prm.sign == 0 ? 2.0 : 1.0);
  }
  double primes_left() const {
    if (cur >= prm.pmax) return 0;
    return (double)(prm.pmax - cur) / log((double)cur);
  }
  double work_left() const { return primes_left() * per_prime(); }
};

}  // namespace

struct ppsd_server {
  ppsd_ctx *ctx = nullptr;
  ppsd_serve_config cfg{};
  Clock::time_point t0;
  uint64_t pool_bytes = 0, ring_bytes = 0;

  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  bool stop = false;
  std::thread worker;

  uint64_t next_id = 1;
  std::deque<Request *> queue;                 // waiting for admission (FIFO)
  std::vector<Request *> active;               // admitted, not complete
  std::deque<Request *> done;                  // complete, not yet taken
  std::map<uint64_t, Request *> taken;         // taken, not yet released
  uint64_t mem_in_use = 0;
  ppsd_serve_stats st{};
  double rate = 4e12;                          // window-steps/s, updated per slice

  double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }
};

namespace {

// Device memory a request holds while admitted: its share of the prime pool
// (at most a quarter of it: its slices take p-pieces no larger) plus a
// hit-ring reservation.
uint64_t request_need(const ppsd_server *s, const Request &r) {
  const double primes = r.primes_left() * 1.05 + 64;
  const uint64_t pool = (uint64_t)std::min<double>(primes * sizeof(uint64_t), (double)s->pool_bytes / 4);
  return pool + (64u << 10);
}

void admit_locked(ppsd_server *s) {
  while (!s->queue.empty() && s->active.size() < s->cfg.max_active) {
    Request *r = s->queue.front();
    if (s->mem_in_use + r->need > s->cfg.mem_budget && !s->active.empty()) break;
    s->queue.pop_front();
    r->t_admit = s->now();
    s->mem_in_use += r->need;
    s->st.mem_peak = std::max(s->st.mem_peak, s->mem_in_use);
    s->st.admitted++;
    s->active.push_back(r);
  }
}

void finish_locked(ppsd_server *s, Request *r) {
  r->t_done = s->now();
  s->mem_in_use -= r->need;
  s->st.completed++;
  s->done.push_back(r);
}

// Pieces of one slice: the slice's work budget is shared equally by the
// active requests, smallest remaining work first, a request's unused share
// passing on to the larger ones; the pieces' primes fit the pool.  While
// requests wait for admission the slice is twice as long: each slice costs
// a sieve, a launch and a result copy besides its walk, and under a backlog
// throughput is what shortens the wait (measured at 200 requests/s of the
// mixed workload: 4.9e12 window-steps/s at 50 ms, 5.2e12 at 100 ms, with
// more requests completed; longer slices completed fewer).
void plan_slice(ppsd_server *s, std::vector<Request *> &act, std::vector<Piece> &pieces,
                std::vector<const ppsd_params *> &par, bool backlog) {
  std::sort(act.begin(), act.end(), [](const Request *a, const Request *b) { return a->work_left() < b->work_left(); });
  double budget = s->cfg.slice_ms * (backlog ? 2.0 : 1.0) * 1e-3 * s->rate;
  const double pool_room = 0.85 * (double)s->ctx->b_pool_cap;
  double pool_left = pool_room;
  const size_t cap = s->ctx->b_piece_cap;
  pieces.clear();
  par.clear();
  for (size_t i = 0; i < act.size() && pieces.size() < cap; ++i) {
    Request *r = act[i];
    const double share = budget / (double)(act.size() - i);
    const double pool_share = pool_left / (double)(act.size() - i);
    const double pp = r->per_prime();
    double want = pp > 0 ? std::min(r->work_left(), share) / pp : r->primes_left();
    want = std::min(want, std::min(pool_share, pool_room / 4));
    want = std::max(want, 64.0);
    const double dens = 1.05 / log((double)r->cur);
    uint64_t hi = r->prm.pmax;
    if (want < r->primes_left()) {
      const double width = want / dens;
      hi = std::min<uint64_t>(r->prm.pmax, r->cur + std::max<uint64_t>(2, (uint64_t)width));
    }
    const double est = (double)(hi - r->cur) * dens + 64;
    budget -= std::min(r->work_left(), est * pp);
    pool_left -= est;
    pieces.push_back(Piece{(uint32_t)par.size(), r->cur, hi, est});
    par.push_back(&r->prm);
  }
}

// A planned slice: one piece per request, sieved into pool `buf`.
struct Slice {
  std::vector<Request *> reqs;
  std::vector<Piece> pieces;
  std::vector<const ppsd_params *> par;
  int buf = 0;
  int rc = 0;
  bool live = false;
};

// Walk a sieved slice and hand its results to its requests.
void finish_slice(ppsd_server *s, Slice &sl) {
  ppsd_ctx *c = s->ctx;
  const size_t np = sl.pieces.size();
  const double t0 = s->now();
  (void)hipStreamSynchronize(c->copy_stream);  // no copy of an earlier, failed slice still in flight
  c->b_result.clear();
  c->b_out_n = 0;
  c->b_unit_primes.assign(np, 0);
  c->b_unit_sums.assign(np, 0);
  int rc = sl.rc;
  if (!rc) rc = walk_round(c, sl.pieces, 0, np, sl.par, s->cfg.raw != 0, sl.buf, c->ev2[sl.buf][0],
                           c->ev2[sl.buf][1]);
  if (rc == 1 && np > 1) {  // more primes than the pool holds (rare): the pieces in two halves
    rc = (int)hipStreamSynchronize(c->stream2) ? set_err(PPSD_ERR_FAILURE, "stream wait failed") : 0;
    const size_t mid = np / 2;
    for (const auto &[x, y] : {std::make_pair((size_t)0, mid), std::make_pair(mid, np)}) {
      if (!rc) rc = sieve_round(c, sl.pieces, x, y, sl.buf, c->stream, c->ev[2], c->ev[3]);
      if (!rc) rc = walk_round(c, sl.pieces, x, y, sl.par, s->cfg.raw != 0, sl.buf, c->ev[2], c->ev[3]);
      if (rc == 1) rc = set_err(PPSD_ERR_FAILURE, "prime pool overflow in a slice");
    }
  } else if (rc == 1) {
    rc = set_err(PPSD_ERR_FAILURE, "one piece overflowed the prime pool");
  }
  if (!rc) rc = batch_flush(c);
  const double dt = s->now() - t0;
  double steps = 0;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->st.slices++;
    s->st.busy_s += dt;
    for (size_t i = 0; i < np; ++i) {
      Request *r = sl.reqs[i];
      r->slices++;
      r->inflight--;
      if (rc) {
        r->status = rc;
        r->cur = r->prm.pmax;  // plan nothing more for it
        continue;
      }
      const uint64_t np_i = c->b_unit_primes[i];
      const uint64_t w = (uint64_t)(r->per_prime() + 0.5) * np_i;
      r->primes += np_i;
      r->sum += c->b_unit_sums[i];
      r->steps += w;
      steps += (double)w;
      s->st.primes += np_i;
      s->st.window_steps += w;
    }
    if (!rc)
      for (const BHit &h : c->b_result) sl.reqs[h.piece]->hits.push_back(ppsd_factor{h.p, h.k, h.n, h.c});
    if (dt > 0 && steps > 1e9) s->rate = 0.7 * s->rate + 0.3 * (steps / dt);
    for (Request *r : sl.reqs) {
      if (r->cur >= r->prm.pmax && r->inflight == 0) {
        auto it = std::find(s->active.begin(), s->active.end(), r);
        if (it != s->active.end()) {
          s->active.erase(it);
          finish_locked(s, r);
        }
      }
    }
  }
  s->cv_done.notify_all();
  sl.live = false;
}

// Two slices in flight: while one is walked on the context's stream, the
// next one -- planned from what is admitted at that moment -- is sieved on
// the second stream into the other pool.
void serve_loop(ppsd_server *s) {
  ppsd_ctx *c = s->ctx;
  (void)hipSetDevice(c->device);
  std::vector<Request *> act;
  Slice slice[2];
  int nxt = 0;  // the slice being planned; slice[nxt ^ 1] is the one to walk
  for (;;) {
    Slice &cur = slice[nxt ^ 1];
    Slice &next = slice[nxt];
    bool backlog = false;
    {
      // admit whatever arrived; sleep only while nothing is admitted and
      // nothing is in flight (a request may arrive at any point of this loop)
      std::unique_lock<std::mutex> lk(s->mu);
      for (;;) {
        admit_locked(s);
        if (s->stop) return;
        act.clear();
        for (Request *r : s->active)
          if (r->cur < r->prm.pmax) act.push_back(r);
        if (!act.empty() || cur.live) break;
        s->cv_work.wait(lk, [&] { return s->stop || !s->queue.empty(); });
      }
      backlog = !s->queue.empty();
    }
    if (!act.empty()) {
      plan_slice(s, act, next.pieces, next.par, backlog);
      const size_t np = next.pieces.size();
      next.reqs.assign(act.begin(), act.begin() + np);
      next.buf = nxt;
      const double t = s->now();
      for (size_t i = 0; i < np; ++i) {
        Request *r = next.reqs[i];
        if (r->t_start < 0) r->t_start = t;
        r->cur = next.pieces[i].hi;
        r->inflight++;
      }
      next.rc = sieve_round(c, next.pieces, 0, np, next.buf, c->stream2, c->ev2[next.buf][0], c->ev2[next.buf][1]);
      next.live = true;
    }
    if (cur.live) finish_slice(s, cur);
    nxt ^= 1;
  }
}

void fill_result(const Request *r, ppsd_serve_result *res) {
  res->id = r->id;
  res->status = r->status;
  res->slices = r->slices;
  res->primes = r->primes;
  res->prime_sum = r->sum;
  res->factors = r->hits.size();
  res->window_steps = r->steps;
  res->t_arrival = r->t_arrival;
  res->t_admit = r->t_admit;
  res->t_start = r->t_start;
  res->t_done = r->t_done;
}

}  // namespace

int ppsd_serve_open(int device, const ppsd_serve_config *cfg, ppsd_server **out) {
  if (!out) return set_err(PPSD_ERR_INVALID_PARAM, "null argument");
  *out = nullptr;
  ppsd_serve_config c{};
  if (cfg) c = *cfg;
  if (c.slice_ms < 0 || c.slice_ms > 60000) return set_err(PPSD_ERR_INVALID_PARAM, "slice_ms out of range");
  if (c.slice_ms == 0) c.slice_ms = 50;
  if (c.max_active == 0) c.max_active = 4096;
  if (c.max_active > 65536) return set_err(PPSD_ERR_INVALID_PARAM, "max_active must be <= 65536");
  // a context whose own search is empty; the slices bring their units
  ppsd_params p{};
  p.pmin = ((uint64_t)1 << 62) - 64;
  p.pmax = (uint64_t)1 << 62;
  p.kmin = p.kmax = 1;
  p.nmin = p.nmax = 64;
  p.kstep = 2;
  p.koffset = 1;
  p.sign = 1;
  p.nstep = p.nstep_ref = 1;
  ppsd_ctx *ctx = nullptr;
  int rc = ppsd_open(device, &p, 1u << 16, &ctx);
  if (rc) return rc;
  rc = batch_alloc(ctx);
  if (!rc) rc = ensure_sieving(ctx, p.pmax);
  if (rc) {
    std::string keep = ppsd_last_error();
    ppsd_close(ctx);
    return set_err(rc, "%s", keep.c_str());
  }
  ppsd_server *s = new ppsd_server;
  s->ctx = ctx;
  s->pool_bytes = ctx->b_pool_cap * sizeof(uint64_t);
  s->ring_bytes = ctx->b_hit_alloc * sizeof(BHit) * 2;
  if (c.mem_budget == 0) c.mem_budget = s->pool_bytes + s->ring_bytes;
  s->cfg = c;
  s->t0 = Clock::now();
  s->worker = std::thread(serve_loop, s);
  *out = s;
  return 0;
}

int ppsd_serve_submit(ppsd_server *s, const ppsd_params *req, uint64_t *id) {
  if (!s || !req) return set_err(PPSD_ERR_INVALID_PARAM, "null argument");
  if (!walk_domain_ok(req)) return set_err(PPSD_ERR_INVALID_PARAM, "request parameters outside the walk's domain");
  if (req->qmax && req->qmax < isqrt64(req->pmax - 1))
    return set_err(PPSD_ERR_INVALID_PARAM, "requests are sieved completely (qmax must be 0)");
  Request *r = new Request;
  r->prm = *req;
  r->cur = req->pmin;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    r->id = s->next_id++;
    r->t_arrival = s->now();
    r->need = request_need(s, *r);
    s->queue.push_back(r);
    s->st.submitted++;
  }
  s->cv_work.notify_one();
  if (id) *id = r->id;
  return 0;
}

int ppsd_serve_next(ppsd_server *s, int timeout_ms, ppsd_serve_result *res) {
  if (!s || !res) return set_err(PPSD_ERR_INVALID_PARAM, "null argument");
  std::unique_lock<std::mutex> lk(s->mu);
  if (!s->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms < 0 ? 0 : timeout_ms),
                           [&] { return !s->done.empty(); }))
    return 1;
  Request *r = s->done.front();
  s->done.pop_front();
  s->taken[r->id] = r;
  fill_result(r, res);
  return 0;
}

int ppsd_serve_factors(ppsd_server *s, uint64_t id, ppsd_factor *out, uint64_t cap, uint64_t *n_out) {
  if (!s) return set_err(PPSD_ERR_INVALID_PARAM, "null argument");
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->taken.find(id);
  if (it == s->taken.end()) return set_err(PPSD_ERR_STATE, "request %llu is not a taken result", (unsigned long long)id);
  const std::vector<ppsd_factor> &h = it->second->hits;
  if (n_out) *n_out = h.size();
  if (out) memcpy(out, h.data(), std::min<uint64_t>(cap, h.size()) * sizeof(ppsd_factor));
  return 0;
}

int ppsd_serve_release(ppsd_server *s, uint64_t id) {
  if (!s) return set_err(PPSD_ERR_INVALID_PARAM, "null argument");
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->taken.find(id);
  if (it == s->taken.end()) return set_err(PPSD_ERR_STATE, "request %llu is not a taken result", (unsigned long long)id);
  delete it->second;
  s->taken.erase(it);
  return 0;
}

int ppsd_serve_get_stats(ppsd_server *s, ppsd_serve_stats *out) {
  if (!s || !out) return set_err(PPSD_ERR_INVALID_PARAM, "null argument");
  std::lock_guard<std::mutex> lk(s->mu);
  *out = s->st;
  out->queued = s->queue.size();
  out->active = s->active.size();
  out->mem_in_use = s->mem_in_use;
  out->mem_budget = s->cfg.mem_budget;
  out->now = s->now();
  return 0;
}

int ppsd_serve_close(ppsd_server *s) {
  if (!s) return 0;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  if (s->worker.joinable()) s->worker.join();
  for (Request *r : s->queue) delete r;
  for (Request *r : s->active) delete r;
  for (Request *r : s->done) delete r;
  for (auto &kv : s->taken) delete kv.second;
  ppsd_close(s->ctx);
  delete s;
  return 0;
}
