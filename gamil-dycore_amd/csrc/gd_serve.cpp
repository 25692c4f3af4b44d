// gd_serve.cpp -- continuous batching of many independent runs (ensemble
// members, sensitivity runs) of uneven resolution and length on one GPU.
//
// A server takes runs as they arrive (gd_server_submit) and hands back
// results when they are done (gd_server_poll).  Runs of one (resolution,
// time step, diffusion) class share a pool: one handle whose members are
// slots, stepped together by the captured step graph, only its active runs
// stepped.  A class gets a pool when it has runs waiting and none live, with
// as many slots (up to max_slots) as the live footprint allows: the HBM
// budget less the bytes of the pools in use, and no more than the free
// device memory; a pool that has been idle for a while is released and its
// bytes go back to the budget, where waiting classes take them.  Each pool is driven by its own host thread
// on its own stream, so pools of different resolutions overlap on the
// device.  A pool admits waiting runs of its class into free slots between
// chunks of steps; a chunk ends when the first of its active runs completes
// (or after chunk_steps), the run's result is taken and its slot reused.  A
// run whose total mass or energy becomes NaN (diag_mod.F90:79-87) is evicted
// with GD_ERR_NAN and its slot reused.  Members are computed independently
// (tiles and reductions are per member), so a run gives the same bits alone
// and in any mix.
//
// gd_serve replays a trace of arrival times through a server.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdint>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <tuple>
#include <vector>

#include "gamil_dycore.h"

namespace {

thread_local std::string g_serve_error;

int serve_fail(int code, const std::string &m) {
  g_serve_error = m;
  return code;
}

using Clock = std::chrono::steady_clock;

double seconds_since(Clock::time_point t0) {
  return std::chrono::duration<double>(Clock::now() - t0).count();
}

// FNV-1a over the bytes of the final u, v, gd (identity of results)
uint64_t fnv1a(uint64_t h, const double *p, size_t n) {
  const unsigned char *b = (const unsigned char *)p;
  for (size_t k = 0; k < n * sizeof(double); ++k) {
    h ^= b[k];
    h *= 1099511628211ull;
  }
  return h;
}

struct PoolKey {
  int N, M;
  double dt, coef;
  bool operator<(const PoolKey &o) const {
    return std::tie(N, M, dt, coef) < std::tie(o.N, o.M, o.dt, o.coef);
  }
};

// A submitted run: its own copy of the initial state and its result.
struct Run {
  gd_serve_request q{};
  std::vector<double> u, v, gd, ghs;
  gd_serve_result res{};
  bool done = false;
};

struct Class;

struct Pool {
  PoolKey key;
  Class *cls = nullptr;
  int slots = 0;
  int64_t bytes = 0;
  gd_handle *h = nullptr;
  std::vector<int32_t> cutoff;
  int free_slots = 0;  // slots not holding a run (guarded by the server mutex)
  bool exited = false;
  std::thread th;
};

struct Class {
  std::deque<int64_t> waiting;  // run ids in arrival order
  std::vector<Pool *> pools;
  double per_member = 0;        // device bytes of one member
};

}  // namespace

struct gd_server {
  gd_serve_config cfg{};
  gd_config base{};
  Clock::time_point t0;
  std::mutex mu;
  std::condition_variable cv;       // a result is ready, or a pool exited
  std::vector<std::unique_ptr<Run>> runs;
  std::map<PoolKey, Class> classes;
  std::vector<std::unique_ptr<Pool>> pools;
  int64_t used = 0;                 // device bytes of the live pools
  bool closing = false;
  int status = GD_OK;
  std::string err;
};

namespace {

// How long an idle pool waits for more runs of its class before it is
// released (its bytes go back to the budget).
constexpr double kIdleSeconds = 0.05;

gd_config class_config(const gd_server *s, const PoolKey &k, std::vector<int32_t> &cutoff) {
  gd_config c = s->base;
  c.num_lon = k.N;
  c.num_lat = k.M;
  c.time_step_size = k.dt;
  c.diffusion_coef = k.coef;
  cutoff.assign(std::max(1, k.M), 0);
  int nc = 0;
  if (c.use_zonal_tend_filter && s->cfg.filter_lat0 > 0)
    nc = gd_cutoff_from_lat0(k.N, k.M, s->cfg.filter_lat0, cutoff.data(), (int32_t)cutoff.size());
  c.n_cutoff = nc < 0 ? 0 : nc;
  c.cutoff = cutoff.data();
  return c;
}

void drive(gd_server *s, Pool *p);

// Makes a pool for each class with waiting runs and no live pool, with as
// many slots as the budget and the free device memory allow.  Called with
// the mutex held.
int dispatch(gd_server *s) {
  for (auto &kv : s->classes) {
    const PoolKey &key = kv.first;
    Class &c = kv.second;
    bool live = false;
    for (Pool *p : c.pools) live = live || !p->exited;
    const int need = (int)c.waiting.size();
    if (live || need <= 0) continue;
    std::unique_ptr<Pool> p(new Pool());
    p->key = key;
    gd_config cc = class_config(s, key, p->cutoff);
    if (c.per_member <= 0) {
      gd_handle *probe = nullptr;
      int rc = gd_create(&cc, 1, &probe);
      if (rc != GD_OK) return serve_fail(rc, std::string("pool: ") + gd_last_error(nullptr));
      c.per_member = (double)gd_device_bytes(probe);
      gd_destroy(probe);
    }
    // the live footprint: what the budget has left, and no more than the
    // device has free (keeping 1 GiB for everything else)
    double room = (double)s->cfg.hbm_budget_bytes - (double)s->used;
    size_t dfree = 0, dtotal = 0;
    if (hipMemGetInfo(&dfree, &dtotal) == hipSuccess) room = std::min(room, (double)dfree - (double)(1ll << 30));
    const int fit = (int)std::floor(room / c.per_member);
    // room for the runs that arrive while these run: max_slots, not the
    // runs waiting now
    const int slots = std::min(s->cfg.max_slots, fit);
    if (slots < 1) {
      if (s->used > 0) continue;  // waits until a pool is released
      // nothing to release: the budget cannot hold one run of this class
      for (int64_t id : c.waiting) {
        Run *r = s->runs[id].get();
        r->res.status = GD_ERR_INVALID_ARG;
        r->res.finish_s = seconds_since(s->t0);
        r->done = true;
      }
      c.waiting.clear();
      s->cv.notify_all();
      continue;
    }
    int rc = gd_create(&cc, slots, &p->h);
    if (rc != GD_OK) return serve_fail(rc, std::string("pool: ") + gd_last_error(nullptr));
    p->slots = p->free_slots = slots;
    p->cls = &c;
    p->bytes = gd_device_bytes(p->h);
    s->used += p->bytes;
    Pool *raw = p.get();
    c.pools.push_back(raw);
    s->pools.push_back(std::move(p));
    raw->th = std::thread(drive, s, raw);
  }
  return GD_OK;
}

// One pool's loop: admit waiting runs of its class, step to the next
// completion, harvest.  Active runs occupy slots 0..k-1 and only those are
// stepped; a finished run's slot is filled by moving the last active run
// into it (its device state copied whole, so it continues bit for bit).
void drive(gd_server *s, Pool *p) {
  const int N = p->key.N, M = p->key.M, S = p->slots;
  std::vector<Run *> run(S, nullptr);
  std::vector<int> left(S, 0);
  std::vector<double> mass, energy, fu((size_t)M * N), fv((size_t)(M - 1) * N), fg((size_t)M * N);
  int k = 0;  // active runs, in slots 0..k-1
  const int chunk = std::max(1, s->cfg.chunk_steps);
  Class &cls = *p->cls;
  auto finish = [&](Run *r, int status) {
    std::lock_guard<std::mutex> lk(s->mu);
    r->res.status = status;
    r->res.finish_s = seconds_since(s->t0);
    r->done = true;
    s->cv.notify_all();
  };
  Clock::time_point idle_since = Clock::now();
  for (;;) {
    // admission: waiting runs of the class into the free slots k, k+1, ...
    std::vector<Run *> admit;
    {
      std::unique_lock<std::mutex> lk(s->mu);
      while (k + (int)admit.size() < S && !cls.waiting.empty()) {
        admit.push_back(s->runs[cls.waiting.front()].get());
        cls.waiting.pop_front();
      }
      p->free_slots = S - k - (int)admit.size();
      if (k == 0 && admit.empty()) {
        // idle: wait for runs of the class, or leave
        if (s->closing || seconds_since(idle_since) >= kIdleSeconds) {
          p->exited = true;
          s->used -= p->bytes;
          p->free_slots = 0;
          dispatch(s);  // the budget went back: classes waiting for room
          s->cv.notify_all();
          break;
        }
        s->cv.wait_for(lk, std::chrono::duration<double>(kIdleSeconds / 4));
        continue;
      }
    }
    for (Run *r : admit) {
      r->res.start_s = seconds_since(s->t0);
      r->res.slot = k;
      const int rc = gd_set_state(p->h, k, r->u.data(), r->v.data(), r->gd.data(), r->ghs.data(), N);
      if (rc != GD_OK || r->q.nsteps <= 0) {
        finish(r, rc);
        std::lock_guard<std::mutex> lk(s->mu);
        ++p->free_slots;
        continue;
      }
      run[k] = r;
      left[k] = r->q.nsteps;
      ++k;
    }
    if (k == 0) continue;
    int n = chunk;
    for (int q = 0; q < k; ++q) n = std::min(n, left[q]);
    int rc = gd_set_active_members(p->h, k);
    mass.assign((size_t)n * S, 0.0);
    energy.assign((size_t)n * S, 0.0);
    int32_t nan_step = 0;
    if (rc == GD_OK) rc = gd_step(p->h, n, mass.data(), energy.data(), nullptr, &nan_step);
    if (rc != GD_OK && rc != GD_ERR_NAN) {
      std::lock_guard<std::mutex> lk(s->mu);
      if (s->status == GD_OK) {
        s->status = rc;
        s->err = gd_last_error(p->h);
      }
      for (int q = 0; q < k; ++q) {
        run[q]->res.status = rc;
        run[q]->res.finish_s = seconds_since(s->t0);
        run[q]->done = true;
      }
      k = 0;
      s->cv.notify_all();
      continue;
    }
    for (int q = 0; q < k; ++q) left[q] -= n;
    // harvest finished and failed runs (from the top, so moves never pass a
    // slot not yet seen)
    for (int q = k - 1; q >= 0; --q) {
      const double m_last = mass[(size_t)(n - 1) * S + q], e_last = energy[(size_t)(n - 1) * S + q];
      const bool bad = std::isnan(m_last) || std::isnan(e_last);
      if (left[q] > 0 && !bad) continue;
      Run *r = run[q];
      r->res.total_mass = m_last;
      r->res.total_energy = e_last;
      if (!bad && gd_get_state(p->h, q, fu.data(), fv.data(), fg.data(), N) == GD_OK) {
        uint64_t hsh = 1469598103934665603ull;
        hsh = fnv1a(hsh, fu.data(), fu.size());
        hsh = fnv1a(hsh, fv.data(), fv.size());
        hsh = fnv1a(hsh, fg.data(), fg.size());
        r->res.state_hash = hsh;
      }
      const int last = k - 1;
      if (q != last) {
        if ((rc = gd_copy_member(p->h, q, last)) != GD_OK) {
          finish(r, rc);
          continue;
        }
        run[q] = run[last];
        left[q] = left[last];
      }
      run[last] = nullptr;
      --k;
      finish(r, bad ? GD_ERR_NAN : GD_OK);
    }
    idle_since = Clock::now();
  }
  gd_destroy(p->h);
  p->h = nullptr;
}

}  // namespace

extern "C" {

const char *gd_serve_last_error(void) { return g_serve_error.c_str(); }

int gd_server_create(const gd_serve_config *cfg, gd_server **out) {
  if (!cfg || !cfg->base || !out) return serve_fail(GD_ERR_INVALID_ARG, "null argument");
  if (cfg->max_slots < 1 || cfg->hbm_budget_bytes <= 0)
    return serve_fail(GD_ERR_INVALID_ARG, "max_slots and hbm_budget_bytes must be positive");
  gd_server *s = new gd_server();
  s->cfg = *cfg;
  s->base = *cfg->base;
  s->cfg.base = &s->base;
  s->t0 = Clock::now();
  *out = s;
  return GD_OK;
}

int gd_server_submit(gd_server *s, const gd_serve_request *q, int64_t *id) {
  if (!s || !q || !id) return serve_fail(GD_ERR_INVALID_ARG, "null argument");
  if (q->num_lon < 8 || q->num_lat < 5 || !(q->time_step_size > 0) || !q->u || !q->v || !q->gd || !q->ghs)
    return serve_fail(GD_ERR_INVALID_ARG, "bad request");
  std::unique_ptr<Run> r(new Run());
  r->q = *q;
  const size_t full = (size_t)q->num_lat * q->num_lon, half = (size_t)(q->num_lat - 1) * q->num_lon;
  r->u.assign(q->u, q->u + full);
  r->v.assign(q->v, q->v + half);
  r->gd.assign(q->gd, q->gd + full);
  r->ghs.assign(q->ghs, q->ghs + full);
  r->q.u = r->u.data();
  r->q.v = r->v.data();
  r->q.gd = r->gd.data();
  r->q.ghs = r->ghs.data();
  std::lock_guard<std::mutex> lk(s->mu);
  if (s->closing) return serve_fail(GD_ERR_INVALID_ARG, "server is closing");
  r->res.arrival_s = seconds_since(s->t0);
  r->res.status = -1;
  r->res.slot = -1;
  *id = (int64_t)s->runs.size();
  const PoolKey key{q->num_lon, q->num_lat, q->time_step_size, q->diffusion_coef};
  s->runs.push_back(std::move(r));
  s->classes[key].waiting.push_back(*id);
  s->cv.notify_all();
  const int rc = dispatch(s);
  return rc;
}

int gd_server_poll(gd_server *s, int64_t id, int32_t wait, gd_serve_result *res, int32_t *done) {
  if (!s) return serve_fail(GD_ERR_INVALID_ARG, "null argument");
  std::unique_lock<std::mutex> lk(s->mu);
  if (id < 0 || id >= (int64_t)s->runs.size()) return serve_fail(GD_ERR_INVALID_ARG, "unknown run id");
  Run *r = s->runs[id].get();
  while (wait && !r->done) s->cv.wait_for(lk, std::chrono::milliseconds(5));
  if (done) *done = r->done ? 1 : 0;
  if (res && r->done) *res = r->res;
  return GD_OK;
}

int gd_server_destroy(gd_server *s) {
  if (!s) return GD_OK;
  {
    std::unique_lock<std::mutex> lk(s->mu);
    // every submitted run finishes first
    for (auto &r : s->runs)
      while (!r->done) s->cv.wait_for(lk, std::chrono::milliseconds(5));
    s->closing = true;
    s->cv.notify_all();
  }
  for (auto &p : s->pools)
    if (p->th.joinable()) p->th.join();
  const int rc = s->status;
  const std::string err = s->err;
  delete s;
  return rc == GD_OK ? GD_OK : serve_fail(rc, err);
}

int gd_serve(const gd_serve_config *cfg, const gd_serve_request *req, int32_t n, gd_serve_result *res,
             double *wall_seconds) {
  if (!cfg || !cfg->base || n < 0 || (n > 0 && (!req || !res)))
    return serve_fail(GD_ERR_INVALID_ARG, "null argument");
  if (cfg->max_slots < 1 || cfg->hbm_budget_bytes <= 0)
    return serve_fail(GD_ERR_INVALID_ARG, "max_slots and hbm_budget_bytes must be positive");
  for (int r = 0; r < n; ++r) {
    const gd_serve_request &q = req[r];
    if (q.num_lon < 8 || q.num_lat < 5 || !(q.time_step_size > 0) || !q.u || !q.v || !q.gd || !q.ghs ||
        !(q.arrival_s >= 0))
      return serve_fail(GD_ERR_INVALID_ARG, "bad request " + std::to_string(r));
  }
  gd_server *s = nullptr;
  int rc = gd_server_create(cfg, &s);
  if (rc != GD_OK) return rc;
  // submit in arrival order, each at its arrival time
  std::vector<int> order(n);
  for (int r = 0; r < n; ++r) order[r] = r;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return req[a].arrival_s < req[b].arrival_s; });
  std::vector<int64_t> ids(n, -1);
  for (int r : order) {
    const double wait = req[r].arrival_s - seconds_since(s->t0);
    if (wait > 0) std::this_thread::sleep_for(std::chrono::duration<double>(wait));
    if ((rc = gd_server_submit(s, &req[r], &ids[r])) != GD_OK) break;
  }
  for (int r = 0; r < n && rc == GD_OK; ++r) {
    int32_t done = 0;
    gd_server_poll(s, ids[r], 1, &res[r], &done);
    res[r].arrival_s = req[r].arrival_s;  // the trace's arrival time
  }
  const double wall = seconds_since(s->t0);
  const std::string err = g_serve_error;
  const int rc2 = gd_server_destroy(s);
  if (wall_seconds) *wall_seconds = wall;
  if (rc != GD_OK) return serve_fail(rc, err);
  return rc2;
}

}  // extern "C"
