// gd_serve.cpp -- continuous batching of many independent runs (ensemble
// members, sensitivity runs) of uneven resolution and length on one GPU.
//
// Runs of one (resolution, time step, diffusion) class share a pool: one
// handle whose members are slots, stepped together by the captured step
// graph, only its active runs stepped.  A pool's slot count is fixed by its
// share of the HBM budget (admission by footprint).  Each pool is driven by its own host thread on its
// own stream, so pools of different resolutions overlap on the device.  A
// request is admitted as soon as it has arrived and a slot of its class is
// free; the pool advances in chunks that end when the first of its active
// runs completes, the run's result is taken and the slot is reused.  Members
// are computed independently (tiles and reductions are per member), so a run
// gives the same bits alone and in any mix.
//
// Host code over the public C ABI only.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
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

struct Pool {
  PoolKey key;
  std::vector<int> queue;       // request indices, by arrival
  int slots = 0;
  gd_handle *h = nullptr;
  std::vector<int32_t> cutoff;
  int status = GD_OK;
  std::string err;
};

// One pool's loop: admit, step to the next completion, harvest.  Active runs
// occupy slots 0..k-1 and only those are stepped; a finished run's slot is
// filled by moving the last active run into it (its device state copied
// whole, so it continues bit for bit).
void drive(Pool &p, const gd_serve_config &cfg, const gd_serve_request *req, gd_serve_result *res,
           Clock::time_point t0) {
  const int N = p.key.N, M = p.key.M, S = p.slots;
  std::vector<int> run(S, -1), left(S, 0);
  std::vector<double> mass, energy, fu((size_t)M * N), fv((size_t)(M - 1) * N), fg((size_t)M * N);
  size_t next = 0;
  int k = 0;  // active runs, in slots 0..k-1
  const int chunk = std::max(1, cfg.chunk_steps);
  auto fail_all = [&](int rc) {
    p.status = rc;
    p.err = gd_last_error(p.h);
    for (int s = 0; s < k; ++s) {
      res[run[s]].status = rc;
      res[run[s]].finish_s = seconds_since(t0);
    }
  };
  while (next < p.queue.size() || k > 0) {
    // admission: arrived requests into the free slots k, k+1, ...
    while (k < S && next < p.queue.size() && req[p.queue[next]].arrival_s <= seconds_since(t0)) {
      const int r = p.queue[next++];
      const gd_serve_request &q = req[r];
      gd_serve_result &o = res[r];
      o.start_s = seconds_since(t0);
      o.slot = k;
      const int rc = gd_set_state(p.h, k, q.u, q.v, q.gd, q.ghs, N);
      if (rc != GD_OK || q.nsteps <= 0) {
        o.status = rc;
        o.finish_s = seconds_since(t0);
        continue;
      }
      run[k] = r;
      left[k] = q.nsteps;
      ++k;
    }
    if (k == 0) {
      if (next < p.queue.size()) {
        const double wait = req[p.queue[next]].arrival_s - seconds_since(t0);
        if (wait > 0) std::this_thread::sleep_for(std::chrono::duration<double>(std::min(wait, 0.01)));
      }
      continue;
    }
    int n = chunk;
    for (int s = 0; s < k; ++s) n = std::min(n, left[s]);
    int rc = gd_set_active_members(p.h, k);
    if (rc != GD_OK) return fail_all(rc);
    mass.assign((size_t)n * S, 0.0);
    energy.assign((size_t)n * S, 0.0);
    int32_t nan_step = 0;
    rc = gd_step(p.h, n, mass.data(), energy.data(), nullptr, &nan_step);
    if (rc != GD_OK && rc != GD_ERR_NAN) return fail_all(rc);
    for (int s = 0; s < k; ++s) left[s] -= n;
    // harvest finished runs (from the top, so moves never pass a slot not yet seen)
    for (int s = k - 1; s >= 0; --s) {
      const double m_last = mass[(size_t)(n - 1) * S + s], e_last = energy[(size_t)(n - 1) * S + s];
      const bool bad = std::isnan(m_last) || std::isnan(e_last);
      if (left[s] > 0 && !bad) continue;
      gd_serve_result &o = res[run[s]];
      o.total_mass = m_last;
      o.total_energy = e_last;
      o.status = bad ? GD_ERR_NAN : GD_OK;
      if (!bad && gd_get_state(p.h, s, fu.data(), fv.data(), fg.data(), N) == GD_OK) {
        uint64_t hsh = 1469598103934665603ull;
        hsh = fnv1a(hsh, fu.data(), fu.size());
        hsh = fnv1a(hsh, fv.data(), fv.size());
        hsh = fnv1a(hsh, fg.data(), fg.size());
        o.state_hash = hsh;
      }
      o.finish_s = seconds_since(t0);
      const int last = k - 1;
      if (s != last) {
        if ((rc = gd_copy_member(p.h, s, last)) != GD_OK) return fail_all(rc);
        run[s] = run[last];
        left[s] = left[last];
      }
      run[last] = -1;
      --k;
    }
  }
}

}  // namespace

extern "C" {

const char *gd_serve_last_error(void) { return g_serve_error.c_str(); }

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
  // classes of runs that can share a pool
  std::map<PoolKey, int> index;
  std::vector<Pool> pools;
  for (int r = 0; r < n; ++r) {
    const PoolKey k{req[r].num_lon, req[r].num_lat, req[r].time_step_size, req[r].diffusion_coef};
    auto it = index.find(k);
    if (it == index.end()) {
      it = index.emplace(k, (int)pools.size()).first;
      pools.emplace_back();
      pools.back().key = k;
    }
    pools[it->second].queue.push_back(r);
  }
  for (Pool &p : pools)
    std::stable_sort(p.queue.begin(), p.queue.end(),
                     [&](int a, int b) { return req[a].arrival_s < req[b].arrival_s; });
  for (int r = 0; r < n; ++r) {
    res[r] = gd_serve_result{};
    res[r].arrival_s = req[r].arrival_s;
    res[r].status = -1;
    res[r].slot = -1;
  }
  // admission by footprint: the budget is shared in proportion to each
  // class's grid size x requests; a pool holds at least one run
  double weight = 0;
  for (const Pool &p : pools) weight += (double)p.key.N * p.key.M * p.queue.size();
  int rc = GD_OK;
  for (Pool &p : pools) {
    gd_config c = *cfg->base;
    c.num_lon = p.key.N;
    c.num_lat = p.key.M;
    c.time_step_size = p.key.dt;
    c.diffusion_coef = p.key.coef;
    p.cutoff.assign(std::max(1, p.key.M), 0);
    int nc = 0;
    if (c.use_zonal_tend_filter && cfg->filter_lat0 > 0)
      nc = gd_cutoff_from_lat0(p.key.N, p.key.M, cfg->filter_lat0, p.cutoff.data(), (int32_t)p.cutoff.size());
    if (nc < 0) nc = 0;
    c.n_cutoff = nc;
    c.cutoff = p.cutoff.data();
    // bytes of one member, from a one-member handle
    gd_handle *probe = nullptr;
    rc = gd_create(&c, 1, &probe);
    if (rc != GD_OK) {
      const std::string m = gd_last_error(nullptr);
      for (Pool &q : pools) gd_destroy(q.h);
      return serve_fail(rc, "pool " + std::to_string(p.key.N) + "x" + std::to_string(p.key.M) + ": " + m);
    }
    const double per = (double)gd_device_bytes(probe);
    gd_destroy(probe);
    const double share = cfg->hbm_budget_bytes * ((double)p.key.N * p.key.M * p.queue.size() / weight);
    const int fit = (int)std::floor(share / per);
    p.slots = std::max(1, std::min({cfg->max_slots, fit, (int)p.queue.size()}));
    rc = gd_create(&c, p.slots, &p.h);
    if (rc != GD_OK) {
      const std::string m = gd_last_error(nullptr);
      for (Pool &q : pools) gd_destroy(q.h);
      return serve_fail(rc, "pool " + std::to_string(p.key.N) + "x" + std::to_string(p.key.M) + ": " + m);
    }
  }
  const Clock::time_point t0 = Clock::now();
  std::vector<std::thread> th;
  for (Pool &p : pools) th.emplace_back([&p, cfg, req, res, t0] { drive(p, *cfg, req, res, t0); });
  for (std::thread &t : th) t.join();
  const double wall = seconds_since(t0);
  if (wall_seconds) *wall_seconds = wall;
  std::string err;
  for (Pool &p : pools) {
    if (p.status != GD_OK && rc == GD_OK) {
      rc = p.status;
      err = p.err;
    }
    gd_destroy(p.h);
  }
  if (rc != GD_OK) return serve_fail(rc, err);
  return GD_OK;
}

}  // extern "C"
