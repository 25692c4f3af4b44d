// gd_api.hip -- C ABI (include/gamil_dycore.h): handle set-up, device memory,
// the per-step launch schedule (captured once into a hipGraph) and the
// operator seams used by the Fortran facade and the tests.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "gamil_dycore.h"
#include "gd_device.hpp"
#include "gd_geom.hpp"
#include "gd_launch.hpp"

using namespace gd;

namespace {

thread_local std::string g_create_error;

// Staging rows of gd_filter_rows_noalloc: one array per device, resolved to
// its device address once per device (g_scratch_ptr) when the first handle on
// that device is created.  In host code the name `scratch` is only a host-side
// stand-in for the symbol; kernels and copies use the resolved address.
constexpr size_t kScratchBytes = (size_t)4 << 20;
static __device__ __align__(256) double scratch[kScratchBytes / sizeof(double)];
std::mutex g_scratch_mu;
std::map<int, double *> g_scratch_ptr;

struct ProfEntry {
  std::string family;
  hipEvent_t e0, e1;
};

}  // namespace

struct gd_handle {
  gd_config cfg{};
  std::vector<int32_t> cutoff;
  Geometry g;
  int nmem = 1, prec = 0, N = 0, M = 0, P = 0;
  int filt_mm = 0;               // momentum filter rows on MFMA: 0 no (FFT), 1 FP32, 2 BF16 inputs
  int nact = 1;                  // members 0..nact-1 are stepped
  int device = 0;
  hipStream_t stream = nullptr;
  std::vector<void *> allocs;
  int64_t bytes = 0;
  char *arena = nullptr;         // all members, mstride bytes apart
  size_t mstride = 0, fs = 0, cs = 0, off_part = 0, off_dpart = 0, off_sx = 0, off_beta = 0, off_sy = 0;
  RowCoef *dcoef = nullptr;
  std::vector<StageSlot> slots;
  StageSlot *dslots = nullptr;
  std::vector<FiltRow> frows;    // filter workgroups: (member, latitude) with staged rows
  FiltRow *dfrows = nullptr;
  std::vector<int32_t> mslots;   // staging slots of the momentum rows (MFMA filter)
  int32_t *dmslots = nullptr;
  void *dmm_phi = nullptr;       // basis matrices of the MFMA filter
  double *dtwid = nullptr;       // [N/2+1][2] cos, sin of 2 pi k/N
  double *dghost = nullptr;      // [2][P]: 1.0 row, 0.0 row (stencil rows beyond the grid)
  int fft_L = 1, fft_half = 1;
  std::vector<int> fft_rad;
  // fused phase launches (k_phase): the filter rows inside the tendency
  // launch; the row tiles cover the row ranges far_lo/far_cnt.  fused_ok:
  // the shape allows them; wide_ok: also one workgroup per row (k_wphase);
  // fused_mode: -1 by launch size, 0 never, 1 always k_phase, 2 always
  // k_wphase where wide_ok
  bool fused_ok = false, wide_ok = false;
  int fused_mode = -1;
  int upd_narrow = -1;  // k_update rounds of four points and partials per thread: -1 by band size, 0 no, 1 yes
  int ncu = 256;  // compute units of the device
  int nfar_rng = 0, nfar_rows = 0;
  int far_lo[kMaxFarRng] = {}, far_cnt[kMaxFarRng] = {};
  int32_t *dstep = nullptr;
  double *ddiag = nullptr;  // [nmem][cap][3]
  int diag_cap = 512;
  std::map<int, hipGraphExec_t> graphs;  // captured steps per (active-member count, steps per replay)
  bool graph_failed = false;
  std::string err;
  std::vector<ProfEntry> *prof = nullptr;
  // gd_filter_rows_noalloc: this device's scratch array, and per row kind
  // (0 full rows, 1 half rows) member 0's filtered rows in staging order with
  // their filter workgroups (slot = position in the scratch array)
  double *scratch = nullptr;
  std::string scratch_err;
  // autotuned launch shapes per active-member count (gd_autotune): phase
  // launch form (0 separate, 1 k_phase, 2 k_wphase), k_update workgroups
  // per member
  std::map<int, int> tuned_fused, tuned_wpm;
  std::vector<int> prow[2];
  FiltRow *dprow[2] = {nullptr, nullptr};
};

namespace {

int fail(gd_handle *h, int code, const std::string &msg) {
  if (h)
    h->err = msg;
  else
    g_create_error = msg;
  return code;
}

int hip_fail(gd_handle *h, hipError_t e, const char *what) {
  return fail(h, GD_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

#define GD_HIP(h, call)                                  \
  do {                                                   \
    hipError_t e_ = (call);                              \
    if (e_ != hipSuccess) return hip_fail((h), e_, #call); \
  } while (0)

void *dalloc(gd_handle *h, size_t bytes, hipError_t &e) {
  void *p = nullptr;
  e = hipMalloc(&p, bytes);
  if (e != hipSuccess) return nullptr;
  h->allocs.push_back(p);
  h->bytes += (int64_t)bytes;
  return p;
}

Args base_args(const gd_handle *h) {
  Args a{};
  a.arena = h->arena;
  a.mstride = h->mstride;
  a.fs = h->fs;
  a.cs = h->cs;
  a.off_part = h->off_part;
  a.off_dpart = h->off_dpart;
  a.off_sx = h->off_sx;
  a.off_beta = h->off_beta;
  a.off_sy = h->off_sy;
  a.mslots = h->dmslots;
  a.mm_phi = h->dmm_phi;
  a.diag = h->ddiag;
  a.coef = h->dcoef;
  a.slots = h->dslots;
  a.frows = h->dfrows;
  a.twid = h->dtwid;
  a.ghost = h->dghost;
  a.fft_L = h->fft_L;
  a.fft_half = h->fft_half;
  a.fft_nrad = (int)h->fft_rad.size();
  for (int k = 0; k < a.fft_nrad; ++k) a.fft_rad[k] = h->fft_rad[k];
  a.step_ctr = h->dstep;
  a.N = h->N;
  a.M = h->M;
  a.pitch = h->P;
  a.nmem = h->nact;
  a.m0 = 0;
  a.nseg = (h->N + kSeg - 1) / kSeg;
  a.ntiles = h->nact * h->M * a.nseg;  // (member, row, segment) tiles
  a.nslots = (int)h->slots.size();
  // k_update: two rows per workgroup for a single run (each workgroup forms
  // its member's beta from all the partials, so fewer workgroups read less;
  // one row each or four rows each measured slower at 720x361); ensembles
  // share about 512 workgroups
  a.nmem_launch = h->nact;
  a.upd_wpm = std::max(1, std::min((h->M + 1) / 2, 512 / h->nact));
  {
    auto it = h->tuned_wpm.find(h->nact);
    if (it != h->tuned_wpm.end()) a.upd_wpm = it->second;
  }
  // four points and partials per thread per round (half the registers of
  // eight: more workgroups resident) except where a workgroup's band fits
  // one round of eight but not of four, which would add a round trip.
  // Measured on MI355X (mixed, same bits): one run of 360x181 4% faster,
  // 1440x721 2.6%, ensembles 0.5-1%; 720x361 (1,440 points per band) equal.
  {
    const int cnt = ((h->M + a.upd_wpm - 1) / a.upd_wpm) * h->N;
    a.upd_narrow = h->upd_narrow >= 0 ? h->upd_narrow : !(cnt > 4 * kBlock && cnt <= 8 * kBlock);
  }
  a.upwind = h->cfg.uv_adv_scheme == GD_ADV_UPWIND;
  a.beta_lon = h->cfg.uv_adv_upwind_lon_beta;
  a.beta_lat = h->cfg.uv_adv_upwind_lat_beta;
  a.qcon = h->cfg.qcon_modified != 0;
  a.diag_cap = h->diag_cap;
  a.dlon = h->g.dlon;
  a.dlat = h->g.dlat;
  return a;
}

Member hmember(const gd_handle *h, int m) { return member_at(base_args(h), m); }

// Runs one launch, bracketed by events when profiling.
template <typename F>
hipError_t run(gd_handle *h, const char *family, F &&f) {
  if (!h->prof) return f();
  ProfEntry pe;
  pe.family = family;
  hipEventCreate(&pe.e0);
  hipEventCreate(&pe.e1);
  hipEventRecord(pe.e0, h->stream);
  hipError_t e = f();
  hipEventRecord(pe.e1, h->stream);
  h->prof->push_back(pe);
  return e;
}

hipError_t enqueue_filter(gd_handle *h, Args a) {
  if (h->frows.empty()) return hipSuccess;
  if (h->filt_mm && !h->mslots.empty()) {
    // the momentum rows of every active member projected on the matrix
    // cores, one workgroup per (slot, 16 members); k_filter then finishes them
    Args m = a;
    m.ntiles = (int)h->mslots.size() * ((h->nact + 15) / 16);
    hipError_t e = run(h, "filter_mm", [&] { return launch_filter_mm(m, h->filt_mm == 2, h->stream); });
    if (e != hipSuccess) return e;
    a.mm = 1;
  }
  // member-major: the first nact members' rows
  a.ntiles = (int)(h->frows.size() / h->nmem) * h->nact;
  return run(h, "filter", [&] { return launch_filter(a, h->prec, h->stream); });
}

// The k_phase tiles of members 0..nm-1: their filter rows, then the row
// tiles of the far-row ranges.
void set_fused(const gd_handle *h, Args &a, int nm) {
  a.nnear = (int)(h->frows.size() / h->nmem) * nm;
  a.nnear_pad = ((a.nnear + kXcd - 1) / kXcd) * kXcd;
  a.nfar_rng = h->nfar_rng;
  a.nfar_rows = h->nfar_rows;
  for (int k = 0; k < kMaxFarRng; ++k) {
    a.far_lo[k] = h->far_lo[k];
    a.far_cnt[k] = h->far_cnt[k];
  }
  a.ntiles = nm * h->nfar_rows * a.nseg;
}

// Fused phase launches pay off while a launch is a few rounds of resident
// tiles (latency bound); past that the lower occupancy of k_phase (five
// tiles per CU against seven or eight for k_tend) costs more than the
// launches it saves.  Measured on MI355X (mixed, aggregate model-days/s):
// fused still ahead at 16 runs of 720x361 (17,328 tiles), 64 members of
// 360x181 (23,168) and 128 of 180x91 (11,648); even at 32 runs of 720x361
// (34,656).
constexpr int64_t kFusedMaxTiles = 24576;
// Rows past 1024 longitudes (eight per thread in the filter tiles; three
// tiles per CU): a 1440x721 run (4,326 row tiles) measured 4% faster fused;
// larger launches are left to the autotuner.
constexpr int64_t kFusedMaxTilesWide = 6144;

// Form of the phase launches: 0 k_tend + k_filter, 1 k_phase, 2 k_wphase.
// k_wphase for rows of two segments by default while its workgroups -- one
// per filter row and one per far row -- fit two per CU: measured on MI355X
// (mixed, one or two runs of 288x145, 360x181, 480x241) 3-10% faster than
// k_phase, and 4-20% slower once they take a second round.  Rows of three
// segments (one run of 720x361: 0.83-0.85 of k_phase) take it only when
// asked for or tuned.
int phase_form(const gd_handle *h) {
  if (!h->fused_ok || h->fused_mode == 0) return 0;
  if (h->fused_mode == 1) return 1;
  if (h->fused_mode == 2) return h->wide_ok ? 2 : 1;
  auto it = h->tuned_fused.find(h->nact);
  if (it != h->tuned_fused.end()) return it->second == 2 && !h->wide_ok ? 1 : it->second;
  if (h->wide_ok && (h->N + kSeg - 1) / kSeg == 2) {
    const int64_t wgs = (int64_t)h->nact * ((int64_t)(h->frows.size() / h->nmem) + h->nfar_rows);
    if (wgs <= 2 * (int64_t)h->ncu) return 2;
  }
  const int64_t tiles = (int64_t)h->nact * h->M * ((h->N + kSeg - 1) / kSeg);
  return tiles <= (h->N <= 4 * kBlock ? kFusedMaxTiles : kFusedMaxTilesWide) ? 1 : 0;
}

bool use_fused(const gd_handle *h) { return phase_form(h) != 0; }

// One tendency phase (a.phase) of the active members with its filter rows:
// one k_phase launch, or k_tend then k_filter.
hipError_t enqueue_phase(gd_handle *h, Args a, int pass, bool in_full) {
  a.ntiles = h->nact * h->M * a.nseg;
  const int form = phase_form(h);
  if (form == 0) {
    hipError_t e = run(h, "tend", [&] { return launch_tend(a, h->prec, pass, in_full, h->stream); });
    if (e != hipSuccess) return e;
    return enqueue_filter(h, a);
  }
  set_fused(h, a, h->nact);
  if (form == 2) return run(h, "wphase", [&] { return launch_wphase(a, h->prec, pass, in_full, h->stream); });
  return run(h, "phase", [&] { return launch_phase(a, h->prec, pass, in_full, h->stream); });
}

// predict_correct(dt, s0 -> s_out, pass) (dycore_mod.F90:754-792)
hipError_t enqueue_call(gd_handle *h, double dt, int pass, int s0, int s_out, int qcon) {
  Args a = base_args(h);
  a.pass = pass;
  a.qcon = qcon;
  a.dt = dt;
  a.hdt = dt * 0.5;
  hipError_t e;
  // first predictor: T = L(S0), S1 = S0 + dt/2 T
  a.phase = PH_A; a.s_in = s0; a.s0 = s0; a.s_out = SP1;
  if ((e = enqueue_phase(h, a, pass, s0 == SX)) != hipSuccess) return e;
  // second predictor: T = L(S1), S2 = S0 + dt/2 T, keep T
  a.phase = PH_B; a.s_in = SP1; a.s_out = SP2;
  if ((e = enqueue_phase(h, a, pass, false)) != hipSuccess) return e;
  // corrector tendency T = L(S2) and the beta inner-product partials
  a.phase = PH_C; a.s_in = SP2;
  if ((e = enqueue_phase(h, a, pass, false)) != hipSuccess) return e;
  a.s_out = s_out;
  return run(h, "update", [&] { return launch_update(a, h->prec, h->stream); });
}

hipError_t enqueue_diffusion(gd_handle *h, int s_in, bool in_full, double dt) {
  Args a = base_args(h);
  a.phase = PH_D;
  a.s_in = s_in;
  a.in_full = in_full;
  const int norder = h->cfg.diffusion_order / 2;
  const int sign = ((norder + 1) % 2 == 0) ? 1 : -1;  // (-1)**(order/2+1)
  a.kdiff = sign * dt * h->cfg.diffusion_coef;
  // diffusion_order/2 applications of the operator; intermediate results
  // ping-pong between the predictor slots (free after the integrator)
  hipError_t e;
  int src = -1;
  for (int k = 1; k <= norder; ++k) {
    a.lap_src = src >= 0;
    a.s_lap = src >= 0 ? src : 0;
    a.lap_final = k == norder;
    a.s_lapout = src == SP1 ? SP2 : SP1;
    if (a.lap_final && use_fused(h)) {
      // the last application with its filter rows in one launch
      Args f = a;
      set_fused(h, f, h->nact);
      return run(h, "diffuse", [&] { return launch_dphase(f, h->prec, h->stream); });
    }
    e = run(h, "diffuse", [&] { return launch_diffuse(a, h->prec, h->stream); });
    if (e != hipSuccess) return e;
    src = a.s_lapout;
  }
  return enqueue_filter(h, a);
}

hipError_t enqueue_finish(gd_handle *h, int mode, int s_in, int m0, int nm) {
  Args a = base_args(h);
  a.s_in = s_in;
  a.m0 = m0;
  a.ntiles = nm * h->M * a.nseg;
  hipError_t e = run(h, "finish", [&] { return launch_finish(a, mode, h->stream); });
  if (e != hipSuccess) return e;
  return run(h, "diag", [&] { return launch_diag(a, nm, h->stream); });
}

// space_operators(state s_in, tend, pass) (dycore_mod.F90:184-365) into Tc,
// filter and rescale included.
hipError_t enqueue_tendency(gd_handle *h, int pass, int s_in) {
  Args a = base_args(h);
  a.pass = pass;
  a.phase = PH_T;
  a.s_in = s_in;
  a.s0 = s_in;
  return enqueue_phase(h, a, pass, s_in == SX);
}

// one pointwise isp_splitting operation (k_isp mode)
hipError_t enqueue_isp_op(gd_handle *h, int mode, int s_in, int s_out, double hdt, double scale) {
  Args a = base_args(h);
  a.isp_mode = mode;
  a.s_in = s_in;
  a.s_out = s_out;
  a.hdt = hdt;
  a.isp_scale = scale;
  return run(h, "isp", [&] { return launch_isp(a, h->prec, h->stream); });
}

// isp_splitting (dycore_mod.F90:689-752).  The saved state F^n is X itself
// (never written during the call); the subcycle states ping-pong between Y
// and Z, the slow-pass state is P1.  Returns the slot of the new state.
hipError_t enqueue_isp_step(gd_handle *h, int *last) {
  const gd_config &c = h->cfg;
  const double dt = c.time_step_size;
  const double fast_dt = dt / c.subcycles, half_dt = dt * 0.5;
  const int F = GD_PASS_FAST, SL = GD_PASS_SLOW;
  hipError_t e;
#define GD_Q(x) do { if ((e = (x)) != hipSuccess) return e; } while (0)
  GD_Q(enqueue_tendency(h, SL, SX));                                  // L2(F^n)
  GD_Q(enqueue_isp_op(h, 6, SX, SX, 0.0, 0.0));                       // accum = 0
  GD_Q(enqueue_isp_op(h, 0, SX, SX, 0.0, 0.0));                       // slow = L2(F^n)
  int t1 = SX, t2 = SY;
  for (int k = 0; k < c.subcycles; ++k) {
    GD_Q(enqueue_tendency(h, F, t1));
    GD_Q(enqueue_isp_op(h, 1, t1, t2, fast_dt * 0.5, 0.0));           // P1^{n,k+1/2}
    GD_Q(enqueue_tendency(h, F, t2));
    GD_Q(enqueue_isp_op(h, 1, t1, t2, fast_dt * 0.5, 0.0));           // P2^{n,k+1/2}
    GD_Q(enqueue_tendency(h, F, t2));
    GD_Q(enqueue_isp_op(h, 2, t1, t2, fast_dt, 0.0));                 // P^{n,k+1}
    const int nt = t2 == SY ? SZ : SY;
    t1 = t2;
    t2 = nt;
  }
  GD_Q(enqueue_isp_op(h, 3, SX, SX, 0.0, 2.0 / c.subcycles));         // accum *= 2/m
  GD_Q(enqueue_tendency(h, SL, t1));                                  // L2(Q^n)
  GD_Q(enqueue_isp_op(h, 4, t1, SP1, half_dt, 0.0));                  // Q1^{n+1}
  GD_Q(enqueue_tendency(h, SL, SP1));                                 // L2(Q1^{n+1})
  GD_Q(enqueue_isp_op(h, 4, t1, SP1, half_dt, 0.0));                  // Q2^{n+1}
  GD_Q(enqueue_tendency(h, SL, SP1));                                 // L2(Q2^{n+1})
  GD_Q(enqueue_isp_op(h, 5, SX, SX, 0.0, 0.0));                       // R2^n, beta partials
#undef GD_Q
  Args a = base_args(h);
  a.upd_isp = 1;
  a.qcon = c.qcon_modified;
  a.dt = dt;
  a.s0 = SX;
  a.s_out = t1 == SY ? SZ : SY;
  *last = a.s_out;
  return run(h, "update", [&] { return launch_update(a, h->prec, h->stream); });
}

// time_integrate (dycore_mod.F90:654-669) + diag_run; leaves the new state in X.
hipError_t enqueue_step(gd_handle *h) {
  const gd_config &c = h->cfg;
  const double dt = c.time_step_size;
  int last;
  hipError_t e;
  if (c.split_scheme == GD_SPLIT_CSP2) {
    // csp2_splitting (dycore_mod.F90:671-687)
    const double fast_dt = dt / c.subcycles;
    if ((e = enqueue_call(h, 0.5 * dt, GD_PASS_SLOW, SX, SY, c.qcon_modified)) != hipSuccess) return e;
    int t1 = SY, t2 = SZ;
    for (int k = 0; k < c.subcycles; ++k) {
      if ((e = enqueue_call(h, fast_dt, GD_PASS_FAST, t1, t2, c.qcon_modified)) != hipSuccess) return e;
      std::swap(t1, t2);
    }
    if ((e = enqueue_call(h, 0.5 * dt, GD_PASS_SLOW, t1, t2, c.qcon_modified)) != hipSuccess) return e;
    last = t2;
  } else if (c.split_scheme == GD_SPLIT_ISP) {
    if ((e = enqueue_isp_step(h, &last)) != hipSuccess) return e;
  } else {
    // no split (and csp1, which time_integrate does not dispatch)
    if ((e = enqueue_call(h, dt, GD_PASS_ALL, SX, SY, c.qcon_modified)) != hipSuccess) return e;
    last = SY;
  }
  if (c.use_diffusion) {
    if ((e = enqueue_diffusion(h, last, false, dt)) != hipSuccess) return e;
    return enqueue_finish(h, 0, SX, 0, h->nact);
  }
  return enqueue_finish(h, 1, last, 0, h->nact);
}

// Steps per replay of the multi-step graph.  Consecutive replays leave a
// gap between them (about 9 us under a kernel trace of the 720x361 step);
// eight steps per replay measured 1.2% faster at 720x361, 1.6% at 360x181
// and 2.1% at 180x91 than one (mixed, same bits).
constexpr int kGraphSteps = 8;

// The captured run of k steps (1 or kGraphSteps) of the current
// active-member count (nullptr: launch eagerly).
hipGraphExec_t step_graph(gd_handle *h, int k = 1) {
  if (h->graph_failed) return nullptr;
  const int key = h->nact * 64 + k;
  auto it = h->graphs.find(key);
  if (it != h->graphs.end()) return it->second;
  // GD_NO_GRAPH=1 launches every kernel eagerly (tracing tools that do not
  // follow graph launches)
  const char *ng = std::getenv("GD_NO_GRAPH");
  if (ng && ng[0] == '1') {
    h->graph_failed = true;
    return nullptr;
  }
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) { h->graph_failed = true; return nullptr; }
  hipError_t e2 = hipSuccess;
  for (int st = 0; st < k && e2 == hipSuccess; ++st) e2 = enqueue_step(h);
  e = hipStreamEndCapture(h->stream, &g);
  if (e2 != hipSuccess || e != hipSuccess) {
    if (g) hipGraphDestroy(g);
    hipGetLastError();
    h->graph_failed = true;
    return nullptr;
  }
  hipGraphExec_t x = nullptr;
  e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
  hipGraphDestroy(g);
  if (e != hipSuccess) {
    h->graph_failed = true;
    hipGetLastError();
    return nullptr;
  }
  h->graphs[key] = x;
  return x;
}

// Staging slots, filter workgroups and the per-row coefficient table.
void build_tables(gd_handle *h, std::vector<RowCoef> &rc) {
  const Geometry &g = h->g;
  const int M = h->M, N = h->N;
  rc.assign(M, RowCoef{});
  h->slots.clear();
  for (int j = 0; j < M; ++j) {
    RowCoef &r = rc[j];
    r.fcos = g.full_cos[j];
    r.f = g.full_f[j];
    r.c = g.full_c[j];
    r.fdlon = g.full_dlon[j];
    r.fdlat = g.full_dlat[j];
    r.hcos = j <= M - 2 ? g.half_cos[j] : 0.0;
    r.hdlon = j <= M - 2 ? g.half_dlon[j] : 0.0;
    r.hdlat = j <= M - 2 ? g.half_dlat[j] : 0.0;
    const bool interior = j >= 1 && j <= M - 2;
    r.k_uadv_lon = interior ? 0.25 / r.fdlon : 0.0;
    r.k_uadv_lat = interior ? 0.25 / r.fdlat : 0.0;
    r.k_vadv_lon = j <= M - 2 ? 0.25 / r.hdlon : 0.0;
    r.k_vadv_lat = j <= M - 2 ? 0.25 / r.hdlat : 0.0;
    r.c1 = interior ? g.half_cos[j - 1] / r.fcos : 0.0;
    r.c2 = interior ? g.half_cos[j] / r.fcos : 0.0;
    r.fdlon2 = r.fdlon * r.fdlon;
    r.fdlat2 = r.fdlat * r.fdlat;
    r.hdlon2 = r.hdlon * r.hdlon;
    r.hdlat2 = r.hdlat * r.hdlat;
    r.mass_w = r.fcos * g.dlon * g.dlat;
    r.stage[0] = r.stage[1] = r.stage[2] = -1;
    r.cut[0] = r.cut[1] = r.cut[2] = 0;
  }
  auto add = [&](int comp, int row, int c) {
    StageSlot s{comp, row, kept_slots(N, c), 0};
    rc[row].stage[comp] = (int)h->slots.size();
    rc[row].cut[comp] = c;
    h->slots.push_back(s);
  };
  for (int j = 1; j <= M - 2; ++j)
    if (g.full_flag[j]) add(CDU, j, g.full_cut[j]);
  for (int j = 0; j <= M - 2; ++j)
    if (g.half_flag[j]) add(CDV, j, g.half_cut[j]);
  for (int j = 0; j < M; ++j)
    if (g.full_flag[j]) add(CDGD, j, g.full_cut[j]);
  // filter workgroups: one per staged row of each member, in (row,
  // component) order so that k_phase can give the components of a row and
  // its neighbours to one XCD.  A row's arithmetic does not depend on any
  // other row, so a member gives the same bits alone and inside an ensemble.
  std::vector<int> order(h->slots.size());
  for (size_t k = 0; k < order.size(); ++k) order[k] = (int)k;
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
    return h->slots[x].row != h->slots[y].row ? h->slots[x].row < h->slots[y].row
                                              : h->slots[x].comp < h->slots[y].comp;
  });
  h->frows.clear();
  for (int m = 0; m < h->nmem; ++m)
    for (int sl : order) {
      const StageSlot &ss = h->slots[sl];
      FiltRow f{};
      f.member = m;
      f.slot = sl;
      f.comp = ss.comp;
      f.row = ss.row;
      f.cut = (ss.kept - 2) / 2;
      f.wrow = ss.comp == CDV ? rc[ss.row].hcos : rc[ss.row].fcos;
      h->frows.push_back(f);
    }
}

// Radix passes of the complex FFT of length L: 8s first, then 4, 2, 3, 5,
// 7 and the remaining prime factors.
std::vector<int> fft_plan(int L) {
  std::vector<int> r;
  while (L % 8 == 0) { r.push_back(8); L /= 8; }
  while (L % 4 == 0) { r.push_back(4); L /= 4; }
  while (L % 2 == 0) { r.push_back(2); L /= 2; }
  for (int p = 3; L > 1; p += 2)
    while (L % p == 0) { r.push_back(p); L /= p; }
  return r;
}

std::string validate(const gd_config &c, int nmem) {
  if (nmem < 1 || nmem > 4096) return "n_members must be in [1, 4096]";
  if (!(c.time_step_size > 0)) return "time_step_size must be > 0";
  if (c.split_scheme < 0 || c.split_scheme > 3) return "unknown split_scheme";
  if ((c.split_scheme == GD_SPLIT_CSP2 || c.split_scheme == GD_SPLIT_ISP) && c.subcycles < 1)
    return "subcycles must be >= 1";
  if (c.uv_adv_scheme < 0 || c.uv_adv_scheme > 2) return "unknown uv_adv_scheme";
  if (c.precision < 0 || c.precision > 3) return "unknown precision";
  return "";
}

int check_member(gd_handle *h, int32_t member) {
  if (!h) return GD_ERR_INVALID_ARG;
  if (member < 0 || member >= h->nmem) return fail(h, GD_ERR_INVALID_ARG, "member out of range");
  return GD_OK;
}

// host rows (ld apart) <-> device rows (pitch apart)
hipError_t h2d_rows(gd_handle *h, double *dst, const double *src, int rows, int64_t ld) {
  return hipMemcpy2DAsync(dst, (size_t)h->P * 8, src, (size_t)ld * 8, (size_t)h->N * 8, rows,
                          hipMemcpyHostToDevice, h->stream);
}
hipError_t d2h_rows(gd_handle *h, double *dst, const double *src, int rows, int64_t ld) {
  return hipMemcpy2DAsync(dst, (size_t)ld * 8, src, (size_t)h->P * 8, (size_t)h->N * 8, rows,
                          hipMemcpyDeviceToHost, h->stream);
}

// Graphs captured with other launch shapes are dropped when the shapes change.
void drop_graphs(gd_handle *h) {
  hipStreamSynchronize(h->stream);
  for (auto &kv : h->graphs) hipGraphExecDestroy(kv.second);
  h->graphs.clear();
  h->graph_failed = false;
}

// Autotuner cache key: everything that decides the launch shapes' timing.
std::string tune_key(const gd_handle *h, const char *arch) {
  const gd_config &c = h->cfg;
  char b[320];
  std::snprintf(b, sizeof b,
                "arch=%s N=%d M=%d members=%d active=%d prec=%d split=%d sub=%d adv=%d diff=%d filt=%zu",
                arch, h->N, h->M, h->nmem, h->nact, h->prec, c.split_scheme, c.subcycles, c.uv_adv_scheme,
                c.use_diffusion ? c.diffusion_order : 0, h->frows.size() / (size_t)h->nmem);
  return b;
}

// Cache lines: "<key> | fused=F wpm=W ms=X"; the last line of a key wins.
bool tune_lookup(const char *path, const std::string &key, int &fused, int &wpm, double &ms) {
  FILE *f = std::fopen(path, "r");
  if (!f) return false;
  char line[512];
  bool found = false;
  while (std::fgets(line, sizeof line, f)) {
    const char *bar = std::strstr(line, " | ");
    if (!bar || std::string(line, bar - line) != key) continue;
    int fu, w;
    double x;
    if (std::sscanf(bar + 3, "fused=%d wpm=%d ms=%lf", &fu, &w, &x) == 3) {
      fused = fu;
      wpm = w;
      ms = x;
      found = true;
    }
  }
  std::fclose(f);
  return found;
}

int read_tend(gd_handle *h, const Member &mb, int comp, double *out, int rows, int64_t ld) {
  if (!out) return GD_OK;
  GD_HIP(h, d2h_rows(h, out, (const double *)mb.tc(comp), rows, ld));
  GD_HIP(h, hipStreamSynchronize(h->stream));
  return GD_OK;
}

}  // namespace

// ---------------------------------------------------------------------------

extern "C" {

const char *gd_last_error(const gd_handle *h) {
  return h ? h->err.c_str() : g_create_error.c_str();
}

int gd_create(const gd_config *cfg, int32_t n_members, gd_handle **out) {
  if (!cfg || !out) return fail(nullptr, GD_ERR_INVALID_ARG, "null argument");
  *out = nullptr;
  std::string v = validate(*cfg, n_members);
  if (!v.empty()) return fail(nullptr, GD_ERR_INVALID_ARG, v);
  if (cfg->uv_adv_scheme == GD_ADV_WENO)
    return fail(nullptr, GD_ERR_NOT_SUPPORTED, "uv_adv_scheme 'weno' is not provided on the device path");
  if (cfg->use_diffusion && cfg->diffusion_order < 2)
    return fail(nullptr, GD_ERR_INVALID_ARG, "diffusion_order must be >= 2");
  gd_handle *h = new gd_handle();
  h->cfg = *cfg;
  h->cutoff.assign(cfg->cutoff, cfg->cutoff + std::max(0, cfg->n_cutoff));
  h->cfg.cutoff = h->cutoff.data();
  h->nmem = n_members;
  h->nact = n_members;
  h->prec = cfg->precision == GD_PREC_FP64 ? 0 : 1;
  h->filt_mm = cfg->precision == GD_PREC_MIXED_MFMA ? 1 : (cfg->precision == GD_PREC_MIXED_BF16 ? 2 : 0);
  h->N = cfg->num_lon;
  h->M = cfg->num_lat;
  h->P = ((h->N + 31) / 32) * 32;
  h->device = cfg->device;
  std::string ge = build_geometry(h->N, h->M, cfg->use_zonal_tend_filter != 0, h->cutoff.data(),
                                  (int)h->cutoff.size(), h->g);
  if (!ge.empty()) {
    delete h;
    return fail(nullptr, GD_ERR_INVALID_ARG, ge);
  }
  hipError_t e = hipSetDevice(h->device);
  if (e != hipSuccess) {
    std::string m = std::string("hipSetDevice: ") + hipGetErrorString(e);
    delete h;
    return fail(nullptr, GD_ERR_HIP, m);
  }
  hipDeviceProp_t prop;
  e = hipGetDeviceProperties(&prop, h->device);
  if (e != hipSuccess) {
    delete h;
    return fail(nullptr, GD_ERR_HIP, "hipGetDeviceProperties failed");
  }
  {
    // the scratch array's address on this device, resolved once per device
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    auto it = g_scratch_ptr.find(h->device);
    if (it == g_scratch_ptr.end()) {
      void *sp = nullptr;
      const hipError_t se = hipGetSymbolAddress(&sp, HIP_SYMBOL(scratch));
      if (se == hipSuccess && sp) {
        it = g_scratch_ptr.emplace(h->device, (double *)sp).first;
      } else {
        h->scratch_err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(se);
        hipGetLastError();
      }
    }
    if (it != g_scratch_ptr.end()) h->scratch = it->second;
  }
  std::vector<RowCoef> rc;
  build_tables(h, rc);
  h->fft_half = h->N % 2 == 0;
  h->fft_L = h->fft_half ? h->N / 2 : h->N;
  h->fft_rad = fft_plan(h->fft_L);
  if (!h->frows.empty() && (filter_lds_bytes(h->N, h->fft_L) > prop.sharedMemPerBlock ||
                            h->N > 16 * kBlock || (int)h->fft_rad.size() > kMaxRad)) {
    delete h;
    return fail(nullptr, GD_ERR_NOT_SUPPORTED, "num_lon too large for the polar filter (at most 4096)");
  }
  // fused phase launches: every row with a component outside the filter
  // (and the pole rows, whose du = 0 goes through the row tiles' epilogue)
  // in at most kMaxFarRng contiguous ranges.  GD_FUSED=0 keeps k_tend +
  // k_filter, GD_FUSED=1 fuses at any launch size.
  {
    std::vector<int> far;
    for (int j = 0; j < h->M; ++j) {
      const bool pole = j == 0 || j == h->M - 1;
      const bool un = rc[j].stage[CDGD] < 0 || (j >= 1 && j <= h->M - 2 && rc[j].stage[CDU] < 0) ||
                      (j <= h->M - 2 && rc[j].stage[CDV] < 0);
      if (pole || un) far.push_back(j);
    }
    int nr = 0;
    bool fits = true;
    for (size_t k = 0; k < far.size(); ++k) {
      if (k == 0 || far[k] != far[k - 1] + 1) {
        if (nr == kMaxFarRng) { fits = false; break; }
        h->far_lo[nr] = far[k];
        h->far_cnt[nr] = 0;
        ++nr;
      }
      ++h->far_cnt[nr - 1];
    }
    h->nfar_rng = nr;
    h->nfar_rows = (int)far.size();
    const char *env = std::getenv("GD_FUSED");
    h->fused_mode = env && env[0] >= '0' && env[0] <= '2' ? env[0] - '0' : -1;
    const char *nar = std::getenv("GD_UPD_NARROW");
    h->upd_narrow = nar && (nar[0] == '0' || nar[0] == '1') ? nar[0] - '0' : -1;
    // (the MFMA filter runs as its own launch between k_tend and k_filter)
    h->fused_ok = fits && !h->frows.empty() && h->N <= 8 * kBlock && !h->filt_mm &&
                  phase_lds_bytes(h->N, h->fft_L, h->prec) <= prop.sharedMemPerBlock;
    const int nseg = (h->N + kSeg - 1) / kSeg;
    h->ncu = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
    h->wide_ok = h->fused_ok && (nseg == 2 || nseg == 3) &&
                 wphase_lds_bytes(h->N, h->fft_L, h->prec) <= prop.sharedMemPerBlock;
  }
  if (h->filt_mm && (h->N < 64 || h->N > 1440 || filter_mm_lds_bytes(h->N, h->filt_mm == 2) > prop.sharedMemPerBlock)) {
    delete h;
    return fail(nullptr, GD_ERR_NOT_SUPPORTED, "the MFMA polar filter takes 64 <= num_lon <= 1440");
  }
  configure_kernels((int)prop.sharedMemPerBlock);
  hipGetLastError();
  e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete h;
    return fail(nullptr, GD_ERR_HIP, "hipStreamCreate failed");
  }
  // ---- device memory: one arena per member ----
  const size_t fld = (size_t)h->M * h->P;
  auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
  const size_t nslots = h->slots.size();
  const size_t a_nseg = (h->N + kSeg - 1) / kSeg;
  h->fs = al(fld * 8);
  h->cs = al(fld * 8);
  h->off_part = (size_t)(NSLOT * NFLD + 5) * h->fs + 4 * h->cs;
  h->off_dpart = h->off_part + al((size_t)h->M * 3 * a_nseg * 2 * 8);
  h->off_sx = h->off_dpart + al((size_t)h->M * a_nseg * 4 * 8);
  h->off_beta = h->off_sx + al(std::max<size_t>(nslots, 1) * h->P * 8);
  h->off_sy = h->off_beta + al(8);
  h->mstride = h->off_sy + (h->filt_mm ? al(std::max<size_t>(nslots, 1) * h->P * 8) : 0);
  h->arena = (char *)dalloc(h, h->mstride * h->nmem, e);
  if (!h->arena) { gd_destroy(h); return fail(nullptr, GD_ERR_HIP, "device allocation failed"); }
  e = hipMemsetAsync(h->arena, 0, h->mstride * h->nmem, h->stream);
  h->ddiag = (double *)dalloc(h, (size_t)h->nmem * h->diag_cap * 3 * 8, e);
  h->dcoef = (RowCoef *)dalloc(h, sizeof(RowCoef) * h->M, e);
  h->dslots = (StageSlot *)dalloc(h, sizeof(StageSlot) * std::max<size_t>(nslots, 1), e);
  h->dfrows = (FiltRow *)dalloc(h, sizeof(FiltRow) * std::max<size_t>(h->frows.size(), 1), e);
  h->dtwid = (double *)dalloc(h, (size_t)tw_padded(h->N) * 2 * 8, e);
  h->dstep = (int32_t *)dalloc(h, sizeof(int32_t) * h->nmem, e);
  h->dghost = (double *)dalloc(h, (size_t)2 * h->P * 8, e);
  // filter workgroups of gd_filter_rows_noalloc: member 0's rows of one kind
  // in row order, staged at consecutive rows of the scratch array
  std::vector<FiltRow> prow_tab[2];
  for (const FiltRow &f : h->frows) {
    if (f.member != 0 || f.comp == CDU) continue;
    const int kind = f.comp == CDV ? 1 : 0;
    FiltRow g = f;
    g.slot = (int)h->prow[kind].size();
    h->prow[kind].push_back(f.row);
    prow_tab[kind].push_back(g);
  }
  for (int k = 0; k < 2; ++k)
    h->dprow[k] = (FiltRow *)dalloc(h, sizeof(FiltRow) * std::max<size_t>(prow_tab[k].size(), 1), e);
  h->mslots.clear();
  for (size_t k = 0; k < h->slots.size(); ++k)
    if (h->slots[k].comp != CDGD) h->mslots.push_back((int32_t)k);
  h->dmslots = (int32_t *)dalloc(h, sizeof(int32_t) * std::max<size_t>(h->mslots.size(), 1), e);
  std::vector<unsigned char> mm_basis;
  if (h->filt_mm) {
    mm_basis_fill(h->N, h->filt_mm == 2, mm_basis);
    h->dmm_phi = dalloc(h, mm_basis.size(), e);
    if (!h->dmm_phi) {
      gd_destroy(h);
      return fail(nullptr, GD_ERR_HIP, "device allocation failed");
    }
  }
  if (!h->ddiag || !h->dcoef || !h->dslots || !h->dfrows || !h->dtwid ||
      !h->dstep || !h->dghost || !h->dprow[0] || !h->dprow[1] || !h->dmslots) {
    gd_destroy(h);
    return fail(nullptr, GD_ERR_HIP, "device allocation failed");
  }
  // twiddle table of the filter FFT: cos and sin of 2 pi k/N for k =
  // 0..N/2, evaluated in extended precision and rounded once to FP64 (the
  // kernels take W_N^k for k > N/2 as the conjugate of W_N^(N-k))
  std::vector<double> tw((size_t)tw_padded(h->N) * 2, 0.0);
  for (int k = 0; k < tw_entries(h->N); ++k) {
    const long double ang = 2.0L * 3.141592653589793238462643383279502884L * (long double)k / (long double)h->N;
    tw[2 * k] = (double)cosl(ang);
    tw[2 * k + 1] = (double)sinl(ang);
  }
  hipMemcpyAsync(h->dcoef, rc.data(), sizeof(RowCoef) * h->M, hipMemcpyHostToDevice, h->stream);
  if (nslots) hipMemcpyAsync(h->dslots, h->slots.data(), sizeof(StageSlot) * nslots, hipMemcpyHostToDevice, h->stream);
  if (!h->frows.empty())
    hipMemcpyAsync(h->dfrows, h->frows.data(), sizeof(FiltRow) * h->frows.size(), hipMemcpyHostToDevice, h->stream);
  hipMemcpyAsync(h->dtwid, tw.data(), tw.size() * 8, hipMemcpyHostToDevice, h->stream);
  if (!h->mslots.empty())
    hipMemcpyAsync(h->dmslots, h->mslots.data(), sizeof(int32_t) * h->mslots.size(), hipMemcpyHostToDevice,
                   h->stream);
  if (h->dmm_phi) hipMemcpyAsync(h->dmm_phi, mm_basis.data(), mm_basis.size(), hipMemcpyHostToDevice, h->stream);
  for (int k = 0; k < 2; ++k)
    if (!prow_tab[k].empty())
      hipMemcpyAsync(h->dprow[k], prow_tab[k].data(), sizeof(FiltRow) * prow_tab[k].size(), hipMemcpyHostToDevice,
                     h->stream);
  std::vector<double> ghost((size_t)2 * h->P, 0.0);
  std::fill(ghost.begin(), ghost.begin() + h->P, 1.0);
  hipMemcpyAsync(h->dghost, ghost.data(), ghost.size() * 8, hipMemcpyHostToDevice, h->stream);
  hipMemsetAsync(h->dstep, 0, sizeof(int32_t) * h->nmem, h->stream);
  e = hipStreamSynchronize(h->stream);
  if (e != hipSuccess) {
    std::string m = std::string("set-up: ") + hipGetErrorString(e);
    gd_destroy(h);
    return fail(nullptr, GD_ERR_HIP, m);
  }
  *out = h;
  return GD_OK;
}

int gd_destroy(gd_handle *h) {
  if (!h) return GD_OK;
  hipSetDevice(h->device);
  if (h->stream) hipStreamSynchronize(h->stream);
  for (auto &kv : h->graphs) hipGraphExecDestroy(kv.second);
  for (void *p : h->allocs) hipFree(p);
  if (h->stream) hipStreamDestroy(h->stream);
  delete h;
  return GD_OK;
}

int gd_set_state(gd_handle *h, int32_t member, const double *u, const double *v, const double *gdp,
                 const double *ghs, int64_t ld) {
  int rc = check_member(h, member);
  if (rc) return rc;
  if (!u || !v || !gdp || !ghs || ld < h->N) return fail(h, GD_ERR_INVALID_ARG, "bad state arguments");
  const int N = h->N, M = h->M;
  for (int i = 0; i < N; ++i)
    if (u[i] != 0.0 || u[(size_t)(M - 1) * ld + i] != 0.0)
      return fail(h, GD_ERR_INVALID_ARG, "u must be zero on the pole rows");
  for (int j = 0; j < M; ++j)
    for (int i = 0; i < N; ++i)
      if (!(gdp[(size_t)j * ld + i] > 0.0)) return fail(h, GD_ERR_INVALID_ARG, "gd must be positive");
  hipSetDevice(h->device);
  const Member mb = hmember(h, member);
  GD_HIP(h, h2d_rows(h, mb.st(SX, Fu), u, M, ld));
  GD_HIP(h, h2d_rows(h, mb.st(SX, Fv), v, M - 1, ld));
  GD_HIP(h, h2d_rows(h, mb.st(SX, FGD), gdp, M, ld));
  GD_HIP(h, h2d_rows(h, mb.ghs(), ghs, M, ld));
  // iap_transform(state(old)) + diag (dycore_mod.F90:123-125)
  GD_HIP(h, enqueue_finish(h, 0, SX, member, 1));
  GD_HIP(h, hipStreamSynchronize(h->stream));
  return GD_OK;
}

int gd_diag(gd_handle *h, double *mass, double *energy) {
  if (!h) return GD_ERR_INVALID_ARG;
  hipSetDevice(h->device);
  GD_HIP(h, hipMemsetAsync(h->dstep, 0, sizeof(int32_t) * h->nmem, h->stream));
  GD_HIP(h, enqueue_finish(h, 2, SX, 0, h->nmem));
  std::vector<double> buf((size_t)h->nmem * h->diag_cap * 3);
  GD_HIP(h, hipMemcpyAsync(buf.data(), h->ddiag, buf.size() * 8, hipMemcpyDeviceToHost, h->stream));
  GD_HIP(h, hipStreamSynchronize(h->stream));
  for (int m = 0; m < h->nmem; ++m) {
    if (mass) mass[m] = buf[(size_t)m * h->diag_cap * 3 + 0];
    if (energy) energy[m] = buf[(size_t)m * h->diag_cap * 3 + 1];
  }
  return GD_OK;
}

int gd_step(gd_handle *h, int32_t nsteps, double *mass, double *energy, double *beta, int32_t *nan_step) {
  if (!h || nsteps < 0) return GD_ERR_INVALID_ARG;
  hipSetDevice(h->device);
  if (nan_step) *nan_step = 0;
  hipGraphExec_t graph = h->prof ? nullptr : step_graph(h);
  hipGraphExec_t graph8 = graph && nsteps >= kGraphSteps ? step_graph(h, kGraphSteps) : nullptr;
  std::vector<double> buf((size_t)h->nmem * h->diag_cap * 3);
  int first_nan = 0;
  for (int done = 0; done < nsteps;) {
    const int n = std::min(h->diag_cap, nsteps - done);
    GD_HIP(h, hipMemsetAsync(h->dstep, 0, sizeof(int32_t) * h->nmem, h->stream));
    for (int k = 0; k < n;) {
      if (graph8 && n - k >= kGraphSteps) {
        GD_HIP(h, hipGraphLaunch(graph8, h->stream));
        k += kGraphSteps;
      } else if (graph) {
        GD_HIP(h, hipGraphLaunch(graph, h->stream));
        ++k;
      } else {
        GD_HIP(h, enqueue_step(h));
        ++k;
      }
    }
    GD_HIP(h, hipMemcpyAsync(buf.data(), h->ddiag, buf.size() * 8, hipMemcpyDeviceToHost, h->stream));
    GD_HIP(h, hipStreamSynchronize(h->stream));
    for (int k = 0; k < n; ++k)
      for (int m = 0; m < h->nmem; ++m) {
        const double *d = &buf[((size_t)m * h->diag_cap + k) * 3];
        const size_t o = (size_t)(done + k) * h->nmem + m;
        if (mass) mass[o] = d[0];
        if (energy) energy[o] = d[1];
        if (beta) beta[o] = d[2];
        if (m < h->nact && !first_nan && (std::isnan(d[0]) || std::isnan(d[1]))) first_nan = done + k + 1;
      }
    done += n;
  }
  if (first_nan) {
    if (nan_step) *nan_step = first_nan;
    return fail(h, GD_ERR_NAN, "Total mass or energy is NaN at step " + std::to_string(first_nan));
  }
  return GD_OK;
}

int gd_get_state(gd_handle *h, int32_t member, double *u, double *v, double *gdp, int64_t ld) {
  int rc = check_member(h, member);
  if (rc) return rc;
  if (ld < h->N) return fail(h, GD_ERR_INVALID_ARG, "ld < num_lon");
  hipSetDevice(h->device);
  const Member mb = hmember(h, member);
  if (u) GD_HIP(h, d2h_rows(h, u, mb.st(SX, Fu), h->M, ld));
  if (v) GD_HIP(h, d2h_rows(h, v, mb.st(SX, Fv), h->M - 1, ld));
  if (gdp) GD_HIP(h, d2h_rows(h, gdp, mb.st(SX, FGD), h->M, ld));
  GD_HIP(h, hipStreamSynchronize(h->stream));
  return GD_OK;
}

int gd_get_diag_fields(gd_handle *h, int32_t member, double *div, double *vor, int64_t ld) {
  int rc = check_member(h, member);
  if (rc) return rc;
  if (ld < h->N) return fail(h, GD_ERR_INVALID_ARG, "ld < num_lon");
  hipSetDevice(h->device);
  Args a = base_args(h);
  a.m0 = member;
  a.ntiles = h->M;
  GD_HIP(h, launch_divvor(a, h->stream));
  const Member mb = hmember(h, member);
  if (div) GD_HIP(h, d2h_rows(h, div, mb.div(), h->M, ld));
  if (vor) GD_HIP(h, d2h_rows(h, vor, mb.vor(), h->M - 1, ld));
  GD_HIP(h, hipStreamSynchronize(h->stream));
  return GD_OK;
}

int gd_space_operators(gd_handle *h, int32_t member, int32_t pass, double *du, double *dv, double *dgd,
                       int64_t ld) {
  int rc = check_member(h, member);
  if (rc) return rc;
  if (pass < 0 || pass > 2 || ld < h->N) return fail(h, GD_ERR_INVALID_ARG, "bad pass or ld");
  // the launch evaluates members 0..nact-1 only
  if (member >= h->nact) return fail(h, GD_ERR_INVALID_ARG, "member is not active");
  hipSetDevice(h->device);
  Args a = base_args(h);
  a.pass = pass;
  a.phase = PH_T;
  a.s_in = SX;
  a.s0 = SX;
  GD_HIP(h, enqueue_phase(h, a, pass, true));
  const Member mb = hmember(h, member);
  if ((rc = read_tend(h, mb, CDU, du, h->M, ld))) return rc;
  if ((rc = read_tend(h, mb, CDV, dv, h->M - 1, ld))) return rc;
  if ((rc = read_tend(h, mb, CDGD, dgd, h->M, ld))) return rc;
  return GD_OK;
}

int gd_check_antisymmetry(gd_handle *h, int32_t member, double *terms, double *magnitudes, double *sums) {
  int rc = check_member(h, member);
  if (rc) return rc;
  hipSetDevice(h->device);
  Args a = base_args(h);
  a.m0 = member;
  a.s_in = SX;
  GD_HIP(h, launch_terms(a, h->stream));
  double out[20];
  GD_HIP(h, hipMemcpyAsync(out, hmember(h, member).tb(CDV), sizeof(out), hipMemcpyDeviceToHost, h->stream));
  GD_HIP(h, hipStreamSynchronize(h->stream));
  for (int k = 0; k < 10; ++k) {
    if (terms) terms[k] = out[k];
    if (magnitudes) magnitudes[k] = out[10 + k];
  }
  if (sums) {
    // the four sums check_antisymmetry prints (dycore_mod.F90:846-850)
    sums[0] = out[0] + out[4] + out[1] + out[5];
    sums[1] = out[3] + out[8];
    sums[2] = out[2] - out[6];
    sums[3] = out[7] + out[9];
  }
  return GD_OK;
}

int gd_predict_correct(gd_handle *h, double dt, int32_t pass, int32_t qcon, double *beta) {
  if (!h || pass < 0 || pass > 2) return GD_ERR_INVALID_ARG;
  hipSetDevice(h->device);
  GD_HIP(h, enqueue_call(h, dt, pass, SX, SY, qcon));
  // only the active members were stepped: the others keep their state
  GD_HIP(h, enqueue_finish(h, 1, SY, 0, h->nact));
  if (beta)
    for (int m = 0; m < h->nact; ++m)
      GD_HIP(h, hipMemcpyAsync(beta + m, hmember(h, m).beta(), 8, hipMemcpyDeviceToHost, h->stream));
  GD_HIP(h, hipStreamSynchronize(h->stream));
  return GD_OK;
}

int gd_ordinary_diffusion(gd_handle *h, double dt) {
  if (!h) return GD_ERR_INVALID_ARG;
  if (h->cfg.diffusion_order < 2) return fail(h, GD_ERR_INVALID_ARG, "diffusion_order must be >= 2");
  hipSetDevice(h->device);
  const size_t fb = (size_t)h->M * h->P * 8;
  for (int m = 0; m < h->nact; ++m)
    for (int f = 0; f < NFLD; ++f)
      GD_HIP(h, hipMemcpyAsync(hmember(h, m).st(SY, f), hmember(h, m).st(SX, f), fb, hipMemcpyDeviceToDevice, h->stream));
  GD_HIP(h, enqueue_diffusion(h, SY, true, dt));
  GD_HIP(h, enqueue_finish(h, 0, SX, 0, h->nact));
  GD_HIP(h, hipStreamSynchronize(h->stream));
  return GD_OK;
}

int gd_filter_rows(gd_handle *h, int32_t half, double *x, int64_t ld) {
  if (!h || !x || ld < h->N) return GD_ERR_INVALID_ARG;
  hipSetDevice(h->device);
  const int comp = half ? CDV : CDGD;
  const Member mb = hmember(h, 0);
  const size_t rowb = (size_t)h->P * 8;
  for (size_t s = 0; s < h->slots.size(); ++s) {
    if (h->slots[s].comp != comp) continue;
    GD_HIP(h, hipMemcpy(mb.sx() + s * h->P, x + (size_t)h->slots[s].row * ld, (size_t)h->N * 8,
                        hipMemcpyHostToDevice));
  }
  Args a = base_args(h);
  a.phase = PH_P;
  GD_HIP(h, enqueue_filter(h, a));
  GD_HIP(h, hipStreamSynchronize(h->stream));
  for (size_t s = 0; s < h->slots.size(); ++s) {
    if (h->slots[s].comp != comp) continue;
    GD_HIP(h, hipMemcpy(x + (size_t)h->slots[s].row * ld, mb.sx() + s * h->P, (size_t)h->N * 8,
                        hipMemcpyDeviceToHost));
  }
  (void)rowb;
  return GD_OK;
}

int gd_filter_rows_noalloc(gd_handle *h, int32_t half, double *x, int64_t ld, void *stream) {
  if (!h) return GD_ERR_INVALID_ARG;
  if (!x || ld < h->N || (half != 0 && half != 1)) return fail(h, GD_ERR_INVALID_ARG, "bad filter_rows arguments");
  if (!h->scratch) return fail(h, GD_ERR_HIP, "no scratch array on this device: " + h->scratch_err);
  const std::vector<int> &rows = h->prow[half];
  const size_t need = rows.size() * (size_t)h->P * sizeof(double);
  if (need > kScratchBytes)
    return fail(h, GD_ERR_INVALID_ARG,
                "gd_filter_rows_noalloc: " + std::to_string(rows.size()) + " filtered rows of " +
                    std::to_string(h->P) + " doubles need " + std::to_string(need) + " bytes of scratch, more than " +
                    std::to_string(kScratchBytes));
  if (rows.empty()) return GD_OK;
  hipSetDevice(h->device);
  hipStream_t s = stream ? (hipStream_t)stream : h->stream;
  const size_t P = h->P, N = h->N;
  // the flagged rows, one copy per run of consecutive rows, into consecutive
  // scratch rows and back
  auto copy_runs = [&](bool in) -> hipError_t {
    for (size_t k = 0; k < rows.size();) {
      size_t n = 1;
      while (k + n < rows.size() && rows[k + n] == rows[k] + (int)n) ++n;
      double *sp = h->scratch + k * P, *xp = x + (size_t)rows[k] * ld;
      hipError_t e = in ? hipMemcpy2DAsync(sp, P * 8, xp, (size_t)ld * 8, N * 8, n, hipMemcpyDeviceToDevice, s)
                        : hipMemcpy2DAsync(xp, (size_t)ld * 8, sp, P * 8, N * 8, n, hipMemcpyDeviceToDevice, s);
      if (e != hipSuccess) return e;
      k += n;
    }
    return hipSuccess;
  };
  GD_HIP(h, copy_runs(true));
  Args a = base_args(h);
  a.phase = PH_P;
  a.sx_ext = h->scratch;
  a.frows = h->dprow[half];
  a.ntiles = (int)rows.size();
  GD_HIP(h, launch_filter(a, h->prec, s));
  GD_HIP(h, copy_runs(false));
  if (!stream) GD_HIP(h, hipStreamSynchronize(s));
  return GD_OK;
}

int gd_filter_table(const gd_handle *h, int32_t *full_flag, int32_t *half_flag, int32_t *full_cut,
                    int32_t *half_cut) {
  if (!h) return GD_ERR_INVALID_ARG;
  for (int j = 0; j < h->M; ++j) {
    if (full_flag) full_flag[j] = h->g.full_flag[j];
    if (full_cut) full_cut[j] = h->g.full_cut[j];
  }
  for (int j = 0; j < h->M - 1; ++j) {
    if (half_flag) half_flag[j] = h->g.half_flag[j];
    if (half_cut) half_cut[j] = h->g.half_cut[j];
  }
  return GD_OK;
}

int gd_profile_kernels(gd_handle *h, int32_t nsteps, int32_t max_families, int32_t *n_families,
                       int64_t *launches, double *ms, char *names, int32_t name_cap) {
  if (!h || nsteps < 1) return GD_ERR_INVALID_ARG;
  hipSetDevice(h->device);
  std::vector<ProfEntry> entries;
  h->prof = &entries;
  int rc = gd_step(h, nsteps, nullptr, nullptr, nullptr, nullptr);
  h->prof = nullptr;
  hipStreamSynchronize(h->stream);
  std::vector<std::string> order;
  std::map<std::string, std::pair<int64_t, double>> acc;
  for (auto &pe : entries) {
    float t = 0;
    hipEventElapsedTime(&t, pe.e0, pe.e1);
    if (!acc.count(pe.family)) order.push_back(pe.family);
    acc[pe.family].first += 1;
    acc[pe.family].second += t;
    hipEventDestroy(pe.e0);
    hipEventDestroy(pe.e1);
  }
  if (rc && rc != GD_ERR_NAN) return rc;
  int nf = std::min<int>((int)order.size(), max_families);
  std::string joined;
  for (int k = 0; k < nf; ++k) {
    if (launches) launches[k] = acc[order[k]].first;
    if (ms) ms[k] = acc[order[k]].second;
    if (k) joined += ",";
    joined += order[k];
  }
  if (n_families) *n_families = nf;
  if (names && name_cap > 0) {
    std::snprintf(names, name_cap, "%s", joined.c_str());
  }
  return GD_OK;
}

int64_t gd_device_bytes(const gd_handle *h) { return h ? h->bytes : 0; }

int gd_set_active_members(gd_handle *h, int32_t n_active) {
  if (!h) return GD_ERR_INVALID_ARG;
  if (n_active < 1 || n_active > h->nmem) return fail(h, GD_ERR_INVALID_ARG, "n_active out of range");
  h->nact = n_active;
  return GD_OK;
}

int gd_copy_member(gd_handle *h, int32_t dst, int32_t src) {
  int rc = check_member(h, dst);
  if (rc) return rc;
  if ((rc = check_member(h, src))) return rc;
  if (dst == src) return GD_OK;
  hipSetDevice(h->device);
  GD_HIP(h, hipMemcpyAsync(h->arena + (size_t)dst * h->mstride, h->arena + (size_t)src * h->mstride,
                           h->mstride, hipMemcpyDeviceToDevice, h->stream));
  GD_HIP(h, hipStreamSynchronize(h->stream));
  return GD_OK;
}

int gd_fused_phases(const gd_handle *h) { return h ? phase_form(h) : 0; }

int gd_autotune(gd_handle *h, int32_t steps, const char *cache_path, int32_t *fused_out, int32_t *wpm_out,
                double *ms_out) {
  if (!h) return GD_ERR_INVALID_ARG;
  if (steps < 1 || steps > h->diag_cap) return fail(h, GD_ERR_INVALID_ARG, "autotune steps must be in [1, 512]");
  hipSetDevice(h->device);
  hipDeviceProp_t prop;
  GD_HIP(h, hipGetDeviceProperties(&prop, h->device));
  const std::string key = tune_key(h, prop.gcnArchName);
  const int nact = h->nact;
  int bf = -1, bw = -1;
  double bms = 0.0;
  if (cache_path && !cache_path[0]) cache_path = nullptr;  // "" : no cache
  bool cached = cache_path && tune_lookup(cache_path, key, bf, bw, bms);
  if (cached && (bw < 1 || bw > h->M || bf < 0 || bf > 2 || (bf && !h->fused_ok) || (bf == 2 && !h->wide_ok)))
    cached = false;
  if (!cached) {
    // fused or separate phase launches (both give the same bits; a forced
    // GD_FUSED keeps its form) ...
    std::vector<int> fopts;
    if (h->fused_mode >= 0 || !h->fused_ok)
      fopts = {phase_form(h)};
    else if (h->wide_ok)
      fopts = {0, 1, 2};
    else
      fopts = {0, 1};
    // ... and k_update rows per workgroup (each workgroup forms its member's
    // beta from all of the member's partials in one order: the same bits)
    std::vector<int> wopts;
    for (int rpw : {1, 2, 3, 4, 6, 8, 12, 16, 24, 32, 48, 64, 96, 128, 192, 256}) {
      const int w = (h->M + rpw - 1) / rpw;
      if (w >= 1 && (int64_t)w * nact <= 65536 && std::find(wopts.begin(), wopts.end(), w) == wopts.end())
        wopts.push_back(w);
    }
    const int wdef = base_args(h).upd_wpm;
    if (std::find(wopts.begin(), wopts.end(), wdef) == wopts.end()) wopts.push_back(wdef);
    // every candidate steps from the same saved state, which is put back at
    // the end (a set-up call: the copy is allocated here)
    const size_t nb = (size_t)nact * h->mstride;
    void *save = nullptr;
    GD_HIP(h, hipMalloc(&save, nb));
    auto restore = [&]() { return hipMemcpyAsync(h->arena, save, nb, hipMemcpyDeviceToDevice, h->stream); };
    hipError_t e = hipMemcpyAsync(save, h->arena, nb, hipMemcpyDeviceToDevice, h->stream);
    int rc = e == hipSuccess ? GD_OK : hip_fail(h, e, "autotune save");
    for (int f : fopts) {
      for (int w : wopts) {
        if (rc) break;
        h->tuned_fused[nact] = f;
        h->tuned_wpm[nact] = w;
        drop_graphs(h);
        if ((e = restore()) != hipSuccess) { rc = hip_fail(h, e, "autotune restore"); break; }
        if ((rc = gd_step(h, kGraphSteps + 1, nullptr, nullptr, nullptr, nullptr))) break;  // capture + warm
        if ((e = restore()) != hipSuccess) { rc = hip_fail(h, e, "autotune restore"); break; }
        if ((e = hipStreamSynchronize(h->stream)) != hipSuccess) { rc = hip_fail(h, e, "autotune"); break; }
        const auto t0 = std::chrono::steady_clock::now();
        if ((rc = gd_step(h, steps, nullptr, nullptr, nullptr, nullptr))) break;
        const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count() / steps;
        if (bw < 0 || ms < bms) {
          bf = f;
          bw = w;
          bms = ms;
        }
      }
    }
    if ((e = restore()) != hipSuccess && !rc) rc = hip_fail(h, e, "autotune restore");
    hipStreamSynchronize(h->stream);
    hipFree(save);
    if (rc) {
      h->tuned_fused.erase(nact);
      h->tuned_wpm.erase(nact);
      drop_graphs(h);
      return rc;
    }
  }
  h->tuned_fused[nact] = bf;
  h->tuned_wpm[nact] = bw;
  drop_graphs(h);
  if (!cached && cache_path) {
    if (FILE *f = std::fopen(cache_path, "a")) {
      std::fprintf(f, "%s | fused=%d wpm=%d ms=%.17g\n", key.c_str(), bf, bw, bms);
      std::fclose(f);
    }
  }
  if (fused_out) *fused_out = bf;
  if (wpm_out) *wpm_out = bw;
  if (ms_out) *ms_out = bms;
  return GD_OK;
}

int gd_time_kernel(gd_handle *h, int32_t kernel, int32_t reps, double *us) {
  if (!h || reps < 1 || !us || kernel < GD_TIME_TEND || kernel > GD_TIME_FILTER_MM) return GD_ERR_INVALID_ARG;
  if (kernel == GD_TIME_FILTER && h->frows.empty()) return fail(h, GD_ERR_INVALID_ARG, "no filtered rows");
  if (kernel == GD_TIME_FILTER_MM && (!h->filt_mm || h->mslots.empty()))
    return fail(h, GD_ERR_NOT_SUPPORTED, "the handle does not filter on the matrix cores");
  if (kernel == GD_TIME_PHASE && !use_fused(h))
    return fail(h, GD_ERR_NOT_SUPPORTED, "the handle does not use fused phase launches");
  hipSetDevice(h->device);
  // arguments of the second predictor of a fast-pass call (predict_correct,
  // dycore_mod.F90:774-777): S2 = S0 + dt/2 L(S1)
  const gd_config &c = h->cfg;
  const double dt = c.split_scheme == GD_SPLIT_CSP2 ? c.time_step_size / c.subcycles : c.time_step_size;
  const int pass = c.split_scheme == GD_SPLIT_CSP2 ? GD_PASS_FAST : GD_PASS_ALL;
  Args a = base_args(h);
  a.pass = pass;
  a.dt = dt;
  a.hdt = 0.5 * dt;
  a.ntiles = h->nmem * h->M * a.nseg;
  a.phase = PH_B;
  a.s_in = SP1;
  a.s0 = SY;
  a.s_out = SP2;
  if (kernel == GD_TIME_UPDATE) {
    a.phase = PH_C;
    a.s_in = SP2;
    a.s_out = SZ;
  }
  hipEvent_t e0, e1;
  GD_HIP(h, hipEventCreate(&e0));
  GD_HIP(h, hipEventCreate(&e1));
  auto one = [&]() -> hipError_t {
    if (kernel == GD_TIME_TEND) return launch_tend(a, h->prec, pass, false, h->stream);
    if (kernel == GD_TIME_FILTER) {
      Args f = a;
      f.ntiles = (int)h->frows.size();
      return launch_filter(f, h->prec, h->stream);
    }
    if (kernel == GD_TIME_PHASE) {
      Args f = a;
      set_fused(h, f, h->nmem);
      if (phase_form(h) == 2) return launch_wphase(f, h->prec, pass, false, h->stream);
      return launch_phase(f, h->prec, pass, false, h->stream);
    }
    if (kernel == GD_TIME_FILTER_MM) {
      Args f = a;
      // the kernel's member blocks follow a.nmem (the active members)
      f.ntiles = (int)h->mslots.size() * ((f.nmem + 15) / 16);
      return launch_filter_mm(f, h->filt_mm == 2, h->stream);
    }
    return launch_update(a, h->prec, h->stream);
  };
  GD_HIP(h, one());  // warm
  GD_HIP(h, hipEventRecord(e0, h->stream));
  for (int k = 0; k < reps; ++k) GD_HIP(h, one());
  GD_HIP(h, hipEventRecord(e1, h->stream));
  GD_HIP(h, hipEventSynchronize(e1));
  float ms = 0;
  GD_HIP(h, hipEventElapsedTime(&ms, e0, e1));
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  *us = (double)ms * 1e3 / reps;
  return GD_OK;
}

}  // extern "C"
