// gd_kernels.hip -- CDNA4 (gfx950) kernels of the barotropic time step.
//
// Row kernels: one workgroup (256 threads, 4 wave64) per (member, latitude
// row, 256-longitude segment); tiles are remapped so each XCD owns a
// contiguous band of rows and the +-1 row stencil reuse stays inside that
// XCD's L2.  Derived fields (G = sqrt(gd), u, v) of the rows a tile needs
// are built once in LDS.  Rows of the polar filter are staged to HBM and
// finished by the filter kernel (gd_filter.hip).
//
// FP64 mode reproduces the reference's expressions operation by operation
// (the library is built with -ffp-contract=off); mixed mode (CT = float)
// evaluates the momentum stencils in FP32 and keeps the state update, mass
// flux divergence, pole sums, geopotential differences and all inner
// products in FP64.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "gd_device.hpp"
#include "gd_geom.hpp"
#include "gd_common.hpp"
#include "gd_launch.hpp"
#include "gd_tend.hpp"
#include "gd_diffuse.hpp"

namespace gd {

// ---------------------------------------------------------------------------
// tendency (space_operators, dycore_mod.F90:184-598) + predictor epilogue
// ---------------------------------------------------------------------------

// One tile per (member, row, segment), remapped so each XCD owns a
// contiguous band of rows (tend_tile, gd_tend.hpp).
template <typename CT, int PASS, bool IN_FULL>
__global__ void __launch_bounds__(kBlock) k_tend(const Args a) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ double red[4 * 6];
  const int t = xcd_tile(a.ntiles);
  if (t < 0) return;
  tend_tile<CT, PASS, IN_FULL, false>(a, t, smem, red);
}

// ---------------------------------------------------------------------------
// beta + corrector update (predict_correct, dycore_mod.F90:784-790)
// ---------------------------------------------------------------------------

// upd_wpm workgroups per member.  Each forms the member's beta from the
// phase-C partials (same loads, same order in every workgroup, so the same
// bits), then applies S_out = S0 + dt beta Tc to its band of rows.  U:
// points per thread per batch of the update, PU: partials per thread per
// load round.  A thread adds its partials k = tid, tid + kBlock, ... in that
// order whatever PU is, so every (U, PU) gives the same bits; U = PU = 8
// holds two 720-point rows in one batch (one round trip for a single run),
// U = PU = 4 about half the registers (more workgroups resident for
// batches of members).
template <typename CT, int U, int PU>
__global__ void __launch_bounds__(kBlock) k_update(const Args a) {
  const int wpm = a.upd_wpm;
  const int m = a.m0 + blockIdx.x / wpm, w = blockIdx.x - (blockIdx.x / wpm) * wpm;
  const Member mb = member_at(a, m);
  __shared__ double red[4 * 6];
  const int tid = threadIdx.x;
  const int N = a.N, M = a.M, P = a.pitch, nseg = a.nseg;
  const int rpw = (M + wpm - 1) / wpm;
  const int r0 = w * rpw, r1 = min(M, r0 + rpw);
  const int cnt = (r1 - r0) * N;
  const double *s0g = mb.st(a.s0, FGD), *s0u = mb.st(a.s0, FU), *s0v = mb.st(a.s0, FV);
  // the corrector tendency: Tc, or for isp_splitting the combined R held in
  // the FP64 accumulator (slot SP2 u, v fields and the Tb dgd field)
  const bool isp = a.upd_isp != 0;
  const double *tcg = isp ? (const double *)mb.tb(CDGD) : (const double *)mb.tc(CDGD);
  const double *tcu = mb.tc(CDU), *tcv = mb.tc(CDV);
  const double *rcu = mb.st(SP2, Fu), *rcv = mb.st(SP2, Fv);
  double *og = mb.st(a.s_out, FGD), *ou = mb.st(a.s_out, FU), *ov = mb.st(a.s_out, FV);
  size_t off[U];
  bool on[U], hv[U];
  double g0[U], u0[U], v0[U], tg[U], tu[U], tv[U];
  // Every lane issues every load (points off the band read offset 0, inside
  // every field), so the number of loads in flight is the same on every
  // path and the waits below can let the later ones stay in flight.
  const double *tu_p = isp ? rcu : tcu, *tv_p = isp ? rcv : tcv;
  auto load = [&](int base) {
#pragma unroll
    for (int k = 0; k < U; ++k) {
      const int e = base + k * kBlock + tid;
      on[k] = e < cnt;
      const int row = r0 + (on[k] ? e / N : 0), i = on[k] ? e - (e / N) * N : 0;
      hv[k] = on[k] && row <= M - 2;
      off[k] = (size_t)row * P + i;
      const size_t ou_ = on[k] ? off[k] : 0, ov_ = hv[k] ? off[k] : 0;
      const double xg = s0g[ou_], xu = s0u[ou_], xtg = tcg[ou_], xtu = tu_p[ou_];
      const double xv = s0v[ov_], xtv = tv_p[ov_];
      g0[k] = on[k] ? xg : 0.0;
      u0[k] = on[k] ? xu : 0.0;
      tg[k] = on[k] ? xtg : 0.0;
      tu[k] = on[k] ? xtu : 0.0;
      v0[k] = hv[k] ? xv : 0.0;
      tv[k] = hv[k] ? xtv : 0.0;
    }
  };
  // beta from the phase-C partials: inner_product(tend(old), tend(new)) and
  // (tend(new), tend(new)) (types_mod.F90:347-371) summed du, then dv, then dgd
  const int per_comp = M * nseg;
  double q[6] = {0, 0, 0, 0, 0, 0};
  auto pload = [&](int k0, double(&x)[PU][6]) {
#pragma unroll
    for (int u = 0; u < PU; ++u) {
      const int k = k0 + u * kBlock + tid;
      const bool in = k < per_comp;
      const int kk = in ? k : 0;  // unconditional loads, as in load()
      const int row = kk / nseg, sg = kk - (kk / nseg) * nseg;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const double2 v = *(const double2 *)(mb.part() + ((size_t)(row * 3 + c) * nseg + sg) * 2);
        x[u][2 * c] = in ? v.x : 0.0;
        x[u][2 * c + 1] = in ? v.y : 0.0;
      }
    }
  };
  auto pacc = [&](const double(&x)[PU][6]) {
#pragma unroll
    for (int u = 0; u < PU; ++u)
#pragma unroll
      for (int c = 0; c < 6; ++c) q[c] += x[u][c];
  };
  {
    // the first partials, then the band's first batch, which does not depend
    // on beta: memory returns in issue order, so the reduction waits only
    // for the partials and overlaps the batch's loads
    double x[PU][6];
    pload(0, x);
    load(0);
    pacc(x);
  }
  for (int k0 = PU * kBlock; k0 < per_comp; k0 += PU * kBlock) {
    double x[PU][6];
    pload(k0, x);
    pacc(x);
  }
  block_sum_n<6>(q, red);
  // du, then dv, then dgd
  const double ip1 = (q[0] + q[2]) + q[4];
  const double ip2 = (q[1] + q[3]) + q[5];
  double beta = (a.qcon && ip1 != 0.0 && ip2 != 0.0) ? ip1 / ip2 : 1.0;
  // isp_splitting: beta*4/dt, applied as (dt/2)*beta (dycore_mod.F90:745-750)
  if (isp) beta = beta * 4.0 / a.dt;
  if (w == 0 && tid == 0) mb.beta()[0] = beta;
  const double dtb = isp ? (a.dt * 0.5) * beta : a.dt * beta;
  for (int base = 0; base < cnt; base += U * kBlock) {
    if (base) load(base);
#pragma unroll
    for (int k = 0; k < U; ++k) {
      if (!on[k]) continue;
      og[off[k]] = g0[k] + dtb * tg[k];
      ou[off[k]] = u0[k] + dtb * tu[k];
      if (hv[k]) ov[off[k]] = v0[k] + dtb * tv[k];
    }
  }
}

// ---------------------------------------------------------------------------
// isp_splitting tendency algebra (dycore_mod.F90:689-752), pointwise
// ---------------------------------------------------------------------------

// The slow tendency L2(F^n) lives in slot SP2 (dgd, du, dv in its gd, U, V
// fields) and the accumulated fast tendency in SP2's u, v fields (du, dv)
// and the Tb dgd field, all FP64.  T is the Tc tendency of the last
// space_operators launch.  Modes:
//   ISP_SLOW_SAVE   slow <- T
//   ISP_FAST        T <- T + slow; S_out <- S_in + h T
//   ISP_FAST_ACC    acc <- acc + T; T <- T + slow; S_out <- S_in + h T
//   ISP_ACC_SCALE   acc <- (2/subcycles) acc
//   ISP_SLOW        T <- T - slow; S_out <- S_in + h T
//   ISP_FINAL       acc <- (T + slow) + acc; beta partials <acc, F^n> and
//                   <acc, acc> against the saved state X
//   ISP_ACC_ZERO    acc <- 0
// (add_data(a, b) is b <- b + a, types_mod.F90:284-304; update_state,
// dycore_mod.F90:600-652, with u, v derived where they are read).
enum IspMode { ISP_SLOW_SAVE = 0, ISP_FAST = 1, ISP_FAST_ACC = 2, ISP_ACC_SCALE = 3, ISP_SLOW = 4, ISP_FINAL = 5,
               ISP_ACC_ZERO = 6 };

template <typename CT>
__global__ void __launch_bounds__(kBlock) k_isp(const Args a) {
  __shared__ double red[4 * 6];
  const int t = xcd_tile(a.ntiles);
  if (t < 0) return;
  const int N = a.N, M = a.M, P = a.pitch, nseg = a.nseg, tid = threadIdx.x;
  const int mloc = t / (M * nseg), rem = t - mloc * (M * nseg);
  const int j = rem / nseg, g = rem - (rem / nseg) * nseg;
  const Member mb = member_at(a, a.m0 + mloc);
  const int i0 = g * kSeg, W = min(N - i0, kSeg);
  const bool hv = j <= M - 2;
  const int mode = a.isp_mode;
  const size_t o = (size_t)j * P + i0 + tid;
  double *sl_g = mb.st(SP2, FGD), *sl_u = mb.st(SP2, FU), *sl_v = mb.st(SP2, FV);
  double *ac_g = (double *)mb.tb(CDGD), *ac_u = mb.st(SP2, Fu), *ac_v = mb.st(SP2, Fv);
  double p[6] = {0, 0, 0, 0, 0, 0};
  if (tid < W) {
    const double tg = ((const double *)mb.tc(CDGD))[o];
    const double tu = mb.tc(CDU)[o];
    const double tv = hv ? mb.tc(CDV)[o] : 0.0;
    if (mode == ISP_SLOW_SAVE) {
      sl_g[o] = tg;
      sl_u[o] = tu;
      if (hv) sl_v[o] = tv;
    } else if (mode == ISP_ACC_ZERO) {
      ac_g[o] = 0.0;
      ac_u[o] = 0.0;
      if (hv) ac_v[o] = 0.0;
    } else if (mode == ISP_ACC_SCALE) {
      ac_g[o] = a.isp_scale * ac_g[o];
      ac_u[o] = a.isp_scale * ac_u[o];
      if (hv) ac_v[o] = a.isp_scale * ac_v[o];
    } else if (mode == ISP_FINAL) {
      const double rg = (tg + sl_g[o]) + ac_g[o];
      const double ru = (tu + sl_u[o]) + ac_u[o];
      const double rv = hv ? (tv + sl_v[o]) + ac_v[o] : 0.0;
      ac_g[o] = rg;
      ac_u[o] = ru;
      if (hv) ac_v[o] = rv;
      // inner_product(tend, state) and (tend, tend) (types_mod.F90:347-397)
      const RowCoef cj = a.coef[j];
      if (j >= 1 && j <= M - 2) {
        p[0] = ru * mb.st(SX, FU)[o] * cj.fcos;
        p[1] = ru * ru * cj.fcos;
      }
      if (hv) {
        p[2] = rv * mb.st(SX, FV)[o] * cj.hcos;
        p[3] = rv * rv * cj.hcos;
      }
      p[4] = rg * mb.st(SX, FGD)[o] * cj.fcos;
      p[5] = rg * rg * cj.fcos;
    } else {
      double Tg, Tu, Tv;
      if (mode == ISP_SLOW) {
        Tg = tg - sl_g[o];
        Tu = tu - sl_u[o];
        Tv = hv ? tv - sl_v[o] : 0.0;
      } else {
        if (mode == ISP_FAST_ACC) {
          ac_g[o] = ac_g[o] + tg;
          ac_u[o] = ac_u[o] + tu;
          if (hv) ac_v[o] = ac_v[o] + tv;
        }
        Tg = tg + sl_g[o];
        Tu = tu + sl_u[o];
        Tv = hv ? tv + sl_v[o] : 0.0;
      }
      const double h = a.hdt;
      mb.st(a.s_out, FGD)[o] = mb.st(a.s_in, FGD)[o] + h * Tg;
      mb.st(a.s_out, FU)[o] = mb.st(a.s_in, FU)[o] + h * Tu;
      if (hv) mb.st(a.s_out, FV)[o] = mb.st(a.s_in, FV)[o] + h * Tv;
    }
  }
  if (mode == ISP_FINAL) {
    block_sum_n<6>(p, red);
    if (tid < 3 && !(tid == CDV && !hv)) {
      double *pp = mb.part() + ((size_t)(j * 3 + tid) * nseg + g) * 2;
      pp[0] = p[2 * tid];
      pp[1] = p[2 * tid + 1];
    }
  }
}

// ---------------------------------------------------------------------------
// ordinary_diffusion (diffusion_mod.F90:74-217), order 2, FP64 in every mode
// ---------------------------------------------------------------------------

// One tile per (member, row, segment) (dif_tile, gd_diffuse.hpp).  One launch
// is one application of the second-order operator (diffusion_mod.F90:107-179):
// diffusion_order/2 applications run in sequence, the first on the state
// s_in (lap_src = 0), later ones on the previous result held in slot s_lap
// (lap_src = 1).  Intermediate applications write slot s_lapout
// (lap_final = 0); the last stages the flagged rows for the filter and adds
// sign*dt*nu times the result to the state s_in into X for the others.
template <typename CT>
__global__ void __launch_bounds__(kBlock) k_diffuse(const Args a) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ double red[16];
  const int t = xcd_tile(a.ntiles);
  if (t < 0) return;
  dif_tile<false>(a, t, smem, red);
}

// ---------------------------------------------------------------------------
// end of step: iap_transform (types_mod.F90:399-426) or derivation of u, v,
// then diag_run mass/energy (diag_mod.F90:42-121)
// ---------------------------------------------------------------------------

// Tile = (member, full row j, segment), one point per thread.  Each tile
// writes its four partial sums; k_diag adds them up.
template <int MODE>  // 0: iap_transform after diffusion, 1: derive u,v from an IAP slot, 2: diag only
__global__ void __launch_bounds__(kBlock) k_finish(const Args a) {
  const int t = xcd_tile(a.ntiles);
  if (t < 0) return;
  const int N = a.N, M = a.M, P = a.pitch, nseg = a.nseg;
  const int mloc = t / (M * nseg), rem = t - mloc * (M * nseg);
  const int j = rem / nseg, g = rem - (rem / nseg) * nseg;
  const Member mb = member_at(a, a.m0 + mloc);
  __shared__ double red[4 * 4];
  const int tid = threadIdx.x;
  const int i0 = g * kSeg, W = min(N - i0, kSeg);
  const RowCoef cj = a.coef[j];
  const size_t o = (size_t)j * P;
  double *Xgd = mb.st(SX, FGD), *XU = mb.st(SX, FU), *XV = mb.st(SX, FV);
  double *Xu = mb.st(SX, Fu), *Xv = mb.st(SX, Fv);
  const double *sgd = MODE != 1 ? Xgd : mb.st(a.s_in, FGD);
  const double *sU = mb.st(a.s_in, FU), *sV = mb.st(a.s_in, FV);
  double e4[4] = {0, 0, 0, 0};
  if (tid < W) {
    const int i = i0 + tid;
    const int ip = i == N - 1 ? 0 : i + 1;
    const bool hv = j <= M - 2;
    // loads
    const double g0 = sgd[o + i];
    const double hs = mb.ghs()[o + i];
    double g1 = 1.0, gn = 1.0, a0 = 0.0, b0 = 0.0;
    if (MODE != 2) {
      g1 = sgd[o + ip];
      if (hv) gn = sgd[o + P + i];
    }
    if (MODE == 0) {
      a0 = Xu[o + i];
      if (hv) b0 = Xv[o + i];
    } else if (MODE == 1) {
      a0 = sU[o + i];
      if (hv) b0 = sV[o + i];
    } else {
      a0 = XU[o + i];
      if (hv) b0 = XV[o + i];
    }
    double U = a0, V = b0;
    if (MODE == 0) {
      const double G0 = sqrt(g0), G1 = sqrt(g1), Gn = hv ? sqrt(gn) : 0.0;
      U = 0.5 * (G0 + G1) * a0;
      XU[o + i] = U;
      if (hv) {
        V = 0.5 * (G0 + Gn) * b0;
        XV[o + i] = V;
      }
    } else if (MODE == 1) {
      const double G0 = sqrt(g0), G1 = sqrt(g1), Gn = hv ? sqrt(gn) : 0.0;
      Xgd[o + i] = g0;
      XU[o + i] = U;
      Xu[o + i] = (j == 0 || j == M - 1) ? 0.0 : U * 2.0 / (G0 + G1);
      if (hv) {
        XV[o + i] = V;
        Xv[o + i] = V * 2.0 / (G0 + Gn);
      }
    }
    e4[0] = cj.mass_w * g0;
    if (j >= 1 && j <= M - 2) e4[1] = U * U * cj.fcos;
    if (hv) e4[2] = V * V * cj.hcos;
    const double h = g0 + hs;
    e4[3] = h * h * cj.fcos;
  }
  block_sum_n<4>(e4, red);
  if (tid == 0) {
    double *dp = mb.dpart() + ((size_t)j * nseg + g) * 4;
    dp[0] = e4[0]; dp[1] = e4[1]; dp[2] = e4[2]; dp[3] = e4[3];
  }
}

// diag_run totals (diag_mod.F90:56-87) from the k_finish partials, one
// workgroup per member; appends (mass, energy, beta) to the member's ring.
__global__ void __launch_bounds__(kBlock) k_diag(const Args a) {
  const int m = a.m0 + blockIdx.x;
  const Member mb = member_at(a, m);
  __shared__ double red[4 * 4];
  const int tid = threadIdx.x;
  const int np = a.M * a.nseg;
  double q[4] = {0, 0, 0, 0};
  for (int r0 = 0; r0 < np; r0 += 4 * kBlock) {
    double x[4][4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int r = r0 + u * kBlock + tid;
#pragma unroll
      for (int k = 0; k < 4; ++k) x[u][k] = r < np ? mb.dpart()[(size_t)r * 4 + k] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int k = 0; k < 4; ++k) q[k] += x[u][k];
  }
  block_sum_n<4>(q, red);
  if (tid == 0) {
    const int s = a.step_ctr[m];
    double *out = mb.diag_out + (size_t)(s % a.diag_cap) * 3;
    out[0] = q[0] * (kRadius * kRadius);
    out[1] = q[1] + q[2] + q[3];
    out[2] = mb.beta()[0];
    a.step_ctr[m] = s + 1;
  }
}

// diag_run div and vor of state X (diag_mod.F90:49-77), history only
__global__ void __launch_bounds__(kBlock) k_divvor(const Args a) {
  const int t = xcd_tile(a.ntiles);
  if (t < 0) return;
  const int N = a.N, M = a.M, P = a.pitch;
  const int m = a.m0 + t / M, j = t % M;
  const Member mb = member_at(a, m);
  const RowCoef cj = a.coef[j];
  const double *u = mb.st(SX, Fu), *v = mb.st(SX, Fv);
  const size_t o = (size_t)j * P;
  for (int i = threadIdx.x; i < N; i += kBlock) {
    const int im = i == 0 ? N - 1 : i - 1, ip = i == N - 1 ? 0 : i + 1;
    if (j >= 1 && j <= M - 2) {
      const double um1 = u[o + im], up1 = u[o + i];
      const double vm1 = v[o - P + i] * a.coef[j - 1].hcos, vp1 = v[o + i] * cj.hcos;
      mb.div()[o + i] = (up1 - um1) / cj.fdlon + (vp1 - vm1) / cj.fdlat;
    } else {
      mb.div()[o + i] = 0.0;
    }
    if (j <= M - 2) {
      const double um1 = u[o + i], up1 = u[o + P + i];
      const double vm1 = v[o + i], vp1 = v[o + ip];
      mb.vor()[o + i] = (vp1 - vm1) / cj.hdlon - (up1 - um1) / cj.hdlat;
    }
  }
}

// ---------------------------------------------------------------------------
// check_antisymmetry (dycore_mod.F90:794-851): the ten inner products of the
// operator terms of state X with the IAP state they act on
// ---------------------------------------------------------------------------

// One workgroup per full row j of member a.m0; thread tid takes longitudes
// tid, tid + kBlock, ...  Each term is the reference operator's expression
// (dycore_mod.F90:367-598, as the oracle restates it) in FP64 from the
// state's u, v, U, V, gd and ghs, and enters as term * weight * cos
// (u terms: iap u and full cos; v terms: iap v and half cos; mass terms:
// gd + ghs and full cos).  Row j's ten signed sums and ten sums of
// magnitudes go to scratch row j (Tb du field): [j][20].
__global__ void __launch_bounds__(kBlock) k_terms(const Args a) {
  const int j = blockIdx.x;
  const int N = a.N, M = a.M, P = a.pitch;
  const Member mb = member_at(a, a.m0);
  __shared__ double red[4 * 20];
  const double *gd = mb.st(SX, FGD), *U = mb.st(SX, FU), *V = mb.st(SX, FV);
  const double *u = mb.st(SX, Fu), *v = mb.st(SX, Fv), *ghs = mb.ghs();
  const RowCoef cj = a.coef[j];
  const bool upwind = a.upwind != 0;
  const double bl = a.beta_lon, bt = a.beta_lat;
  // pole rows: the meridional mass divergence is the ring sum at every longitude
  const bool pole = j == 0 || j == M - 1;
  const double pflux = pole ? pole_flux(a, mb, j, red) : 0.0;
  auto G = [&](int jj, int ii) { return sqrt(gd[(size_t)jj * P + ii]); };
  auto at = [&](const double *f, int jj, int ii) { return f[(size_t)jj * P + ii]; };
  double q[20];
#pragma unroll
  for (int k = 0; k < 20; ++k) q[k] = 0.0;
  auto add = [&](int k, double t, double w, double c) {
    const double x = t * w * c;
    q[k] += x;
    q[10 + k] += fabs(x);
  };
  for (int i = threadIdx.x; i < N; i += kBlock) {
    const int im = i == 0 ? N - 1 : i - 1, ip = i == N - 1 ? 0 : i + 1;
    // ---- u terms, full row j (interior rows; zero at the poles) ----
    if (j >= 1 && j <= M - 2) {
      const double U0 = at(U, j, i), Up = at(U, j, ip), Um = at(U, j, im);
      const double u1 = at(u, j, i) + at(u, j, im), u2 = at(u, j, i) + at(u, j, ip);
      const double ual = upwind ? 0.25 / cj.fdlon * (u2 * (U0 + Up) - bl * fabs(u2) * (Up - U0) - u1 * (U0 + Um) +
                                                     bl * fabs(u1) * (U0 - Um) - (u2 - u1) * U0)
                                : 0.25 / cj.fdlon * (u2 * Up - u1 * Um);
      const double hcm = a.coef[j - 1].hcos;
      const double v1 = (at(v, j - 1, i) + at(v, j - 1, ip)) * hcm, v2 = (at(v, j, i) + at(v, j, ip)) * cj.hcos;
      const double Ujp = at(U, j + 1, i), Ujm = at(U, j - 1, i);
      const double uat = upwind ? 0.25 / cj.fdlat * (v2 * (U0 + Ujp) - bt * fabs(v2) * (Ujp - U0) - v1 * (U0 + Ujm) +
                                                     bt * fabs(v1) * (U0 - Ujm) - (v2 - v1) * U0)
                                : 0.25 / cj.fdlat * (v2 * Ujp - v1 * Ujm);
      const double fv = 0.25 * (cj.f + cj.c * at(u, j, i)) *
                        (cj.c1 * (at(V, j - 1, i) + at(V, j - 1, ip)) + cj.c2 * (at(V, j, i) + at(V, j, ip)));
      const double upgf = 0.5 * (G(j, i) + G(j, ip)) / cj.fdlon *
                          (at(gd, j, ip) + at(ghs, j, ip) - at(gd, j, i) - at(ghs, j, i));
      add(0, ual, U0, cj.fcos);
      add(1, uat, U0, cj.fcos);
      add(2, fv, U0, cj.fcos);
      add(3, upgf, U0, cj.fcos);
    }
    // ---- v terms, half row j ----
    if (j <= M - 2) {
      const RowCoef cn = a.coef[j + 1];
      const double V0 = at(V, j, i), Vp = at(V, j, ip), Vm = at(V, j, im);
      const double u1 = at(u, j, im) + at(u, j + 1, im), u2 = at(u, j, i) + at(u, j + 1, i);
      const double val = upwind ? 0.25 / cj.hdlon * (u2 * (V0 + Vp) - bl * fabs(u2) * (Vp - V0) - u1 * (V0 + Vm) +
                                                     bl * fabs(u1) * (V0 - Vm) - (u2 - u1) * V0)
                                : 0.25 / cj.hdlon * (u2 * Vp - u1 * Vm);
      // v and V beyond the half rows are zero, as are the cosines there
      const double hcm = j >= 1 ? a.coef[j - 1].hcos : 0.0, hcp = j + 1 <= M - 2 ? cn.hcos : 0.0;
      const double vm = j >= 1 ? at(v, j - 1, i) : 0.0, vp = j + 1 <= M - 2 ? at(v, j + 1, i) : 0.0;
      const double Vjm = j >= 1 ? at(V, j - 1, i) : 0.0, Vjp = j + 1 <= M - 2 ? at(V, j + 1, i) : 0.0;
      const double v1 = at(v, j, i) * cj.hcos + vm * hcm, v2 = at(v, j, i) * cj.hcos + vp * hcp;
      const double vat = upwind ? 0.25 / cj.hdlat * (v2 * (V0 + Vjp) - bt * fabs(v2) * (Vjp - V0) - v1 * (V0 + Vjm) +
                                                     bt * fabs(v1) * (V0 - Vjm) - (v2 - v1) * V0)
                                : 0.25 / cj.hdlat * (v2 * Vjp - v1 * Vjm);
      const double fu = 0.25 * ((cj.f + cj.c * at(u, j, i)) * at(U, j, i) + (cj.f + cj.c * at(u, j, im)) * at(U, j, im) +
                                (cn.f + cn.c * at(u, j + 1, i)) * at(U, j + 1, i) +
                                (cn.f + cn.c * at(u, j + 1, im)) * at(U, j + 1, im));
      const double vpgf = 0.5 * (G(j, i) + G(j + 1, i)) / cj.hdlat * cj.hcos *
                          (at(gd, j + 1, i) + at(ghs, j + 1, i) - at(gd, j, i) - at(ghs, j, i));
      add(4, val, V0, cj.hcos);
      add(5, vat, V0, cj.hcos);
      add(6, fu, V0, cj.hcos);
      add(7, vpgf, V0, cj.hcos);
    }
    // ---- mass divergence, full row j ----
    {
      double mdl = 0.0, mdp = pflux;
      if (!pole) {
        const double Gi = G(j, i);
        mdl = ((Gi + G(j, ip)) * at(U, j, i) - (Gi + G(j, im)) * at(U, j, im)) * 0.5 / cj.fdlon;
        mdp = ((Gi + G(j + 1, i)) * at(V, j, i) * cj.hcos - (Gi + G(j - 1, i)) * at(V, j - 1, i) * a.coef[j - 1].hcos) *
              0.5 / cj.fdlat;
      }
      const double wg = at(gd, j, i) + at(ghs, j, i);
      add(8, mdl, wg, cj.fcos);
      add(9, mdp, wg, cj.fcos);
    }
  }
  block_sum_n<20>(q, red);
  if (threadIdx.x < 20) mb.tb(CDU)[(size_t)j * 20 + threadIdx.x] = q[threadIdx.x];
}

// The rows' sums added in row order, one thread per value: the twenty
// totals to the first entries of the Tb dv field.
__global__ void __launch_bounds__(64) k_terms_sum(const Args a) {
  const Member mb = member_at(a, a.m0);
  const int k = threadIdx.x;
  if (k >= 20) return;
  double s = 0.0;
  for (int j = 0; j < a.M; ++j) s += mb.tb(CDU)[(size_t)j * 20 + k];
  mb.tb(CDV)[k] = s;
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------

static inline int row_grid(int ntiles) { return ((ntiles + kXcd - 1) / kXcd) * kXcd; }

template <typename CT>
static hipError_t launch_tend_t(const Args &a, int pass, bool in_full, hipStream_t s) {
  dim3 g(row_grid(a.ntiles)), b(kBlock);
  const size_t lds = tend_lds_bytes(a, sizeof(CT));
#define GD_TEND(P, F) hipLaunchKernelGGL((k_tend<CT, P, F>), g, b, lds, s, a)
  if (pass == 0) { if (in_full) GD_TEND(0, true); else GD_TEND(0, false); }
  else if (pass == 1) { if (in_full) GD_TEND(1, true); else GD_TEND(1, false); }
  else { if (in_full) GD_TEND(2, true); else GD_TEND(2, false); }
#undef GD_TEND
  return hipGetLastError();
}

size_t tend_lds_bytes(const Args &a, size_t cts) {
  (void)a;
  return cts == 8 ? tend_stencil_bytes<double>() : tend_stencil_bytes<float>();
}

hipError_t launch_tend(const Args &a, int prec, int pass, bool in_full, hipStream_t s) {
  return prec == 0 ? launch_tend_t<double>(a, pass, in_full, s) : launch_tend_t<float>(a, pass, in_full, s);
}


hipError_t launch_isp(const Args &a, int prec, hipStream_t s) {
  dim3 g(row_grid(a.ntiles)), b(kBlock);
  if (prec == 0)
    hipLaunchKernelGGL(k_isp<double>, g, b, 0, s, a);
  else
    hipLaunchKernelGGL(k_isp<float>, g, b, 0, s, a);
  return hipGetLastError();
}

hipError_t launch_update(const Args &a, int prec, hipStream_t s) {
  dim3 g(a.upd_wpm * a.nmem_launch), b(kBlock);
  if (a.upd_narrow) {
    if (prec == 0)
      hipLaunchKernelGGL((k_update<double, 4, 4>), g, b, 0, s, a);
    else
      hipLaunchKernelGGL((k_update<float, 4, 4>), g, b, 0, s, a);
  } else {
    if (prec == 0)
      hipLaunchKernelGGL((k_update<double, 8, 8>), g, b, 0, s, a);
    else
      hipLaunchKernelGGL((k_update<float, 8, 8>), g, b, 0, s, a);
  }
  return hipGetLastError();
}


hipError_t launch_diffuse(const Args &a, int prec, hipStream_t s) {
  dim3 g(row_grid(a.ntiles)), b(kBlock);
  const size_t lds = dif_stencil_bytes();
  if (prec == 0)
    hipLaunchKernelGGL(k_diffuse<double>, g, b, lds, s, a);
  else
    hipLaunchKernelGGL(k_diffuse<float>, g, b, lds, s, a);
  return hipGetLastError();
}

hipError_t launch_finish(const Args &a, int mode, hipStream_t s) {
  dim3 g(row_grid(a.ntiles)), b(kBlock);
  if (mode == 0)
    hipLaunchKernelGGL(k_finish<0>, g, b, 0, s, a);
  else if (mode == 1)
    hipLaunchKernelGGL(k_finish<1>, g, b, 0, s, a);
  else
    hipLaunchKernelGGL(k_finish<2>, g, b, 0, s, a);
  return hipGetLastError();
}

hipError_t launch_diag(const Args &a, int nmem, hipStream_t s) {
  hipLaunchKernelGGL(k_diag, dim3(nmem), dim3(kBlock), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_terms(const Args &a, hipStream_t s) {
  hipLaunchKernelGGL(k_terms, dim3(a.M), dim3(kBlock), 0, s, a);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_terms_sum, dim3(1), dim3(64), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_divvor(const Args &a, hipStream_t s) {
  dim3 g(row_grid(a.ntiles)), b(kBlock);
  hipLaunchKernelGGL(k_divvor, g, b, 0, s, a);
  return hipGetLastError();
}

// Raise the dynamic-LDS limit of the kernels that size LDS by num_lon.
hipError_t configure_kernels(int max_lds) {
  hipError_t e = configure_filter_mm(max_lds);
  if (e != hipSuccess) return e;
  e = configure_filter(max_lds);
  return e != hipSuccess ? e : configure_phase(max_lds);
}

}  // namespace gd
