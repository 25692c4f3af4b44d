// gd_tend.hpp -- the tendency of one row segment (space_operators,
// dycore_mod.F90:184-598) and the row tile around it, shared by k_tend
// (gd_kernels.hip) and the fused phase launch k_phase (gd_phase.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "gd_common.hpp"
#include "gd_device.hpp"
#include "gd_geom.hpp"

namespace gd {

// LDS of one segment: the rows whose longitude neighbours i-1, i+1 are read,
// over the segment plus one halo point on both sides: G row j, U rows j,
// j+1, V half rows j-1, j (double), u rows j, j+1 and v half rows j-1, j
// (compute type).  The rows read only at a thread's own longitude (G rows
// j-1, j+1, U row j-1, V and v half row j+1) stay in its registers, which
// keeps the tile at 14.6 KB (mixed) / 18.7 KB (FP64) of LDS.
constexpr int kLW = kBlock + 4;  // LDS slot li = i - i0 + 2 holds longitude i
template <typename CT>
__host__ __device__ constexpr size_t tend_stencil_bytes() {
  return (size_t)(5 * sizeof(double) + 4 * sizeof(CT)) * kLW;
}

// Tendency values of one point and the epilogue inputs loaded with them.
struct TendPoint {
  double du, dv, dgd;
  double pre_u, pre_v, pre_g;
};

// Components a tend_segment call evaluates (CM bits); TC_RING: a pole row
// forms its ring sum itself (without it the pole-row dgd is not evaluated).
constexpr int TC_DU = 1, TC_DV = 2, TC_DGD = 4, TC_RING = 8, TC_ALL = 15;

// Pole mass flux of pole row j (dycore_mod.F90:572-596): the ring sum over
// the row, in the same order in every caller.  Ends with barriers (block_sum).
// This thread's part of the ring sum of pole row j: longitudes q = tid,
// tid + kBlock, ... in that order.  The loads go in blocks of four, each
// issued without a condition: past N a lane reads the ghost rows (gd 1,
// V 0), whose term (1 + 1) * 0 = +0 leaves the sum as it is.
__device__ __forceinline__ double ring_part(const Args &a, const Member &mb, int j) {
  const int N = a.N, M = a.M, P = a.pitch, tid = ltid();
  const double *gd = mb.st(a.s_in, FGD);
  const double *Vg = mb.st(a.s_in, FV);
  const double *one = a.ghost, *zero = a.ghost + P;
  const int jn = j == 0 ? 1 : M - 2, hr = j == 0 ? 0 : M - 2;
  double pacc = 0.0;
  for (int q0 = 0; q0 < N; q0 += 4 * kBlock) {
    double g0[4], g1[4], v[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int q = q0 + k * kBlock + tid;
      const bool in = q < N;
      const int qc = in ? q : 0;
      g0[k] = (in ? gd + (size_t)j * P : one)[qc];
      g1[k] = (in ? gd + (size_t)jn * P : one)[qc];
      v[k] = (in ? Vg + (size_t)hr * P : zero)[qc];
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) pacc += (sqrt(g0[k]) + sqrt(g1[k])) * v[k];
  }
  return pacc;
}

__device__ __forceinline__ double pole_flux(const Args &a, const Member &mb, int j, double *red) {
  const int N = a.N, M = a.M;
  double sum = block_sum(ring_part(a, mb, j), red);
  if (j == M - 1) sum = -sum;
  return sum * 2.0 / N / kRadius / a.dlat;
}

// Global loads of one segment (see tend_segment): the column of the
// thread's longitude (own point for tid < W, halo i0-1 for tid == W, halo
// i0+W for tid == W+1) and the own point's geopotential terms.
struct SegLoads {
  double lg[4], lgn[2], lU[3], lV[3], lfu[2], lfv[3];
  double g_i, g_ip, h_i, h_ip, g_n, h_n;
};

template <int PASS, bool IN_FULL>
__device__ __forceinline__ void seg_load(const Args &a, const Member &mb, int j, int i0, int W, SegLoads &L) {
  const int N = a.N, M = a.M, P = a.pitch, tid = ltid();
  const double *gd = mb.st(a.s_in, FGD);
  const double *Ug = mb.st(a.s_in, FU);
  const double *Vg = mb.st(a.s_in, FV);
  const double *ghs = mb.ghs();
  auto wrap = [N](int i) { return i < 0 ? i + N : (i >= N ? i - N : i); };
  const int lcol = tid < W ? i0 + tid : (tid == W ? i0 - 1 : i0 + W);
  const int p = wrap(lcol), pn = wrap(lcol + 1);
  // Every lane loads, whatever its column (lanes past the right halo load
  // its column and are not used), and a row beyond the grid is read from a
  // ghost row of the value the stencil takes there (gd 1, winds 0).  No load
  // then sits under a condition: the compiler would make that a branch,
  // move the value's use into it and wait for each load in turn.
  const double *one = a.ghost, *zero = a.ghost + P;
  auto in = [](int row, int last) { return row >= 0 && row <= last; };
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int row = j - 1 + r;
    L.lg[r] = (in(row, M - 1) ? gd + (size_t)row * P : one)[p];
  }
#pragma unroll
  for (int r = 0; r < 2; ++r) {
    const int row = j + r;
    L.lgn[r] = (in(row, M - 1) ? gd + (size_t)row * P : one)[pn];
    L.lfu[r] = IN_FULL ? (in(row, M - 1) ? mb.st(a.s_in, Fu) + (size_t)row * P : zero)[p] : 0.0;
  }
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    const int row = j - 1 + r;
    L.lU[r] = (in(row, M - 1) ? Ug + (size_t)row * P : zero)[p];
    L.lV[r] = (in(row, M - 2) ? Vg + (size_t)row * P : zero)[p];
    L.lfv[r] = IN_FULL ? (in(row, M - 2) ? mb.st(a.s_in, Fv) + (size_t)row * P : zero)[p] : 0.0;
  }
  // the own point's geopotential: gd at (j, i), (j, i+1), (j+1, i) is the
  // column load already made (lg[1], lgn[0], lg[2]); only ghs is loaded
  // (lanes tid >= W load column i0 and are not used)
  const int i = tid < W ? i0 + tid : i0;
  const int ip = i + 1 >= N ? i + 1 - N : i + 1;
  const size_t o = (size_t)j * P;
  L.g_i = L.lg[1];
  L.g_ip = L.lgn[0];
  L.g_n = L.lg[2];
  L.h_i = L.h_ip = L.h_n = 0.0;
  if (PASS != 2) {
    L.h_i = ghs[o + i];
    L.h_ip = ghs[o + ip];
    L.h_n = (j <= M - 2 ? ghs + o + P : zero)[i];
  }
}

// Coefficients of the neighbouring rows j-1 and j+1 (zero beyond the grid).
// Loaded once per tile, before its first barrier: a barrier orders memory,
// so loads written after it are issued after it.
struct RowNbr {
  double hc_m, hc_p, f_p, c_p;
};

__device__ __forceinline__ RowNbr row_nbr(const Args &a, int j) {
  const int M = a.M;
  // read through the constant address space (no kernel writes the table):
  // scalar loads, which the vector-memory waits do not count, so no wait
  // for them also waits for the loads in flight before them
  const __attribute__((address_space(4))) RowCoef *coef = (const __attribute__((address_space(4))) RowCoef *)a.coef;
  RowNbr nb;
  nb.hc_m = j >= 1 ? coef[j - 1].hcos : 0.0;
  nb.hc_p = j + 1 <= M - 2 ? coef[j + 1].hcos : 0.0;
  nb.f_p = j + 1 <= M - 1 ? coef[j + 1].f : 0.0;
  nb.c_p = j + 1 <= M - 1 ? coef[j + 1].c : 0.0;
  return nb;
}

// The rest of a segment once its loads are issued: the LDS rows, one
// barrier (or the pole ring sum pacc's block reduction when pole_row) and
// the components CM of the own point, into out.  The LDS rows only other
// components read are not written, so their loads drop out.  bar2: a tile
// sharing its workgroup with a pole-row tile takes two barriers as that
// tile's reduction does, so every tile of the workgroup takes the same
// barriers.
template <typename CT, int PASS, bool IN_FULL, int CM>
__device__ __forceinline__ TendPoint seg_eval(const Args &a, const RowCoef &cj, const RowNbr &nb, int j, int i0,
                                              int W, const SegLoads &L, bool pole_row, double pacc, TendPoint out,
                                              unsigned char *smem, double *red, bool bar2 = false) {
  const int N = a.N, M = a.M;
  double *sG1 = (double *)smem;                          // G row j
  double(*sU)[kLW] = (double(*)[kLW])(smem + 1 * kLW * sizeof(double));  // U rows j, j+1
  double(*sV)[kLW] = (double(*)[kLW])(smem + 3 * kLW * sizeof(double));  // V half rows j-1, j
  CT(*su)[kLW] = (CT(*)[kLW])(smem + 5 * kLW * sizeof(double));          // u rows j, j+1
  CT(*sv)[kLW] = (CT(*)[kLW])(smem + 5 * kLW * sizeof(double) + 2 * kLW * sizeof(CT));  // v j-1, j
  const int tid = ltid();
  constexpr bool DU = (CM & TC_DU) != 0, DV = (CM & TC_DV) != 0, DG = (CM & TC_DGD) != 0;
  const bool has_u = j >= 1 && j <= M - 2;
  const bool has_v = j <= M - 2;
  const double(&lg)[4] = L.lg;
  const double(&lgn)[2] = L.lgn;
  const double(&lU)[3] = L.lU;
  const double(&lV)[3] = L.lV;
  const double(&lfu)[2] = L.lfu;
  const double(&lfv)[3] = L.lfv;
  const double g_i = L.g_i, g_ip = L.g_ip, h_i = L.h_i, h_ip = L.h_ip, g_n = L.g_n, h_n = L.h_n;
  const int lslot = tid < W ? tid + 2 : (tid == W ? 1 : W + 2);
  const bool own = tid < W;
  // LDS rows the components of this pass read at i-1 / i+1 (PASS 1: the
  // fast terms fv, upgf, fu, vpgf and the mass flux; PASS 2: the advection
  // terms); rows nobody reads are not written, so their loads and the u, v
  // they would derive drop out
  constexpr bool W_G1 = PASS != 2 && (DU || DG);                            // upgf, mass flux
  constexpr bool W_U0 = (DU && PASS != 1) || ((DV || DG) && PASS != 2);     // ual; fu, mass flux
  constexpr bool W_U1 = DV && PASS != 2;                                    // fu
  constexpr bool W_V0 = DU && PASS != 2;                                    // fv
  constexpr bool W_V1 = (DU && PASS != 2) || (DV && PASS != 1);             // fv; val
  constexpr bool W_u0 = (DU && PASS != 1) || DV;                            // ual; fu, val
  constexpr bool W_u1 = DV;                                                 // fu, val
  constexpr bool W_v = DU && PASS != 1;                                     // uat
  // ---- LDS rows; the own-longitude rows stay in registers ----
  double G[4];
  CT u_own[2] = {(CT)0, (CT)0}, v_own3[3] = {(CT)0, (CT)0, (CT)0};  // u rows j, j+1; v half rows j-1..j+1
  if (tid < W + 2) {
    const int li = lslot;
#pragma unroll
    for (int r = 0; r < 4; ++r) G[r] = sqrt(lg[r]);
    if (W_G1) sG1[li] = G[1];
    if (W_U0) sU[0][li] = lU[1];
    if (W_U1) sU[1][li] = lU[2];
    if (W_V0) sV[0][li] = lV[0];
    if (W_V1) sV[1][li] = lV[1];
#pragma unroll
    for (int r = 0; r < 2; ++r) {
      const int row = j + r;
      CT val = (CT)0;
      if (IN_FULL) {
        val = (CT)lfu[r];
      } else if (row >= 1 && row <= M - 2) {
        // update_state: u = U * 2 / (G(i) + G(i+1))  (dycore_mod.F90:636-640)
        val = ((CT)lU[r + 1] * (CT)2) / ((CT)G[r + 1] + (CT)sqrt(lgn[r]));
      }
      if ((r == 0 && W_u0) || (r == 1 && W_u1)) su[r][li] = val;
      u_own[r] = val;
    }
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      const int hrow = j - 1 + r;
      CT val = (CT)0;
      if (hrow >= 0 && hrow <= M - 2) {
        if (IN_FULL)
          val = (CT)lfv[r];
        else  // v = V * 2 / (G(j) + G(j+1))  (dycore_mod.F90:641-645)
          val = ((CT)lV[r] * (CT)2) / ((CT)G[r] + (CT)G[r + 1]);
      }
      if (r < 2 && W_v) sv[r][li] = val;
      v_own3[r] = val;
    }
  }
  double pole = 0.0;
  if (pole_row) {
    double sum = block_sum(pacc, red);  // its barriers also publish the LDS rows
    if (j == M - 1) sum = -sum;
    pole = sum * 2.0 / N / kRadius / a.dlat;
  } else {
    __syncthreads();
    if (bar2) __syncthreads();
  }
  if (!own) return out;

  const double hc_m = nb.hc_m, hc_0 = has_v ? cj.hcos : 0.0, hc_p = nb.hc_p, f_p = nb.f_p, c_p = nb.c_p;
  const CT bl = (CT)a.beta_lon, bt = (CT)a.beta_lat;
  const bool upwind = a.upwind != 0;
  const int li = tid + 2;
  // ---- u equation (full row j) ----
  if (DU && has_u) {
    const CT u0 = u_own[0], um = su[0][li - 1], up = su[0][li + 1];
    CT ual = 0, uat = 0, fv = 0, upgf = 0;
    if (PASS != 1) {
      const CT U0 = (CT)lU[1], Um = (CT)sU[0][li - 1], Up = (CT)sU[0][li + 1];
      const CT u1 = u0 + um, u2 = u0 + up;
      const CT kl = (CT)cj.k_uadv_lon;
      if (!upwind)
        ual = kl * (u2 * Up - u1 * Um);
      else
        ual = kl * (u2 * (U0 + Up) - bl * fabs(u2) * (Up - U0) - u1 * (U0 + Um) +
                    bl * fabs(u1) * (U0 - Um) - (u2 - u1) * U0);
      const CT v1 = (v_own3[0] + sv[0][li + 1]) * (CT)hc_m;
      const CT v2 = (v_own3[1] + sv[1][li + 1]) * (CT)hc_0;
      const CT Ujp = (CT)lU[2], Ujm = (CT)lU[0];
      const CT kt = (CT)cj.k_uadv_lat;
      if (!upwind)
        uat = kt * (v2 * Ujp - v1 * Ujm);
      else
        uat = kt * (v2 * (U0 + Ujp) - bt * fabs(v2) * (Ujp - U0) - v1 * (U0 + Ujm) +
                    bt * fabs(v1) * (U0 - Ujm) - (v2 - v1) * U0);
    }
    if (PASS != 2) {
      const CT Vm0 = (CT)lV[0], Vm1 = (CT)sV[0][li + 1];
      const CT V00 = (CT)lV[1], V01 = (CT)sV[1][li + 1];
      fv = (CT)0.25 * ((CT)cj.f + (CT)cj.c * u0) * ((CT)cj.c1 * (Vm0 + Vm1) + (CT)cj.c2 * (V00 + V01));
      const double dphi = g_ip + h_ip - g_i - h_i;
      upgf = (CT)0.5 * ((CT)G[1] + (CT)sG1[li + 1]) / (CT)cj.fdlon * (CT)dphi;
    }
    CT d;
    if (PASS == 0)
      d = -ual - uat + fv - upgf;
    else if (PASS == 1)
      d = fv - upgf;
    else
      d = -ual - uat;
    out.du = (double)d;
  }
  // ---- v equation (half row j) ----
  if (DV && has_v) {
    CT val = 0, vat = 0, fu = 0, vpgf = 0;
    if (PASS != 1) {
      const CT u1 = su[0][li - 1] + su[1][li - 1];
      const CT u2 = u_own[0] + u_own[1];
      const CT V0 = (CT)lV[1], Vm = (CT)sV[1][li - 1], Vp = (CT)sV[1][li + 1];
      const CT kl = (CT)cj.k_vadv_lon;
      if (!upwind)
        val = kl * (u2 * Vp - u1 * Vm);
      else
        val = kl * (u2 * (V0 + Vp) - bl * fabs(u2) * (Vp - V0) - u1 * (V0 + Vm) +
                    bl * fabs(u1) * (V0 - Vm) - (u2 - u1) * V0);
      const CT vj = v_own3[1];
      const CT v1 = vj * (CT)hc_0 + v_own3[0] * (CT)hc_m;
      const CT v2 = vj * (CT)hc_0 + v_own3[2] * (CT)hc_p;
      const CT Vjp = (CT)lV[2];  // zero beyond the last half row
      const CT Vjm = (CT)lV[0];  // zero before the first half row
      const CT kt = (CT)cj.k_vadv_lat;
      if (!upwind)
        vat = kt * (v2 * Vjp - v1 * Vjm);
      else
        vat = kt * (v2 * (V0 + Vjp) - bt * fabs(v2) * (Vjp - V0) - v1 * (V0 + Vjm) +
                    bt * fabs(v1) * (V0 - Vjm) - (v2 - v1) * V0);
    }
    if (PASS != 2) {
      const CT f0 = (CT)cj.f, c0 = (CT)cj.c, fp = (CT)f_p, cp = (CT)c_p;
      const CT a1 = (f0 + c0 * u_own[0]) * (CT)lU[1];
      const CT a2 = (f0 + c0 * su[0][li - 1]) * (CT)sU[0][li - 1];
      const CT a3 = (fp + cp * u_own[1]) * (CT)lU[2];
      const CT a4 = (fp + cp * su[1][li - 1]) * (CT)sU[1][li - 1];
      fu = (CT)0.25 * (a1 + a2 + a3 + a4);
      const double dphi = g_n + h_n - g_i - h_i;
      vpgf = (CT)0.5 * ((CT)G[1] + (CT)G[2]) / (CT)cj.hdlat * (CT)cj.hcos * (CT)dphi;
    }
    CT d;
    if (PASS == 0)
      d = -val - vat - fu - vpgf;
    else if (PASS == 1)
      d = -fu - vpgf;
    else
      d = -val - vat;
    out.dv = (double)d;
  }
  // ---- mass continuity (full row j), FP64 in every mode ----
  if (DG && PASS != 2) {
    double mdl = 0.0, mdp;
    if (j == 0 || j == M - 1) {
      mdp = pole;
    } else {
      const double Gi = G[1], Gip = sG1[li + 1], Gim = sG1[li - 1];
      mdl = ((Gi + Gip) * lU[1] - (Gi + Gim) * sU[0][li - 1]) * 0.5 / cj.fdlon;
      mdp = ((Gi + G[2]) * lV[1] * hc_0 - (Gi + G[0]) * lV[0] * hc_m) * 0.5 / cj.fdlat;
    }
    out.dgd = -mdl - mdp;
  }
  return out;
}

// Segment i0 .. i0+W-1 of full row j: thread tid < W loads the column of
// longitude i0+tid and evaluates that point; threads W and W+1 load the halo
// columns i0-1 and i0+W.  Every global load is issued before the first
// barrier (neighbour rows to LDS, the point's own rows, geopotential terms
// and epilogue inputs to registers).  CM selects the components.  smem holds
// tend_stencil_bytes<CT>() bytes, red 24 doubles.
template <typename CT, int PASS, bool IN_FULL, int CM = TC_ALL>
__device__ __forceinline__ TendPoint tend_segment(const Args &a, const Member &mb, const RowCoef &cj, int j,
                                                  int i0, int W, bool stg_u, bool stg_v, bool stg_g,
                                                  unsigned char *smem, double *red, bool bar2 = false) {
  const int M = a.M, P = a.pitch, tid = ltid();
  // pole mass flux sums over the whole ring (dycore_mod.F90:572-596); every
  // segment of a pole row forms the same sum in the same order (ring_part,
  // as pole_flux).  Pole tiles finish these loads before issuing the halo
  // loads, so the two sets of registers are never live together.
  const bool pole_row = (CM & TC_RING) != 0 && PASS != 2 && (j == 0 || j == M - 1);
  const double pacc = pole_row ? ring_part(a, mb, j) : 0.0;
  const RowNbr nb = row_nbr(a, j);
  SegLoads L;
  seg_load<PASS, IN_FULL>(a, mb, j, i0, W, L);
  // the epilogue inputs of the own point
  TendPoint out{0, 0, 0, 0, 0, 0};
  if (tid < W) {
    const size_t o = (size_t)j * P + i0 + tid;
    if (!stg_u) out.pre_u = phase_pre<CT>(a, mb, CDU, o);
    if (!stg_v && j <= M - 2) out.pre_v = phase_pre<CT>(a, mb, CDV, o);
    if (!stg_g) out.pre_g = phase_pre<CT>(a, mb, CDGD, o);
  }
  return seg_eval<CT, PASS, IN_FULL, CM>(a, cj, nb, j, i0, W, L, pole_row, pacc, out, smem, red, bar2);
}

// Row tile t = (member, row, segment of kSeg longitudes) of a phase launch.
// FUSED = false (k_tend): rows 0..M-1; components of flagged rows are
// staged to HBM for the filter kernel, which finishes their epilogue.
// FUSED = true (k_phase): rows from the far-row ranges; flagged components
// are left to the launch's whole-row filter tiles.
__device__ __forceinline__ int far_row(const Args &a, int r) {
#pragma unroll
  for (int k = 0; k < kMaxFarRng; ++k) {
    if (k < a.nfar_rng && r < a.far_cnt[k]) return a.far_lo[k] + r;
    r -= a.far_cnt[k];
  }
  return 0;
}

// The row of fused-launch row tile t.
__device__ __forceinline__ int far_tile_row(const Args &a, int t) {
  const int per = a.nfar_rows * a.nseg;
  return far_row(a, (t - (t / per) * per) / a.nseg);
}

// bar2: see seg_eval.  live = false: the tile's work and barriers without
// its stores (a workgroup's spare tile past the last one).
template <typename CT, int PASS, bool IN_FULL, bool FUSED>
__device__ __forceinline__ void tend_tile(const Args &a, int t, unsigned char *smem, double *red, bool bar2 = false,
                                          bool live = true) {
  const int M = a.M, P = a.pitch, nseg = a.nseg, tid = ltid();
  const int nrows = FUSED ? a.nfar_rows : M;
  const int mloc = t / (nrows * nseg), rem = t - mloc * (nrows * nseg);
  const int r = rem / nseg, g = rem - (rem / nseg) * nseg;
  const int j = FUSED ? far_row(a, r) : r;
  const Member mb = member_at(a, a.m0 + mloc);
  const RowCoef cj = a.coef[j];
  const bool has_u = j >= 1 && j <= M - 2, has_v = j <= M - 2;
  const int st_u = has_u ? cj.stage[CDU] : -1;
  const int st_v = has_v ? cj.stage[CDV] : -1;
  const int st_g = cj.stage[CDGD];
  const int i0 = g * kSeg, W = min(a.N - i0, kSeg);
  const TendPoint pt = tend_segment<CT, PASS, IN_FULL>(a, mb, cj, j, i0, W, st_u >= 0, st_v >= 0, st_g >= 0,
                                                       smem, red, bar2);
  double p1[3] = {0, 0, 0}, p2[3] = {0, 0, 0};
  if (live && tid < W) {
    const int i = i0 + tid;
    const size_t o = (size_t)j * P + i;
    // (pole rows: du = 0 goes through the epilogue like any other value)
    if (st_u < 0)
      phase_store<CT>(a, mb, CDU, o, pt.du, pt.pre_u, cj.fcos, p1[0], p2[0]);
    else if (!FUSED)
      mb.sx()[(size_t)st_u * P + i] = pt.du;
    if (has_v) {
      if (st_v < 0)
        phase_store<CT>(a, mb, CDV, o, pt.dv, pt.pre_v, cj.hcos, p1[1], p2[1]);
      else if (!FUSED)
        mb.sx()[(size_t)st_v * P + i] = pt.dv;
    }
    if (st_g < 0)
      phase_store<CT>(a, mb, CDGD, o, pt.dgd, pt.pre_g, cj.fcos, p1[2], p2[2]);
    else if (!FUSED)
      mb.sx()[(size_t)st_g * P + i] = pt.dgd;
  }
  if (a.phase == PH_C) {
    double pq[6] = {p1[0], p2[0], p1[1], p2[1], p1[2], p2[2]};
    block_sum_n<6>(pq, red);
    if (live && tid == 0) {
      const int st[3] = {st_u, st_v, st_g};
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        if (st[c] >= 0 || (c == CDV && !has_v)) continue;
        double *pp = mb.part() + ((size_t)(j * 3 + c) * nseg + g) * 2;
        pp[0] = pq[2 * c];
        pp[1] = pq[2 * c + 1];
      }
    }
  }
}

}  // namespace gd
