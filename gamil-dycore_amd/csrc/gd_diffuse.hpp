// gd_diffuse.hpp -- one application of the second-order diffusion operator
// (ordinary_diffusion, diffusion_mod.F90:74-217) on one row segment, shared
// by k_diffuse (gd_kernels.hip) and the fused diffusion launch k_dphase
// (gd_phase.hip).  FP64 in every precision mode.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "gd_common.hpp"
#include "gd_device.hpp"
#include "gd_tend.hpp"

namespace gd {

// LDS of one segment: gd, u and v rows j-1 .. j+1 over the segment plus two
// halo slots on each side (slot li = i - i0 + 2 holds longitude i).
constexpr int kDW = kBlock + 4;
__host__ __device__ constexpr size_t dif_stencil_bytes() { return (size_t)9 * kDW * sizeof(double); }

struct DiffPoint {
  double gdd, ud, vd;
};

// This thread's part of the zonal sum of the gd differences next to pole
// row j: longitudes q = tid, tid + kBlock, ... in that order, loaded in
// blocks of four without conditions (past N a lane reads the gd ghost row
// twice, and 1 - 1 = +0 leaves the sum as it is).
__device__ __forceinline__ double dif_ring_part(const Args &a, const double *lgd, int j) {
  const int N = a.N, M = a.M, P = a.pitch, tid = threadIdx.x;
  const int ja = j == 0 ? 1 : M - 1, jb = j == 0 ? 0 : M - 2;
  double pacc = 0.0;
  for (int q0 = 0; q0 < N; q0 += 4 * kBlock) {
    double x[4], y[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int q = q0 + k * kBlock + tid;
      const bool in = q < N;
      const int qc = in ? q : 0;
      x[k] = (in ? lgd + (size_t)ja * P : a.ghost)[qc];
      y[k] = (in ? lgd + (size_t)jb * P : a.ghost)[qc];
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) pacc += x[k] - y[k];
  }
  return pacc;
}

// Pole-row value of gdd (diffusion_mod.F90:116-133): the zonal sum of the
// operand's gd differences next to the pole, in the same order in every
// caller.  Ends with barriers (block_sum).
__device__ __forceinline__ double dif_pole(const Args &a, const Member &mb, int j, double *red) {
  const int N = a.N;
  const double *lgd = a.lap_src != 0 ? mb.st(a.s_lap, FGD) : mb.st(a.s_in, FGD);
  const double pacc = dif_ring_part(a, lgd, j);
  const RowCoef cj = a.coef[j];
  const double s = block_sum(pacc, red);
  if (j == 0) return s * cj.hcos / cj.fdlat2 * cj.fcos / N;
  return -s * a.coef[j - 1].hcos / cj.fdlat2 * cj.fcos / N;
}

// Segment i0 .. i0+W-1 of full row j: thread e loads point i0-1+e (e < W+2);
// every global load is issued before the barrier.  CM (TC_DU, TC_DV, TC_DGD
// bits) selects the components; with TC_RING a pole row forms its ring sum
// for gdd itself.  The operand is the state s_in (lap_src = 0; u, v derived
// from U, V unless in_full) or the previous application's result in slot
// s_lap (lap_src = 1).  smem holds dif_stencil_bytes(), red 16 doubles.
template <int CM>
__device__ __forceinline__ DiffPoint dif_segment(const Args &a, const Member &mb, const RowCoef &cj, int j, int i0,
                                                 int W, unsigned char *smem, double *red) {
  const int N = a.N, M = a.M, P = a.pitch;
  double(*sgd)[kDW] = (double(*)[kDW])smem;                          // gd rows j-1 .. j+1
  double(*su)[kDW] = (double(*)[kDW])(smem + 3 * kDW * sizeof(double));  // u rows (0 on and beyond the poles)
  double(*sv)[kDW] = (double(*)[kDW])(smem + 6 * kDW * sizeof(double));  // v half rows (0 outside)
  constexpr bool DU = (CM & TC_DU) != 0, DV = (CM & TC_DV) != 0, DG = (CM & TC_DGD) != 0;
  const int tid = threadIdx.x;
  const double *gd = mb.st(a.s_in, FGD);
  const double *Ug = mb.st(a.s_in, FU);
  const double *Vg = mb.st(a.s_in, FV);
  const double *ug = mb.st(a.s_in, Fu), *vg = mb.st(a.s_in, Fv);
  const bool prev = a.lap_src != 0;
  const bool full = a.in_full || prev;
  const double *lgd = prev ? mb.st(a.s_lap, FGD) : gd;
  const double *lu = prev ? mb.st(a.s_lap, Fu) : (a.in_full ? ug : Ug);
  const double *lv = prev ? mb.st(a.s_lap, Fv) : (a.in_full ? vg : Vg);
  auto wrap = [N](int i) { return i < 0 ? i + N : (i >= N ? i - N : i); };
  double lg[4], lgn[3], lU[3], lV[3];
  {
    // every lane loads (lanes e >= W+2 are not used); a row beyond the grid
    // reads the ghost row of the value the stencil takes there (gd 1,
    // winds 0), so no load sits under a condition that would make it wait
    // alone
    const double *one = a.ghost, *zero = a.ghost + P;
    const int e = tid;
    const int p = wrap(i0 - 1 + e), pn = wrap(i0 + e);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int row = j - 1 + r;
      lg[r] = (row >= 0 && row <= M - 1 ? lgd + (size_t)row * P : one)[p];
    }
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      const int row = j - 1 + r;
      const bool urow = row >= 1 && row <= M - 2;
      const bool vrow = row >= 0 && row <= M - 2;
      lgn[r] = (urow && !full ? gd + (size_t)row * P : one)[pn];
      lU[r] = (urow ? lu + (size_t)row * P : zero)[p];
      lV[r] = (vrow ? lv + (size_t)row * P : zero)[p];
    }
  }
  // pole rows of gdd: zonal sums (diffusion_mod.F90:116-133)
  const bool pole_row = (CM & TC_RING) != 0 && (j == 0 || j == M - 1);
  const double pacc = pole_row ? dif_ring_part(a, lgd, j) : 0.0;
  if (tid < W + 2) {
    const int li = tid + 1;
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      const int row = j - 1 + r;
      if (DG) sgd[r][li] = lg[r];
      double uv = 0.0, vv = 0.0;
      if (row >= 1 && row <= M - 2)
        uv = full ? lU[r] : lU[r] * 2.0 / (sqrt(lg[r]) + sqrt(lgn[r]));
      if (row >= 0 && row <= M - 2)
        vv = full ? lV[r] : lV[r] * 2.0 / (sqrt(lg[r]) + sqrt(lg[r + 1]));
      if (DU) su[r][li] = uv;
      if (DV) sv[r][li] = vv;
    }
  }
  const bool has_u = j >= 1 && j <= M - 2;
  const bool has_v = j <= M - 2;
  const double hc_m = j >= 1 ? a.coef[j - 1].hcos : 0.0;
  const double hc_0 = has_v ? cj.hcos : 0.0;
  const double fc_p = j + 1 <= M - 1 ? a.coef[j + 1].fcos : 0.0;
  double pole = 0.0;
  if (pole_row) {
    const double s = block_sum(pacc, red);  // its barriers also publish the LDS rows
    if (j == 0)
      pole = s * hc_0 / cj.fdlat2 * cj.fcos / N;
    else
      pole = -s * hc_m / cj.fdlat2 * cj.fcos / N;
  } else {
    __syncthreads();
  }
  DiffPoint out{0.0, 0.0, 0.0};
  if (tid >= W) return out;
  const int li = tid + 2;
  if (DG) {
    const double g0 = sgd[1][li];
    if (j == 0 || j == M - 1) {
      out.gdd = pole;
    } else {
      out.gdd = (sgd[1][li + 1] - 2 * g0 + sgd[1][li - 1]) / cj.fdlon2 +
                ((sgd[2][li] - g0) * hc_0 - (g0 - sgd[0][li]) * hc_m) / cj.fdlat2 * cj.fcos;
    }
  }
  if (DU && has_u) {
    const double u0 = su[1][li];
    out.ud = (su[1][li + 1] - 2 * u0 + su[1][li - 1]) / cj.fdlon2 +
             ((su[2][li] - u0) * hc_0 - (u0 - su[0][li]) * hc_m) / cj.fdlat2 * cj.fcos;
  }
  if (DV && has_v) {
    const double v0 = sv[1][li];
    double vd = (sv[1][li + 1] - 2 * v0 + sv[1][li - 1]) / cj.hdlon2;
    if (j >= 1 && j <= M - 3)
      vd = vd + ((sv[2][li] - v0) * fc_p - (v0 - sv[0][li]) * cj.fcos) / cj.hdlat2 * cj.hcos;
    else if (j == 0)
      vd = vd + (sv[2][li] - v0) * fc_p / cj.hdlat2 * cj.hcos;
    else
      vd = vd - (v0 - sv[0][li]) * cj.fcos / cj.hdlat2 * cj.hcos;
    out.vd = vd;
  }
  return out;
}

// Row tile t = (member, row, segment) of one application.  Intermediate
// applications write slot s_lapout; the last one adds kdiff times the result
// to the state s_in into X for unflagged components, and stages the flagged
// ones for k_filter (FUSED = false) or leaves them to the launch's filter
// tiles (FUSED = true: rows from the far-row ranges).  The reference's pole
// rows of u stay zero (diffusion_mod.F90:172-179).
template <bool FUSED>
__device__ __forceinline__ void dif_tile(const Args &a, int t, unsigned char *smem, double *red) {
  const int N = a.N, M = a.M, P = a.pitch, nseg = a.nseg, tid = threadIdx.x;
  const int nrows = FUSED ? a.nfar_rows : M;
  const int mloc = t / (nrows * nseg), rem = t - mloc * (nrows * nseg);
  const int r = rem / nseg, g = rem - (rem / nseg) * nseg;
  const int j = FUSED ? far_row(a, r) : r;
  const Member mb = member_at(a, a.m0 + mloc);
  const int i0 = g * kSeg, W = min(N - i0, kSeg);
  const RowCoef cj = a.coef[j];
  const double *gd = mb.st(a.s_in, FGD);
  const double *Ug = mb.st(a.s_in, FU);
  const double *Vg = mb.st(a.s_in, FV);
  const double *ug = mb.st(a.s_in, Fu), *vg = mb.st(a.s_in, Fv);
  const bool prev = a.lap_src != 0;
  // the state at the tile's own point for the final update when the operand
  // is a previous result (u, v derived from U, V exactly as update_state)
  double x_g = 0.0, x_u = 0.0, x_v = 0.0;
  if (prev && a.lap_final && tid < W) {
    const int i = i0 + tid, ip = i + 1 == N ? 0 : i + 1;
    const size_t o = (size_t)j * P + i;
    x_g = gd[o];
    if (j >= 1 && j <= M - 2)
      x_u = a.in_full ? ug[o] : Ug[o] * 2.0 / (sqrt(gd[o]) + sqrt(gd[o - i + ip]));
    if (j <= M - 2) x_v = a.in_full ? vg[o] : Vg[o] * 2.0 / (sqrt(gd[o]) + sqrt(gd[o + P]));
  }
  const DiffPoint d = dif_segment<TC_ALL>(a, mb, cj, j, i0, W, smem, red);
  if (tid >= W) return;
  const bool has_u = j >= 1 && j <= M - 2;
  const bool has_v = j <= M - 2;
  const int st_u = has_u ? cj.stage[CDU] : -1;
  const int st_v = has_v ? cj.stage[CDV] : -1;
  const int st_g = cj.stage[CDGD];
  const double kd = a.kdiff;
  const int i = i0 + tid;
  const size_t o = (size_t)j * P;
  if (!a.lap_final) {
    mb.st(a.s_lapout, FGD)[o + i] = d.gdd;
    mb.st(a.s_lapout, Fu)[o + i] = d.ud;
    if (has_v) mb.st(a.s_lapout, Fv)[o + i] = d.vd;
    return;
  }
  // the operand's own values (the state s_in when lap_src = 0)
  double *sgd1 = (double *)smem;  // gd row j of the operand at slot li
  double(*su)[kDW] = (double(*)[kDW])(smem + 3 * kDW * sizeof(double));
  double(*sv)[kDW] = (double(*)[kDW])(smem + 6 * kDW * sizeof(double));
  const int li = tid + 2;
  const double sg = prev ? x_g : sgd1[kDW + li];
  const double su0 = prev ? x_u : su[1][li];
  const double sv0 = prev ? x_v : sv[1][li];
  double *sx = mb.sx();
  double *Xgd = mb.st(SX, FGD), *Xu = mb.st(SX, Fu), *Xv = mb.st(SX, Fv);
  if (st_g >= 0) {
    if (!FUSED) sx[(size_t)st_g * P + i] = d.gdd;
  } else {
    Xgd[o + i] = sg + kd * d.gdd;
  }
  if (st_u >= 0) {
    if (!FUSED) sx[(size_t)st_u * P + i] = d.ud;
  } else {
    Xu[o + i] = has_u ? su0 + kd * d.ud : 0.0;
  }
  if (has_v) {
    if (st_v >= 0) {
      if (!FUSED) sx[(size_t)st_v * P + i] = d.vd;
    } else {
      Xv[o + i] = sv0 + kd * d.vd;
    }
  }
}

}  // namespace gd
