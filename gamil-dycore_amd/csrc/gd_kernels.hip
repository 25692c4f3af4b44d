// gd_kernels.hip -- CDNA4 (gfx950) kernels of the barotropic time step.
//
// Row kernels: one workgroup (256 threads, 4 wave64) per (member, latitude
// row); tiles are remapped so each XCD owns a contiguous band of rows and
// the +-1 row stencil reuse stays inside that XCD's L2.  Derived fields
// (G = sqrt(gd), u, v) of the rows a tile needs are built once in LDS.
//
// Filter kernels: the polar zonal filter (filter_mod.F90) as a batched
// truncated real-DFT product on MFMA, C = X B^T (coefficients) then
// Y = C' B (reconstruction), over tiles of 8 staged rows.
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
#include "gd_launch.hpp"

namespace gd {

typedef double d4_t __attribute__((ext_vector_type(4)));
typedef float f4_t __attribute__((ext_vector_type(4)));

// ---------------------------------------------------------------------------
// helpers
// ---------------------------------------------------------------------------

__device__ __forceinline__ int xcd_tile(int ntiles) {
  const int b = blockIdx.x, nb = gridDim.x;  // nb is a multiple of kXcd
  const int per = nb / kXcd;
  const int t = (b % kXcd) * per + b / kXcd;
  return t < ntiles ? t : -1;
}

// Deterministic workgroup sum (fixed butterfly + fixed cross-wave order);
// every thread returns the same value.
__device__ __forceinline__ double block_sum(double v, double *red) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  const int w = threadIdx.x >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[w] = v;
  __syncthreads();
  double s = red[0];
  for (int k = 1; k < kBlock / 64; ++k) s += red[k];
  return s;
}

// Sum over the 16 lanes of a lane group (lane % 16 varies), deterministic.
__device__ __forceinline__ double group16_sum(double v) {
  for (int off = 8; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

__device__ __forceinline__ double fld_base(const Member &mb, int slot, int comp, size_t o) {
  return mb.st[slot][comp == CDU ? FU : (comp == CDV ? FV : FGD)][o];
}

template <typename CT>
__device__ __forceinline__ double tend_load(const Member &mb, bool tb, int comp, size_t o) {
  void *p = tb ? mb.tb[comp] : mb.tc[comp];
  if (comp == CDGD) return ((const double *)p)[o];
  return (double)((const CT *)p)[o];
}

template <typename CT>
__device__ __forceinline__ void tend_store(const Member &mb, bool tb, int comp, size_t o, double v) {
  void *p = tb ? mb.tb[comp] : mb.tc[comp];
  if (comp == CDGD)
    ((double *)p)[o] = v;
  else
    ((CT *)p)[o] = (CT)v;
}

// Epilogue of one tendency value of a predict/correct phase (not diffusion).
// Returns the (Tb*T, T*T) products of phase C for the beta partials.
template <typename CT>
__device__ __forceinline__ void phase_epilogue(const Args &a, const Member &mb, int comp, int row,
                                               int i, double T, double w, double &p1, double &p2) {
  const size_t o = (size_t)row * a.pitch + i;
  const int f = comp == CDU ? FU : (comp == CDV ? FV : FGD);
  if (a.phase == PH_A || a.phase == PH_B) {
    mb.st[a.s_out][f][o] = mb.st[a.s0][f][o] + a.hdt * T;
    if (a.phase == PH_B) tend_store<CT>(mb, true, comp, o, T);
  } else {
    tend_store<CT>(mb, false, comp, o, T);
    if (a.phase == PH_C) {
      double tbv = tend_load<CT>(mb, true, comp, o);
      p1 += tbv * T * w;
      p2 += T * T * w;
    }
  }
}

// ---------------------------------------------------------------------------
// tendency (space_operators, dycore_mod.F90:184-598) + predictor epilogue
// ---------------------------------------------------------------------------

// Tile = (member, full row j, segment of kBlock longitudes); one thread per
// longitude.  LDS rows cover the segment plus a halo of two points so the
// zonal neighbours i-1, i+1 wrap periodically.
template <typename CT, int PASS, bool IN_FULL>
__global__ void __launch_bounds__(kBlock) k_tend(const Args a) {
  const int t = xcd_tile(a.ntiles);
  if (t < 0) return;
  const int N = a.N, M = a.M, P = a.pitch, nseg = a.nseg;
  const int mloc = t / (M * nseg), rem = t - mloc * (M * nseg);
  const int j = rem / nseg, g = rem - (rem / nseg) * nseg;
  const int m = a.m0 + mloc;
  const int i0 = g * kBlock, W = min(N - i0, kBlock);
  constexpr int LW = kBlock + 4;  // LDS index li = i - i0 + 2
  __shared__ double sG[4][LW];    // G rows j-1..j+2, points i0-1 .. i0+W+1
  __shared__ CT su[2][LW];        // u rows j, j+1, points i0-1 .. i0+W
  __shared__ CT sv[3][LW];        // v half rows j-1, j, j+1, points i0 .. i0+W
  __shared__ double red[16];
  const Member &mb = a.mem[m];
  const int tid = threadIdx.x;
  const double *gd = mb.st[a.s_in][FGD];
  const double *Ug = mb.st[a.s_in][FU];
  const double *Vg = mb.st[a.s_in][FV];
  const double *ghs = mb.ghs;
  auto wrap = [N](int i) { return i < 0 ? i + N : (i >= N ? i - N : i); };

#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int row = j - 1 + r;
    const bool ok = row >= 0 && row <= M - 1;
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      const int e = tid + k * kBlock;  // point i0-1+e
      if (ok && e < W + 3) sG[r][e + 1] = sqrt(gd[(size_t)row * P + wrap(i0 - 1 + e)]);
    }
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < 2; ++r) {
    const int row = j + r;
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      const int e = tid + k * kBlock;  // point i0-1+e, li = e+1
      if (row > M - 1 || e >= W + 2) continue;
      const int i = wrap(i0 - 1 + e);
      CT val;
      if (IN_FULL) {
        val = (CT)mb.st[a.s_in][Fu][(size_t)row * P + i];
      } else if (row == 0 || row == M - 1) {
        val = (CT)0;
      } else {
        // update_state: u = U * 2 / (G(i) + G(i+1))  (dycore_mod.F90:636-640)
        val = ((CT)Ug[(size_t)row * P + i] * (CT)2) / ((CT)sG[r + 1][e + 1] + (CT)sG[r + 1][e + 2]);
      }
      su[r][e + 1] = val;
    }
  }
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    const int hrow = j - 1 + r;
    const bool ok = hrow >= 0 && hrow <= M - 2;
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      const int e = tid + k * kBlock;  // point i0+e, li = e+2
      if (e > W) continue;
      CT val = (CT)0;
      if (ok) {
        const int i = wrap(i0 + e);
        if (IN_FULL)
          val = (CT)mb.st[a.s_in][Fv][(size_t)hrow * P + i];
        else  // v = V * 2 / (G(j) + G(j+1))  (dycore_mod.F90:641-645)
          val = ((CT)Vg[(size_t)hrow * P + i] * (CT)2) / ((CT)sG[r][e + 2] + (CT)sG[r + 1][e + 2]);
      }
      sv[r][e + 2] = val;
    }
  }
  __syncthreads();

  const RowCoef cj = a.coef[j];
  const bool has_u = j >= 1 && j <= M - 2;
  const bool has_v = j <= M - 2;
  const double hc_m = j >= 1 ? a.coef[j - 1].hcos : 0.0;
  const double hc_0 = has_v ? cj.hcos : 0.0;
  const double hc_p = j + 1 <= M - 2 ? a.coef[j + 1].hcos : 0.0;
  const double f_p = j + 1 <= M - 1 ? a.coef[j + 1].f : 0.0;
  const double c_p = j + 1 <= M - 1 ? a.coef[j + 1].c : 0.0;
  const CT bl = (CT)a.beta_lon, bt = (CT)a.beta_lat;
  const bool upwind = a.upwind != 0;

  // pole mass flux sums over the whole ring (dycore_mod.F90:572-596); every
  // segment of a pole row forms the same sum in the same order
  double pole = 0.0;
  if (PASS != 2 && (j == 0 || j == M - 1)) {
    double acc = 0.0;
    const int jn = j == 0 ? 1 : M - 2, hr = j == 0 ? 0 : M - 2;
    for (int i = tid; i < N; i += kBlock)
      acc += (sqrt(gd[(size_t)j * P + i]) + sqrt(gd[(size_t)jn * P + i])) * Vg[(size_t)hr * P + i];
    double sum = block_sum(acc, red);
    if (j == M - 1) sum = -sum;
    pole = sum * 2.0 / N / kRadius / a.dlat;
  }

  const int st_du = has_u ? cj.stage[CDU] : -1;
  const int st_dv = has_v ? cj.stage[CDV] : -1;
  const int st_dgd = cj.stage[CDGD];
  double p1[3] = {0, 0, 0}, p2[3] = {0, 0, 0};
  double *sx = mb.sx;

  if (tid < W) {
    const int li = tid + 2;
    const int i = i0 + tid;
    const int im = i == 0 ? N - 1 : i - 1, ip = i == N - 1 ? 0 : i + 1;
    const size_t o = (size_t)j * P;
    // ---- u equation (full row j) ----
    double du = 0.0;
    if (has_u) {
      const CT u0 = su[0][li], um = su[0][li - 1], up = su[0][li + 1];
      CT ual = 0, uat = 0, fv = 0, upgf = 0;
      if (PASS != 1) {
        const CT U0 = (CT)Ug[o + i], Um = (CT)Ug[o + im], Up = (CT)Ug[o + ip];
        const CT u1 = u0 + um, u2 = u0 + up;
        const CT kl = (CT)cj.k_uadv_lon;
        if (!upwind)
          ual = kl * (u2 * Up - u1 * Um);
        else
          ual = kl * (u2 * (U0 + Up) - bl * fabs(u2) * (Up - U0) - u1 * (U0 + Um) +
                      bl * fabs(u1) * (U0 - Um) - (u2 - u1) * U0);
        const CT v1 = (sv[0][li] + sv[0][li + 1]) * (CT)hc_m;
        const CT v2 = (sv[1][li] + sv[1][li + 1]) * (CT)hc_0;
        const CT Ujp = (CT)Ug[o + P + i], Ujm = (CT)Ug[o - P + i];
        const CT kt = (CT)cj.k_uadv_lat;
        if (!upwind)
          uat = kt * (v2 * Ujp - v1 * Ujm);
        else
          uat = kt * (v2 * (U0 + Ujp) - bt * fabs(v2) * (Ujp - U0) - v1 * (U0 + Ujm) +
                      bt * fabs(v1) * (U0 - Ujm) - (v2 - v1) * U0);
      }
      if (PASS != 2) {
        const CT Vm0 = (CT)Vg[o - P + i], Vm1 = (CT)Vg[o - P + ip];
        const CT V00 = (CT)Vg[o + i], V01 = (CT)Vg[o + ip];
        fv = (CT)0.25 * ((CT)cj.f + (CT)cj.c * u0) * ((CT)cj.c1 * (Vm0 + Vm1) + (CT)cj.c2 * (V00 + V01));
        const double dphi = gd[o + ip] + ghs[o + ip] - gd[o + i] - ghs[o + i];
        upgf = (CT)0.5 * ((CT)sG[1][li] + (CT)sG[1][li + 1]) / (CT)cj.fdlon * (CT)dphi;
      }
      CT d;
      if (PASS == 0)
        d = -ual - uat + fv - upgf;
      else if (PASS == 1)
        d = fv - upgf;
      else
        d = -ual - uat;
      du = (double)d;
    }
    // ---- v equation (half row j) ----
    double dv = 0.0;
    if (has_v) {
      CT val = 0, vat = 0, fu = 0, vpgf = 0;
      if (PASS != 1) {
        const CT u1 = su[0][li - 1] + su[1][li - 1];
        const CT u2 = su[0][li] + su[1][li];
        const CT V0 = (CT)Vg[o + i], Vm = (CT)Vg[o + im], Vp = (CT)Vg[o + ip];
        const CT kl = (CT)cj.k_vadv_lon;
        if (!upwind)
          val = kl * (u2 * Vp - u1 * Vm);
        else
          val = kl * (u2 * (V0 + Vp) - bl * fabs(u2) * (Vp - V0) - u1 * (V0 + Vm) +
                      bl * fabs(u1) * (V0 - Vm) - (u2 - u1) * V0);
        const CT vj = sv[1][li];
        const CT v1 = vj * (CT)hc_0 + sv[0][li] * (CT)hc_m;
        const CT v2 = vj * (CT)hc_0 + sv[2][li] * (CT)hc_p;
        const CT Vjp = j + 1 <= M - 2 ? (CT)Vg[o + P + i] : (CT)0;
        const CT Vjm = j >= 1 ? (CT)Vg[o - P + i] : (CT)0;
        const CT kt = (CT)cj.k_vadv_lat;
        if (!upwind)
          vat = kt * (v2 * Vjp - v1 * Vjm);
        else
          vat = kt * (v2 * (V0 + Vjp) - bt * fabs(v2) * (Vjp - V0) - v1 * (V0 + Vjm) +
                      bt * fabs(v1) * (V0 - Vjm) - (v2 - v1) * V0);
      }
      if (PASS != 2) {
        const CT f0 = (CT)cj.f, c0 = (CT)cj.c, fp = (CT)f_p, cp = (CT)c_p;
        const CT a1 = (f0 + c0 * su[0][li]) * (CT)Ug[o + i];
        const CT a2 = (f0 + c0 * su[0][li - 1]) * (CT)Ug[o + im];
        const CT a3 = (fp + cp * su[1][li]) * (CT)Ug[o + P + i];
        const CT a4 = (fp + cp * su[1][li - 1]) * (CT)Ug[o + P + im];
        fu = (CT)0.25 * (a1 + a2 + a3 + a4);
        const double dphi = gd[o + P + i] + ghs[o + P + i] - gd[o + i] - ghs[o + i];
        vpgf = (CT)0.5 * ((CT)sG[1][li] + (CT)sG[2][li]) / (CT)cj.hdlat * (CT)cj.hcos * (CT)dphi;
      }
      CT d;
      if (PASS == 0)
        d = -val - vat - fu - vpgf;
      else if (PASS == 1)
        d = -fu - vpgf;
      else
        d = -val - vat;
      dv = (double)d;
    }
    // ---- mass continuity (full row j), FP64 in every mode ----
    double dgd = 0.0;
    if (PASS != 2) {
      double mdl = 0.0, mdp;
      if (j == 0 || j == M - 1) {
        mdp = pole;
      } else {
        const double Gi = sG[1][li], Gip = sG[1][li + 1], Gim = sG[1][li - 1];
        mdl = ((Gi + Gip) * Ug[o + i] - (Gi + Gim) * Ug[o + im]) * 0.5 / cj.fdlon;
        mdp = ((Gi + sG[2][li]) * Vg[o + i] * hc_0 - (Gi + sG[0][li]) * Vg[o - P + i] * hc_m) * 0.5 / cj.fdlat;
      }
      dgd = -mdl - mdp;
    }
    // ---- filter staging or epilogue ----
    if (st_du >= 0)
      sx[(size_t)st_du * P + i] = du;
    else
      phase_epilogue<CT>(a, mb, CDU, j, i, du, cj.fcos, p1[0], p2[0]);
    if (has_v) {
      if (st_dv >= 0)
        sx[(size_t)st_dv * P + i] = dv;
      else
        phase_epilogue<CT>(a, mb, CDV, j, i, dv, cj.hcos, p1[1], p2[1]);
    }
    if (st_dgd >= 0)
      sx[(size_t)st_dgd * P + i] = dgd;
    else
      phase_epilogue<CT>(a, mb, CDGD, j, i, dgd, cj.fcos, p1[2], p2[2]);
  }
  if (a.phase == PH_C) {
    for (int c = 0; c < 3; ++c) {
      const bool staged = c == CDU ? st_du >= 0 : (c == CDV ? st_dv >= 0 : st_dgd >= 0);
      if (staged || (c == CDV && !has_v)) continue;
      double s1v = block_sum(p1[c], red);
      double s2v = block_sum(p2[c], red);
      if (tid == 0) {
        double *pp = mb.part + ((size_t)(j * 3 + c) * a.nchunk_row + i0 / kChunk) * 2;
        pp[0] = s1v;
        pp[1] = s2v;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// beta of predict_correct (dycore_mod.F90:784-787), one workgroup per member
// ---------------------------------------------------------------------------

__global__ void __launch_bounds__(kBlock) k_beta(const Args a) {
  const int m = a.m0 + blockIdx.x;
  const Member &mb = a.mem[m];
  __shared__ double red[16];
  const int tid = threadIdx.x;
  const int M = a.M, nc = a.nchunk_row;
  // inner_product(tend(old), tend(new)) and (tend(new), tend(new)):
  // du rows, then dv rows, then dgd rows (types_mod.F90:347-371)
  double ip1 = 0, ip2 = 0;
  const int per_comp = M * nc;
  for (int c = 0; c < 3; ++c) {
    double q1 = 0, q2 = 0;
    for (int k0 = 0; k0 < per_comp; k0 += 4 * kBlock) {
      double x1[4], x2[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int k = k0 + u * kBlock + tid;
        x1[u] = x2[u] = 0.0;
        if (k < per_comp) {
          const int row = k / nc, ch = k - (k / nc) * nc;
          const double *pp = mb.part + ((size_t)(row * 3 + c) * nc + ch) * 2;
          x1[u] = pp[0];
          x2[u] = pp[1];
        }
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        q1 += x1[u];
        q2 += x2[u];
      }
    }
    ip1 += block_sum(q1, red);
    ip2 += block_sum(q2, red);
  }
  if (tid == 0) mb.beta[0] = (a.qcon && ip1 != 0.0 && ip2 != 0.0) ? ip1 / ip2 : 1.0;
}

// ---------------------------------------------------------------------------
// corrector update (predict_correct, dycore_mod.F90:789-790), pointwise
// ---------------------------------------------------------------------------

template <typename CT>
__global__ void __launch_bounds__(kBlock) k_correct(const Args a) {
  const int t = xcd_tile(a.ntiles);
  if (t < 0) return;
  const int N = a.N, M = a.M, P = a.pitch, nseg = a.nseg;
  const int mloc = t / (M * nseg), rem = t - mloc * (M * nseg);
  const int j = rem / nseg, g = rem - (rem / nseg) * nseg;
  const Member &mb = a.mem[a.m0 + mloc];
  const int i = g * kBlock + threadIdx.x;
  if (i >= N) return;
  const double dtb = a.dt * mb.beta[0];
  const size_t o = (size_t)j * P + i;
  mb.st[a.s_out][FGD][o] = mb.st[a.s0][FGD][o] + dtb * ((const double *)mb.tc[CDGD])[o];
  mb.st[a.s_out][FU][o] = mb.st[a.s0][FU][o] + dtb * (double)((const CT *)mb.tc[CDU])[o];
  if (j <= M - 2) mb.st[a.s_out][FV][o] = mb.st[a.s0][FV][o] + dtb * (double)((const CT *)mb.tc[CDV])[o];
}

// ---------------------------------------------------------------------------
// ordinary_diffusion (diffusion_mod.F90:74-217), order 2, FP64 in every mode
// ---------------------------------------------------------------------------

template <typename CT>
__global__ void __launch_bounds__(kBlock) k_diffuse(const Args a) {
  const int t = xcd_tile(a.ntiles);
  if (t < 0) return;
  const int N = a.N, M = a.M, P = a.pitch;
  const int m = a.m0 + t / M, j = t % M;
  const Member &mb = a.mem[m];
  extern __shared__ __align__(16) unsigned char smem[];
  double *sG = (double *)smem;   // G rows j-1..j+2
  double *su = sG + 4 * N;       // u rows j-1, j, j+1
  double *sv = su + 3 * N;       // v half rows j-1, j, j+1
  double *red = sv + 3 * N;      // 16
  const int tid = threadIdx.x;
  const double *gd = mb.st[a.s_in][FGD];
  const double *Ug = mb.st[a.s_in][FU];
  const double *Vg = mb.st[a.s_in][FV];
  for (int r = 0; r < 4; ++r) {
    const int row = j - 1 + r;
    if (row < 0 || row > M - 1) continue;
    for (int i = tid; i < N; i += kBlock) sG[r * N + i] = sqrt(gd[(size_t)row * P + i]);
  }
  __syncthreads();
  const double *ug = mb.st[a.s_in][Fu], *vg = mb.st[a.s_in][Fv];
  for (int r = 0; r < 3; ++r) {
    const int row = j - 1 + r;
    for (int i = tid; i < N; i += kBlock) {
      double val = 0.0;
      if (row >= 1 && row <= M - 2) {
        const int ip = i + 1 == N ? 0 : i + 1;
        val = a.in_full ? ug[(size_t)row * P + i]
                        : Ug[(size_t)row * P + i] * 2.0 / (sG[r * N + i] + sG[r * N + ip]);
      }
      su[r * N + i] = val;
    }
    const int hrow = row;
    for (int i = tid; i < N; i += kBlock) {
      double val = 0.0;
      if (hrow >= 0 && hrow <= M - 2)
        val = a.in_full ? vg[(size_t)hrow * P + i]
                        : Vg[(size_t)hrow * P + i] * 2.0 / (sG[r * N + i] + sG[(r + 1) * N + i]);
      sv[r * N + i] = val;
    }
  }
  __syncthreads();
  const RowCoef cj = a.coef[j];
  const bool has_u = j >= 1 && j <= M - 2;
  const bool has_v = j <= M - 2;
  const double hc_m = j >= 1 ? a.coef[j - 1].hcos : 0.0;
  const double hc_0 = has_v ? cj.hcos : 0.0;
  const double fc_p = j + 1 <= M - 1 ? a.coef[j + 1].fcos : 0.0;
  // pole rows of gdd: zonal sums (diffusion_mod.F90:116-133)
  double pole = 0.0;
  if (j == 0 || j == M - 1) {
    double acc = 0.0;
    const int ja = j == 0 ? 1 : M - 1, jb = j == 0 ? 0 : M - 2;
    for (int i = tid; i < N; i += kBlock) acc += gd[(size_t)ja * P + i] - gd[(size_t)jb * P + i];
    double s = block_sum(acc, red);
    if (j == 0)
      pole = s * hc_0 / cj.fdlat2 * cj.fcos / N;
    else
      pole = -s * hc_m / cj.fdlat2 * cj.fcos / N;
  }
  const int st_u = has_u ? cj.stage[CDU] : -1;
  const int st_v = has_v ? cj.stage[CDV] : -1;
  const int st_g = cj.stage[CDGD];
  double *sx = mb.sx;
  const double kd = a.kdiff;
  double *Xgd = mb.st[SX][FGD], *Xu = mb.st[SX][Fu], *Xv = mb.st[SX][Fv];
  for (int i = tid; i < N; i += kBlock) {
    const int im = i == 0 ? N - 1 : i - 1, ip = i == N - 1 ? 0 : i + 1;
    const size_t o = (size_t)j * P;
    double gdd;
    if (j == 0 || j == M - 1) {
      gdd = pole;
    } else {
      const double g0 = gd[o + i];
      gdd = (gd[o + ip] - 2 * g0 + gd[o + im]) / cj.fdlon2 +
            ((gd[o + P + i] - g0) * hc_0 - (g0 - gd[o - P + i]) * hc_m) / cj.fdlat2 * cj.fcos;
    }
    double ud = 0.0;
    if (has_u) {
      const double u0 = su[N + i];
      ud = (su[N + ip] - 2 * u0 + su[N + im]) / cj.fdlon2 +
           ((su[2 * N + i] - u0) * hc_0 - (u0 - su[i]) * hc_m) / cj.fdlat2 * cj.fcos;
    }
    double vd = 0.0;
    if (has_v) {
      const double v0 = sv[N + i];
      vd = (sv[N + ip] - 2 * v0 + sv[N + im]) / cj.hdlon2;
      if (j >= 1 && j <= M - 3)
        vd = vd + ((sv[2 * N + i] - v0) * fc_p - (v0 - sv[i]) * cj.fcos) / cj.hdlat2 * cj.hcos;
      else if (j == 0)
        vd = vd + (sv[2 * N + i] - v0) * fc_p / cj.hdlat2 * cj.hcos;
      else
        vd = vd - (v0 - sv[i]) * cj.fcos / cj.hdlat2 * cj.hcos;
    }
    if (st_g >= 0)
      sx[(size_t)st_g * P + i] = gdd;
    else
      Xgd[o + i] = gd[o + i] + kd * gdd;
    if (st_u >= 0)
      sx[(size_t)st_u * P + i] = ud;
    else
      Xu[o + i] = has_u ? su[N + i] + kd * ud : 0.0;
    if (has_v) {
      if (st_v >= 0)
        sx[(size_t)st_v * P + i] = vd;
      else
        Xv[o + i] = sv[N + i] + kd * vd;
    }
  }
}

// ---------------------------------------------------------------------------
// end of step: iap_transform (types_mod.F90:399-426) or derivation of u, v,
// then diag_run mass/energy (diag_mod.F90:42-121)
// ---------------------------------------------------------------------------

template <int MODE>  // 0: iap_transform after diffusion, 1: derive u,v from an IAP slot, 2: diag only
__global__ void __launch_bounds__(kBlock) k_finish(const Args a) {
  const int t = xcd_tile(a.ntiles);
  if (t < 0) return;
  const int N = a.N, M = a.M, P = a.pitch;
  const int m = a.m0 + t / M, j = t % M;
  const Member &mb = a.mem[m];
  __shared__ double red[16];
  __shared__ int last;
  const int tid = threadIdx.x;
  const RowCoef cj = a.coef[j];
  const size_t o = (size_t)j * P;
  double *Xgd = mb.st[SX][FGD], *XU = mb.st[SX][FU], *XV = mb.st[SX][FV];
  double *Xu = mb.st[SX][Fu], *Xv = mb.st[SX][Fv];
  const double *sgd = MODE != 1 ? Xgd : mb.st[a.s_in][FGD];
  const double *sU = mb.st[a.s_in][FU], *sV = mb.st[a.s_in][FV];
  double em = 0, eu = 0, ev = 0, eg = 0;
  for (int i = tid; i < N; i += kBlock) {
    const int ip = i == N - 1 ? 0 : i + 1;
    const double g0 = sgd[o + i];
    const double G0 = sqrt(g0), G1 = sqrt(sgd[o + ip]);
    const double Gn = j <= M - 2 ? sqrt(sgd[o + P + i]) : 0.0;
    double U, V = 0.0;
    if (MODE == 2) {
      U = XU[o + i];
      if (j <= M - 2) V = XV[o + i];
    } else if (MODE == 0) {
      U = 0.5 * (G0 + G1) * Xu[o + i];
      XU[o + i] = U;
      if (j <= M - 2) {
        V = 0.5 * (G0 + Gn) * Xv[o + i];
        XV[o + i] = V;
      }
    } else {
      U = sU[o + i];
      Xgd[o + i] = g0;
      XU[o + i] = U;
      Xu[o + i] = (j == 0 || j == M - 1) ? 0.0 : U * 2.0 / (G0 + G1);
      if (j <= M - 2) {
        V = sV[o + i];
        XV[o + i] = V;
        Xv[o + i] = V * 2.0 / (G0 + Gn);
      }
    }
    em += cj.mass_w * g0;
    if (j >= 1 && j <= M - 2) eu += U * U * cj.fcos;
    if (j <= M - 2) ev += V * V * cj.hcos;
    const double h = g0 + mb.ghs[o + i];
    eg += h * h * cj.fcos;
  }
  em = block_sum(em, red);
  eu = block_sum(eu, red);
  ev = block_sum(ev, red);
  eg = block_sum(eg, red);
  if (tid == 0) {
    double *dp = mb.dpart + (size_t)j * 4;
    dp[0] = em; dp[1] = eu; dp[2] = ev; dp[3] = eg;
    __threadfence();
    const uint32_t prev = atomicAdd(&a.done[m], 1u);
    last = prev == (uint32_t)(M - 1);
  }
  __syncthreads();
  if (!last) return;
  __threadfence();
  double q[4] = {0, 0, 0, 0};
  for (int r = tid; r < M; r += kBlock)
    for (int k = 0; k < 4; ++k) q[k] += __builtin_nontemporal_load(mb.dpart + (size_t)r * 4 + k);
  for (int k = 0; k < 4; ++k) q[k] = block_sum(q[k], red);
  if (tid == 0) {
    const int s = a.step_ctr[m];
    double *out = mb.diag_out + (size_t)(s % a.diag_cap) * 3;
    out[0] = q[0] * (kRadius * kRadius);
    out[1] = q[1] + q[2] + q[3];
    out[2] = mb.beta[0];
    a.step_ctr[m] = s + 1;
    a.done[m] = 0;
  }
}

// diag_run div and vor of state X (diag_mod.F90:49-77), history only
__global__ void __launch_bounds__(kBlock) k_divvor(const Args a) {
  const int t = xcd_tile(a.ntiles);
  if (t < 0) return;
  const int N = a.N, M = a.M, P = a.pitch;
  const int m = a.m0 + t / M, j = t % M;
  const Member &mb = a.mem[m];
  const RowCoef cj = a.coef[j];
  const double *u = mb.st[SX][Fu], *v = mb.st[SX][Fv];
  const size_t o = (size_t)j * P;
  for (int i = threadIdx.x; i < N; i += kBlock) {
    const int im = i == 0 ? N - 1 : i - 1, ip = i == N - 1 ? 0 : i + 1;
    if (j >= 1 && j <= M - 2) {
      const double um1 = u[o + im], up1 = u[o + i];
      const double vm1 = v[o - P + i] * a.coef[j - 1].hcos, vp1 = v[o + i] * cj.hcos;
      mb.div[o + i] = (up1 - um1) / cj.fdlon + (vp1 - vm1) / cj.fdlat;
    } else {
      mb.div[o + i] = 0.0;
    }
    if (j <= M - 2) {
      const double um1 = u[o + i], up1 = u[o + P + i];
      const double vm1 = v[o + i], vp1 = v[o + ip];
      mb.vor[o + i] = (vp1 - vm1) / cj.hdlon - (up1 - um1) / cj.hdlat;
    }
  }
}

// ---------------------------------------------------------------------------
// polar filter on MFMA
// ---------------------------------------------------------------------------

template <typename CT> struct Mfma;
template <> struct Mfma<double> {
  typedef d4_t acc_t;
  static __device__ __forceinline__ acc_t run(double x, double y, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(x, y, c, 0, 0, 0);
  }
};
template <> struct Mfma<float> {
  typedef f4_t acc_t;
  static __device__ __forceinline__ acc_t run(float x, float y, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(x, y, c, 0, 0, 0);
  }
};

// Accumulator layout of the 16x16x4 MFMAs: lane l (q = l/16) register r holds
// D[acc_row(q, r)][l % 16].  The FP64 form interleaves rows (q + 4r), the
// FP32 form blocks them (4q + r); checked on the device by k_mfma_layout.
template <typename CT> __device__ __forceinline__ int acc_row(int q, int r);
template <> __device__ __forceinline__ int acc_row<double>(int q, int r) { return q + 4 * r; }
template <> __device__ __forceinline__ int acc_row<float>(int q, int r) { return 4 * q + r; }

// FFTPACK5 real basis of slot s at longitude index l (0 outside [0,N)).
template <typename CT>
__device__ __forceinline__ CT basis(const CT *ctab, const CT *stab, int N, int s, int l) {
  if (s >= N || l >= N) return (CT)0;
  if (s == 0) return (CT)1;
  const int k = (s + 1) >> 1;
  const int q = (k * l) % N;
  return (s & 1) ? ctab[q] : stab[q];
}

__device__ __forceinline__ double slot_weight(int N, int s) {
  return (s == 0 || ((N & 1) == 0 && s == N - 1)) ? 1.0 / N : 2.0 / N;
}

// F1: coefficients C = X B^T over one tile x one chunk of 64 slots, masked
// and weighted; also the s2 partial sum_s w_s C_x[s] C_w[s] of each row, which
// equals <P x, w> because the projection P is symmetric.  Operands are staged
// through LDS in 64-wide blocks: the tile's 8 staged rows and their 8 weight
// rows (16 x 64) and the basis block Bt[l][s] (64 x 64).
template <typename CT>
__device__ __forceinline__ void filter_coef_body(const Args &a, const FiltTile *ft, int tile, int chunk,
                                                 const CT *Bt, unsigned char *smem) {
  const int N = a.N, P = a.pitch;
  const int slo = ft->lo + chunk * kChunk;
  const int hi = ft->hi, lo = ft->lo, nrows = ft->nrows, comp_mode = ft->complement;
  constexpr int LB = 64, BP = 65;
  const double **rp = (const double **)smem;  // [16] source row, [16] second term (ghs)
  const double **rq = rp + 16;
  int *kept = (int *)(rq + 16);               // [8]
  CT *xs = (CT *)(smem + 512);                // [16][BP]
  CT *bs = xs + 16 * BP;                      // [64][BP]
  double *xd = (double *)(smem + 512 + (size_t)(16 + 64) * BP * 8);  // [16][64] FP64 copy (chunk 0)
  __shared__ double s2w[4][kFiltRows];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const bool want_w = a.phase <= PH_T;
  if (tid < 16) {
    const int rr = tid & 7;
    const double *p = nullptr, *q = nullptr;
    if (rr < nrows) {
      const Member &mb = a.mem[ft->member[rr]];
      const StageSlot ss = a.slots[ft->slot[rr]];
      if (tid < 8) {
        p = mb.sx + (size_t)ft->slot[rr] * P;
        kept[rr] = ss.kept;
      } else if (want_w) {
        const size_t o = (size_t)ss.row * P;
        p = mb.st[a.s_in][ss.comp == CDU ? FU : (ss.comp == CDV ? FV : FGD)] + o;
        q = ss.comp == CDGD ? mb.ghs + o : nullptr;
      }
    } else if (tid < 8) {
      kept[rr] = 0;
    }
    rp[tid] = p;
    rq[tid] = q;
  }
  typename Mfma<CT>::acc_t acc = {0, 0, 0, 0};
  const int s_mine = slo + wave * 16 + (lane & 15);
  const bool s_ok = s_mine < hi;
  // register double buffer: each thread owns 4 values of the 16 x 64 operand
  // block (row tid/16, columns 4*(tid%16)..+3) and 16 values of the 64 x 64
  // basis block (row tid/4, columns 16*(tid%4)..+15); the loads of block b+1
  // are in flight while the MFMAs consume block b from LDS.
  const int xr = tid >> 4, xc = (tid & 15) * 4;
  const int br = tid >> 2, bc = (tid & 3) * 16;
  __syncthreads();  // row pointers visible
  const double *xp = rp[xr], *xq = rq[xr];
  double xv[4];
  CT bv[16];
  auto load_block = [&](int l0) {
    const int l = l0 + xc;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      double v = 0.0;
      if (xp && l + k < N) v = xq ? xp[l + k] + xq[l + k] : xp[l + k];
      xv[k] = v;
    }
    const int lb = l0 + br;
    if (lb < N) {
      const CT *src = Bt + (size_t)lb * a.spad + slo + bc;
#pragma unroll
      for (int k = 0; k < 16; ++k) bv[k] = src[k];
    } else {
#pragma unroll
      for (int k = 0; k < 16; ++k) bv[k] = (CT)0;
    }
  };
  // chunk 0 also forms s1 = <x, w> in FP64, the row inner product before
  // filtering (dycore_mod.F90:213)
  const bool do_s1 = chunk == 0 && want_w;
  double s1p = 0.0;
  load_block(0);
  for (int l0 = 0; l0 < N; l0 += LB) {
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 4; ++k) xs[xr * BP + xc + k] = (CT)xv[k];
#pragma unroll
    for (int k = 0; k < 16; ++k) bs[br * BP + bc + k] = bv[k];
    if (do_s1)
#pragma unroll
      for (int k = 0; k < 4; ++k) xd[xr * 64 + xc + k] = xv[k];
    __syncthreads();
    if (do_s1 && xr < 8)
#pragma unroll
      for (int k = 0; k < 4; ++k) s1p += xd[xr * 64 + xc + k] * xd[(xr + 8) * 64 + xc + k];
    if (l0 + LB < N) load_block(l0 + LB);
#pragma unroll 4
    for (int kk = 0; kk < LB; kk += 4) {
      const int k = kk + (lane >> 4);
      acc = Mfma<CT>::run(xs[(lane & 15) * BP + k], bs[k * BP + wave * 16 + (lane & 15)], acc);
    }
  }
  if (do_s1) {
    // the 16 threads of row xr are 16 consecutive lanes of one wave
    s1p = group16_sum(s1p);
    if ((tid & 15) == 0 && xr < nrows) a.mem[ft->member[xr]].s1[ft->slot[xr]] = s1p;
  }
  // acc[r] = C[acc_row(q, r)][slot s_mine]; weight row = staged row + 8
  const int qd = lane >> 4;
  double part[4];
  double *fc = a.fcoef + (size_t)tile * kFiltRows * a.kpad;
  const double ws = s_ok ? slot_weight(N, s_mine) : 0.0;
  double accd[4], accw[4];
  for (int r = 0; r < 4; ++r) accd[r] = (double)acc[r];
  for (int r = 0; r < 4; ++r) {
    if constexpr (sizeof(CT) == 8)
      accw[r] = accd[(r + 2) & 3];            // FP64: row + 8 is register r + 2
    else
      accw[r] = __shfl_xor(accd[r], 32, 64);  // FP32: row + 8 is 32 lanes away
  }
  for (int r = 0; r < 4; ++r) {
    const int row = acc_row<CT>(qd, r);  // 0..15
    part[r] = 0.0;
    if (row >= kFiltRows) continue;
    const int sl = s_mine - lo;
    if (sl >= a.kpad) continue;
    double cval = 0.0;
    if (row < nrows && s_ok) {
      const bool keep = comp_mode ? s_mine >= kept[row] : s_mine < kept[row];
      cval = keep ? ws * accd[r] : 0.0;
      part[r] = cval * accw[r];
    }
    fc[(size_t)row * a.kpad + sl] = cval;
  }
  for (int r = 0; r < 4; ++r) part[r] = group16_sum(part[r]);
  if ((lane & 15) == 0)
    for (int r = 0; r < 4; ++r) {
      const int row = acc_row<CT>(qd, r);
      if (row < kFiltRows) s2w[wave][row] = part[r];
    }
  __syncthreads();
  if (tid < kFiltRows) {
    const double sv = ((s2w[0][tid] + s2w[1][tid]) + s2w[2][tid]) + s2w[3][tid];
    a.fs2[((size_t)tile * a.nchunk1 + chunk) * kFiltRows + tid] = sv;
  }
}

template <typename CT>
__global__ void __launch_bounds__(kBlock) k_filter_coef(const Args a) {
  const int wi = blockIdx.x;
  if (wi >= a.ntiles) return;
  const int tile = wi / a.nchunk1, chunk = wi % a.nchunk1;
  const FiltTile *ft = a.tiles + tile;
  if (ft->lo + chunk * kChunk >= ft->hi) return;
  extern __shared__ __align__(16) unsigned char smem[];
  if (sizeof(CT) == 4 && ft->fp64)
    filter_coef_body<double>(a, ft, tile, chunk, (const double *)a.basis_t64, smem);
  else
    filter_coef_body<CT>(a, ft, tile, chunk, (const CT *)a.basis_t, smem);
}

// F2: reconstruction Y = C' B over one tile x 64 columns, the rescale
// tend*s1/s2 (dycore_mod.F90:213-217) and the phase epilogue.
template <typename CT, typename ST>
__device__ __forceinline__ void filter_apply_body(const Args &a, const FiltTile *ft, int tile, int chunk,
                                                  const CT *B, unsigned char *smem) {
  const int N = a.N, P = a.pitch;
  const int lo = ft->lo, hi = ft->hi, nrows = ft->nrows, comp_mode = ft->complement;
  constexpr int LB = 64, BP = 65;
  CT *bs = (CT *)smem;          // [64][BP] basis block B[s][l]
  CT *cs = bs + LB * BP;        // [8][kpad]
  __shared__ double red[4][2][kFiltRows];
  __shared__ double s2row[kFiltRows];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int ks = hi - lo;
  const double *fc = a.fcoef + (size_t)tile * kFiltRows * a.kpad;
  for (int e = tid; e < kFiltRows * a.kpad; e += kBlock) cs[e] = (CT)fc[e];
  if (tid < kFiltRows) {
    double sv = 0.0;
    const int nch = (ks + kChunk - 1) / kChunk;
    for (int c = 0; c < nch; ++c) sv += a.fs2[((size_t)tile * a.nchunk1 + c) * kFiltRows + tid];
    s2row[tid] = sv;
  }
  const int col0 = chunk * kChunk;
  const int col = col0 + wave * 16 + (lane & 15);
  typename Mfma<CT>::acc_t acc = {0, 0, 0, 0};
  // register double buffer of the 64 x 64 basis block (row tid/4, columns
  // 16*(tid%4)..+15), see k_filter_coef
  const int br = tid >> 2, bc = (tid & 3) * 16;
  CT bv[16];
  auto load_block = [&](int s0) {
    const int sidx = s0 + br;
    if (sidx < ks) {
      const CT *src = B + (size_t)(lo + sidx) * a.lpad + col0 + bc;
#pragma unroll
      for (int k = 0; k < 16; ++k) bv[k] = src[k];
    } else {
#pragma unroll
      for (int k = 0; k < 16; ++k) bv[k] = (CT)0;
    }
  };
  load_block(0);
  for (int s0 = 0; s0 < ks; s0 += LB) {
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 16; ++k) bs[br * BP + bc + k] = bv[k];
    __syncthreads();
    if (s0 + LB < ks) load_block(s0 + LB);
    const int kend = ks - s0 < LB ? ((ks - s0 + 3) & ~3) : LB;
    for (int kk = 0; kk < kend; kk += 4) {
      const int k = kk + (lane >> 4);
      const int row = lane & 15;
      const CT av = (row < kFiltRows && s0 + k < ks) ? cs[row * a.kpad + s0 + k] : (CT)0;
      acc = Mfma<CT>::run(av, bs[k * BP + wave * 16 + (lane & 15)], acc);
    }
  }
  __syncthreads();
  const int qd = lane >> 4;
  double p1[4] = {0, 0, 0, 0}, p2[4] = {0, 0, 0, 0};
  if (col < N) {
    for (int r = 0; r < 4; ++r) {
      const int row = acc_row<CT>(qd, r);
      if (row >= nrows) continue;
      const Member &mb = a.mem[ft->member[row]];
      const int slot = ft->slot[row];
      const StageSlot ss = a.slots[slot];
      const double x = mb.sx[(size_t)slot * P + col];
      const double y = comp_mode ? x - (double)acc[r] : (double)acc[r];
      double T;
      if (a.phase <= PH_T) {
        const double s1 = mb.s1[slot];
        const double s2 = comp_mode ? s1 - s2row[row] : s2row[row];
        T = fabs(s1) > 1.0e-16 ? y * s1 / s2 : x;
      } else {
        T = y;
      }
      const size_t o = (size_t)ss.row * P + col;
      if (a.phase <= PH_T) {
        const double w = ss.comp == CDV ? a.coef[ss.row].hcos : a.coef[ss.row].fcos;
        phase_epilogue<ST>(a, mb, ss.comp, ss.row, col, T, w, p1[r], p2[r]);
      } else if (a.phase == PH_D) {
        // ordinary_diffusion update of the filtered component
        const double *sgd = mb.st[a.s_in][FGD];
        if (ss.comp == CDGD) {
          mb.st[SX][FGD][o] = sgd[o] + a.kdiff * T;
        } else if (ss.comp == CDU) {
          const int ip = col + 1 == N ? 0 : col + 1;
          const double u = a.in_full ? mb.st[a.s_in][Fu][o]
                                     : mb.st[a.s_in][FU][o] * 2.0 / (sqrt(sgd[o]) + sqrt(sgd[o - col + ip]));
          mb.st[SX][Fu][o] = u + a.kdiff * T;
        } else {
          const double v = a.in_full ? mb.st[a.s_in][Fv][o]
                                     : mb.st[a.s_in][FV][o] * 2.0 / (sqrt(sgd[o]) + sqrt(sgd[o + P]));
          mb.st[SX][Fv][o] = v + a.kdiff * T;
        }
      } else {  // PH_P: projection only, result back into the staging row
        mb.sx[(size_t)slot * P + col] = T;
      }
    }
  }
  if (a.phase == PH_C) {
    for (int r = 0; r < 4; ++r) {
      p1[r] = group16_sum(p1[r]);
      p2[r] = group16_sum(p2[r]);
    }
    if ((lane & 15) == 0)
      for (int r = 0; r < 4; ++r) {
        const int row = acc_row<CT>(qd, r);
        if (row < kFiltRows) {
          red[wave][0][row] = p1[r];
          red[wave][1][row] = p2[r];
        }
      }
    __syncthreads();
    if (tid < nrows) {
      const Member &mb = a.mem[ft->member[tid]];
      const StageSlot ss = a.slots[ft->slot[tid]];
      double *pp = mb.part + ((size_t)(ss.row * 3 + ss.comp) * a.nchunk_row + chunk) * 2;
      pp[0] = ((red[0][0][tid] + red[1][0][tid]) + red[2][0][tid]) + red[3][0][tid];
      pp[1] = ((red[0][1][tid] + red[1][1][tid]) + red[2][1][tid]) + red[3][1][tid];
    }
  }
}

template <typename CT>
__global__ void __launch_bounds__(kBlock) k_filter_apply(const Args a) {
  const int wi = blockIdx.x;
  if (wi >= a.ntiles) return;
  const int tile = wi / a.nchunk2, chunk = wi % a.nchunk2;
  const FiltTile *ft = a.tiles + tile;
  extern __shared__ __align__(16) unsigned char smem[];
  if (sizeof(CT) == 4 && ft->fp64)
    filter_apply_body<double, CT>(a, ft, tile, chunk, (const double *)a.basis64, smem);
  else
    filter_apply_body<CT, CT>(a, ft, tile, chunk, (const CT *)a.basis, smem);
}

// MFMA operand-layout check used by the tests: D = A B for 16x4 A, 4x16 B.
__global__ void k_mfma_layout(const double *A, const double *B, double *D, const float *Af,
                              const float *Bf, float *Df) {
  const int lane = threadIdx.x;
  d4_t acc = {0, 0, 0, 0};
  acc = __builtin_amdgcn_mfma_f64_16x16x4f64(A[(lane & 15) * 4 + (lane >> 4)],
                                             B[(lane >> 4) * 16 + (lane & 15)], acc, 0, 0, 0);
  f4_t accf = {0, 0, 0, 0};
  accf = __builtin_amdgcn_mfma_f32_16x16x4f32(Af[(lane & 15) * 4 + (lane >> 4)],
                                              Bf[(lane >> 4) * 16 + (lane & 15)], accf, 0, 0, 0);
  for (int r = 0; r < 4; ++r) {
    D[acc_row<double>(lane >> 4, r) * 16 + (lane & 15)] = acc[r];
    Df[acc_row<float>(lane >> 4, r) * 16 + (lane & 15)] = accf[r];
  }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------

static inline int row_grid(int ntiles) { return ((ntiles + kXcd - 1) / kXcd) * kXcd; }

template <typename CT>
static hipError_t launch_tend_t(const Args &a, int pass, bool in_full, hipStream_t s) {
  dim3 g(row_grid(a.ntiles)), b(kBlock);
#define GD_TEND(P, F) hipLaunchKernelGGL((k_tend<CT, P, F>), g, b, 0, s, a)
  if (pass == 0) { if (in_full) GD_TEND(0, true); else GD_TEND(0, false); }
  else if (pass == 1) { if (in_full) GD_TEND(1, true); else GD_TEND(1, false); }
  else { if (in_full) GD_TEND(2, true); else GD_TEND(2, false); }
#undef GD_TEND
  return hipGetLastError();
}

hipError_t launch_tend(const Args &a, int prec, int pass, bool in_full, hipStream_t s) {
  return prec == 0 ? launch_tend_t<double>(a, pass, in_full, s) : launch_tend_t<float>(a, pass, in_full, s);
}

hipError_t launch_beta(const Args &a, int nmem, hipStream_t s) {
  hipLaunchKernelGGL(k_beta, dim3(nmem), dim3(kBlock), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_correct(const Args &a, int prec, hipStream_t s) {
  dim3 g(row_grid(a.ntiles)), b(kBlock);
  if (prec == 0)
    hipLaunchKernelGGL(k_correct<double>, g, b, 0, s, a);
  else
    hipLaunchKernelGGL(k_correct<float>, g, b, 0, s, a);
  return hipGetLastError();
}

hipError_t launch_diffuse(const Args &a, int prec, hipStream_t s) {
  dim3 g(row_grid(a.ntiles)), b(kBlock);
  const size_t lds = diffuse_lds_bytes(a.N);
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

hipError_t launch_divvor(const Args &a, hipStream_t s) {
  dim3 g(row_grid(a.ntiles)), b(kBlock);
  hipLaunchKernelGGL(k_divvor, g, b, 0, s, a);
  return hipGetLastError();
}

// Raise the dynamic-LDS limit of every kernel that may need more than the
// default (large num_lon).
hipError_t configure_kernels(int max_lds) {
  const void *fns[] = {
      (const void *)k_diffuse<double>,       (const void *)k_diffuse<float>,
      (const void *)k_filter_coef<double>,   (const void *)k_filter_coef<float>,
      (const void *)k_filter_apply<double>,  (const void *)k_filter_apply<float>};
  for (const void *f : fns) {
    hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, max_lds);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

hipError_t launch_filter_coef(const Args &a, int prec, hipStream_t s) {
  dim3 g(a.ntiles), b(kBlock);
  if (prec == 0)
    hipLaunchKernelGGL(k_filter_coef<double>, g, b, filt_coef_lds_bytes(), s, a);
  else
    hipLaunchKernelGGL(k_filter_coef<float>, g, b, filt_coef_lds_bytes(), s, a);
  return hipGetLastError();
}

hipError_t launch_filter_apply(const Args &a, int prec, hipStream_t s) {
  dim3 g(a.ntiles), b(kBlock);
  if (prec == 0)
    hipLaunchKernelGGL(k_filter_apply<double>, g, b, filt_apply_lds_bytes(a.kpad), s, a);
  else
    hipLaunchKernelGGL(k_filter_apply<float>, g, b, filt_apply_lds_bytes(a.kpad), s, a);
  return hipGetLastError();
}

hipError_t launch_mfma_layout(const double *A, const double *B, double *D, const float *Af,
                              const float *Bf, float *Df, hipStream_t s) {
  hipLaunchKernelGGL(k_mfma_layout, dim3(1), dim3(64), 0, s, A, B, D, Af, Bf, Df);
  return hipGetLastError();
}

}  // namespace gd
