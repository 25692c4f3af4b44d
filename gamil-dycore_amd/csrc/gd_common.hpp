// gd_common.hpp -- device helpers shared by the row kernels and the polar
// filter: XCD-aware tile mapping, deterministic workgroup sums, tendency
// loads/stores in the compute type and the predictor/corrector epilogue.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "gd_device.hpp"

namespace gd {

// ---------------------------------------------------------------------------
// helpers
// ---------------------------------------------------------------------------

__device__ __forceinline__ int xcd_tile(int ntiles) {
  const int b = blockIdx.x, nb = gridDim.x;  // nb is a multiple of kXcd
  const int per = nb / kXcd;
  const int t = (b % kXcd) * per + b / kXcd;
  return t < ntiles ? t : -1;
}

// Thread index within a tile of kBlock threads.  A workgroup of the wide
// phase launch holds several such tiles side by side (tile threadIdx.x /
// kBlock); in every other launch the workgroup is one tile.
__device__ __forceinline__ int ltid() { return (int)(threadIdx.x & (kBlock - 1)); }

// Deterministic sum over the kBlock threads of a tile (fixed butterfly +
// fixed cross-wave order); every thread of the tile returns the same value.
// The barriers are the workgroup's, so every tile of a workgroup calls it
// at the same point, each with its own `red`.
__device__ __forceinline__ double block_sum(double v, double *red) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  const int w = (threadIdx.x >> 6) & (kBlock / 64 - 1);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[w] = v;
  __syncthreads();
  double s = red[0];
  for (int k = 1; k < kBlock / 64; ++k) s += red[k];
  return s;
}

// Deterministic tile sums of K values at once (one pair of barriers);
// `red` holds 4*K doubles.  Every thread of the tile returns the same sums.
template <int K>
__device__ __forceinline__ void block_sum_n(double (&v)[K], double *red) {
#pragma unroll
  for (int q = 0; q < K; ++q)
    for (int off = 32; off > 0; off >>= 1) v[q] += __shfl_xor(v[q], off, 64);
  const int w = (threadIdx.x >> 6) & (kBlock / 64 - 1);
  __syncthreads();
  if ((threadIdx.x & 63) == 0)
#pragma unroll
    for (int q = 0; q < K; ++q) red[w * K + q] = v[q];
  __syncthreads();
#pragma unroll
  for (int q = 0; q < K; ++q) {
    double s = red[q];
    for (int k = 1; k < kBlock / 64; ++k) s += red[k * K + q];
    v[q] = s;
  }
}

__device__ __forceinline__ double fld_base(const Member &mb, int slot, int comp, size_t o) {
  return mb.st(slot, comp == CDU ? FU : (comp == CDV ? FV : FGD))[o];
}

// Stored tendencies Tb, Tc are FP64 in every precision mode: the FP32
// values of mixed mode widen exactly, filtered rows keep the FP64 rescale,
// and the wider loads measured faster than FP32 ones in the corrector.
__device__ __forceinline__ double tend_load(const Member &mb, bool tb, int comp, size_t o) {
  return (tb ? mb.tb(comp) : mb.tc(comp))[o];
}

__device__ __forceinline__ void tend_store(const Member &mb, bool tb, int comp, size_t o, double v) {
  (tb ? mb.tb(comp) : mb.tc(comp))[o] = v;
}

// The epilogue of one tendency value of a predict/correct phase (not
// diffusion), split in two so a kernel can issue the load early: the value
// phase_store needs (state s0 for the predictors, Tb for phase C) ...
template <typename CT>
__device__ __forceinline__ double phase_pre(const Args &a, const Member &mb, int comp, size_t o) {
  if (a.phase == PH_A || a.phase == PH_B) return fld_base(mb, a.s0, comp, o);
  if (a.phase == PH_C) return tend_load(mb, true, comp, o);
  return 0.0;
}

// ... and the stores given that value: S_out = s0 + dt/2 T for the
// predictors (phase B also keeps T as Tb), T as Tc for phase C with the
// (Tb T w, T T w) products of the beta partials.
template <typename CT>
__device__ __forceinline__ void phase_store(const Args &a, const Member &mb, int comp, size_t o, double T,
                                            double pre, double w, double &p1, double &p2) {
  const int f = comp == CDU ? FU : (comp == CDV ? FV : FGD);
  if (a.phase == PH_A || a.phase == PH_B) {
    mb.st(a.s_out, f)[o] = pre + a.hdt * T;
    if (a.phase == PH_B) tend_store(mb, true, comp, o, T);
  } else {
    tend_store(mb, false, comp, o, T);
    if (a.phase == PH_C) {
      p1 += pre * T * w;
      p2 += T * T * w;
    }
  }
}

}  // namespace gd
