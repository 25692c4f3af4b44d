// gd_filter_mm.hip -- the polar filter of the momentum rows as a batched
// dense product on the matrix cores (MFMA), for the mixed precision modes
// that ask for it (mixed_mfma: FP32 inputs; mixed_bf16: BF16 inputs, each
// operand split into a high and a low BF16 part and the product formed as
// hi.hi + hi.lo + lo.hi, since one BF16 product (8-bit mantissas) left a
// 5e-3 error in the projection that the inner-product rescale amplifies;
// both accumulate in FP32).
//
// filter_array_at_full_lat / _half_lat (filter_mod.F90:105-167) is
// FFTPACK5's forward transform, the 0/1 mask of the row and the backward
// transform (rfftf1.f:87-108).  Its result is the orthogonal projection
//   y = sum_j w_j phi_j (phi_j . x)
// onto the kept basis of the row's cutoff c: phi_0 = 1, phi_{2k-1} =
// cos(k theta), phi_{2k} = sin(k theta) for k = 1..c, and phi_{2c+1} =
// cos((c+1) theta) (the mask ends on a_{c+1}); w = 1/N for phi_0 and for the
// Nyquist cosine (2(c+1) = N), 2/N otherwise.  Every member's row of one
// staging slot (one latitude, one component) shares that projector, so the
// members' rows form the batch of two products:
//   C = X Phi^T   ([members x N] . [N x K]),   Y = (C W) Phi   ([members x K] . [K x N]).
// One workgroup takes one slot and a block of 16 members; the 16 x N rows
// sit in LDS, the basis operands are read from matrices built once per
// handle (Phi[j][n] for the forward product, Phi^T[n][j] for the backward
// one, so that a lane's operand elements are contiguous: one 16-byte load
// for eight BF16 values), the forward product's accumulators stay in
// registers until every wave is done with X, and the backward product writes
// Y (FP64) to the member's second staging area.  k_filter then rescales and applies
// the phase epilogue to those rows exactly as to its own FFT rows.
//
// The basis matrices (mm_basis_bytes, filled by mm_basis_fill on the host)
// hold phi_j(n) for j < KB = N + 2 rounded up to 32 and n < N (zero up to N
// rounded up to 32), in FP32, or as the high and low BF16 parts of the FP32
// value; the columns past a row's K are masked out of C.
//
// MFMA lane layouts (checked on the device): v_mfma_f32_16x16x4_f32 takes
// A[i][k] from lane i + 16k and B[k][j] from lane j + 16k;
// v_mfma_f32_16x16x32_bf16 takes A[i][8q+e] (e = 0..7) from lane i + 16q and
// B[8q+e][j] from lane j + 16q; both leave D[4q+r][j] in register r of lane
// j + 16q.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <type_traits>
#include <vector>

#include "gd_common.hpp"
#include "gd_device.hpp"
#include "gd_launch.hpp"

namespace gd {

typedef float v4f __attribute__((ext_vector_type(4)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

constexpr int kMmRows = 16;   // members per workgroup (one MFMA row tile)
constexpr int kMmTpw = 24;    // forward column tiles per wave (K <= 4 * 24 * 16)

__host__ __device__ inline int mm_kpad(int K) { return (K + 31) / 32 * 32; }
__host__ __device__ inline int mm_npad(int N) { return (N + 31) / 32 * 32; }
// row stride of the LDS tile (elements): the wider of the padded N and K,
// plus 8 so that the 16 rows do not start on one bank
__host__ __device__ inline int mm_ld(int N, int K) {
  const int a = mm_npad(N), b = mm_kpad(K);
  return (a > b ? a : b) + 8;
}

// The tile holds FP32 values, or the high and low BF16 parts of each value
// (two arrays of BF16): 4 bytes per element either way.
// The tile holds FP32 values, or the high and low BF16 parts of each value
// (two arrays of BF16): 4 bytes per element either way.
size_t filter_mm_lds_bytes(int N, bool bf16) {
  (void)bf16;
  const int K = N + 2;  // the widest kept basis
  return (size_t)kMmRows * mm_ld(N, K) * 4;
}

// x = hi + lo with hi = bf16(x), lo = bf16(x - hi): 16 mantissa bits.
__device__ __forceinline__ void mm_split(float x, __bf16 &hi, __bf16 &lo) {
  hi = (__bf16)x;
  lo = (__bf16)(x - (float)hi);
}

template <bool BF16>
__global__ void __launch_bounds__(kBlock) k_filter_mm(const Args a) {
  typedef typename std::conditional<BF16, __bf16, float>::type T;
  // BF16: the products use the split operands, hi.hi + hi.lo + lo.hi (the
  // lo.lo term is below FP32 accumulation); xl holds the low parts
  const int nblk = (a.nmem + kMmRows - 1) / kMmRows;
  const int g = blockIdx.x;
  if (g >= a.ntiles) return;
  const StageSlot ss = a.slots[a.mslots[g / nblk]];
  const int slot = a.mslots[g / nblk];
  const int m0 = (g - (g / nblk) * nblk) * kMmRows;
  const int N = a.N, P = a.pitch, tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int c = (ss.kept - 2) / 2, K = ss.kept;  // kept basis: 2c + 2 functions
  const int Kp = mm_kpad(K), Np = mm_npad(N), LD = mm_ld(N, K);
  const int KB = mm_kpad(N + 2);  // rows of the basis matrix (every kept basis fits)
  extern __shared__ __align__(16) unsigned char smem[];
  T *xs = (T *)smem;
  T *xl = xs + (size_t)kMmRows * mm_ld(N, N + 2);  // BF16 only
  // basis matrices: Phi[j][n] (KB x Np) then Phi^T[n][j] (Np x KB); BF16:
  // high parts then low parts of each
  const T *phi = (const T *)a.mm_phi;
  const T *phi_lo = phi + (size_t)KB * Np;
  const T *phit = BF16 ? phi + (size_t)2 * KB * Np : phi + (size_t)KB * Np;
  const T *phit_lo = phit + (size_t)KB * Np;

  // the block's rows (zero past the last active member and past N)
  for (int e = tid; e < kMmRows * Np; e += kBlock) {
    const int b = e / Np, n = e - b * Np;
    const int m = m0 + b;
    double v = 0.0;
    if (m < a.nmem && n < N) v = member_at(a, m).sx()[(size_t)slot * P + n];
    if (BF16) {
      __bf16 hi, lo;
      mm_split((float)v, hi, lo);
      xs[b * LD + n] = hi;
      xl[b * LD + n] = lo;
    } else {
      xs[b * LD + n] = (T)v;
    }
  }
  __syncthreads();

  // ---- forward product C = X Phi^T: wave wv takes column tiles wv, wv+4, ...
  const int ntile = Kp / 16;
  const int q = lane >> 4, col = lane & 15;
  v4f acc[kMmTpw];
#pragma unroll
  for (int t = 0; t < kMmTpw; ++t) acc[t] = v4f{0.f, 0.f, 0.f, 0.f};
  auto col_j = [&](int t) { return (wv + 4 * t) * 16 + col; };
  constexpr int kStep = BF16 ? 32 : 4;
  for (int n0 = 0; n0 < Np; n0 += kStep) {
    if (BF16) {
      v8bf ah, al;
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        ah[e] = (__bf16)xs[col * LD + n0 + 8 * q + e];
        al[e] = (__bf16)xl[col * LD + n0 + 8 * q + e];
      }
#pragma unroll
      for (int t = 0; t < kMmTpw; ++t) {
        if (wv + 4 * t < ntile) {
          // longitudes past N carry zero rows
          const size_t o = (size_t)col_j(t) * Np + n0 + 8 * q;
          const v8bf bh = *(const v8bf *)(phi + o), bl = *(const v8bf *)(phi_lo + o);
          acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh, acc[t], 0, 0, 0);
          acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl, acc[t], 0, 0, 0);
          acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh, acc[t], 0, 0, 0);
        }
      }
    } else {
      const float av = (float)xs[col * LD + n0 + q];
#pragma unroll
      for (int t = 0; t < kMmTpw; ++t) {
        if (wv + 4 * t < ntile) {
          const float bv = (float)phi[(size_t)col_j(t) * Np + n0 + q];
          acc[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(av, bv, acc[t], 0, 0, 0);
        }
      }
    }
  }
  __syncthreads();  // every wave is done with X
  // C W into the tile: column j of row 4q + r; columns past K are zero
  const float wlast = (2 * (c + 1) == N) ? 1.0f / N : 2.0f / N;
#pragma unroll
  for (int t = 0; t < kMmTpw; ++t) {
    if (wv + 4 * t >= ntile) continue;
    const int j = col_j(t);
    const float w = j == 0 ? 1.0f / N : (j == K - 1 ? wlast : 2.0f / N);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const float cw = j < K ? acc[t][r] * w : 0.0f;
      if (BF16) {
        __bf16 hi, lo;
        mm_split(cw, hi, lo);
        xs[(4 * q + r) * LD + j] = hi;
        xl[(4 * q + r) * LD + j] = lo;
      } else {
        xs[(4 * q + r) * LD + j] = (T)cw;
      }
    }
  }
  __syncthreads();

  // ---- backward product Y = (C W) Phi: wave wv takes output tiles of 16
  // longitudes, four at a time
  const int nt_all = Np / 16;
  for (int nt0 = wv * 4; nt0 < nt_all; nt0 += 16) {
    v4f y[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) y[u] = v4f{0.f, 0.f, 0.f, 0.f};
    for (int j0 = 0; j0 < Kp; j0 += kStep) {
      if (BF16) {
        v8bf ah, al;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
          ah[e] = (__bf16)xs[col * LD + j0 + 8 * q + e];
          al[e] = (__bf16)xl[col * LD + j0 + 8 * q + e];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          if (nt0 + u >= nt_all) continue;
          const size_t o = (size_t)((nt0 + u) * 16 + col) * KB + j0 + 8 * q;
          const v8bf bh = *(const v8bf *)(phit + o), bl = *(const v8bf *)(phit_lo + o);
          y[u] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh, y[u], 0, 0, 0);
          y[u] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl, y[u], 0, 0, 0);
          y[u] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh, y[u], 0, 0, 0);
        }
      } else {
        const float av = (float)xs[col * LD + j0 + q];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          if (nt0 + u >= nt_all) continue;
          const float bv = (float)phit[(size_t)((nt0 + u) * 16 + col) * KB + j0 + q];
          y[u] = __builtin_amdgcn_mfma_f32_16x16x4f32(av, bv, y[u], 0, 0, 0);
        }
      }
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int n = (nt0 + u) * 16 + col;
      if (nt0 + u >= nt_all || n >= N) continue;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = m0 + 4 * q + r;
        if (m < a.nmem) member_at(a, m).sy()[(size_t)slot * P + n] = (double)y[u][r];
      }
    }
  }
}

template <bool BF16>
static hipError_t launch_mm_t(const Args &a, hipStream_t s) {
  dim3 g(a.ntiles), b(kBlock);
  hipLaunchKernelGGL((k_filter_mm<BF16>), g, b, filter_mm_lds_bytes(a.N, BF16), s, a);
  return hipGetLastError();
}

hipError_t launch_filter_mm(const Args &a, bool bf16, hipStream_t s) {
  if (a.ntiles <= 0) return hipSuccess;
  return bf16 ? launch_mm_t<true>(a, s) : launch_mm_t<false>(a, s);
}

size_t mm_basis_bytes(int N) { return (size_t)2 * mm_kpad(N + 2) * mm_npad(N) * 4; }

// BF16 with round to nearest even, as the device conversion does.
static uint16_t mm_bf16(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
static float mm_bf16f(uint16_t h) {
  const uint32_t u = (uint32_t)h << 16;
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}

void mm_basis_fill(int N, bool bf16, std::vector<unsigned char> &out) {
  const int KB = mm_kpad(N + 2), Np = mm_npad(N);
  // cos, sin of 2 pi m/N in extended precision, rounded once
  std::vector<double> cs(N), sn(N);
  for (int m = 0; m < N; ++m) {
    const long double ang = 2.0L * 3.141592653589793238462643383279502884L * (long double)m / (long double)N;
    cs[m] = (double)cosl(ang);
    sn[m] = (double)sinl(ang);
  }
  std::vector<float> phi((size_t)KB * Np, 0.0f);
  for (int j = 0; j < KB; ++j) {
    const long long k = (j + 1) >> 1;
    const bool sine = j > 0 && !(j & 1);
    for (int n = 0; n < N; ++n) {
      const int m = (int)((k * n) % N);
      phi[(size_t)j * Np + n] = (float)(sine ? sn[m] : cs[m]);
    }
  }
  out.assign(mm_basis_bytes(N), 0);
  if (!bf16) {
    float *a = (float *)out.data(), *t = a + (size_t)KB * Np;
    for (int j = 0; j < KB; ++j)
      for (int n = 0; n < Np; ++n) {
        a[(size_t)j * Np + n] = phi[(size_t)j * Np + n];
        t[(size_t)n * KB + j] = phi[(size_t)j * Np + n];
      }
    return;
  }
  uint16_t *h = (uint16_t *)out.data(), *l = h + (size_t)KB * Np, *th = l + (size_t)KB * Np,
           *tl = th + (size_t)KB * Np;
  for (int j = 0; j < KB; ++j)
    for (int n = 0; n < Np; ++n) {
      const float x = phi[(size_t)j * Np + n];
      const uint16_t hi = mm_bf16(x), lo = mm_bf16(x - mm_bf16f(hi));
      h[(size_t)j * Np + n] = th[(size_t)n * KB + j] = hi;
      l[(size_t)j * Np + n] = tl[(size_t)n * KB + j] = lo;
    }
}

hipError_t configure_filter_mm(int max_lds) {
  const void *fns[] = {(const void *)k_filter_mm<false>, (const void *)k_filter_mm<true>};
  for (const void *f : fns) {
    hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, max_lds);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

}  // namespace gd
