// gd_fft.hpp -- the polar zonal filter of one tendency row
// (filter_mod.F90:105-172) fused with the rescale of space_operators
// (dycore_mod.F90:210-217) and the phase epilogue; used by k_filter (staged
// rows) and by the filter tiles of k_phase and k_dphase (rows evaluated in
// LDS by the tile itself).
//
// FFTPACK5's rfft1f -> mask -> rfft1b of one row is the orthogonal projection
// onto the Fourier modes the mask keeps: wavenumbers 0..c and the cosine part
// of wavenumber c+1 (the mask keeps array slots 1..2c+2, and slot 2c+2 is
// a_{c+1} in FFTPACK5's [a0, a1, b1, a2, b2, ...] layout).  The kernel computes
// it with a complex FFT in LDS:
//   even N: the real row is read as N/2 complex points z_m = x_2m + i x_2m+1,
//           transformed, split into the real spectrum X_0..X_{N/2}, masked,
//           recombined and transformed back (the standard half-length real
//           FFT);
//   odd N:  the row is transformed as N complex points with zero imaginary
//           part, masked symmetrically and transformed back.
// The complex FFT is a Stockham autosort FFT (out-of-place passes between two
// LDS buffers, so every pass is one read, one write and one barrier) with
// radix-8/4/2/3/5 butterflies written out with their exact constants and a
// direct O(R^2) pass for other prime factors (7 included: a dedicated
// radix-7 butterfly needs about 130 VGPRs, which would cap every kernel that
// holds the filter at three waves per SIMD).  Twiddles W_N^k come from an FP64 table of cos, sin of 2 pi k/N
// for k = 0..N/2, loaded into LDS; W_N^k for k > N/2 is the conjugate of
// W_N^(N-k).  Everything is FP64 in every precision mode.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "gd_common.hpp"
#include "gd_device.hpp"

namespace gd {

__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// The twiddle table in LDS: t[k] = (cos, sin)(2 pi k/n) for k = 0..n/2.
struct Tw {
  const double2 *t;
  int n;
};

// W_n^m = exp(-2 pi i m/n) (forward) or its conjugate (inverse), 0 <= m < n.
__device__ __forceinline__ double2 twf(Tw tw, int m, bool inv) {
  const bool hi = 2 * m > tw.n;
  const double2 c = tw.t[hi ? tw.n - m : m];
  const double s = hi ? -c.y : c.y;  // sin(2 pi m/n)
  return make_double2(c.x, inv ? s : -s);
}

// i * z (inverse) or -i * z (forward)
__device__ __forceinline__ double2 rot(double2 z, bool inv) {
  return inv ? make_double2(-z.y, z.x) : make_double2(z.y, -z.x);
}

// In-register DFT of R points.  Radix 2, 3, 4, 5 and 8 use their closed
// forms; other radices use W_R = W_N^rstep from the table.
template <int R>
__device__ __forceinline__ void dft_small(double2 (&v)[R], Tw tw, int rstep, bool inv) {
  if constexpr (R == 2) {
    const double2 a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  } else if constexpr (R == 3) {
    constexpr double s3 = 0.86602540378443864676;  // sin(2 pi/3)
    const double2 t1 = cadd(v[1], v[2]);
    const double2 t2 = make_double2(v[0].x - 0.5 * t1.x, v[0].y - 0.5 * t1.y);
    const double2 d = csub(v[1], v[2]);
    const double2 t3 = rot(make_double2(s3 * d.x, s3 * d.y), inv);
    v[0] = cadd(v[0], t1);
    v[1] = cadd(t2, t3);
    v[2] = csub(t2, t3);
  } else if constexpr (R == 4) {
    const double2 t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
    const double2 t2 = cadd(v[1], v[3]);
    const double2 t3 = rot(csub(v[1], v[3]), inv);
    v[0] = cadd(t0, t2);
    v[1] = cadd(t1, t3);
    v[2] = csub(t0, t2);
    v[3] = csub(t1, t3);
  } else if constexpr (R == 5) {
    constexpr double c1 = 0.30901699437494742410;   // cos(2 pi/5)
    constexpr double c2 = -0.80901699437494742410;  // cos(4 pi/5)
    constexpr double s1 = 0.95105651629515357212;   // sin(2 pi/5)
    constexpr double s2 = 0.58778525229247312917;   // sin(4 pi/5)
    const double2 t1 = cadd(v[1], v[4]), t2 = cadd(v[2], v[3]);
    const double2 t3 = csub(v[1], v[4]), t4 = csub(v[2], v[3]);
    const double2 a1 = make_double2(v[0].x + c1 * t1.x + c2 * t2.x, v[0].y + c1 * t1.y + c2 * t2.y);
    const double2 a2 = make_double2(v[0].x + c2 * t1.x + c1 * t2.x, v[0].y + c2 * t1.y + c1 * t2.y);
    const double2 b1 = rot(make_double2(s1 * t3.x + s2 * t4.x, s1 * t3.y + s2 * t4.y), inv);
    const double2 b2 = rot(make_double2(s2 * t3.x - s1 * t4.x, s2 * t3.y - s1 * t4.y), inv);
    v[0] = cadd(v[0], cadd(t1, t2));
    v[1] = cadd(a1, b1);
    v[4] = csub(a1, b1);
    v[2] = cadd(a2, b2);
    v[3] = csub(a2, b2);
  } else if constexpr (R == 8) {
    constexpr double r2 = 0.70710678118654752440;   // 1/sqrt(2)
    // two radix-4 transforms of the even and odd points, then W_8^q
    double2 e[4] = {v[0], v[2], v[4], v[6]}, o[4] = {v[1], v[3], v[5], v[7]};
    dft_small<4>(e, tw, rstep, inv);
    dft_small<4>(o, tw, rstep, inv);
    // W_8 = (1 - i)/sqrt2 forward, (1 + i)/sqrt2 inverse
    const double2 w1 = inv ? make_double2(r2 * (o[1].x - o[1].y), r2 * (o[1].x + o[1].y))
                           : make_double2(r2 * (o[1].x + o[1].y), r2 * (o[1].y - o[1].x));
    const double2 w2 = rot(o[2], inv);
    const double2 w3 = inv ? make_double2(-r2 * (o[3].x + o[3].y), r2 * (o[3].x - o[3].y))
                           : make_double2(r2 * (o[3].y - o[3].x), -r2 * (o[3].x + o[3].y));
    v[0] = cadd(e[0], o[0]);
    v[4] = csub(e[0], o[0]);
    v[1] = cadd(e[1], w1);
    v[5] = csub(e[1], w1);
    v[2] = cadd(e[2], w2);
    v[6] = csub(e[2], w2);
    v[3] = cadd(e[3], w3);
    v[7] = csub(e[3], w3);
  } else {
    double2 o[R];
#pragma unroll
    for (int q = 0; q < R; ++q) {
      double2 acc = v[0];
#pragma unroll
      for (int r = 1; r < R; ++r) acc = cadd(acc, cmul(v[r], twf(tw, ((r * q) % R) * rstep, inv)));
      o[q] = acc;
    }
#pragma unroll
    for (int q = 0; q < R; ++q) v[q] = o[q];
  }
}

// One Stockham pass of radix R over a row of L points: butterfly j reads
// points j + q L/R, twiddles them by W_{Ns R}^{q (j mod Ns)} and writes its
// outputs to ((j - j mod Ns) R + j mod Ns) + q Ns.
// LDS slot of point e in a buffer written with the bank swizzle: the low
// three index bits XOR the next three (a permutation within each group of
// eight points, its own inverse).
template <bool SW>
__device__ __forceinline__ unsigned swz(unsigned e) {
  return SW ? e ^ ((e >> 3) & 7u) : e;
}

// SWI / SWO: the source / destination buffer uses the bank swizzle.  The
// first pass of a radix-8 plan writes each butterfly's eight outputs to
// consecutive points, a 128-byte stride across the lanes (eight-way bank
// conflicts); swizzled, the lanes cover all banks.
// NT: threads of the workgroup taking part (butterfly j goes to thread j mod
// NT); the butterflies' arithmetic does not depend on it.
template <int R, bool SWI = false, bool SWO = false, int NT = kBlock>
__device__ __forceinline__ void fft_pass(const double2 *src, double2 *dst, int L, int Ns, Tw tw,
                                         int tstride, bool inv) {
  const unsigned Lr = L / R;
  const int tws = (L / (Ns * R)) * tstride;
  const int rstep = (int)Lr * tstride;
  for (unsigned j = threadIdx.x; j < Lr; j += NT) {
    const unsigned k = j % (unsigned)Ns;
    double2 v[R];
#pragma unroll
    for (int q = 0; q < R; ++q) v[q] = src[swz<SWI>(j + q * Lr)];
    if (Ns > 1) {
#pragma unroll
      for (int q = 1; q < R; ++q) v[q] = cmul(v[q], twf(tw, q * (int)k * tws, inv));
    }
    dft_small<R>(v, tw, rstep, inv);
    const unsigned o = (j - k) * R + k;
#pragma unroll
    for (int q = 0; q < R; ++q) dst[swz<SWO>(o + q * Ns)] = v[q];
  }
}

// The same pass for any radix R (prime factors above 7), O(R^2) per butterfly.
template <int NT = kBlock>
__device__ __forceinline__ void fft_pass_any(const double2 *src, double2 *dst, int L, int Ns, int R, Tw tw,
                                             int tstride, bool inv) {
  const int Lr = L / R;
  const int tws = (L / (Ns * R)) * tstride;
  const int rstep = Lr * tstride;
  for (int j = threadIdx.x; j < Lr; j += NT) {
    const int k = j % Ns;
    const int o = (j - k) * R + k;
    for (int q = 0; q < R; ++q) {
      double2 acc = src[j];
      for (int p = 1; p < R; ++p) {
        double2 v = src[j + p * Lr];
        if (Ns > 1) v = cmul(v, twf(tw, p * k * tws, inv));
        acc = cadd(acc, cmul(v, twf(tw, ((p * q) % R) * rstep, inv)));
      }
      dst[o + q * Ns] = acc;
    }
  }
}

// Runs the plan's passes starting from buffer `cur`; returns the buffer that
// holds the result.  Ends with a barrier.
template <int NT = kBlock>
__device__ __forceinline__ int fft_run(double2 *b0, double2 *b1, int cur, const Args &a, Tw tw, bool inv) {
  const int L = a.fft_L, tstride = a.fft_half ? 2 : 1;
  int Ns = 1;
  for (int p = 0; p < a.fft_nrad; ++p) {
    const int R = a.fft_rad[p];
    const double2 *src = cur ? b1 : b0;
    double2 *dst = cur ? b0 : b1;
    switch (R) {
      case 2: fft_pass<2, false, false, NT>(src, dst, L, Ns, tw, tstride, inv); break;
      case 3: fft_pass<3, false, false, NT>(src, dst, L, Ns, tw, tstride, inv); break;
      case 4: fft_pass<4, false, false, NT>(src, dst, L, Ns, tw, tstride, inv); break;
      case 5: fft_pass<5, false, false, NT>(src, dst, L, Ns, tw, tstride, inv); break;
      case 8: fft_pass<8, false, false, NT>(src, dst, L, Ns, tw, tstride, inv); break;
      default: fft_pass_any<NT>(src, dst, L, Ns, R, tw, tstride, inv); break;
    }
    __syncthreads();
    Ns *= R;
    cur ^= 1;
  }
  return cur;
}

// The same passes with the plan fixed at compile time: L, the stride and
// every Ns are constants, so the index arithmetic of each pass (j mod Ns,
// the twiddle strides) folds to shifts and constant multiplies.  The
// arithmetic on the data is that of fft_run, so both give the same bits.
// The first radix-8 pass of a length divisible by 8 writes swizzled and the
// second pass reads swizzled (only between those two passes does a buffer
// hold the swizzled layout).
template <int NT, int L, int TS, int NS, bool SWI, int R, int... REST>
__device__ __forceinline__ int fft_fixed_p(double2 *b0, double2 *b1, int cur, Tw tw, bool inv) {
  constexpr bool SWO = NS == 1 && R == 8 && L % 8 == 0 && sizeof...(REST) > 0;
  fft_pass<R, SWI, SWO, NT>(cur ? b1 : b0, cur ? b0 : b1, L, NS, tw, TS, inv);
  __syncthreads();
  if constexpr (sizeof...(REST) > 0)
    return fft_fixed_p<NT, L, TS, NS * R, SWO, REST...>(b0, b1, cur ^ 1, tw, inv);
  else
    return cur ^ 1;
}

template <int NT, int L, int TS, int NS, int... RS>
__device__ __forceinline__ int fft_fixed(double2 *b0, double2 *b1, int cur, Tw tw, bool inv) {
  return fft_fixed_p<NT, L, TS, NS, false, RS...>(b0, b1, cur, tw, inv);
}

// The plans of the even row lengths 720, 360 and 180 (fft_plan of their
// half lengths) run fixed, and 1440 (half length 720 = 8 2 3 3 5) where the
// row takes eight or more points per thread (KP); others run through
// fft_run.
template <int KP, int NT = kBlock>
__device__ __forceinline__ int fft_any(double2 *b0, double2 *b1, int cur, const Args &a, Tw tw, bool inv) {
  if constexpr (KP >= 8) {
    if (a.fft_half && a.fft_L == 720) return fft_fixed<NT, 720, 2, 1, 8, 2, 3, 3, 5>(b0, b1, cur, tw, inv);
  }
  if (a.fft_half && a.fft_L == 360) return fft_fixed<NT, 360, 2, 1, 8, 3, 3, 5>(b0, b1, cur, tw, inv);
  if (a.fft_half && a.fft_L == 180) return fft_fixed<NT, 180, 2, 1, 4, 3, 3, 5>(b0, b1, cur, tw, inv);
  if (a.fft_half && a.fft_L == 90) return fft_fixed<NT, 90, 2, 1, 2, 3, 3, 5>(b0, b1, cur, tw, inv);
  return fft_run<NT>(b0, b1, cur, a, tw, inv);
}

// FFTPACK5 mask of cutoff c on the complex spectrum: wavenumbers 0..c kept,
// only the cosine (real) part of wavenumber c+1 kept, the rest removed.
__device__ __forceinline__ double2 fmask(double2 X, int k, int c) {
  if (k <= c) return X;
  if (k == c + 1) return make_double2(X.x, 0.0);
  return make_double2(0.0, 0.0);
}

// Even N: Z = FFT_L(z) of the packed row (L = N/2) -> real spectrum X_k =
// E_k + W_N^k O_k with E_k = (Z_k + conj Z_{L-k})/2, O_k = -i (Z_k - conj
// Z_{L-k})/2 -> mask -> the packed spectrum of the filtered row, Z'_k =
// E'_k + i O'_k with E'_k = (Y_k + conj Y_{L-k})/2, O'_k = (Y_k - conj
// Y_{L-k}) W_N^{-k}/2, scaled by 1/L for the unnormalised inverse.  Pairs
// (k, L-k) are handled together; k = 0 carries X_0 and the Nyquist X_L.
template <int NT = kBlock>
__device__ __forceinline__ void real_mid(const double2 *s, double2 *d, int L, int c, Tw tw, double sc) {
  const int H = L / 2 + 1;
  for (int k = threadIdx.x; k < H; k += NT) {
    if (k == 0) {
      const double2 z = s[0];
      const double X0 = z.x + z.y, XL = z.x - z.y;
      const double YL = L <= c + 1 ? XL : 0.0;
      d[0] = make_double2(0.5 * (X0 + YL) * sc, 0.5 * (X0 - YL) * sc);
      continue;
    }
    const int km = L - k;
    const double2 zk = s[k], zm = s[km];
    const double2 E = make_double2(0.5 * (zk.x + zm.x), 0.5 * (zk.y - zm.y));
    const double2 O = make_double2(0.5 * (zk.y + zm.y), -0.5 * (zk.x - zm.x));
    const double2 wk = twf(tw, k, false);
    const double2 wO = cmul(wk, O);
    const double2 Xk = cadd(E, wO);
    const double2 Xm = make_double2(E.x - wO.x, -(E.y - wO.y));  // X_{L-k} = conj(E - W^k O)
    const double2 Yk = fmask(Xk, k, c), Ym = fmask(Xm, km, c);
    const double2 Ep = make_double2(0.5 * (Yk.x + Ym.x), 0.5 * (Yk.y - Ym.y));
    const double2 Dh = make_double2(0.5 * (Yk.x - Ym.x), 0.5 * (Yk.y + Ym.y));
    const double2 Op = cmul(Dh, make_double2(wk.x, -wk.y));
    d[k] = make_double2((Ep.x - Op.y) * sc, (Ep.y + Op.x) * sc);
    if (km != k) d[km] = make_double2((Ep.x + Op.y) * sc, (Op.x - Ep.y) * sc);
  }
}

// Odd N: mask the full complex spectrum symmetrically, scaled by 1/N.
template <int NT = kBlock>
__device__ __forceinline__ void cplx_mid(const double2 *s, double2 *d, int L, int c, double sc) {
  for (int k = threadIdx.x; k < L; k += NT) {
    const int kk = k <= L - k ? k : L - k;
    const double2 Y = fmask(s[k], kk, c);
    d[k] = make_double2(Y.x * sc, Y.y * sc);
  }
}

// Filters one tendency row and applies the rescale and the phase epilogue.
// XIN = false: the row is staging slot fr.slot of member mb; the function
// fills the twiddle table and the row buffer b0 itself.  XIN = true: the
// caller has written the row to b0 in the FFT layout (even N: N reals;
// odd N: N complex points with zero imaginary parts) and the twiddle table
// to tw, followed by a barrier.  zero (XIN only, rescaled phases only): the
// row is known to be zero and b0 is not read; with s1 = 0 the row stays
// unfiltered, so the epilogue gets T = +0 with no transform, as the
// transform path would give it.  KP: longitudes per thread (N <= KP *
// kBlock).  tw holds tw_entries(N) entries, b0 and b1 fft_L complex points
// each; red holds 8 doubles.  Ends with a barrier.
template <typename CT, int KP, bool XIN>
__device__ __forceinline__ void filter_row(const Args &a, const Member &mb, const FiltRow &fr, double2 *tw,
                                           double2 *b0, double2 *b1, double *red, bool zero = false) {
  const int N = a.N, P = a.pitch, L = a.fft_L, tid = threadIdx.x;
  const bool rescale = a.phase <= PH_T;
  const double sc = 1.0 / (double)L;
  const int comp = fr.comp;
  const size_t o = (size_t)fr.row * P;
  const int ystep = a.fft_half ? 1 : 2;
  double *rowbuf = (double *)b0;

  // the row's rescale weight (state%iap%u, %v, gd + ghs) and the epilogue
  // inputs (state s0 or Tb), and for a staged row the row itself, all loads
  // issued together; gd + ghs is added after the forward transform, when
  // the loads have long arrived
  double *const sxb = a.sx_ext ? a.sx_ext : mb.sx();
  const double *xp = sxb + (size_t)fr.slot * P;
  const double *wp = mb.st(a.s_in, comp == CDU ? FU : (comp == CDV ? FV : FGD)) + o;
  const bool gdrow = comp == CDGD;
  const double *wq = mb.ghs() + o;
  double x[KP], w[KP], wg[KP], pre[KP];
#pragma unroll
  for (int k = 0; k < KP; ++k) {
    const int l = tid + k * kBlock;
    double xv = 0.0, wv = 0.0, gv = 0.0, pv = 0.0;
    if (l < N) {
      if (!XIN) xv = xp[l];
      if (rescale) {
        if (!zero) {
          wv = wp[l];
          if (gdrow) gv = wq[l];
        }
        pv = phase_pre<CT>(a, mb, comp, o + l);
      }
    }
    x[k] = xv;
    w[k] = wv;
    wg[k] = gv;
    pre[k] = pv;
  }
  if (!XIN) {
    for (int k = tid; k < tw_entries(N); k += kBlock) tw[k] = make_double2(a.twid[2 * k], a.twid[2 * k + 1]);
#pragma unroll
    for (int k = 0; k < KP; ++k) {
      const int l = tid + k * kBlock;
      if (l >= N) continue;
      if (a.fft_half) {
        rowbuf[l] = x[k];
      } else {
        rowbuf[2 * l] = x[k];
        rowbuf[2 * l + 1] = 0.0;
      }
    }
    __syncthreads();
  } else {
    // the row values are read before the first pass; b0 is first written
    // by the second pass, after a barrier
#pragma unroll
    for (int k = 0; k < KP; ++k) {
      const int l = tid + k * kBlock;
      x[k] = l < N && !zero ? rowbuf[(size_t)l * ystep] : 0.0;
    }
  }

  // s1 = sum(tend * weight) before and s2 = sum(filtered tend * weight)
  // after filtering (dycore_mod.F90:211, 214), reduced together
  double yv[KP];
  double s12[2] = {0.0, 0.0};
  // MFMA modes: a momentum row arrives projected (k_filter_mm) in the
  // member's second staging area; only the rescale and epilogue remain
  const bool proj = !XIN && a.mm && comp != CDGD;
  if (proj) {
    const double *yp = mb.sy() + (size_t)fr.slot * P;
#pragma unroll
    for (int k = 0; k < KP; ++k) {
      const int l = tid + k * kBlock;
      yv[k] = l < N ? yp[l] : 0.0;
      s12[0] += x[k] * w[k];
      s12[1] += yv[k] * w[k];
    }
    if (rescale) block_sum_n<2>(s12, red);
  } else if (!zero) {
    const Tw twt{tw, N};
    int cur = fft_any<KP>(b0, b1, 0, a, twt, false);
    if (gdrow) {
#pragma unroll
      for (int k = 0; k < KP; ++k) w[k] = w[k] + wg[k];
    }
    if (a.fft_half)
      real_mid(cur ? b1 : b0, cur ? b0 : b1, L, fr.cut, twt, sc);
    else
      cplx_mid(cur ? b1 : b0, cur ? b0 : b1, L, fr.cut, sc);
    __syncthreads();
    cur = fft_any<KP>(b0, b1, cur ^ 1, a, twt, true);
    const double *y = (const double *)(cur ? b1 : b0);
#pragma unroll
    for (int k = 0; k < KP; ++k) {
      const int l = tid + k * kBlock;
      yv[k] = l < N ? y[(size_t)l * ystep] : 0.0;
      s12[0] += x[k] * w[k];
      s12[1] += yv[k] * w[k];
    }
    if (rescale) block_sum_n<2>(s12, red);
  } else {
#pragma unroll
    for (int k = 0; k < KP; ++k) yv[k] = 0.0;
  }
  const double s1 = s12[0], s2 = s12[1];
  // rows with |s1| <= filter_inner_product_threshold stay unfiltered
  const bool filt = !rescale || fabs(s1) > 1.0e-16;

  double p[2] = {0.0, 0.0};
#pragma unroll
  for (int k = 0; k < KP; ++k) {
    const int l = tid + k * kBlock;
    if (l >= N) continue;
    const double T = rescale ? (filt ? yv[k] * s1 / s2 : x[k]) : yv[k];
    if (a.phase <= PH_T) {
      phase_store<CT>(a, mb, comp, o + l, T, pre[k], fr.wrow, p[0], p[1]);
    } else if (a.phase == PH_D) {
      // ordinary_diffusion update of the filtered component
      const double *sgd = mb.st(a.s_in, FGD);
      if (comp == CDGD) {
        mb.st(SX, FGD)[o + l] = sgd[o + l] + a.kdiff * T;
      } else if (comp == CDU) {
        const int ip = l + 1 == N ? 0 : l + 1;
        const double u = a.in_full ? mb.st(a.s_in, Fu)[o + l]
                                   : mb.st(a.s_in, FU)[o + l] * 2.0 / (sqrt(sgd[o + l]) + sqrt(sgd[o + ip]));
        mb.st(SX, Fu)[o + l] = u + a.kdiff * T;
      } else {
        const double v = a.in_full ? mb.st(a.s_in, Fv)[o + l]
                                   : mb.st(a.s_in, FV)[o + l] * 2.0 / (sqrt(sgd[o + l]) + sqrt(sgd[o + P + l]));
        mb.st(SX, Fv)[o + l] = v + a.kdiff * T;
      }
    } else {  // PH_P: projection only, back into the staging row
      sxb[(size_t)fr.slot * P + l] = T;
    }
  }
  if (a.phase == PH_C) {
    block_sum_n<2>(p, red);
    // the whole row's products go to segment 0 of its partials, the other
    // segments of the row hold zero
    double *pp = mb.part() + ((size_t)(fr.row * 3 + comp) * a.nseg) * 2;
    if (tid == 0) {
      pp[0] = p[0];
      pp[1] = p[1];
    } else if (tid < a.nseg) {
      pp[2 * tid] = 0.0;
      pp[2 * tid + 1] = 0.0;
    }
  }
  __syncthreads();  // LDS free for the caller
}

// Sums of K values over the first kBlock threads of a workgroup of NT
// threads, in block_sum_n's order (so with its bits); the other threads
// pass through the barriers, and every thread returns the sums.
template <int K>
__device__ __forceinline__ void head_sum_n(double (&v)[K], double *red) {
#pragma unroll
  for (int q = 0; q < K; ++q)
    for (int off = 32; off > 0; off >>= 1) v[q] += __shfl_xor(v[q], off, 64);
  const int w = threadIdx.x >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0 && w < kBlock / 64)
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

// filter_row<CT, KP, true> (the row already in b0, a rescaled phase) for a
// workgroup of NT threads with KW longitudes each (l = thread + k NT, N <=
// KW NT): the loads, the transforms' butterflies and the epilogue spread
// over all NT threads.  The sums keep filter_row's order: thread t < kBlock
// adds the products of longitudes t, t + kBlock, ... in turn (staged through
// scr, 2 N doubles), then head_sum_n reduces those kBlock partial sums as
// block_sum_n does, so both give the same bits.  Ends with a barrier.
template <typename CT, int NT, int KW>
__device__ __forceinline__ void filter_row_w(const Args &a, const Member &mb, const FiltRow &fr, double2 *tw,
                                             double2 *b0, double2 *b1, double *scr, double *red, bool zero) {
  const int N = a.N, P = a.pitch, L = a.fft_L, tid = threadIdx.x;
  const double sc = 1.0 / (double)L;
  const int comp = fr.comp;
  const size_t o = (size_t)fr.row * P;
  const int ystep = a.fft_half ? 1 : 2;
  const double *wp = mb.st(a.s_in, comp == CDU ? FU : (comp == CDV ? FV : FGD)) + o;
  const bool gdrow = comp == CDGD;
  double x[KW], w[KW], wg[KW], pre[KW], yv[KW];
#pragma unroll
  for (int k = 0; k < KW; ++k) {
    const int l = tid + k * NT;
    double wv = 0.0, gv = 0.0, pv = 0.0;
    if (l < N) {
      if (!zero) {
        wv = wp[l];
        if (gdrow) gv = mb.ghs()[o + l];
      }
      pv = phase_pre<CT>(a, mb, comp, o + l);
    }
    w[k] = wv;
    wg[k] = gv;
    pre[k] = pv;
    yv[k] = 0.0;
  }
  // the row values are read before the first pass; b0 is first written by
  // the second pass, after a barrier
#pragma unroll
  for (int k = 0; k < KW; ++k) {
    const int l = tid + k * NT;
    x[k] = l < N && !zero ? ((const double *)b0)[(size_t)l * ystep] : 0.0;
  }
  const int nk = (N + kBlock - 1) / kBlock;  // longitudes per head thread
  double s12[2] = {0.0, 0.0};
  if (!zero) {
    const Tw twt{tw, N};
    int cur = fft_any<4, NT>(b0, b1, 0, a, twt, false);
    if (gdrow) {
#pragma unroll
      for (int k = 0; k < KW; ++k) w[k] = w[k] + wg[k];
    }
    if (a.fft_half)
      real_mid<NT>(cur ? b1 : b0, cur ? b0 : b1, L, fr.cut, twt, sc);
    else
      cplx_mid<NT>(cur ? b1 : b0, cur ? b0 : b1, L, fr.cut, sc);
    __syncthreads();
    cur = fft_any<4, NT>(b0, b1, cur ^ 1, a, twt, true);
    const double *y = (const double *)(cur ? b1 : b0);
#pragma unroll
    for (int k = 0; k < KW; ++k) {
      const int l = tid + k * NT;
      if (l < N) {
        yv[k] = y[(size_t)l * ystep];
        scr[2 * l] = x[k] * w[k];
        scr[2 * l + 1] = yv[k] * w[k];
      }
    }
    __syncthreads();
    if (tid < kBlock)
      for (int k = 0; k < nk; ++k) {
        const int e = tid + k * kBlock;
        if (e < N) {
          s12[0] += scr[2 * e];
          s12[1] += scr[2 * e + 1];
        }
      }
    head_sum_n<2>(s12, red);
  }
  const double s1 = s12[0], s2 = s12[1];
  // rows with |s1| <= filter_inner_product_threshold stay unfiltered
  const bool filt = fabs(s1) > 1.0e-16;
  double pk[KW][2];
#pragma unroll
  for (int k = 0; k < KW; ++k) {
    const int l = tid + k * NT;
    pk[k][0] = pk[k][1] = 0.0;
    if (l < N) {
      const double T = filt ? yv[k] * s1 / s2 : x[k];
      phase_store<CT>(a, mb, comp, o + l, T, pre[k], fr.wrow, pk[k][0], pk[k][1]);
    }
  }
  if (a.phase == PH_C) {
    // (scr was last read before head_sum_n's barriers)
#pragma unroll
    for (int k = 0; k < KW; ++k) {
      const int l = tid + k * NT;
      if (l < N) {
        scr[2 * l] = pk[k][0];
        scr[2 * l + 1] = pk[k][1];
      }
    }
    __syncthreads();
    double p[2] = {0.0, 0.0};
    if (tid < kBlock)
      for (int k = 0; k < nk; ++k) {
        const int e = tid + k * kBlock;
        if (e < N) {
          p[0] += scr[2 * e];
          p[1] += scr[2 * e + 1];
        }
      }
    head_sum_n<2>(p, red);
    // the whole row's products go to segment 0 of its partials, the other
    // segments of the row hold zero
    double *pp = mb.part() + ((size_t)(fr.row * 3 + comp) * a.nseg) * 2;
    if (tid == 0) {
      pp[0] = p[0];
      pp[1] = p[1];
    } else if (tid < a.nseg) {
      pp[2 * tid] = 0.0;
      pp[2 * tid + 1] = 0.0;
    }
  }
  __syncthreads();  // LDS free for the caller
}

inline size_t filter_lds_bytes_h(int N, int L) { return (size_t)tw_entries(N) * 16 + (size_t)2 * L * 16; }

}  // namespace gd
