// gd_phase.hip -- k_phase: one tendency phase of the integrator (the
// tendency, the polar filter of its flagged rows, the rescale and the phase
// epilogue) in a single launch, in place of k_tend followed by k_filter.
//
// Blocks 0..nnear_pad-1 are whole-row filter tiles, one per flagged (member,
// row, component).  Such a tile evaluates its component's tendency over the
// whole row, segment by segment, into an LDS row buffer (the arithmetic of
// the row tiles), then filters that buffer in place (filter_row).  These are
// the longest chains of the launch, so they take the lowest block indices
// and are dispatched first.  The remaining blocks are the row tiles of k_tend
// over the rows that have a component outside the filter.  Every value is
// formed by the same expressions in the same order as with k_tend +
// k_filter, so both paths give the same bits.
//
// k_dphase does the same for the last application of ordinary_diffusion:
// its filter tiles evaluate their component's Laplacian over the row
// (dif_segment, gd_diffuse.hpp) and filter it in phase D; its row tiles are
// those of k_diffuse.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "gd_common.hpp"
#include "gd_device.hpp"
#include "gd_fft.hpp"
#include "gd_launch.hpp"
#include "gd_tend.hpp"
#include "gd_diffuse.hpp"

namespace gd {

// LDS: [stencil rows of a segment, later FFT buffer b1] [row buffer b0]
// [twiddle table].
template <typename CT>
__host__ __device__ inline size_t phase_region0(int L) {
  const size_t st = tend_stencil_bytes<CT>(), fb = (size_t)L * 16;
  return align16(st > fb ? st : fb);
}

template <typename CT>
__host__ __device__ inline size_t phase_lds_t(int N, int L) {
  return phase_region0<CT>(L) + (size_t)L * 16 + (size_t)tw_padded(N) * 16;
}

// The twiddle table straight from HBM to LDS (global_load_lds: no registers
// hold it in flight).  Each wave copies whole 64-entry blocks; a lane moves
// one 16-byte entry, and the wave's block lands contiguously from the LDS
// address it names.  The copy is complete after s_waitcnt vmcnt(0).
template <int NT = kBlock>
__device__ __forceinline__ void tw_to_lds(const Args &a, double2 *tw) {
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63, np = tw_padded(a.N);
  for (int base = w * 64; base < np; base += NT) {
    __builtin_amdgcn_global_load_lds((const void *)(a.twid + 2 * (base + lane)),
                                     (__attribute__((address_space(3))) void *)(tw + base), 16, 0, 0);
  }
}

// Row value x of longitude l into the FFT layout of b0.
__device__ __forceinline__ void row_put(const Args &a, double *rowbuf, int l, double x) {
  if (a.fft_half) {
    rowbuf[l] = x;
  } else {
    rowbuf[2 * l] = x;
    rowbuf[2 * l + 1] = 0.0;
  }
}

// Component CM of full row j (half row j for dv), segment by segment; the
// loads of segment g+1 are issued before segment g is evaluated, so only the
// first segment waits a full load round.
template <typename CT, int PASS, bool IN_FULL, int CM>
__device__ __forceinline__ void near_row(const Args &a, const Member &mb, const RowCoef &cj, int j, double *rowbuf,
                                         unsigned char *smem, double *red) {
  const int N = a.N, tid = threadIdx.x;
  const RowNbr nb = row_nbr(a, j);
  SegLoads cur;
  seg_load<PASS, IN_FULL>(a, mb, j, 0, min(N, kSeg), cur);
  for (int g = 0; g < a.nseg; ++g) {
    const int i0 = g * kSeg, W = min(N - i0, kSeg);
    SegLoads nxt;
    if (g + 1 < a.nseg) seg_load<PASS, IN_FULL>(a, mb, j, i0 + kSeg, min(N - i0 - kSeg, kSeg), nxt);
    if (g) __syncthreads();  // the previous segment's stencil rows have been read
    const TendPoint pt = seg_eval<CT, PASS, IN_FULL, CM>(a, cj, nb, j, i0, W, cur, false, 0.0,
                                                         TendPoint{0, 0, 0, 0, 0, 0}, smem, red);
    if (tid < W) row_put(a, rowbuf, i0 + tid, CM == TC_DU ? pt.du : (CM == TC_DV ? pt.dv : pt.dgd));
    cur = nxt;
  }
}

template <typename CT, int PASS, bool IN_FULL, int KP>
__device__ __forceinline__ void near_tile(const Args &a, const FiltRow fr, unsigned char *smem, double *red) {
  const int N = a.N, L = a.fft_L, tid = threadIdx.x;
  const Member mb = member_at(a, fr.member);
  const int j = fr.row;
  const RowCoef cj = a.coef[j];
  double2 *b1 = (double2 *)smem;
  double2 *b0 = (double2 *)(smem + phase_region0<CT>(L));
  double2 *tw = b0 + L;
  tw_to_lds(a, tw);
  double *rowbuf = (double *)b0;
  if (fr.comp == CDU) {
    near_row<CT, PASS, IN_FULL, TC_DU>(a, mb, cj, j, rowbuf, smem, red);
  } else if (fr.comp == CDV) {
    near_row<CT, PASS, IN_FULL, TC_DV>(a, mb, cj, j, rowbuf, smem, red);
  } else if (PASS != 2 && (j == 0 || j == a.M - 1)) {
    // a pole row's dgd is minus the pole flux at every longitude
    // (dycore_mod.F90:572-596)
    const double pole = pole_flux(a, mb, j, red);
    for (int l = tid; l < N; l += kBlock) row_put(a, rowbuf, l, -pole);
  } else if (PASS == 2) {
    // the slow pass has no mass continuity term: a zero row, which a
    // rescaled phase leaves unfiltered without transforming it
    if (a.phase > PH_T)
      for (int l = tid; l < N; l += kBlock) row_put(a, rowbuf, l, 0.0);
  } else {
    near_row<CT, PASS, IN_FULL, TC_DGD>(a, mb, cj, j, rowbuf, smem, red);
  }
  __builtin_amdgcn_s_waitcnt(0);  // the twiddle copy has landed
  __syncthreads();  // row and twiddles published; the stencil rows are free for b1
  filter_row<CT, KP, true>(a, mb, fr, tw, b0, b1, red, PASS == 2 && fr.comp == CDGD && a.phase <= PH_T);
}

// Four longitudes per thread (rows up to 1024): five waves per SIMD (at most
// 96 VGPRs); with the LDS of an FP64 tile that is five tiles per CU, so a
// 720x361 run's phase fits in one round of resident workgroups.  Eight
// (rows up to 2048): the twiddle table and row buffer of a 1440-point row
// hold the LDS to three tiles per CU, so three waves per SIMD.
template <int KP>
constexpr int phase_waves() { return KP == 4 ? 5 : 3; }

template <typename CT, int PASS, bool IN_FULL, int KP>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(phase_waves<KP>(), 8)))
k_phase(const Args a) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ double red[4 * 6];
  const int b = blockIdx.x;
  if (b < a.nnear_pad) {
    // filter rows are in (member, row, component) order; each XCD takes a
    // contiguous run of them, so the components of a row and its
    // neighbours share that XCD's L2
    const int t = (b % kXcd) * (a.nnear_pad / kXcd) + b / kXcd;
    // (the row descriptor is copied: a reference into global memory would be
    // reloaded after every store of the epilogue)
    if (t < a.nnear) near_tile<CT, PASS, IN_FULL, KP>(a, a.frows[t], smem, red);
    return;
  }
  // row tiles: each XCD owns a contiguous band of them (nnear_pad and the
  // remaining grid are multiples of kXcd, so block b stays on XCD b % kXcd)
  const int bb = b - a.nnear_pad, per = (gridDim.x - a.nnear_pad) / kXcd;
  const int t = (bb % kXcd) * per + bb / kXcd;
  if (t >= a.ntiles) return;
  tend_tile<CT, PASS, IN_FULL, true>(a, t, smem, red);
}

// ---------------------------------------------------------------------------
// k_wphase: k_phase with workgroups of two tiles (kWTiles x kBlock threads).
// A filter workgroup evaluates its row's segments two at a time, one per
// tile with its own stencil rows, instead of one after another, and filters
// the row with one or two longitudes per thread (filter_row_w).  A row
// workgroup runs two consecutive row tiles (t = 2 r, 2 r + 1), of one row or
// of two; when one of them is a pole-row tile the other takes the same two
// barriers (seg_eval's bar2), and a spare tile past the last one does the
// last tile's work without its stores.  Tile t's arithmetic is that of
// k_phase, so both launches give the same bits.
//
// Rows of two segments (255 to 508 longitudes) map one row to a workgroup.
// Measured on MI355X against k_phase: single runs of 288x145, 360x181 and
// 480x241 step 9-10% faster, two runs of 360x181 3-5%.  At 720x361 (three
// segments) a single run steps at 0.83-0.85 of k_phase: the filter
// workgroups' chain is no shorter there (the first round of stencil loads
// dominates the tendency stage, and k_phase already overlaps the later
// segments' loads with evaluation), and the 633 workgroups take a second
// round at two per CU, where k_phase's 1,087 tiles fit five per CU; capped
// at 80 VGPRs (three per CU, with spills) it stepped at 0.82.  Workgroups of
// three tiles (one per segment) were no better.
//
// LDS: kWTiles stencil regions of wphase_region bytes (region 0 later holds
// FFT buffer b1, region 1 the filter's products), then the row buffer b0
// and the twiddle table.
// ---------------------------------------------------------------------------

constexpr int kWTiles = 2;

template <typename CT>
__host__ __device__ inline size_t wphase_region(int L) {
  const size_t st = tend_stencil_bytes<CT>(), fb = (size_t)L * 16;
  return align16(st > fb ? st : fb);
}

template <typename CT>
__host__ __device__ inline size_t wphase_lds(int N, int L) {
  return (size_t)kWTiles * wphase_region<CT>(L) + (size_t)L * 16 + (size_t)tw_padded(N) * 16;
}

// Component CM of row j into the row buffer: tile s of the workgroup takes
// segments s, s + kWTiles, ... with the arithmetic of near_row (the next
// segment's loads issued before the current one is evaluated).  A tile with
// no segment left in a round evaluates an empty one (W = 0: its loads stay
// inside the row, nothing is written) to take the round's barriers.
template <typename CT, int PASS, bool IN_FULL, int CM>
__device__ __forceinline__ void wnear_row(const Args &a, const Member &mb, const RowCoef &cj, int j, int s,
                                          double *rowbuf, unsigned char *st, double *red) {
  const int N = a.N, nseg = a.nseg, tid = ltid();
  const int rounds = (nseg + kWTiles - 1) / kWTiles;
  const RowNbr nb = row_nbr(a, j);
  auto seg_i0 = [&](int g) { return g < nseg ? g * kSeg : 0; };
  auto seg_w = [&](int g) { return g < nseg ? min(N - g * kSeg, kSeg) : 0; };
  SegLoads cur;
  seg_load<PASS, IN_FULL>(a, mb, j, seg_i0(s), seg_w(s), cur);
  for (int k = 0; k < rounds; ++k) {
    const int g = k * kWTiles + s, gn = g + kWTiles;
    SegLoads nxt;
    if (k + 1 < rounds) seg_load<PASS, IN_FULL>(a, mb, j, seg_i0(gn), seg_w(gn), nxt);
    if (k) __syncthreads();  // the previous round's stencil rows have been read
    const int i0 = seg_i0(g), W = seg_w(g);
    const TendPoint pt = seg_eval<CT, PASS, IN_FULL, CM>(a, cj, nb, j, i0, W, cur, false, 0.0,
                                                         TendPoint{0, 0, 0, 0, 0, 0}, st, red);
    if (tid < W) row_put(a, rowbuf, i0 + tid, CM == TC_DU ? pt.du : (CM == TC_DV ? pt.dv : pt.dgd));
    cur = nxt;
  }
}

template <typename CT, int PASS, bool IN_FULL, int KW>
__device__ __forceinline__ void wnear_tile(const Args &a, const FiltRow fr, unsigned char *smem, double *red) {
  constexpr int NT = kWTiles * kBlock;
  const int N = a.N, L = a.fft_L, s = threadIdx.x / kBlock;
  const Member mb = member_at(a, fr.member);
  const int j = fr.row;
  const RowCoef cj = a.coef[j];
  const size_t R = wphase_region<CT>(L);
  double2 *b1 = (double2 *)smem;
  double2 *b0 = (double2 *)(smem + kWTiles * R);
  double2 *tw = b0 + L;
  tw_to_lds<NT>(a, tw);
  double *rowbuf = (double *)b0;
  unsigned char *st = smem + s * R;
  double *rd = red + s * 4 * 6;
  if (fr.comp == CDU) {
    wnear_row<CT, PASS, IN_FULL, TC_DU>(a, mb, cj, j, s, rowbuf, st, rd);
  } else if (fr.comp == CDV) {
    wnear_row<CT, PASS, IN_FULL, TC_DV>(a, mb, cj, j, s, rowbuf, st, rd);
  } else if (PASS != 2 && (j == 0 || j == a.M - 1)) {
    // a pole row's dgd is minus the pole flux at every longitude; each tile
    // forms the same ring sum (dycore_mod.F90:572-596)
    const double pole = pole_flux(a, mb, j, rd);
    for (int l = threadIdx.x; l < N; l += NT) row_put(a, rowbuf, l, -pole);
  } else if (PASS != 2) {
    wnear_row<CT, PASS, IN_FULL, TC_DGD>(a, mb, cj, j, s, rowbuf, st, rd);
  }
  // (the slow pass has no mass continuity term: its dgd rows are zero rows,
  // which filter_row_w leaves unfiltered without reading b0)
  __builtin_amdgcn_s_waitcnt(0);  // the twiddle copy has landed
  __syncthreads();  // row and twiddles published; the stencil rows are free
  filter_row_w<CT, NT, KW>(a, mb, fr, tw, b0, b1, (double *)(smem + R), red, PASS == 2 && fr.comp == CDGD);
}

// Two workgroups per CU (16 of the 20 waves that 96 VGPRs allow).  KW:
// longitudes per thread in the filter (1 up to 512 longitudes, 2 up to 1024).
template <typename CT, int PASS, bool IN_FULL, int KW>
__global__ void __launch_bounds__(kWTiles * kBlock) __attribute__((amdgpu_waves_per_eu(5, 8)))
k_wphase(const Args a) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ double red[kWTiles * 4 * 6];
  const int b = blockIdx.x;
  if (b < a.nnear_pad) {
    const int t = (b % kXcd) * (a.nnear_pad / kXcd) + b / kXcd;
    if (t < a.nnear) wnear_tile<CT, PASS, IN_FULL, KW>(a, a.frows[t], smem, red);
    return;
  }
  // row workgroups: tiles 2 r and 2 r + 1, each XCD owning a contiguous band
  const int bb = b - a.nnear_pad, per = (gridDim.x - a.nnear_pad) / kXcd;
  const int r = (bb % kXcd) * per + bb / kXcd;
  const int t0 = r * kWTiles;
  if (t0 >= a.ntiles) return;
  bool bar2 = false;
  if (PASS != 2) {
#pragma unroll
    for (int q = 0; q < kWTiles; ++q) {
      const int jq = far_tile_row(a, min(t0 + q, a.ntiles - 1));
      bar2 = bar2 || jq == 0 || jq == a.M - 1;
    }
  }
  const int s = threadIdx.x / kBlock;
  const bool live = t0 + s < a.ntiles;
  tend_tile<CT, PASS, IN_FULL, true>(a, live ? t0 + s : a.ntiles - 1, smem + s * wphase_region<CT>(a.fft_L),
                                     red + s * 4 * 6, bar2, live);
}

// ---------------------------------------------------------------------------
// k_dphase: the last application of ordinary_diffusion with its filter rows
// (diffusion_mod.F90:74-217 and the filter of dycore_mod.F90:658-666) in one
// launch, in place of k_diffuse followed by k_filter; same bits.
// ---------------------------------------------------------------------------

__host__ __device__ inline size_t dphase_region0(int L) {
  const size_t st = dif_stencil_bytes(), fb = (size_t)L * 16;
  return align16(st > fb ? st : fb);
}

__host__ __device__ inline size_t dphase_lds(int N, int L) {
  return dphase_region0(L) + (size_t)L * 16 + (size_t)tw_padded(N) * 16;
}

template <int CM>
__device__ __forceinline__ void dnear_row(const Args &a, const Member &mb, const RowCoef &cj, int j, double *rowbuf,
                                          unsigned char *smem, double *red) {
  const int N = a.N, tid = threadIdx.x;
  for (int g = 0; g < a.nseg; ++g) {
    if (g) __syncthreads();  // the previous segment's stencil rows have been read
    const int i0 = g * kSeg, W = min(N - i0, kSeg);
    const DiffPoint d = dif_segment<CM>(a, mb, cj, j, i0, W, smem, red);
    if (tid < W) row_put(a, rowbuf, i0 + tid, CM == TC_DU ? d.ud : (CM == TC_DV ? d.vd : d.gdd));
  }
}

template <typename CT, int KP>
__device__ __forceinline__ void dnear_tile(const Args &a, const FiltRow fr, unsigned char *smem, double *red) {
  const int N = a.N, L = a.fft_L, tid = threadIdx.x;
  const Member mb = member_at(a, fr.member);
  const int j = fr.row;
  const RowCoef cj = a.coef[j];
  double2 *b1 = (double2 *)smem;
  double2 *b0 = (double2 *)(smem + dphase_region0(L));
  double2 *tw = b0 + L;
  tw_to_lds(a, tw);
  double *rowbuf = (double *)b0;
  if (fr.comp == CDU) {
    dnear_row<TC_DU>(a, mb, cj, j, rowbuf, smem, red);
  } else if (fr.comp == CDV) {
    dnear_row<TC_DV>(a, mb, cj, j, rowbuf, smem, red);
  } else if (j == 0 || j == a.M - 1) {
    // a pole row's gdd is the same zonal sum at every longitude
    const double pole = dif_pole(a, mb, j, red);
    for (int l = tid; l < N; l += kBlock) row_put(a, rowbuf, l, pole);
  } else {
    dnear_row<TC_DGD>(a, mb, cj, j, rowbuf, smem, red);
  }
  __builtin_amdgcn_s_waitcnt(0);  // the twiddle copy has landed
  __syncthreads();  // row and twiddles published; the stencil rows are free for b1
  filter_row<CT, KP, true>(a, mb, fr, tw, b0, b1, red);
}

template <typename CT, int KP>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(phase_waves<KP>(), 8)))
k_dphase(const Args a) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ double red[4 * 6];
  const int b = blockIdx.x;
  if (b < a.nnear_pad) {
    const int t = (b % kXcd) * (a.nnear_pad / kXcd) + b / kXcd;
    if (t < a.nnear) dnear_tile<CT, KP>(a, a.frows[t], smem, red);
    return;
  }
  const int bb = b - a.nnear_pad, per = (gridDim.x - a.nnear_pad) / kXcd;
  const int t = (bb % kXcd) * per + bb / kXcd;
  if (t >= a.ntiles) return;
  dif_tile<true>(a, t, smem, red);
}

hipError_t launch_dphase(const Args &a, int prec, hipStream_t s) {
  if (a.N > 8 * kBlock) return hipErrorInvalidValue;
  const int rows = ((a.ntiles + kXcd - 1) / kXcd) * kXcd;
  dim3 g(a.nnear_pad + rows), b(kBlock);
  const size_t lds = dphase_lds(a.N, a.fft_L);
  const bool k8 = a.N > 4 * kBlock;
  if (prec == 0) {
    if (k8) hipLaunchKernelGGL((k_dphase<double, 8>), g, b, lds, s, a);
    else hipLaunchKernelGGL((k_dphase<double, 4>), g, b, lds, s, a);
  } else {
    if (k8) hipLaunchKernelGGL((k_dphase<float, 8>), g, b, lds, s, a);
    else hipLaunchKernelGGL((k_dphase<float, 4>), g, b, lds, s, a);
  }
  return hipGetLastError();
}

size_t phase_lds_bytes(int N, int L, int prec) {
  const size_t p = prec == 0 ? phase_lds_t<double>(N, L) : phase_lds_t<float>(N, L);
  const size_t d = dphase_lds(N, L);
  return p > d ? p : d;
}

template <typename CT, int KP>
static hipError_t launch_phase_k(const Args &a, int pass, bool in_full, hipStream_t s) {
  const int rows = ((a.ntiles + kXcd - 1) / kXcd) * kXcd;
  dim3 g(a.nnear_pad + rows), b(kBlock);
  const size_t lds = phase_lds_t<CT>(a.N, a.fft_L);
#define GD_PH(P, F) hipLaunchKernelGGL((k_phase<CT, P, F, KP>), g, b, lds, s, a)
  if (pass == 0) { if (in_full) GD_PH(0, true); else GD_PH(0, false); }
  else if (pass == 1) { if (in_full) GD_PH(1, true); else GD_PH(1, false); }
  else { if (in_full) GD_PH(2, true); else GD_PH(2, false); }
#undef GD_PH
  return hipGetLastError();
}

hipError_t launch_phase(const Args &a, int prec, int pass, bool in_full, hipStream_t s) {
  if (a.N > 8 * kBlock) return hipErrorInvalidValue;
  if (a.N > 4 * kBlock)
    return prec == 0 ? launch_phase_k<double, 8>(a, pass, in_full, s) : launch_phase_k<float, 8>(a, pass, in_full, s);
  return prec == 0 ? launch_phase_k<double, 4>(a, pass, in_full, s) : launch_phase_k<float, 4>(a, pass, in_full, s);
}

template <typename CT, int KW>
static hipError_t launch_wphase_k(const Args &a, int pass, bool in_full, hipStream_t s) {
  const int nwg = (a.ntiles + kWTiles - 1) / kWTiles;  // row workgroups
  const int rows = ((nwg + kXcd - 1) / kXcd) * kXcd;
  dim3 g(a.nnear_pad + rows), b(kWTiles * kBlock);
  const size_t lds = wphase_lds<CT>(a.N, a.fft_L);
#define GD_PH(P, F) hipLaunchKernelGGL((k_wphase<CT, P, F, KW>), g, b, lds, s, a)
  if (pass == 0) { if (in_full) GD_PH(0, true); else GD_PH(0, false); }
  else if (pass == 1) { if (in_full) GD_PH(1, true); else GD_PH(1, false); }
  else { if (in_full) GD_PH(2, true); else GD_PH(2, false); }
#undef GD_PH
  return hipGetLastError();
}

hipError_t launch_wphase(const Args &a, int prec, int pass, bool in_full, hipStream_t s) {
  // rows of two or three segments: one or two longitudes per filter thread
  if (a.nseg < 2 || a.nseg > 3 || a.N > a.nseg * kSeg) return hipErrorInvalidValue;
  if (a.nseg == 2)
    return prec == 0 ? launch_wphase_k<double, 1>(a, pass, in_full, s) : launch_wphase_k<float, 1>(a, pass, in_full, s);
  return prec == 0 ? launch_wphase_k<double, 2>(a, pass, in_full, s) : launch_wphase_k<float, 2>(a, pass, in_full, s);
}

size_t wphase_lds_bytes(int N, int L, int prec) {
  return prec == 0 ? wphase_lds<double>(N, L) : wphase_lds<float>(N, L);
}

template <typename CT, int KW>
static hipError_t configure_wphase_k(int max_lds) {
  const void *fns[] = {(const void *)k_wphase<CT, 0, true, KW>,  (const void *)k_wphase<CT, 0, false, KW>,
                       (const void *)k_wphase<CT, 1, true, KW>,  (const void *)k_wphase<CT, 1, false, KW>,
                       (const void *)k_wphase<CT, 2, true, KW>,  (const void *)k_wphase<CT, 2, false, KW>};
  for (const void *f : fns) {
    hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, max_lds);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

template <typename CT, int KP>
static hipError_t configure_phase_k(int max_lds) {
  const void *fns[] = {(const void *)k_phase<CT, 0, true, KP>,  (const void *)k_phase<CT, 0, false, KP>,
                       (const void *)k_phase<CT, 1, true, KP>,  (const void *)k_phase<CT, 1, false, KP>,
                       (const void *)k_phase<CT, 2, true, KP>,  (const void *)k_phase<CT, 2, false, KP>};
  for (const void *f : fns) {
    hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, max_lds);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

hipError_t configure_phase(int max_lds) {
  hipError_t e;
  if ((e = configure_phase_k<double, 4>(max_lds)) != hipSuccess) return e;
  if ((e = configure_phase_k<float, 4>(max_lds)) != hipSuccess) return e;
  if ((e = configure_phase_k<double, 8>(max_lds)) != hipSuccess) return e;
  if ((e = configure_phase_k<float, 8>(max_lds)) != hipSuccess) return e;
  if ((e = configure_wphase_k<double, 1>(max_lds)) != hipSuccess) return e;
  if ((e = configure_wphase_k<float, 1>(max_lds)) != hipSuccess) return e;
  if ((e = configure_wphase_k<double, 2>(max_lds)) != hipSuccess) return e;
  if ((e = configure_wphase_k<float, 2>(max_lds)) != hipSuccess) return e;
  const void *fns[] = {(const void *)k_dphase<double, 4>, (const void *)k_dphase<float, 4>,
                       (const void *)k_dphase<double, 8>, (const void *)k_dphase<float, 8>};
  for (const void *f : fns)
    if ((e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, max_lds)) != hipSuccess) return e;
  return hipSuccess;
}

}  // namespace gd
