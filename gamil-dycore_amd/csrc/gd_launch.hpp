// gd_launch.hpp -- launch entry points of the kernels and their LDS sizes.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <vector>

#include "gd_device.hpp"

namespace gd {

__host__ __device__ inline size_t align16(size_t x) { return (x + 15) & ~(size_t)15; }

// k_filter: twiddle table (N complex) + two FFT buffers of L complex points.
size_t filter_lds_bytes(int N, int L);

// k_tend: one segment's stencil rows.  phase_lds_bytes: the larger of the
// fused phase and fused diffusion launches.
size_t tend_lds_bytes(const Args &a, size_t cts);

// prec: 0 = FP64 plain, 1 = mixed (FP32 compute type)
hipError_t launch_tend(const Args &a, int prec, int pass, bool in_full, hipStream_t s);
// beta of predict_correct and the corrector update in one launch
// (upd_wpm workgroups for each of nmem_launch members)
hipError_t launch_update(const Args &a, int prec, hipStream_t s);
// isp_splitting tendency algebra and state updates (mode a.isp_mode)
hipError_t launch_isp(const Args &a, int prec, hipStream_t s);
hipError_t launch_diffuse(const Args &a, int prec, hipStream_t s);
// mode 0: iap_transform of X after diffusion; 1: X <- slot s_in with u, v
// derived; 2: diagnostics of X only.  All modes write the mass/energy
// partials; launch_diag (one workgroup per member) turns them into the
// diag_out entry.
hipError_t launch_finish(const Args &a, int mode, hipStream_t s);
hipError_t launch_diag(const Args &a, int nmem, hipStream_t s);
hipError_t launch_divvor(const Args &a, hipStream_t s);
// check_antisymmetry's ten inner products (and their magnitudes) of member
// a.m0's state X, left in its Tb dv field (entries 0..19; Tb du is scratch)
hipError_t launch_terms(const Args &a, hipStream_t s);
hipError_t configure_kernels(int max_lds);
// polar filter of the staged rows + rescale + phase epilogue (gd_filter.hip)
hipError_t launch_filter(const Args &a, int prec, hipStream_t s);
hipError_t configure_filter(int max_lds);
// one tendency phase with its filter rows in one launch (gd_phase.hip);
// N <= 4 * kBlock
hipError_t launch_phase(const Args &a, int prec, int pass, bool in_full, hipStream_t s);
// the last diffusion application with its filter rows in one launch
// (gd_phase.hip); N <= 4 * kBlock
hipError_t launch_dphase(const Args &a, int prec, hipStream_t s);
size_t phase_lds_bytes(int N, int L, int prec);
// the same phase with workgroups of two tiles (k_wphase, gd_phase.hip): rows
// of two or three segments; a.ntiles counts segment tiles as for launch_phase
hipError_t launch_wphase(const Args &a, int prec, int pass, bool in_full, hipStream_t s);
size_t wphase_lds_bytes(int N, int L, int prec);
// momentum rows of the filter as batched MFMA products (gd_filter_mm.hip):
// one workgroup per (momentum slot in a.mslots, block of 16 members)
hipError_t launch_filter_mm(const Args &a, bool bf16, hipStream_t s);
hipError_t configure_filter_mm(int max_lds);
size_t filter_mm_lds_bytes(int N, bool bf16);
// the basis matrices of k_filter_mm (Phi and Phi^T, FP32 or BF16 high and
// low parts), built once per handle
size_t mm_basis_bytes(int N);
void mm_basis_fill(int N, bool bf16, std::vector<unsigned char> &out);
hipError_t configure_phase(int max_lds);

}  // namespace gd
