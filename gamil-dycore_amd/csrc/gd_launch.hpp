// gd_launch.hpp -- launch entry points of the kernels and their LDS sizes.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

#include "gd_device.hpp"

namespace gd {

__host__ __device__ inline size_t align16(size_t x) { return (x + 15) & ~(size_t)15; }

// k_diffuse: G (4 rows), u (3), v (3), all double, + 16 doubles.
__host__ __device__ inline size_t diffuse_lds_bytes(int N) {
  return (size_t)10 * N * 8 + 16 * sizeof(double);
}
// k_filter_coef: 512 bytes of row pointers, a 16 x 64 operand block and a
// 64 x 64 basis block (pitch 65), sized for FP64 operands, and an FP64 copy
// of the operand block for the s1 inner product.
__host__ __device__ inline size_t filt_coef_lds_bytes() {
  return 512 + (size_t)(16 + 64) * 65 * 8 + (size_t)16 * 64 * 8;
}
// k_filter_apply: a 64 x 64 basis block (pitch 65) + the tile's 8 x kpad
// coefficients, sized for FP64 operands.
__host__ __device__ inline size_t filt_apply_lds_bytes(int kpad) {
  return ((size_t)64 * 65 + (size_t)kFiltRows * kpad) * 8;
}

// prec: 0 = FP64 plain, 1 = mixed (FP32 compute type)
hipError_t launch_tend(const Args &a, int prec, int pass, bool in_full, hipStream_t s);
hipError_t launch_beta(const Args &a, int nmem, hipStream_t s);
hipError_t launch_correct(const Args &a, int prec, hipStream_t s);
hipError_t launch_diffuse(const Args &a, int prec, hipStream_t s);
// mode 0: iap_transform of X after diffusion; 1: X <- slot s_in with u, v
// derived; 2: diagnostics of X only.  All modes write the mass/energy row
// partials and, from the member's last workgroup, the diag_out entry.
hipError_t launch_finish(const Args &a, int mode, hipStream_t s);
hipError_t launch_divvor(const Args &a, hipStream_t s);
hipError_t configure_kernels(int max_lds);
hipError_t launch_filter_coef(const Args &a, int prec, hipStream_t s);
hipError_t launch_filter_apply(const Args &a, int prec, hipStream_t s);
hipError_t launch_mfma_layout(const double *A, const double *B, double *D, const float *Af,
                              const float *Bf, float *Df, hipStream_t s);

}  // namespace gd
