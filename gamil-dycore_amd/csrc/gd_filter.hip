// gd_filter.hip -- k_filter: the polar filter of every staged row as its own
// workgroup (filter_row in gd_fft.hpp), after a k_tend or k_diffuse launch
// that staged the rows.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "gd_common.hpp"
#include "gd_device.hpp"
#include "gd_fft.hpp"
#include "gd_launch.hpp"

namespace gd {

template <typename CT, int KP>
__global__ void __launch_bounds__(kBlock) k_filter(const Args a) {
  const int gi = blockIdx.x;
  if (gi >= a.ntiles) return;
  const FiltRow fr = a.frows[gi];
  const Member mb = member_at(a, fr.member);
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ double red[4 * 2];
  double2 *tw = (double2 *)smem;
  double2 *b0 = tw + tw_entries(a.N);
  filter_row<CT, KP, false>(a, mb, fr, tw, b0, b0 + a.fft_L, red);
}

size_t filter_lds_bytes(int N, int L) { return filter_lds_bytes_h(N, L); }

template <typename CT>
static hipError_t launch_filter_t(const Args &a, hipStream_t s) {
  dim3 g(a.ntiles), b(kBlock);
  const size_t lds = filter_lds_bytes(a.N, a.fft_L);
  if (a.N <= 4 * kBlock)
    hipLaunchKernelGGL((k_filter<CT, 4>), g, b, lds, s, a);
  else if (a.N <= 8 * kBlock)
    hipLaunchKernelGGL((k_filter<CT, 8>), g, b, lds, s, a);
  else
    hipLaunchKernelGGL((k_filter<CT, 16>), g, b, lds, s, a);
  return hipGetLastError();
}

hipError_t launch_filter(const Args &a, int prec, hipStream_t s) {
  if (a.ntiles <= 0) return hipSuccess;
  return prec == 0 ? launch_filter_t<double>(a, s) : launch_filter_t<float>(a, s);
}

hipError_t configure_filter(int max_lds) {
  const void *fns[] = {(const void *)k_filter<double, 4>, (const void *)k_filter<double, 8>,
                       (const void *)k_filter<double, 16>, (const void *)k_filter<float, 4>,
                       (const void *)k_filter<float, 8>,  (const void *)k_filter<float, 16>};
  for (const void *f : fns) {
    hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, max_lds);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

}  // namespace gd
