// FP64 instantiation of the engine kernels.  Built with -ffp-contract=off:
// every multiply and add rounds separately, as in the reference Fortran,
// which keeps the drop-in path within rounding of the FP64 oracle.
#include "uasd_kernels_impl.hpp"
namespace uasd {
UASD_INSTANTIATE(double)

// precision-independent launchers live in this unit
hipError_t launch_accumulate(hipStream_t s, int nseg, int M, const int *seg_n, const double *sums,
                             int nv, double *acc) {
  const int n = nseg * M;
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_accumulate<kAccWidth>, dim3((n + 255) / 256), dim3(256), 0, s, nseg, M, seg_n, sums,
                     nv, acc);
  return hipGetLastError();
}

hipError_t launch_counter(hipStream_t s, int idx, uint64_t *ctr, uint64_t value, int add) {
  if (idx == 0)
    hipLaunchKernelGGL(k_counter<0>, dim3(1), dim3(64), 0, s, ctr, value, add);
  else
    hipLaunchKernelGGL(k_counter<1>, dim3(1), dim3(64), 0, s, ctr, value, add);
  return hipGetLastError();
}
}  // namespace uasd
