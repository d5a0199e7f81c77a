// uasd_kernels_impl.hpp -- definitions of the launchers declared in
// uasd_launch.hpp, instantiated for one arithmetic type by the including
// translation unit.
#pragma once
#include "uasd_launch.hpp"

namespace uasd {

// Grid for n work items, padded to a multiple of the XCD count (the kernels
// remap blockIdx so each XCD takes a contiguous range; padding blocks exit).
static inline unsigned grid_for(int64_t n) {
  const int64_t b = (n + kBlock - 1) / kBlock;
  return (unsigned)((b + kXCD - 1) / kXCD * kXCD);
}

template <typename R>
hipError_t launch_field(hipStream_t s, const HamDev<R> &h, const Quad<R> *src, Quad<R> *beff,
                        double *en) {
  const int64_t n = (int64_t)h.N * h.M;
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_field<R>, dim3(grid_for(n)), dim3(kBlock), 0, s, h, src, beff, en);
  return hipGetLastError();
}

template <typename R, int ALG, int HALF, bool FUSED>
static hipError_t llg_one(hipStream_t s, const HamDev<R> &h, const LLGDev<R> &prm,
                          const Quad<R> *e, Quad<R> *p, Quad<R> *b2, const Quad<R> *beff,
                          Quad<R> *out) {
  const int64_t n = (int64_t)h.N * h.M;
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL((k_llg<R, ALG, HALF, FUSED>), dim3(grid_for(n)), dim3(kBlock), 0, s, h, prm,
                     e, p, b2, beff, out);
  return hipGetLastError();
}

template <typename R>
hipError_t launch_llg(hipStream_t s, int alg, int half, bool fused, const HamDev<R> &h,
                      const LLGDev<R> &prm, const Quad<R> *e, Quad<R> *p, Quad<R> *b2,
                      const Quad<R> *beff, Quad<R> *out) {
#define UASD_LLG_CASE(A, H, F)                              \
  if (alg == A && half == H && fused == F)                  \
    return llg_one<R, A, H, F>(s, h, prm, e, p, b2, beff, out);
  UASD_LLG_CASE(1, 0, true)
  UASD_LLG_CASE(1, 1, true)
  UASD_LLG_CASE(1, 0, false)
  UASD_LLG_CASE(1, 1, false)
  UASD_LLG_CASE(5, 0, true)
  UASD_LLG_CASE(5, 1, true)
  UASD_LLG_CASE(5, 0, false)
  UASD_LLG_CASE(5, 1, false)
#undef UASD_LLG_CASE
  return hipErrorInvalidValue;
}

template <typename R>
hipError_t launch_mc(hipStream_t s, const HamDev<R> &h, const MCParams &mp, Quad<R> *spins) {
  const int64_t n = (int64_t)mp.nsites * h.M;
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_mc_colour<R>, dim3(grid_for(n)), dim3(kBlock), 0, s, h, mp, spins);
  return hipGetLastError();
}

template <typename R>
hipError_t launch_reduce(hipStream_t s, int N, int M, const Quad<R> *spins, const double *en,
                         int nv, double *partial, int nblocks, double *out) {
  if (N == 0 || M == 0) return hipSuccess;
  const int chunk = (N + nblocks - 1) / nblocks;
  dim3 g1(nblocks, M);
  if (nv == 4) {
    hipLaunchKernelGGL((k_reduce_stage1<R, 4>), g1, dim3(kBlock), 0, s, N, M, chunk, spins, en,
                       partial);
    hipLaunchKernelGGL(k_reduce_stage2<4>, dim3((M * 4 + 63) / 64), dim3(64), 0, s, M, nblocks,
                       partial, out);
  } else {
    hipLaunchKernelGGL((k_reduce_stage1<R, 3>), g1, dim3(kBlock), 0, s, N, M, chunk, spins, en,
                       partial);
    hipLaunchKernelGGL(k_reduce_stage2<3>, dim3((M * 3 + 63) / 64), dim3(64), 0, s, M, nblocks,
                       partial, out);
  }
  return hipGetLastError();
}

#define UASD_INSTANTIATE(R)                                                                    \
  template hipError_t launch_field<R>(hipStream_t, const HamDev<R> &, const Quad<R> *,         \
                                      Quad<R> *, double *);                                    \
  template hipError_t launch_llg<R>(hipStream_t, int, int, bool, const HamDev<R> &,            \
                                    const LLGDev<R> &, const Quad<R> *, Quad<R> *, Quad<R> *,  \
                                    const Quad<R> *, Quad<R> *);                               \
  template hipError_t launch_mc<R>(hipStream_t, const HamDev<R> &, const MCParams &,           \
                                   Quad<R> *);                                                 \
  template hipError_t launch_reduce<R>(hipStream_t, int, int, const Quad<R> *, const double *, \
                                       int, double *, int, double *);

}  // namespace uasd
