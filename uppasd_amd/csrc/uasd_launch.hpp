// uasd_launch.hpp -- host-callable launchers for the kernels in
// uasd_kernels.hpp.  Each arithmetic type is instantiated in its own
// translation unit (uasd_kernels_f64.hip / uasd_kernels_f32.hip) so the FP64
// unit can be compiled without multiply-add contraction.
#pragma once
#include "uasd_kernels.hpp"

namespace uasd {

constexpr int kBlock = 256;

template <typename R>
hipError_t launch_field(hipStream_t s, const HamDev<R> &h, const Quad<R> *src, Quad<R> *beff,
                        double *en);

// alg: 1 midpoint / 5 Depondt; half: 0 predictor, 1 corrector
template <typename R>
hipError_t launch_llg(hipStream_t s, int alg, int half, bool fused, const HamDev<R> &h,
                      const LLGDev<R> &prm, const Quad<R> *e, Quad<R> *p, Quad<R> *b2,
                      const Quad<R> *beff, Quad<R> *out);

template <typename R>
hipError_t launch_mc(hipStream_t s, const HamDev<R> &h, const MCParams &mp, Quad<R> *spins);

// FP64 sums per ensemble of emomM (nv = 3) or of emomM and en (nv = 4);
// partial must hold M*nblocks*nv doubles, out M*nv doubles.
template <typename R>
hipError_t launch_reduce(hipStream_t s, int N, int M, const Quad<R> *spins, const double *en,
                         int nv, double *partial, int nblocks, double *out);

}  // namespace uasd
