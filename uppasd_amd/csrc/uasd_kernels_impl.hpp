// uasd_kernels_impl.hpp -- definitions of the launchers declared in
// uasd_launch.hpp, instantiated for one arithmetic type by the including
// translation unit.
#pragma once
#include <type_traits>

#include "uasd_launch.hpp"

namespace uasd {

// Grid for n work items, padded to a multiple of the XCD count (the kernels
// remap blockIdx so each XCD takes a contiguous range; padding blocks exit).
static inline unsigned grid_for(int64_t n, int block) {
  const int64_t b = (n + block - 1) / block;
  return (unsigned)((b + kXCD - 1) / kXCD * kXCD);
}

// Calls f(std::integral_constant<int, G>) for the gather width of t.
template <typename F>
static hipError_t with_gather(const Tune &t, F &&f) {
  switch (t.gather) {
    case 2: return f(std::integral_constant<int, 2>{});
    case 4: return f(std::integral_constant<int, 4>{});
    case 8: return f(std::integral_constant<int, 8>{});
    case 16: return f(std::integral_constant<int, 16>{});
    default: return hipErrorInvalidValue;
  }
}

template <typename R>
hipError_t launch_field(hipStream_t s, const Tune &t, const HamDev<R> &h, const Quad<R> *src,
                        Quad<R> *beff, double *en, Quad<R> *parts) {
  const int64_t n = (int64_t)h.N * h.M;
  if (n == 0) return hipSuccess;
  if (!tune_valid(t)) return hipErrorInvalidValue;
  const HamDev<R> hh = tuned(h, t);
  return with_gather(t, [&](auto g) {
    hipLaunchKernelGGL((k_field<R, decltype(g)::value>), dim3(grid_for(n, t.block)),
                       dim3(t.block), 0, s, hh, src, beff, en, parts);
    return hipGetLastError();
  });
}

template <typename R, int ALG, int HALF, bool FUSED>
static hipError_t llg_one(hipStream_t s, const Tune &t, const HamDev<R> &h,
                          const LLGDev<R> &prm, const Quad<R> *e, Quad<R> *p, Quad<R> *b2,
                          const Quad<R> *beff, Quad<R> *out) {
  const int64_t n = (int64_t)h.N * h.M;
  if (n == 0) return hipSuccess;
  const dim3 grid(grid_for(n, t.block)), block(t.block);
  const HamDev<R> hh = tuned(h, t);
  if (!FUSED) {  // no gather: one instance
    hipLaunchKernelGGL((k_llg<R, ALG, HALF, false, 4>), grid, block, 0, s, hh, prm, e, p, b2,
                       beff, out);
    return hipGetLastError();
  }
  return with_gather(t, [&](auto g) {
    hipLaunchKernelGGL((k_llg<R, ALG, HALF, true, decltype(g)::value>), grid, block, 0, s, hh,
                       prm, e, p, b2, beff, out);
    return hipGetLastError();
  });
}

template <typename R>
hipError_t launch_llg(hipStream_t s, const Tune &t, int alg, int half, bool fused,
                      const HamDev<R> &h, const LLGDev<R> &prm, const Quad<R> *e, Quad<R> *p,
                      Quad<R> *b2, const Quad<R> *beff, Quad<R> *out) {
  if (!tune_valid(t)) return hipErrorInvalidValue;
#define UASD_LLG_CASE(A, H, F)             \
  if (alg == A && half == H && fused == F) \
    return llg_one<R, A, H, F>(s, t, h, prm, e, p, b2, beff, out);
  UASD_LLG_CASE(1, 0, true)
  UASD_LLG_CASE(1, 1, true)
  UASD_LLG_CASE(1, 0, false)
  UASD_LLG_CASE(1, 1, false)
  UASD_LLG_CASE(4, 0, true)
  UASD_LLG_CASE(4, 1, true)
  UASD_LLG_CASE(4, 0, false)
  UASD_LLG_CASE(4, 1, false)
  UASD_LLG_CASE(6, 0, true)
  UASD_LLG_CASE(6, 1, true)
  UASD_LLG_CASE(6, 0, false)
  UASD_LLG_CASE(6, 1, false)
  UASD_LLG_CASE(7, 0, true)
  UASD_LLG_CASE(7, 1, true)
  UASD_LLG_CASE(7, 0, false)
  UASD_LLG_CASE(7, 1, false)
  UASD_LLG_CASE(5, 0, true)
  UASD_LLG_CASE(5, 1, true)
  UASD_LLG_CASE(5, 0, false)
  UASD_LLG_CASE(5, 1, false)
#undef UASD_LLG_CASE
  return hipErrorInvalidValue;
}

template <typename R>
hipError_t launch_mc(hipStream_t s, const Tune &t, const HamDev<R> &h, const MCParams &mp,
                     Quad<R> *spins) {
  const int64_t n = (int64_t)mp.nsites * h.M;
  if (n == 0) return hipSuccess;
  if (!tune_valid(t)) return hipErrorInvalidValue;
  const HamDev<R> hh = tuned(h, t);
  // two adjacent ensembles per thread where M is even
  const bool two = t.pair && h.M % 2 == 0;
  MCParams m = mp;
  m.pdiv = make_fast_div((uint32_t)(two ? h.M / 2 : h.M));
  return with_gather(t, [&](auto g) {
    constexpr int G = decltype(g)::value;
    const dim3 grid(grid_for(two ? n / 2 : n, t.block)), block(t.block);
    if (mp.heat_bath) {
      if (two)
        hipLaunchKernelGGL((k_mc_colour<R, G, 2, true>), grid, block, 0, s, hh, m, spins);
      else
        hipLaunchKernelGGL((k_mc_colour<R, G, 1, true>), grid, block, 0, s, hh, m, spins);
    } else {
      if (two)
        hipLaunchKernelGGL((k_mc_colour<R, G, 2, false>), grid, block, 0, s, hh, m, spins);
      else
        hipLaunchKernelGGL((k_mc_colour<R, G, 1, false>), grid, block, 0, s, hh, m, spins);
    }
    return hipGetLastError();
  });
}

template <typename R>
hipError_t launch_llg_tile(hipStream_t s, TileDev t, int pipe, int alg, int half,
                           const HamDev<R> &h, const LLGDev<R> &prm, const Quad<R> *e,
                           Quad<R> *p, Quad<R> *b2, Quad<R> *out) {
  // work units: (tile, ensemble) pairs
  const int64_t nunits64 = (int64_t)t.ntx * t.nty * t.ntz * h.M;
  if (nunits64 == 0) return hipSuccess;
  if (nunits64 >= (int64_t)1 << 31 || (int64_t)h.N * h.M >= (int64_t)1 << 31)
    return hipErrorInvalidValue;
  const int nunits = (int)nunits64;
  const int threads = t.TX * t.TY * t.TZ * t.NA;
  if (threads > 1024) return hipErrorInvalidValue;
  const int plane = tile_plane(t.NA * t.HV);
  // DM, anisotropy or Zeeman terms present: the full kernel
  const bool full = h.Wd > 0 || h.an_type != nullptr || h.ext_mode != 0;
  // several ensembles: the (tile, ensemble) unit instance
  const bool ens = h.M > 1;
  if (plane == 0) return hipErrorInvalidValue;
  if (t.rs < 1) return hipErrorInvalidValue;  // at least one halo row per round
  const size_t lds = tile_lds_bytes(sizeof(R), plane, t.HX + t.NA * t.HY * t.HZ);
  // one unit per workgroup, or (pipe) as many workgroups as are resident at
  // once (occupancy x CUs, cached per kernel and shape), each looping over a
  // run of consecutive units
  auto tiles_per_wg = [&](const void *kern, size_t &c_lds, int &c_thr, int &c_slots) -> hipError_t {
    t.per = 1;
    if (!pipe) return hipSuccess;
    if (c_lds != lds || c_thr != threads) {
      int dev = 0, ncu = 0, occ = 0;
      hipError_t err = hipGetDevice(&dev);
      if (err == hipSuccess)
        err = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
      if (err == hipSuccess)
        err = hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kern, threads, lds);
      if (err != hipSuccess) return err;
      c_lds = lds;
      c_thr = threads;
      c_slots = std::max(1, occ) * std::max(1, ncu);
    }
    t.per = (nunits + c_slots - 1) / c_slots;
    return hipSuccess;
  };
#define UASD_TILE_CASE(A, H, PL, F, E)                                                          \
  if (alg == A && half == H && plane == PL && full == F && ens == E) {                          \
    static size_t lds_set = 0; /* dynamic LDS above 64 KiB must be requested */                \
    if (lds > lds_set) {                                                                        \
      hipError_t err = hipFuncSetAttribute((const void *)&k_llg_tile<R, A, H, PL, F, E>,        \
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds); \
      if (err != hipSuccess) return err;                                                        \
      lds_set = lds;                                                                            \
    }                                                                                           \
    static size_t c_lds = 0;                                                                    \
    static int c_thr = 0, c_slots = 0;                                                          \
    hipError_t gerr = tiles_per_wg((const void *)&k_llg_tile<R, A, H, PL, F, E>, c_lds, c_thr,  \
                                   c_slots);                                                    \
    if (gerr != hipSuccess) return gerr;                                                        \
    const int nwg = (nunits + t.per - 1) / t.per;                                               \
    const dim3 grid((nwg + kXCD - 1) / kXCD * kXCD), block(threads);                            \
    hipLaunchKernelGGL((k_llg_tile<R, A, H, PL, F, E>), grid, block, lds, s, h, prm, t, e, p, b2, \
                       out);                                                                    \
    return hipGetLastError();                                                                   \
  }
#define UASD_TILE_FULL(A, H, PL)                                                                \
  UASD_TILE_CASE(A, H, PL, false, false)                                                        \
  UASD_TILE_CASE(A, H, PL, true, false)                                                         \
  UASD_TILE_CASE(A, H, PL, false, true)                                                         \
  UASD_TILE_CASE(A, H, PL, true, true)
#define UASD_TILE_PLANES(A, H)                                                                  \
  UASD_TILE_FULL(A, H, kTilePlanes[0])                                                          \
  UASD_TILE_FULL(A, H, kTilePlanes[1])                                                          \
  UASD_TILE_FULL(A, H, kTilePlanes[2])                                                          \
  UASD_TILE_FULL(A, H, kTilePlanes[3])
  static_assert(kNumTilePlanes == 4, "one case per plane size");
  UASD_TILE_PLANES(1, 0)
  UASD_TILE_PLANES(1, 1)
  UASD_TILE_PLANES(4, 0)
  UASD_TILE_PLANES(4, 1)
  UASD_TILE_PLANES(5, 0)
  UASD_TILE_PLANES(5, 1)
  UASD_TILE_PLANES(6, 0)
  UASD_TILE_PLANES(6, 1)
  UASD_TILE_PLANES(7, 0)
  UASD_TILE_PLANES(7, 1)
#undef UASD_TILE_PLANES
#undef UASD_TILE_FULL
#undef UASD_TILE_CASE
  return hipErrorInvalidValue;
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

template <typename R>
hipError_t launch_segsum(hipStream_t s, int M, int nseg, int nchunks, const int2 *chunks,
                         const int *seg_chunk, const Quad<R> *spins, const double *en, int nv,
                         double *partial, double *out) {
  if (nseg == 0 || M == 0 || nchunks == 0) return hipSuccess;
  const bool wide = kBlock % M == 0;
  const dim3 g = wide ? dim3(nchunks) : dim3(nchunks, M);
  if (nv == 4) {
    if (wide)
      hipLaunchKernelGGL((k_segsum_wide<R, 4>), g, dim3(kBlock), 0, s, M, chunks, spins, en,
                         partial);
    else
      hipLaunchKernelGGL((k_segsum_strided<R, 4>), g, dim3(kBlock), 0, s, M, chunks, spins, en,
                         partial);
    hipLaunchKernelGGL(k_segsum_final<4>, dim3((nseg * M * 4 + 255) / 256), dim3(256), 0, s, nseg,
                       M, seg_chunk, partial, out);
  } else {
    if (wide)
      hipLaunchKernelGGL((k_segsum_wide<R, 3>), g, dim3(kBlock), 0, s, M, chunks, spins, en,
                         partial);
    else
      hipLaunchKernelGGL((k_segsum_strided<R, 3>), g, dim3(kBlock), 0, s, M, chunks, spins, en,
                         partial);
    hipLaunchKernelGGL(k_segsum_final<3>, dim3((nseg * M * 3 + 255) / 256), dim3(256), 0, s, nseg,
                       M, seg_chunk, partial, out);
  }
  return hipGetLastError();
}

template <typename R>
hipError_t launch_energy_terms(hipStream_t s, const HamDev<R> &h, int nseg, int nchunks,
                               const int2 *chunks, const int *seg_chunk, const Quad<R> *spins,
                               double *partial, double *out) {
  if (nseg == 0 || h.M == 0 || nchunks == 0) return hipSuccess;
  hipLaunchKernelGGL((k_energy_terms<R>), dim3(nchunks, h.M), dim3(kBlock), 0, s, h, chunks,
                     spins, partial);
  hipLaunchKernelGGL(k_segsum_final<4>, dim3((nseg * h.M * 4 + 255) / 256), dim3(256), 0, s,
                     nseg, h.M, seg_chunk, partial, out);
  return hipGetLastError();
}

#define UASD_INSTANTIATE(R)                                                                   \
  template hipError_t launch_field<R>(hipStream_t, const Tune &, const HamDev<R> &,           \
                                      const Quad<R> *, Quad<R> *, double *, Quad<R> *);       \
  template hipError_t launch_llg<R>(hipStream_t, const Tune &, int, int, bool,                \
                                    const HamDev<R> &, const LLGDev<R> &, const Quad<R> *,    \
                                    Quad<R> *, Quad<R> *, const Quad<R> *, Quad<R> *);        \
  template hipError_t launch_mc<R>(hipStream_t, const Tune &, const HamDev<R> &,              \
                                   const MCParams &, Quad<R> *);                              \
  template hipError_t launch_llg_tile<R>(hipStream_t, TileDev, int, int, int,              \
                                         const HamDev<R> &, const LLGDev<R> &,                \
                                         const Quad<R> *, Quad<R> *, Quad<R> *, Quad<R> *);   \
  template hipError_t launch_reduce<R>(hipStream_t, int, int, const Quad<R> *, const double *, \
                                       int, double *, int, double *);                         \
  template hipError_t launch_segsum<R>(hipStream_t, int, int, int, const int2 *, const int *,  \
                                       const Quad<R> *, const double *, int, double *,        \
                                       double *);                                             \
  template hipError_t launch_energy_terms<R>(hipStream_t, const HamDev<R> &, int, int,        \
                                             const int2 *, const int *, const Quad<R> *,      \
                                             double *, double *);

}  // namespace uasd
