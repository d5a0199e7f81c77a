// uasd_launch.hpp -- host-callable launchers for the kernels in
// uasd_kernels.hpp.  Each arithmetic type is instantiated in its own
// translation unit (uasd_kernels_f64.hip / uasd_kernels_f32.hip) so the FP64
// unit can be compiled without multiply-add contraction.
#pragma once
#include "uasd_kernels.hpp"

namespace uasd {

constexpr int kBlock = 256;  // largest workgroup the kernels are built for

// Launch shape of the gathering kernels (chosen by the autotuner):
//   gather  neighbours whose loads are issued together: 2, 4, 8 or 16
//   block   workgroup size: 64, 128 or 256
//   nt      1: neighbour lists read with non-temporal loads
//   idx32   1: int32 list entries even where packed int16 entries exist
//   tile    fused LLG only: 0 = list gathers, k > 0 = LDS-tiled stencil with
//           tile shape kTileShapes[k-1] (periodic lattices with a stencil)
constexpr int kNumTileShapes = 15;
constexpr int kTileShapes[kNumTileShapes][3] = {
    {8, 8, 4}, {16, 8, 2}, {32, 4, 2}, {16, 4, 4}, {16, 16, 1}, {32, 8, 1}, {64, 4, 1}, {8, 4, 4},
    {4, 4, 4}, {8, 8, 2},  {16, 8, 1}, {8, 8, 1},  {16, 8, 4},  {8, 16, 4}, {8, 8, 8}};
struct Tune {
  int gather = 4;
  int block = 256;
  int nt = 0;
  int idx32 = 0;
  int tile = 0;
  int pipe = 0;  // tile kernel: 1 = resident workgroups loop over several tiles
  int pair = 0;  // MC colour pass: 1 = two adjacent ensembles per thread (M even)
};

inline bool tune_valid(const Tune &t) {
  return (t.gather == 2 || t.gather == 4 || t.gather == 8 || t.gather == 16) &&
         (t.block == 64 || t.block == 128 || t.block == 256) && (t.nt == 0 || t.nt == 1) &&
         (t.idx32 == 0 || t.idx32 == 1) && t.tile >= 0 && t.tile <= kNumTileShapes &&
         (t.pipe == 0 || (t.pipe == 1 && t.tile > 0)) &&
         (t.pair == 0 || (t.pair == 1 && t.tile == 0));
}

// The Hamiltonian view a launch uses under a tuning choice.
template <typename R>
inline HamDev<R> tuned(const HamDev<R> &h, const Tune &t) {
  HamDev<R> hh = h;
  hh.nt = t.nt;
  if (t.idx32) {
    hh.nbr16 = nullptr;
    hh.dnbr16 = nullptr;
  }
  return hh;
}

// parts (nullable) receives beff1 then beff2, N*M quads each.
template <typename R>
hipError_t launch_field(hipStream_t s, const Tune &t, const HamDev<R> &h, const Quad<R> *src,
                        Quad<R> *beff, double *en, Quad<R> *parts);

// alg: 1 midpoint / 5 Depondt; half: 0 predictor, 1 corrector
template <typename R>
hipError_t launch_llg(hipStream_t s, const Tune &t, int alg, int half, bool fused,
                      const HamDev<R> &h, const LLGDev<R> &prm, const Quad<R> *e, Quad<R> *p,
                      Quad<R> *b2, const Quad<R> *beff, Quad<R> *out);

template <typename R>
hipError_t launch_mc(hipStream_t s, const Tune &t, const HamDev<R> &h, const MCParams &mp,
                     Quad<R> *spins);

// Fused LLG half-step with the LDS-tiled stencil kernel, work units (tile,
// ensemble).
template <typename R>
hipError_t launch_llg_tile(hipStream_t s, TileDev t, int pipe, int alg, int half,
                           const HamDev<R> &h, const LLGDev<R> &prm, const Quad<R> *e,
                           Quad<R> *p, Quad<R> *b2, Quad<R> *out);
// LDS bytes the tile kernel needs for a halo tile of hv cells
// Component plane size of k_llg_tile for NA*HV halo values: the smallest of
// kTilePlanes that holds them, or 0.
inline int tile_plane(int nahv) {
  for (int k = 0; k < kNumTilePlanes; ++k)
    if (nahv <= kTilePlanes[k]) return kTilePlanes[k];
  return 0;
}
// LDS of k_llg_tile: three planes of `plane` values (NA*HV used), then two
// staging tables of the halo (ntab = HX + NA*HY*HZ ints each)
inline size_t tile_lds_bytes(size_t rsize, int plane, int ntab) {
  return 3 * (size_t)plane * rsize + 2 * (size_t)ntab * sizeof(int);
}

// FP64 sums per ensemble of emomM (nv = 3) or of emomM and en (nv = 4);
// partial must hold M*nblocks*nv doubles, out M*nv doubles.
template <typename R>
hipError_t launch_reduce(hipStream_t s, int N, int M, const Quad<R> *spins, const double *en,
                         int nv, double *partial, int nblocks, double *out);

// Per-segment FP64 sums (see k_segsum_*): chunks (nchunks), seg_chunk
// (nseg+1, chunk ranges of each segment), partial (nchunks*M*nv doubles),
// out (nseg*M*nv doubles).
template <typename R>
hipError_t launch_segsum(hipStream_t s, int M, int nseg, int nchunks, const int2 *chunks,
                         const int *seg_chunk, const Quad<R> *spins, const double *en, int nv,
                         double *partial, double *out);

// Term-resolved energies (see k_energy_terms) per (segment, ensemble):
// out[(seg*M + ens)*4 + t] for t = exchange, anisotropy, DM, Zeeman, sums
// over the segment's sites in field units; partial holds nchunks*M*4 doubles.
template <typename R>
hipError_t launch_energy_terms(hipStream_t s, const HamDev<R> &h, int nseg, int nchunks,
                               const int2 *chunks, const int *seg_chunk, const Quad<R> *spins,
                               double *partial, double *out);

// ctr[idx] = value (add = 0) or ctr[idx] += value (add = 1), on the stream.
hipError_t launch_counter(hipStream_t s, int idx, uint64_t *ctr, uint64_t value, int add);

// Appends one cumulant sample per (segment, ensemble) from out of launch_segsum.
hipError_t launch_accumulate(hipStream_t s, int nseg, int M, const int *seg_n,
                             const double *sums, int nv, double *acc);

}  // namespace uasd
