// uasd_kernels.hpp -- device code of the spin-update engine, templated on the
// arithmetic type R (double for the plain drop-in, float for the reduced
// precision form).  Included by uasd_kernels_f64.hip (built with
// -ffp-contract=off so every product and sum rounds as the reference's
// Fortran does) and uasd_kernels_f32.hip.
//
// Data layout in HBM (one context = one lattice, M ensembles):
//   spins      quads {x, y, z, m} of R per q = site*M + ens: the unit vector
//              and the moment magnitude, so one 16/32-byte load gathers a
//              neighbour's emomM = e*m (bit-identical to the reference's
//              stored emomM = emom2*mmom2)
//   neighbours slot-major ELL int32 nbr[j*N + site], 0-based
//   couplings  slot-major per Hamiltonian row cpl[j*nrows + row]; rows are
//              the distinct (length, coupling) rows of ham%ncoup, so a
//              periodic lattice with NA basis atoms stores NA rows
//   site class int32 per site -> {exchange row, DM row, anisotropy row}
// Lanes of a wavefront take consecutive q, so with M > 1 the lanes of one
// site share its neighbour indices (one broadcast load) and gather
// neighbouring spins with unit stride.  Workgroups are mapped to the eight
// XCDs in contiguous ranges of sites, so each XCD's L2 holds the spins its
// own neighbour gathers reuse.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

namespace uasd {

// ---------------------------------------------------------------- constants
// source/Parameters/constants.f90
constexpr double kGama = 1.760859644e11;
constexpr double kKBolt = 1.38064852e-23;
constexpr double kMub = 9.274009994e-24;
constexpr double kMry = 2.179872325e-21;
constexpr double kPi = 3.141592653589793;
constexpr double kDblTolerance = (double)1e-14f;  // parameters.f90:12 (default real literal)

constexpr uint32_t kTagLLG = 0x11u;   // Philox tag of the thermal field
constexpr uint32_t kTagLLG2 = 0x12u;  // its second block (FP64 engine: the angles)
constexpr uint32_t kTagMCU = 0x21u;  // MC uniforms
constexpr uint32_t kTagMCG = 0x22u;  // MC Gaussian trial step

constexpr int kXCD = 8;  // MI355X: 8 accelerator complex dies, each with its own L2

template <typename R>
struct alignas(4 * sizeof(R)) Quad {
  R x, y, z, w;
};

struct SiteClass {
  int ex_row, dm_row, an_row, pad;
};

// Division of an index q < 2^31 by a runtime divisor d >= 1 as a multiply and
// a shift (Granlund and Montgomery, 1994, theorem 4.2 with N = 31): with
// l = ceil(log2 d) and m = floor(2^(31+l) / d) + 1, which fits in 32 bits,
// q / d = (q * m) >> (31 + l).  make_fast_div computes (m, 31 + l) on the host.
struct FastDiv {
  uint32_t m;
  int s;
};

inline FastDiv make_fast_div(uint32_t d) {
  int l = 0;
  while ((1ull << l) < d) ++l;
  const uint64_t m = ((uint64_t)1 << (31 + l)) / d + 1;
  return FastDiv{(uint32_t)m, 31 + l};
}

__host__ __device__ __forceinline__ uint32_t fast_div(uint32_t q, FastDiv f) {
  return (uint32_t)(((uint64_t)q * f.m) >> f.s);
}

template <typename R>
struct HamDev {
  int N, M;
  FastDiv mdiv;  // division by M
  // exchange
  int W, ex_rows;
  int ex_full;  // every exchange row holds W neighbours
  const int *__restrict__ nbr;
  const int16_t *__restrict__ nbr16;  // (n - site) mod N differences, or nullptr
  const int *__restrict__ ex_len;
  const R *__restrict__ cpl;
  // DM
  int Wd, dm_rows;
  // every site's DM neighbours are its exchange neighbours in the same order
  // and number at most kFusedW: one gather serves both sums
  int dm_shared;
  const int *__restrict__ dnbr;
  const int16_t *__restrict__ dnbr16;
  const int *__restrict__ dm_len;
  const R *__restrict__ dvx, *__restrict__ dvy, *__restrict__ dvz;
  // anisotropy rows
  const int *__restrict__ an_type;
  const R *__restrict__ an_ex, *__restrict__ an_ey, *__restrict__ an_ez;
  const R *__restrict__ an_k1, *__restrict__ an_k2, *__restrict__ an_sb;
  // site classes
  const int *__restrict__ cls;
  const SiteClass *__restrict__ classes;
  // external field: 0 none, 1 uniform per ensemble (ext_u[3*ens+c]), 2 per q
  int ext_mode;
  // 1: neighbour lists are read with non-temporal loads (streamed past L2,
  // which then keeps the gathered spins); a tuning choice, results identical
  int nt;
  const R *__restrict__ ext_u;
  const Quad<R> *__restrict__ ext_q;
  // Site order on the device: slot d holds reference site
  // (d % order_nc) * order_na + d / order_nc, i.e. the sites of one basis atom
  // are contiguous (basis-major) when order_na = NA; order_na = 1 keeps the
  // reference order.  Counter-based random numbers are keyed by the
  // reference site, so they do not depend on the order.
  int order_na, order_nc;
  int idx_kb;  // low bits of an int16 list entry holding the basis shift
};

template <typename R>
__device__ __forceinline__ uint32_t host_site(const HamDev<R> &h, int d) {
  if (h.order_na == 1) return (uint32_t)d;  // reference order: no division
  return (uint32_t)((d % h.order_nc) * h.order_na + d / h.order_nc);
}

// Block index of this workgroup in a grid laid out so that XCD k (which
// receives workgroups k, k+8, k+16, ...) works on the k-th contiguous eighth
// of the logical blocks.  The launch pads the grid to a multiple of 8.
__device__ __forceinline__ int64_t xcd_block() {
  const uint32_t per = gridDim.x / kXCD;
  const uint32_t b = blockIdx.x;
  return (int64_t)(b % kXCD) * per + (b / kXCD);
}

// ------------------------------------------------------------------- Philox
__device__ __forceinline__ void philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                              uint32_t k0, uint32_t k1, uint32_t &o0,
                                              uint32_t &o1, uint32_t &o2, uint32_t &o3) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    // one 32x32->64 multiply each (v_mad_u64_u32) rather than separate hi/lo
    const uint64_t p0 = (uint64_t)0xD2511F53u * c0, p1 = (uint64_t)0xCD9E8D57u * c2;
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    uint32_t n0 = hi1 ^ c1 ^ k0;
    uint32_t n2 = hi0 ^ c3 ^ k1;
    c0 = n0;
    c1 = lo1;
    c2 = n2;
    c3 = lo0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  o0 = c0;
  o1 = c1;
  o2 = c2;
  o3 = c3;
}

// Counter layout shared with the oracle: (site, ens, step_lo, tag | step_hi << 8),
// key = seed.
__device__ __forceinline__ void counter_words(uint64_t seed, uint32_t tag, uint64_t step,
                                              uint32_t site, uint32_t ens, uint32_t &w0,
                                              uint32_t &w1, uint32_t &w2, uint32_t &w3) {
  philox4x32_10(site, ens, (uint32_t)step, (tag & 0xffu) | ((uint32_t)(step >> 32) << 8),
                (uint32_t)seed, (uint32_t)(seed >> 32), w0, w1, w2, w3);
}

// Three N(0,1) deviates in FP64 for the FP64 MC engine's Gaussian trial
// step: Box-Muller on (u0,u1) and (u2,u3) with libm (the oracle restates it,
// or_noise3_f64, equal to libm ulps).
__device__ __forceinline__ void normal3_f64(uint64_t seed, uint32_t tag, uint64_t step,
                                            uint32_t site, uint32_t ens, double &a, double &b,
                                            double &c) {
  uint32_t w0, w1, w2, w3;
  counter_words(seed, tag, step, site, ens, w0, w1, w2, w3);
  const double s = 2.3283064365386963e-10, two_pi = 6.283185307179586;
  double u0 = ((double)w0 + 0.5) * s, u1 = ((double)w1 + 0.5) * s;
  double u2 = ((double)w2 + 0.5) * s, u3 = ((double)w3 + 0.5) * s;
  double r0 = sqrt(-2.0 * log(u0)), r1 = sqrt(-2.0 * log(u2));
  double t0 = two_pi * u1, t1 = two_pi * u3;
  a = r0 * cos(t0);
  b = r0 * sin(t0);
  c = r1 * cos(t1);
}

// Thermal-field deviates of the FP32 engine's LLG steps.  Box-Muller with
// ln, sin and cos evaluated by
// polynomials fitted here (least squares with Lawson reweighting on Chebyshev
// nodes; relative error below 3e-7): every operation is an IEEE add,
// multiply, fmaf, sqrtf or rint, so the oracle reproduces the deviates bit
// for bit.  Radius from all 32 bits of w0 (tails to 6.8 sigma).
//   ln(u):  u = m 2^e, m in [sqrt(1/2), sqrt(2)], x = m - 1,
//           ln(1+x) = x + x^2 (-1/2 + x P(x))
//   angle:  2 pi u = (pi/2) v, v in (0, 4), q = rint(v), f = v - q in
//           [-1/2, 1/2]: sin((pi/2) f) = f S(f^2), cos((pi/2) f) = C(f^2)
namespace bm {
__device__ __forceinline__ float ln(float u) {
#pragma clang fp contract(off)
  const uint32_t bits = __float_as_uint(u);
  int ex = (int)(bits >> 23) - 127;
  float m = __uint_as_float((bits & 0x7fffffu) | 0x3f800000u);
  if (m > 1.41421354f) {
    m = m * 0.5f;
    ex += 1;
  }
  const float x = m - 1.0f;
  float p = -0.0830537900328636f;
  p = fmaf(p, x, 0.127988979220390f);
  p = fmaf(p, x, -0.129434198141098f);
  p = fmaf(p, x, 0.141902372241020f);
  p = fmaf(p, x, -0.166434884071350f);
  p = fmaf(p, x, 0.200018793344498f);
  p = fmaf(p, x, -0.250002801418304f);
  p = fmaf(p, x, 0.333333253860474f);
  const float t = fmaf(p, x, -0.5f);
  const float x2 = x * x;
  const float lnm = fmaf(t, x2, x);
  return fmaf((float)ex, 0.693147182464599609375f, lnm);
}
// sin and cos of (pi/2) v
__device__ __forceinline__ void sincos_q(float v, float &s, float &c) {
#pragma clang fp contract(off)
  const float qf = rintf(v);
  const float f = v - qf;
  const int q = (int)qf & 3;
  const float f2 = f * f;
  float sp = 0.000158195078256540f;
  sp = fmaf(sp, f2, -0.00468126079067588f);
  sp = fmaf(sp, f2, 0.0796925798058510f);
  sp = fmaf(sp, f2, -0.645964086055756f);
  sp = fmaf(sp, f2, 1.57079637050629f);
  const float sn = sp * f;
  float cs = 0.000903362990356982f;
  cs = fmaf(cs, f2, -0.0208599362522364f);
  cs = fmaf(cs, f2, 0.253669172525406f);
  cs = fmaf(cs, f2, -1.23370051383972f);
  cs = fmaf(cs, f2, 1.0f);
  s = q == 0 ? sn : q == 1 ? cs : q == 2 ? -sn : -cs;
  c = q == 0 ? cs : q == 1 ? -sn : q == 2 ? -cs : sn;
}
}  // namespace bm

__device__ __forceinline__ void thermal3(uint64_t seed, uint32_t tag, uint64_t step, uint32_t site,
                                         uint32_t ens, float &a, float &b, float &c) {
#pragma clang fp contract(off)
  uint32_t w0, w1, w2, w3;
  counter_words(seed, tag, step, site, ens, w0, w1, w2, w3);
  const float s32 = 2.3283064365386963e-10f, s22 = 2.384185791015625e-07f;
  const float u0 = ((float)w0 + 0.5f) * s32, u2 = ((float)w2 + 0.5f) * s32;
  const float v1 = (float)((w1 >> 9) * 2u + 1u) * s22, v3 = (float)((w3 >> 9) * 2u + 1u) * s22;
  const float r0 = sqrtf(-2.0f * bm::ln(u0)), r1 = sqrtf(-2.0f * bm::ln(u2));
  float s0, c0, s1, c1;
  bm::sincos_q(v1, s0, c0);
  bm::sincos_q(v3, s1, c1);
  a = r0 * c0;
  b = r0 * s0;
  c = r1 * c1;
}

// Thermal-field deviates of the FP64 engine's LLG steps, at the resolution
// of the reference's: r4_nor returns real(IEEE64) values
// (randomnumbers.f90:379,389).  Box-Muller in double precision on 52-bit
// uniforms, u = (x + 1/2) 2^-52 with x the top 52 bits of two Philox words
// (exact; radius tails to 8.6 sigma): block kTagLLG gives the two radii,
// block kTagLLG2 the two angles, taken by sincospi / cospi (exact argument
// reduction).  The oracle restates it with libm (or_thermal3_f64; the
// deviates agree to a few ulps).
__device__ __forceinline__ double u52(uint32_t hi, uint32_t lo) {
  const uint64_t x = ((uint64_t)hi << 20) | (uint64_t)(lo >> 12);
  return ((double)x + 0.5) * 2.220446049250313080847e-16;
}

__device__ __forceinline__ void thermal3_f64(uint64_t seed, uint64_t step, uint32_t site,
                                             uint32_t ens, double &a, double &b, double &c) {
  uint32_t w0, w1, w2, w3, x0, x1, x2, x3;
  counter_words(seed, kTagLLG, step, site, ens, w0, w1, w2, w3);
  counter_words(seed, kTagLLG2, step, site, ens, x0, x1, x2, x3);
  const double r0 = sqrt(-2.0 * log(u52(w0, w1))), r1 = sqrt(-2.0 * log(u52(w2, w3)));
  double s0, c0;
  sincospi(2.0 * u52(x0, x1), &s0, &c0);
  const double c1 = cospi(2.0 * u52(x2, x3));
  a = r0 * c0;
  b = r0 * s0;
  c = r1 * c1;
}

// Uniform on (0, 1) from the top 24 bits of a Philox word: exact in FP32.
__device__ __forceinline__ float u24(uint32_t w) {
  return ((float)(w >> 8) + 0.5f) * 5.9604644775390625e-08f;
}

__device__ __forceinline__ void uniform4(uint64_t seed, uint32_t tag, uint64_t step, uint32_t site,
                                         uint32_t ens, double &u0, double &u1, double &u2,
                                         double &u3) {
  uint32_t w0, w1, w2, w3;
  counter_words(seed, tag, step, site, ens, w0, w1, w2, w3);
  const double s = 2.3283064365386963e-10;
  u0 = ((double)w0 + 0.5) * s;
  u1 = ((double)w1 + 0.5) * s;
  u2 = ((double)w2 + 0.5) * s;
  u3 = ((double)w3 + 0.5) * s;
}

// ------------------------------------------------------------- field terms
// hamiltonianactions.f90:431-466 (exchange) and 546-580 (DM), summed in list
// order; 842-879 / 883-919 anisotropy; 183-248 assembly.  The moments are
// gathered as q.xyz * q.w (emomM).  Exchange terms accumulate by fused
// multiply-add, field = fma(J, emomM, field), as the reference's release
// build contracts them (-Ofast -march=native, cmake/SetFortranFlags.cmake);
// everything else rounds each multiply and add on its own.

template <typename R>
__device__ __forceinline__ void moment(const Quad<R> *__restrict__ s, size_t q, R &a, R &b, R &c) {
  const Quad<R> v = s[q];
  a = v.x * v.w;
  b = v.y * v.w;
  c = v.z * v.w;
}

template <typename R>
__device__ __forceinline__ void uniaxial(const HamDev<R> &h, int row, R m0, R m1, R m2, R &f0,
                                         R &f1, R &f2) {
  R ea0 = h.an_ex[row], ea1 = h.an_ey[row], ea2 = h.an_ez[row];
  R k1 = h.an_k1[row], k2 = h.an_k2[row];
  R tt1 = m0 * ea0 + m1 * ea1 + m2 * ea2;
  R tt2 = k1 + (R)2.0 * k2 * ((R)1.0 - tt1 * tt1);
  R tt3 = (R)2.0 * tt1 * tt2;
  f0 = f0 - tt3 * ea0;
  f1 = f1 - tt3 * ea1;
  f2 = f2 - tt3 * ea2;
}

template <typename R>
__device__ __forceinline__ void cubic(const HamDev<R> &h, int row, R m0, R m1, R m2, R &f0, R &f1,
                                      R &f2) {
  R k1 = h.an_k1[row], k2 = h.an_k2[row];
  R s0 = m0 * m0, s1 = m1 * m1, s2 = m2 * m2;
  f0 = f0 + (R)2.0 * k1 * m0 * (s1 + s2) + (R)2.0 * k2 * m0 * (s1 * s2);
  f1 = f1 + (R)2.0 * k1 * m1 * (s2 + s0) + (R)2.0 * k2 * m1 * (s2 * s0);
  f2 = f2 + (R)2.0 * k1 * m2 * (s0 + s1) + (R)2.0 * k2 * m2 * (s0 * s1);
}

// Field parts of one site: s = beff_s (pair terms + uniaxial), qf = beff_q
// (type-7 cubic part times sb), x = external field.
// Decodes the slot-major list entries of one site (device slot d = b*nc + c,
// basis b, cell c).  int32 entries are slots.  int16 entries, used when every
// entry of the list fits, pack the cell difference and the basis shift of
// the neighbour, v = delta * 2^kb + beta: the neighbour is slot
// ((b + beta) mod na)*nc + ((c + delta) mod nc).  With the reference order
// (na = 1, nc = N, kb = 0) that is the difference (n - site) mod N.  A
// periodic lattice numbered cell by cell keeps delta within a few planes of
// cells, so the index stream halves.
struct NbDec {
  int b, c, nc, na, kb;
  template <typename R>
  __device__ __forceinline__ NbDec(const HamDev<R> &h, int d)
      : nc(h.order_nc), na(h.order_na), kb(h.idx_kb) {
    if (na == 1) {  // reference order: one block of nc = N sites
      b = 0;
      c = d;
    } else {
      b = d / nc;
      c = d - b * nc;
    }
  }
  __device__ __forceinline__ int operator()(int v) const { return v; }
  __device__ __forceinline__ int operator()(int16_t v) const {
    const int beta = (int)v & ((1 << kb) - 1);
    const int delta = (int)v >> kb;
    int b2 = b + beta;
    b2 -= b2 >= na ? na : 0;
    int c2 = c + delta;
    c2 += c2 < 0 ? nc : 0;
    c2 -= c2 >= nc ? nc : 0;
    return b2 * nc + c2;
  }
};

// NbDec for the one-block site order (order_na = 1, so kb = 0): the entry is
// the wrapped difference of the slots, decoded without the basis shift.
struct NbDec1 {
  int c, nc;
  template <typename R>
  __device__ __forceinline__ NbDec1(const HamDev<R> &h, int d) : c(d), nc(h.order_nc) {}
  __device__ __forceinline__ int operator()(int v) const { return v; }
  __device__ __forceinline__ int operator()(int16_t v) const {
    int c2 = c + (int)v;
    c2 += c2 < 0 ? nc : 0;
    c2 -= c2 >= nc ? nc : 0;
    return c2;
  }
};

template <bool NT, typename I>
__device__ __forceinline__ I nb_load(const I *p) {
  if constexpr (NT) return __builtin_nontemporal_load(p);
  return *p;
}

// Exchange sum of site_field.  Neighbours are taken G at a time: all index
// loads of a batch are issued, then all spin and coupling loads, then the
// terms are added in neighbour order (the same order, hence the same
// rounding, as one at a time).  A last partial batch masks the slots past len.
// U > 1 sums the fields of the U adjacent ensembles e0 .. e0+U-1 of the site
// together: each neighbour entry and coupling is read once for all of them,
// and each field is summed exactly as it is for U = 1.  JW: each term as
// (J m_j) e_j instead of J (e_j m_j), one multiply fewer per neighbour and
// equal to rounding (the MC energy differences, not the fields of the LLG
// step, which keep the reference's emomM products).
template <typename R, int G, bool NT, typename I, int U, bool JW = false, typename D = NbDec>
__device__ __forceinline__ void exchange_run(const HamDev<R> &h, const I *__restrict__ nb,
                                             const Quad<R> *__restrict__ src, int site, int e0,
                                             int row, int len, R (&f)[U][3]) {
  const int M = h.M, N = h.N;
  const D dec(h, site);
  const R *cp = h.cpl + row;
  const size_t rs = (size_t)h.ex_rows;
  int j = 0;
  for (; j + G <= len; j += G) {
    int n[G];
#pragma unroll
    for (int u = 0; u < G; ++u) n[u] = dec(nb_load<NT>(nb + (size_t)(j + u) * N));
    Quad<R> v[G][U];
    R J[G];
#pragma unroll
    for (int u = 0; u < G; ++u) {
#pragma unroll
      for (int k = 0; k < U; ++k) v[u][k] = src[(size_t)n[u] * M + e0 + k];
      J[u] = cp[(size_t)(j + u) * rs];
    }
#pragma unroll
    for (int u = 0; u < G; ++u) {
#pragma unroll
      for (int k = 0; k < U; ++k) {
        if (JW) {
          const R jw = J[u] * v[u][k].w;
          f[k][0] = fma(jw, v[u][k].x, f[k][0]);
          f[k][1] = fma(jw, v[u][k].y, f[k][1]);
          f[k][2] = fma(jw, v[u][k].z, f[k][2]);
        } else {
          f[k][0] = fma(J[u], v[u][k].x * v[u][k].w, f[k][0]);
          f[k][1] = fma(J[u], v[u][k].y * v[u][k].w, f[k][1]);
          f[k][2] = fma(J[u], v[u][k].z * v[u][k].w, f[k][2]);
        }
      }
    }
  }
  if (j < len) {
    int n[G];
#pragma unroll
    for (int u = 0; u < G; ++u)
      n[u] = j + u < len ? dec(nb_load<NT>(nb + (size_t)(j + u) * N)) : site;
    Quad<R> v[G][U];
    R J[G];
#pragma unroll
    for (int u = 0; u < G; ++u) {
      if (j + u < len) {
#pragma unroll
        for (int k = 0; k < U; ++k) v[u][k] = src[(size_t)n[u] * M + e0 + k];
        J[u] = cp[(size_t)(j + u) * rs];
      }
    }
#pragma unroll
    for (int u = 0; u < G; ++u) {
      if (j + u < len) {
#pragma unroll
        for (int k = 0; k < U; ++k) {
          if (JW) {
            const R jw = J[u] * v[u][k].w;
            f[k][0] = fma(jw, v[u][k].x, f[k][0]);
            f[k][1] = fma(jw, v[u][k].y, f[k][1]);
            f[k][2] = fma(jw, v[u][k].z, f[k][2]);
          } else {
            f[k][0] = fma(J[u], v[u][k].x * v[u][k].w, f[k][0]);
            f[k][1] = fma(J[u], v[u][k].y * v[u][k].w, f[k][1]);
            f[k][2] = fma(J[u], v[u][k].z * v[u][k].w, f[k][2]);
          }
        }
      }
    }
  }
}

// When every row holds W neighbours the loop runs to the launch argument W,
// the same for every lane: its control and the batch tail are scalar.
template <typename R, int G, bool NT, typename I, int U, bool JW = false>
__device__ __forceinline__ void exchange_fields(const HamDev<R> &h, const I *__restrict__ nb,
                                                const Quad<R> *__restrict__ src, int site, int e0,
                                                int row, R (&f)[U][3]) {
  if (h.ex_full)
    exchange_run<R, G, NT, I, U, JW>(h, nb, src, site, e0, row, h.W, f);
  else
    exchange_run<R, G, NT, I, U, JW>(h, nb, src, site, e0, row, h.ex_len[row], f);
}

// The MC energy gather: as exchange_fields with JW, and int16 entries of the
// one-block site order decoded by NbDec1 (the choice made once, outside the
// neighbour loop).
template <typename R, int G, bool NT, typename I, int U>
__device__ __forceinline__ void mc_exchange_fields(const HamDev<R> &h, const I *__restrict__ nb,
                                                   const Quad<R> *__restrict__ src, int site,
                                                   int e0, int row, R (&f)[U][3]) {
  if (std::is_same<I, int16_t>::value && h.order_na == 1) {
    if (h.ex_full)
      exchange_run<R, G, NT, I, U, true, NbDec1>(h, nb, src, site, e0, row, h.W, f);
    else
      exchange_run<R, G, NT, I, U, true, NbDec1>(h, nb, src, site, e0, row, h.ex_len[row], f);
  } else {
    exchange_fields<R, G, NT, I, U, true>(h, nb, src, site, e0, row, f);
  }
}

template <typename R, int G, bool NT, typename I>
__device__ __forceinline__ void exchange_sum(const HamDev<R> &h, const I *__restrict__ nb,
                                             const Quad<R> *__restrict__ src, int site, int ens,
                                             int row, R &s0, R &s1, R &s2) {
  R f[1][3] = {{s0, s1, s2}};
  exchange_fields<R, G, NT, I, 1>(h, nb, src, site, ens, row, f);
  s0 = f[0][0];
  s1 = f[0][1];
  s2 = f[0][2];
}

// DM sum of site_field: field += m_j x D_ij (hamiltonianactions.f90:565-578).
template <typename R, int G, typename I>
__device__ __forceinline__ void dm_sum(const HamDev<R> &h, const I *__restrict__ nb,
                                       const Quad<R> *__restrict__ src, int site, int ens,
                                       int row, R &s0, R &s1, R &s2) {
  const int M = h.M, N = h.N;
  const NbDec dec(h, site);
  const int len = h.dm_len[row];
  const size_t rs = (size_t)h.dm_rows;
  for (int j = 0; j < len; j += G) {
    int n[G];
#pragma unroll
    for (int u = 0; u < G; ++u)
      n[u] = j + u < len ? dec(nb[(size_t)(j + u) * N]) : site;
    Quad<R> v[G];
#pragma unroll
    for (int u = 0; u < G; ++u)
      if (j + u < len) v[u] = src[(size_t)n[u] * M + ens];
#pragma unroll
    for (int u = 0; u < G; ++u) {
      if (j + u < len) {
        const size_t r = (size_t)(j + u) * rs + row;
        const R D0 = h.dvx[r], D1 = h.dvy[r], D2 = h.dvz[r];
        const R m0 = v[u].x * v[u].w, m1 = v[u].y * v[u].w, m2 = v[u].z * v[u].w;
        s0 = s0 + D2 * m1 - D1 * m2;
        s1 = s1 + D0 * m2 - D2 * m0;
        s2 = s2 + D1 * m0 - D0 * m1;
      }
    }
  }
}

// Exchange and DM sums of a site whose DM bonds are its exchange bonds
// (HamDev::dm_shared): each neighbour is gathered once, its moment m_j kept
// in registers for the DM terms.  The exchange terms are added in list
// order as exchange_sum adds them, then the DM terms in list order as dm_sum
// adds them, so the field is bit-identical to the two separate sums.
constexpr int kFusedW = 8;

template <typename R, int G, bool NT, typename I>
__device__ __forceinline__ void exdm_run(const HamDev<R> &h, const I *__restrict__ nb,
                                         const Quad<R> *__restrict__ src, int site, int ens,
                                         int exrow, int dmrow, int len, R &s0, R &s1, R &s2) {
  const int M = h.M, N = h.N;
  const NbDec dec(h, site);
  const R *cp = h.cpl + exrow;
  const size_t rs = (size_t)h.ex_rows;
  R c0[kFusedW], c1[kFusedW], c2[kFusedW];
#pragma unroll
  for (int j0 = 0; j0 < kFusedW; j0 += G) {
    int n[G];
#pragma unroll
    for (int u = 0; u < G; ++u)
      n[u] = j0 + u < len ? dec(nb_load<NT>(nb + (size_t)(j0 + u) * N)) : site;
    Quad<R> v[G];
    R J[G];
#pragma unroll
    for (int u = 0; u < G; ++u) {
      if (j0 + u < len) {
        v[u] = src[(size_t)n[u] * M + ens];
        J[u] = cp[(size_t)(j0 + u) * rs];
      }
    }
#pragma unroll
    for (int u = 0; u < G; ++u) {
      if (j0 + u < kFusedW && j0 + u < len) {
        c0[j0 + u] = v[u].x * v[u].w;
        c1[j0 + u] = v[u].y * v[u].w;
        c2[j0 + u] = v[u].z * v[u].w;
        s0 = fma(J[u], c0[j0 + u], s0);
        s1 = fma(J[u], c1[j0 + u], s1);
        s2 = fma(J[u], c2[j0 + u], s2);
      }
    }
  }
  if (dmrow < 0) return;
  const size_t drs = (size_t)h.dm_rows;
#pragma unroll
  for (int j = 0; j < kFusedW; ++j) {
    if (j < len) {
      const size_t r = (size_t)j * drs + dmrow;
      const R D0 = h.dvx[r], D1 = h.dvy[r], D2 = h.dvz[r];
      s0 = s0 + D2 * c1[j] - D1 * c2[j];
      s1 = s1 + D0 * c2[j] - D2 * c0[j];
      s2 = s2 + D1 * c0[j] - D0 * c1[j];
    }
  }
}

// As exchange_sum: with full rows every guard compares with the launch's W.
template <typename R, int G, bool NT, typename I>
__device__ __forceinline__ void exdm_sum(const HamDev<R> &h, const I *__restrict__ nb,
                                         const Quad<R> *__restrict__ src, int site, int ens,
                                         int exrow, int dmrow, R &s0, R &s1, R &s2) {
  if (h.ex_full)
    exdm_run<R, G, NT, I>(h, nb, src, site, ens, exrow, dmrow, h.W, s0, s1, s2);
  else
    exdm_run<R, G, NT, I>(h, nb, src, site, ens, exrow, dmrow, h.ex_len[exrow], s0, s1, s2);
}

// G = neighbours whose loads are issued together (a tuning parameter).
template <typename R, int G>
__device__ __forceinline__ void site_field(const HamDev<R> &h, const Quad<R> *__restrict__ src,
                                           int site, int ens, R &s0, R &s1, R &s2, R &qf0,
                                           R &qf1, R &qf2, R &x0, R &x1, R &x2) {
  s0 = s1 = s2 = (R)0;
  qf0 = qf1 = qf2 = (R)0;
  const SiteClass sc = h.classes[h.cls[site]];
  const int M = h.M;
  if (h.dm_shared) {
    const int dr = h.Wd > 0 ? sc.dm_row : -1;
    if (h.nbr16) {
      if (h.nt)
        exdm_sum<R, G, true>(h, h.nbr16 + site, src, site, ens, sc.ex_row, dr, s0, s1, s2);
      else
        exdm_sum<R, G, false>(h, h.nbr16 + site, src, site, ens, sc.ex_row, dr, s0, s1, s2);
    } else {
      if (h.nt)
        exdm_sum<R, G, true>(h, h.nbr + site, src, site, ens, sc.ex_row, dr, s0, s1, s2);
      else
        exdm_sum<R, G, false>(h, h.nbr + site, src, site, ens, sc.ex_row, dr, s0, s1, s2);
    }
  } else if (h.nbr16) {
    if (h.nt)
      exchange_sum<R, G, true>(h, h.nbr16 + site, src, site, ens, sc.ex_row, s0, s1, s2);
    else
      exchange_sum<R, G, false>(h, h.nbr16 + site, src, site, ens, sc.ex_row, s0, s1, s2);
  } else {
    if (h.nt)
      exchange_sum<R, G, true>(h, h.nbr + site, src, site, ens, sc.ex_row, s0, s1, s2);
    else
      exchange_sum<R, G, false>(h, h.nbr + site, src, site, ens, sc.ex_row, s0, s1, s2);
  }
  if (!h.dm_shared && h.Wd > 0 && sc.dm_row >= 0) {
    if (h.dnbr16)
      dm_sum<R, G>(h, h.dnbr16 + site, src, site, ens, sc.dm_row, s0, s1, s2);
    else
      dm_sum<R, G>(h, h.dnbr + site, src, site, ens, sc.dm_row, s0, s1, s2);
  }
  if (sc.an_row >= 0) {
    const int t = h.an_type[sc.an_row];
    if (t == 1 || t == 2 || t == 7) {
      R m0, m1, m2;
      moment(src, (size_t)site * M + ens, m0, m1, m2);
      if (t == 1) {
        uniaxial(h, sc.an_row, m0, m1, m2, s0, s1, s2);
      } else if (t == 2) {
        cubic(h, sc.an_row, m0, m1, m2, s0, s1, s2);
      } else {
        R t0 = 0, t1 = 0, t2 = 0;
        uniaxial(h, sc.an_row, m0, m1, m2, s0, s1, s2);
        cubic(h, sc.an_row, m0, m1, m2, t0, t1, t2);
        const R sb = h.an_sb[sc.an_row];
        qf0 = qf0 + t0 * sb;
        qf1 = qf1 + t1 * sb;
        qf2 = qf2 + t2 * sb;
      }
    }
  }
  if (h.ext_mode == 1) {
    x0 = h.ext_u[3 * ens];
    x1 = h.ext_u[3 * ens + 1];
    x2 = h.ext_u[3 * ens + 2];
  } else if (h.ext_mode == 2) {
    const Quad<R> v = h.ext_q[(size_t)site * M + ens];
    x0 = v.x;
    x1 = v.y;
    x2 = v.z;
  } else {
    x0 = x1 = x2 = (R)0;
  }
}

// The pair, DM and anisotropy parts of site_field for the U adjacent
// ensembles e0 .. e0+U-1 of one site: the exchange terms of all U from one
// read of the neighbour entries and couplings, then each ensemble's DM and
// anisotropy terms.  The terms are added in site_field's order (the fused
// exchange/DM gather adds them in the order of the separate sums), so each
// field equals site_field's bit for bit.
template <typename R, int G, int U>
__device__ __forceinline__ void site_field_u(const HamDev<R> &h, const Quad<R> *__restrict__ src,
                                             int site, int e0, R (&sf)[U][3], R (&qf)[U][3]) {
  const SiteClass sc = h.classes[h.cls[site]];
  const int M = h.M, row = sc.ex_row;
#pragma unroll
  for (int k = 0; k < U; ++k) sf[k][0] = sf[k][1] = sf[k][2] = qf[k][0] = qf[k][1] = qf[k][2] = (R)0;
  if (h.nbr16) {
    if (h.nt)
      exchange_fields<R, G, true, int16_t, U>(h, h.nbr16 + site, src, site, e0, row, sf);
    else
      exchange_fields<R, G, false, int16_t, U>(h, h.nbr16 + site, src, site, e0, row, sf);
  } else {
    if (h.nt)
      exchange_fields<R, G, true, int, U>(h, h.nbr + site, src, site, e0, row, sf);
    else
      exchange_fields<R, G, false, int, U>(h, h.nbr + site, src, site, e0, row, sf);
  }
#pragma unroll
  for (int k = 0; k < U; ++k) {
    if (h.Wd > 0 && sc.dm_row >= 0) {
      if (h.dnbr16)
        dm_sum<R, G>(h, h.dnbr16 + site, src, site, e0 + k, sc.dm_row, sf[k][0], sf[k][1],
                     sf[k][2]);
      else
        dm_sum<R, G>(h, h.dnbr + site, src, site, e0 + k, sc.dm_row, sf[k][0], sf[k][1],
                     sf[k][2]);
    }
    if (sc.an_row >= 0) {
      const int t = h.an_type[sc.an_row];
      if (t == 1 || t == 2 || t == 7) {
        R m0, m1, m2;
        moment(src, (size_t)site * M + e0 + k, m0, m1, m2);
        if (t == 1) {
          uniaxial(h, sc.an_row, m0, m1, m2, sf[k][0], sf[k][1], sf[k][2]);
        } else if (t == 2) {
          cubic(h, sc.an_row, m0, m1, m2, sf[k][0], sf[k][1], sf[k][2]);
        } else {
          R t0 = 0, t1 = 0, t2 = 0;
          uniaxial(h, sc.an_row, m0, m1, m2, sf[k][0], sf[k][1], sf[k][2]);
          cubic(h, sc.an_row, m0, m1, m2, t0, t1, t2);
          const R sb = h.an_sb[sc.an_row];
          qf[k][0] = qf[k][0] + t0 * sb;
          qf[k][1] = qf[k][1] + t1 * sb;
          qf[k][2] = qf[k][2] + t2 * sb;
        }
      }
    }
  }
}

// beff = 0 + beff1 + beff2, beff2 = beff_q + ext + 0  (hamiltonianactions.f90:240-243)
template <typename R, int G>
__device__ __forceinline__ void total_field(const HamDev<R> &h, const Quad<R> *__restrict__ src,
                                            int site, int ens, R &b0, R &b1, R &b2) {
  R s0, s1, s2, q0, q1, q2, x0, x1, x2;
  site_field<R, G>(h, src, site, ens, s0, s1, s2, q0, q1, q2, x0, x1, x2);
  b0 = (R)0 + s0 + (q0 + x0 + (R)0);
  b1 = (R)0 + s1 + (q1 + x1 + (R)0);
  b2 = (R)0 + s2 + (q2 + x2 + (R)0);
}

// ------------------------------------------------------------ LLG params
// Solver constants that depend on lam, dt and bn only: computed once on the
// host when lam is uniform, per site otherwise (same operations either way).
template <typename R>
struct LamConst {
  double Dk;       // midpoint_sigma: lam/(1+lam^2) kB/gamma/muB * gamma/bn
  double Dp;       // depondt_sigma: 2 lam kB / (dt gamma muB)
  double Dh;       // heun_sigma: lam/(1+lam^2) kB/muB
  R lam, lldamp;   // lam and 1/(1+lam^2)
  R mdt, msqrtdt;  // midpoint_A: dt bn gamma lldamp and its square root
};

template <typename R>
__host__ __device__ inline LamConst<R> lam_const(double lam, double dt, double bn) {
  LamConst<R> c;
  c.Dk = (lam / (1.0 + lam * lam) * kKBolt / kGama / (kMub)) * (kGama / bn);
  c.Dp = (2.0 * lam * kKBolt) / (dt * kGama * kMub);
  c.Dh = lam / (1.0 + lam * lam) * kKBolt / kMub;
  c.lam = (R)lam;
  c.lldamp = (R)1.0 / ((R)1.0 + c.lam * c.lam);
  c.mdt = (R)dt * (R)bn * (R)kGama * c.lldamp;
  c.msqrtdt = (R)sqrt((double)c.mdt);  // correctly rounded in R as well
  return c;
}

template <typename R>
struct LLGDev {
  const double *lam, *temp, *lande;  // per site or nullptr
  double lam0, temp0, lande0;        // uniform values when the arrays are null
  double dt, temprescale, bn;
  int noise_mode;
  uint64_t seed;
  uint64_t step;              // step counter, or offset from *step_dev
  const uint64_t *step_dev;   // device-held step base (graph replays) or nullptr
  const Quad<R> *noise;  // injected N(0,1) per q
  // Philox deviates of a fused step: stored by the predictor, read back by
  // the corrector (the same values it would draw), or nullptr
  Quad<R> *nsave;
  LamConst<R> lc0;       // lam_const(lam0, dt, bn)
  // uniform lam, temperature and moment: the noise amplitudes, computed on
  // the host by the same functions (sig_uniform = 1)
  int sig_uniform;
  R sigk0, sigp0, sigh0;  // midpoint_sigma, depondt_sigma, heun_sigma
  __device__ __forceinline__ LamConst<R> lamc(int s) const {
    return lam ? lam_const<R>(lam[s], dt, bn) : lc0;
  }
  __device__ __forceinline__ double lambda(int s) const { return lam ? lam[s] : lam0; }
  __device__ __forceinline__ double temperature(int s) const { return temp ? temp[s] : temp0; }
  __device__ __forceinline__ double landeg(int s) const { return lande ? lande[s] : lande0; }
  __device__ __forceinline__ void normals(int site, int ens, size_t q, R &a, R &b, R &c) const {
    if (noise_mode == 1) {
      const uint64_t st = step_dev ? *step_dev + step : step;
      if constexpr (std::is_same<R, double>::value) {
        thermal3_f64(seed, st, (uint32_t)site, (uint32_t)ens, a, b, c);
      } else {
        float fa, fb, fc;
        thermal3(seed, kTagLLG, st, (uint32_t)site, (uint32_t)ens, fa, fb, fc);
        a = fa;
        b = fb;
        c = fc;
      }
    } else if (noise_mode == 2) {
      const Quad<R> v = noise[q];
      a = v.x;
      b = v.y;
      c = v.z;
    } else {
      a = b = c = (R)0;
    }
  }
};

// rannum, llg = 1 (randomnumbers.f90:674-676, 735-745)
template <typename R>
__host__ __device__ inline R midpoint_sigma(double Dk, double mmom, double T,
                                            double temprescale) {
  double mmomi = 1.0 / mmom;
  double D = Dk * mmomi * T * temprescale;
  return (R)sqrt(2.0 * D);
}

// depondt.f90:142-145
template <typename R>
__host__ __device__ inline R depondt_sigma(double Dp, double mmom, double T,
                                           double temprescale) {
  return (R)sqrt(Dp * temprescale * T / mmom);
}

// evolve4p (heun_proper.f90:115-126): D = Dk/mmom*T*temprescale,
// sigma = sqrt(dt*2*D) with dt = delta_t*gama
template <typename R>
__host__ __device__ inline R heun_sigma(double Dh, double mmom, double T, double temprescale,
                                        double delta_t) {
  const double mmomi = 1.0 / mmom;
  const double D = Dh * mmomi * T * temprescale;
  return (R)sqrt(delta_t * kGama * 2.0 * D);
}

template <typename R>
__device__ __forceinline__ R fsqrt(R x) {
  return sqrt(x);
}
template <typename R>
__device__ __forceinline__ R fcos(R x) {
  return cos(x);
}
template <typename R>
__device__ __forceinline__ R fsin(R x) {
  return sin(x);
}

// Rodrigues rotation of e about bd (depondt.f90:156-179)
template <typename R>
__device__ __forceinline__ void rodrigues(R bd0, R bd1, R bd2, R dt, R lldamp, R e0, R e1, R e2,
                                          R &o0, R &o1, R &o2) {
  R Bnorm = bd0 * bd0 + bd1 * bd1 + bd2 * bd2;
  Bnorm = fsqrt(Bnorm) + (R)1.0e-15;
  R hx = bd0 / Bnorm, hy = bd1 / Bnorm, hz = bd2 / Bnorm;
  R v = Bnorm * dt * (R)kGama * lldamp;
  R cosv = fcos(v), sinv = fsin(v), u = (R)1.0 - cosv;
  o0 = hx * hx * u * e0 + cosv * e0 + hx * hy * u * e1 - hz * sinv * e1 + hx * hz * u * e2 +
       hy * sinv * e2;
  o1 = hy * hx * u * e0 + hz * sinv * e0 + hy * hy * u * e1 + cosv * e1 + hy * hz * u * e2 -
       hx * sinv * e2;
  o2 = hx * hz * u * e0 - hy * sinv * e0 + hz * hy * u * e1 + hx * sinv * e1 + hz * hz * u * e2 +
       cosv * e2;
}

// bdup = 0 + bloc + lambda e x bloc (depondt.f90:148-154)
template <typename R>
__device__ __forceinline__ void depondt_bdup(R lam, R e0, R e1, R e2, R bl0, R bl1, R bl2, R &d0,
                                             R &d1, R &d2) {
  d0 = (R)0 + bl0 + lam * e1 * bl2 - lam * e2 * bl1;
  d1 = (R)0 + bl1 + lam * e2 * bl0 - lam * e0 * bl2;
  d2 = (R)0 + bl2 + lam * e0 * bl1 - lam * e1 * bl0;
}

// midpoint A vector (midpoint.f90:127-151 / 270-296)
template <typename R>
__device__ __forceinline__ void midpoint_A(R lam, R dt, R sqrtdt, R lande, R b0, R b1, R b2, R r0,
                                           R r1, R r2, R p0, R p1, R p2, R &A0, R &A1, R &A2) {
  R dtg = dt * lande;
  R sqrtdtg = sqrtdt * lande;
  R a0 = -(R)0.0 - b0 - lam * (p1 * b2 - p2 * b1);
  R a1 = -(R)0.0 - b1 - lam * (p2 * b0 - p0 * b2);
  R a2 = -(R)0.0 - b2 - lam * (p0 * b1 - p1 * b0);
  R s0 = -r0 - lam * (p1 * r2 - p2 * r1);
  R s1 = -r1 - lam * (p2 * r0 - p0 * r2);
  R s2 = -r2 - lam * (p0 * r1 - p1 * r0);
  A0 = (R)0.5 * dtg * a0 + (R)0.5 * sqrtdtg * s0;
  A1 = (R)0.5 * dtg * a1 + (R)0.5 * sqrtdtg * s1;
  A2 = (R)0.5 * dtg * a2 + (R)0.5 * sqrtdtg * s2;
}

// Heun drift a(e, h) = lam h + e x h - e (e . lam h) (heun_proper.f90:319-325,
// 448-455; the spin torques are zero)
template <typename R>
__device__ __forceinline__ void heun_drift(R lam, R ex, R ey, R ez, R hx, R hy, R hz, R &ax,
                                           R &ay, R &az) {
  const R px = lam * hx, py = lam * hy, pz = lam * hz;
  const R prot = ex * px + ey * py + ez * pz;
  ax = px + ez * hy - ey * hz - ex * prot;
  ay = py - ez * hx + ex * hz - ey * prot;
  az = pz + ey * hx - ex * hy - ez * prot;
}

// Heun diffusion B(e) dW with B = lam (I - e e^T) + [e]x^T (heun_proper.f90:
// 327-335); QUIRK reproduces the corrector's b1zz = lam - ez*ez - lam
// (heun_proper.f90:480).
template <typename R, bool QUIRK>
__device__ __forceinline__ void heun_diffusion(R lam, R ex, R ey, R ez, R (&b)[9]) {
  b[0] = lam - ex * ex * lam;
  b[1] = ez - ex * ey * lam;
  b[2] = -ey - ex * ez * lam;
  b[3] = -ez - ey * ex * lam;
  b[4] = lam - ey * ey * lam;
  b[5] = ex - ey * ez * lam;
  b[6] = ey - ez * ex * lam;
  b[7] = -ex - ez * ey * lam;
  b[8] = QUIRK ? lam - ez * ez - lam : lam - ez * ez * lam;
}

// Fixed-point solve of the semi-implicit midpoint solvers (sphmidpoint.f90,
// sibt/sibf 36-477 for SPH = false, sispht/sisphf 485-923 for SPH = true):
// e = e1 + dtg u x a + sqrtdtg u x s with u = (e1 + e0)/2 or
// (e1 + e0)/|e1 + e0|, at most 100 iterations from e0 = e1, stopping when
// every component moved by less than float(1e-12) (a default-real literal
// in the reference).  f = the last iterate, p = the one before it when the
// loop stopped early (else the last).
template <typename R, bool SPH>
__device__ __forceinline__ void sib_solve(R e1x, R e1y, R e1z, R a0, R a1, R a2, R s0, R s1,
                                          R s2, R dtg, R sqrtdtg, R &fx, R &fy, R &fz, R &px,
                                          R &py, R &pz) {
  const R eps = (R)(double)1e-12f;
  R e0x = e1x, e0y = e1y, e0z = e1z;
  fx = e1x;
  fy = e1y;
  fz = e1z;
  for (int k = 0; k < 100; ++k) {
    R ux, uy, uz;
    if (SPH) {
      const R sx = e1x + e0x, sy = e1y + e0y, sz = e1z + e0z;
      const R un = sqrt(sx * sx + sy * sy + sz * sz);
      ux = sx / un;
      uy = sy / un;
      uz = sz / un;
    } else {
      ux = (R)0.5 * (e1x + e0x);
      uy = (R)0.5 * (e1y + e0y);
      uz = (R)0.5 * (e1z + e0z);
    }
    const R dx = dtg * ux, dy = dtg * uy, dz = dtg * uz;
    const R qx = sqrtdtg * ux, qy = sqrtdtg * uy, qz = sqrtdtg * uz;
    fx = e1x + dy * a2 - dz * a1 + qy * s2 - qz * s1;
    fy = e1y + dz * a0 - dx * a2 + qz * s0 - qx * s2;
    fz = e1z + dx * a1 - dy * a0 + qx * s1 - qy * s0;
    const R cx = e0x - fx, cy = e0y - fy, cz = e0z - fz;
    if (sqrt(cx * cx) < eps && sqrt(cy * cy) < eps && sqrt(cz * cz) < eps) break;
    e0x = fx;
    e0y = fy;
    e0z = fz;
  }
  px = e0x;
  py = e0y;
  pz = e0z;
}

// Cramer solve of (I + [A]x) e_t = (I - [A]x) e  (midpoint.f90:153-164)
template <typename R>
__device__ __forceinline__ void midpoint_solve(R A0, R A1, R A2, R e0, R e1, R e2, R &t0, R &t1,
                                               R &t2, R &detAi) {
  detAi = (R)1.0 / ((R)1.0 + (A0 * A0 + A1 * A1 + A2 * A2));
  R a0 = e0 + e1 * A2 - e2 * A1;
  R a1 = e1 + e2 * A0 - e0 * A2;
  R a2 = e2 + e0 * A1 - e1 * A0;
  t0 = a0 * ((R)1 + A0 * A0) + a1 * (A0 * A1 + A2) + a2 * (A0 * A2 - A1);
  t1 = a0 * (A1 * A0 - A2) + a1 * ((R)1 + A1 * A1) + a2 * (A1 * A2 + A0);
  t2 = a0 * (A2 * A0 + A1) + a1 * (A2 * A1 - A0) + a2 * ((R)1 + A2 * A2);
}

// =================================================================== kernels

// effective_field_full over all (site, ens): beff (quads, w unused) and, when
// en != null, the per-site field-routine energy -m.(0.5*(bs+2bq+2bext)) in FP64;
// when parts != null also beff1 = bs into parts[q] and beff2 = bq + bext into
// parts[N*M + q] (hamiltonianactions.f90:240-241).
template <typename R, int G>
__global__ void __launch_bounds__(256) k_field(HamDev<R> h, const Quad<R> *__restrict__ src,
                                               Quad<R> *__restrict__ beff, double *en,
                                               Quad<R> *__restrict__ parts) {
  const int64_t total = (int64_t)h.N * h.M;
  const int64_t q = xcd_block() * blockDim.x + threadIdx.x;
  if (q >= total) return;
  const int site = (int)fast_div((uint32_t)q, h.mdiv), ens = (int)q - site * h.M;
  R s0, s1, s2, q0, q1, q2, x0, x1, x2;
  site_field<R, G>(h, src, site, ens, s0, s1, s2, q0, q1, q2, x0, x1, x2);
  Quad<R> b;
  b.x = (R)0 + s0 + (q0 + x0 + (R)0);
  b.y = (R)0 + s1 + (q1 + x1 + (R)0);
  b.z = (R)0 + s2 + (q2 + x2 + (R)0);
  b.w = (R)0;
  beff[q] = b;
  if (parts) {
    parts[q] = Quad<R>{s0, s1, s2, (R)0};
    parts[total + q] = Quad<R>{q0 + x0 + (R)0, q1 + x1 + (R)0, q2 + x2 + (R)0, (R)0};
  }
  if (en) {
    R m0, m1, m2;
    moment(src, (size_t)q, m0, m1, m2);
    double t0 = 0.5 * ((double)s0 + 2.0 * (double)q0 + 2.0 * (double)x0 + 0.0);
    double t1 = 0.5 * ((double)s1 + 2.0 * (double)q1 + 2.0 * (double)x1 + 0.0);
    double t2 = 0.5 * ((double)s2 + 2.0 * (double)q2 + 2.0 * (double)x2 + 0.0);
    en[q] = -(double)m0 * t0 - (double)m1 * t1 - (double)m2 * t2;
  }
}

// Corrector output.  The FP32 form renormalises the moment direction: the
// rotations conserve |e| only to rounding, and FP32 rounding accumulated over
// thousands of steps is the largest error of that form.  FP64 keeps the
// reference's arithmetic (no renormalisation anywhere).
template <typename R>
__device__ __forceinline__ Quad<R> finish(R o0, R o1, R o2, R m) {
  return Quad<R>{o0, o1, o2, m};
}
template <>
__device__ __forceinline__ Quad<float> finish<float>(float o0, float o1, float o2, float m) {
  const float inv = 1.0f / sqrtf(o0 * o0 + o1 * o1 + o2 * o2);
  return Quad<float>{o0 * inv, o1 * inv, o2 * inv, m};
}

// One half of an LLG step.
//   ALG 1 midpoint (midpoint.f90), 5 Depondt (depondt.f90)
//   HALF 0 predictor (evolve_first), 1 corrector (evolve_second)
//   FUSED: compute the field of the gathered moments in-kernel; otherwise read
//   beff quads written by k_field.
// Buffers: e = emom at step start, p = predicted emom (emom2 after the first
// half), b2 = Depondt's b2eff.  Quads carry mmom in w.  The corrector writes
// the new emom into `out` (e when moment_update is fused, p otherwise).
// What the LLG half-step of one (site, ens) reads besides its field: the
// moment at step start, the predicted moment and Depondt's b2eff (corrector),
// and the thermal deviates.  Split from the update so a kernel can issue
// these loads and draw the deviates before its field is ready.
template <typename R>
struct SiteIn {
  Quad<R> eq, pq, bq;
  R n0, n1, n2;
};

// hsite: the site's reference index (the noise counter), q its quad index.
template <typename R, int ALG, int HALF>
__device__ __forceinline__ SiteIn<R> llg_fetch(const LLGDev<R> &prm, int hsite, int ens,
                                               int64_t q, const Quad<R> *e, const Quad<R> *p,
                                               const Quad<R> *b2) {
  SiteIn<R> in;
  in.eq = e[q];
  if (HALF == 1) in.pq = p[q];
  if ((ALG == 5 || ALG == 4) && HALF == 1) in.bq = b2[q];
  if (HALF == 1 && prm.nsave) {
    const Quad<R> v = prm.nsave[q];
    in.n0 = v.x;
    in.n1 = v.y;
    in.n2 = v.z;
  } else {
    prm.normals(hsite, ens, (size_t)q, in.n0, in.n1, in.n2);
    if (HALF == 0 && prm.nsave) prm.nsave[q] = Quad<R>{in.n0, in.n1, in.n2, (R)0};
  }
  return in;
}

// The LLG half-step of one (site, ens) once its effective field b is known.
template <typename R, int ALG, int HALF>
__device__ __forceinline__ void llg_apply(const LLGDev<R> &prm, int site, int64_t q, R b0, R b1,
                                          R b2v, const SiteIn<R> &in, Quad<R> *p, Quad<R> *b2,
                                          Quad<R> *out) {
  const LamConst<R> lc = prm.lamc(site);
  const double T = prm.temperature(site);
  const Quad<R> eq = in.eq;
  const R m = eq.w;
  const R n0 = in.n0, n1 = in.n1, n2 = in.n2;
  if (ALG == 5) {
    const R sig = prm.sig_uniform ? prm.sigp0 : depondt_sigma<R>(lc.Dp, (double)m, T, prm.temprescale);
    const R t0 = n0 * sig, t1 = n1 * sig, t2 = n2 * sig;
    const R bl0 = b0 + t0, bl1 = b1 + t1, bl2 = b2v + t2;
    if (HALF == 0) {
      R d0, d1, d2, o0, o1, o2;
      depondt_bdup(lc.lam, eq.x, eq.y, eq.z, bl0, bl1, bl2, d0, d1, d2);
      rodrigues(d0, d1, d2, (R)prm.dt, lc.lldamp, eq.x, eq.y, eq.z, o0, o1, o2);
      p[q] = Quad<R>{o0, o1, o2, m};
      b2[q] = Quad<R>{d0, d1, d2, (R)0};
    } else {
      const Quad<R> pq = in.pq;
      const Quad<R> bq = in.bq;
      R d0, d1, d2, o0, o1, o2;
      depondt_bdup(lc.lam, pq.x, pq.y, pq.z, bl0, bl1, bl2, d0, d1, d2);
      d0 = (R)0.5 * d0 + (R)0.5 * bq.x;
      d1 = (R)0.5 * d1 + (R)0.5 * bq.y;
      d2 = (R)0.5 * d2 + (R)0.5 * bq.z;
      rodrigues(d0, d1, d2, (R)prm.dt, lc.lldamp, eq.x, eq.y, eq.z, o0, o1, o2);
      out[q] = finish<R>(o0, o1, o2, m);
    }
  } else if (ALG == 6 || ALG == 7) {
    // SIB (6) and spherical midpoint (7), sphmidpoint.f90: the rannum noise
    // of the midpoint solver, a = -b - lam v x b, s = -f - lam v x f with
    // v = e (predictor) or the predicted midpoint (corrector)
    const R sig = prm.sig_uniform ? prm.sigk0 : midpoint_sigma<R>(lc.Dk, (double)m, T, prm.temprescale);
    const R f0 = n0 * sig, f1 = n1 * sig, f2 = n2 * sig;
    const R lg = (R)prm.landeg(site);
    const R dtg = lc.mdt * lg, sqrtdtg = lc.msqrtdt * lg;
    const R lam = lc.lam;
    R vx = eq.x, vy = eq.y, vz = eq.z;
    if (HALF == 1) {
      vx = in.pq.x;
      vy = in.pq.y;
      vz = in.pq.z;
    }
    const R a0 = -b0 - lam * (vy * b2v - vz * b1), a1 = -b1 - lam * (vz * b0 - vx * b2v),
            a2 = -b2v - lam * (vx * b1 - vy * b0);
    const R s0 = -f0 - lam * (vy * f2 - vz * f1), s1 = -f1 - lam * (vz * f0 - vx * f2),
            s2 = -f2 - lam * (vx * f1 - vy * f0);
    R fx, fy, fz, px, py, pz;
    sib_solve<R, ALG == 7>(eq.x, eq.y, eq.z, a0, a1, a2, s0, s1, s2, dtg, sqrtdtg, fx, fy, fz,
                           px, py, pz);
    if (HALF == 0) {
      if (ALG == 7) {
        const R sx = eq.x + px, sy = eq.y + py, sz = eq.z + pz;
        const R un = sqrt(sx * sx + sy * sy + sz * sz);
        p[q] = Quad<R>{sx / un, sy / un, sz / un, m};
      } else {
        p[q] = Quad<R>{(R)0.5 * (eq.x + fx), (R)0.5 * (eq.y + fy), (R)0.5 * (eq.z + fz), m};
      }
    } else {
      out[q] = finish<R>(fx, fy, fz, m);
    }
  } else if (ALG == 4) {
    // Heun predictor-corrector (heun_proper.f90): dW = ranv*Landeg,
    // dtg = dt*Landeg/(1+lam^2) with dt = delta_t*gama
    const R sig = prm.sig_uniform
                      ? prm.sigh0
                      : heun_sigma<R>(lc.Dh, (double)m, T, prm.temprescale, prm.dt);
    const R lg = (R)prm.landeg(site);
    const R w0 = n0 * sig * lg, w1 = n1 * sig * lg, w2 = n2 * sig * lg;
    const R dtg = (R)(prm.dt * kGama) * lg * lc.lldamp;
    const R lam = lc.lam;
    if (HALF == 0) {
      // heun4p: Euler predictor from the step's start, b2eff = beff
      R ax, ay, az, B[9];
      heun_drift(lam, eq.x, eq.y, eq.z, b0, b1, b2v, ax, ay, az);
      heun_diffusion<R, false>(lam, eq.x, eq.y, eq.z, B);
      const R o0 = eq.x + ax * dtg + B[0] * w0 + B[1] * w1 + B[2] * w2;
      const R o1 = eq.y + ay * dtg + B[3] * w0 + B[4] * w1 + B[5] * w2;
      const R o2 = eq.z + az * dtg + B[6] * w0 + B[7] * w1 + B[8] * w2;
      p[q] = Quad<R>{o0, o1, o2, m};
      b2[q] = Quad<R>{b0, b1, b2v, (R)0};
    } else {
      // heun4f + evolve4f: average of the drift and diffusion at the
      // predictor (field b) and at the start (field b2eff), renormalised
      const Quad<R> pq = in.pq, bq = in.bq;
      R a1x, a1y, a1z, a2x, a2y, a2z, B1[9], B2[9];
      heun_drift(lam, pq.x, pq.y, pq.z, b0, b1, b2v, a1x, a1y, a1z);
      heun_diffusion<R, true>(lam, pq.x, pq.y, pq.z, B1);
      heun_drift(lam, eq.x, eq.y, eq.z, bq.x, bq.y, bq.z, a2x, a2y, a2z);
      heun_diffusion<R, false>(lam, eq.x, eq.y, eq.z, B2);
      R B[9];
#pragma unroll
      for (int j = 0; j < 9; ++j) B[j] = B1[j] + B2[j];
      const R ax = a1x + a2x, ay = a1y + a2y, az = a1z + a2z;
      const R t0 = ax * dtg + B[0] * w0 + B[1] * w1 + B[2] * w2;
      const R t1 = ay * dtg + B[3] * w0 + B[4] * w1 + B[5] * w2;
      const R t2 = az * dtg + B[6] * w0 + B[7] * w1 + B[8] * w2;
      const R o0 = eq.x + (R)0.5 * t0, o1 = eq.y + (R)0.5 * t1, o2 = eq.z + (R)0.5 * t2;
      const R etot = (R)1.0 / sqrt(o0 * o0 + o1 * o1 + o2 * o2);
      out[q] = Quad<R>{o0 * etot, o1 * etot, o2 * etot, m};
    }
  } else {
    const R sig = prm.sig_uniform ? prm.sigk0 : midpoint_sigma<R>(lc.Dk, (double)m, T, prm.temprescale);
    const R r0 = n0 * sig, r1 = n1 * sig, r2 = n2 * sig;
    R A0, A1, A2, t0, t1, t2, detAi;
    if (HALF == 0) {
      midpoint_A(lc.lam, lc.mdt, lc.msqrtdt, (R)prm.landeg(site), b0, b1, b2v, r0, r1, r2, eq.x,
                 eq.y, eq.z, A0, A1, A2);
      midpoint_solve(A0, A1, A2, eq.x, eq.y, eq.z, t0, t1, t2, detAi);
      p[q] = Quad<R>{(R)0.5 * (eq.x + t0 * detAi), (R)0.5 * (eq.y + t1 * detAi),
                     (R)0.5 * (eq.z + t2 * detAi), m};
    } else {
      const Quad<R> pq = in.pq;
      midpoint_A(lc.lam, lc.mdt, lc.msqrtdt, (R)prm.landeg(site), b0, b1, b2v, r0, r1, r2, pq.x,
                 pq.y, pq.z, A0, A1, A2);
      midpoint_solve(A0, A1, A2, eq.x, eq.y, eq.z, t0, t1, t2, detAi);
      out[q] = finish<R>(t0 * detAi, t1 * detAi, t2 * detAi, m);
    }
  }
}

template <typename R, int ALG, int HALF>
__device__ __forceinline__ void llg_update(const HamDev<R> &h, const LLGDev<R> &prm, int site,
                                           int ens, int64_t q, R b0, R b1, R b2v,
                                           const Quad<R> *e, Quad<R> *p, Quad<R> *b2,
                                           Quad<R> *out) {
  const SiteIn<R> in = llg_fetch<R, ALG, HALF>(prm, (int)host_site(h, site), ens, q, e, p, b2);
  llg_apply<R, ALG, HALF>(prm, site, q, b0, b1, b2v, in, p, b2, out);
}

template <typename R, int ALG, int HALF, bool FUSED, int G>
__global__ void __launch_bounds__(256)
    k_llg(HamDev<R> h, LLGDev<R> prm, const Quad<R> *e, Quad<R> *p, Quad<R> *b2,
          const Quad<R> *__restrict__ beff, Quad<R> *out) {
  const int64_t total = (int64_t)h.N * h.M;
  const int64_t q = xcd_block() * blockDim.x + threadIdx.x;
  if (q >= total) return;
  const int site = (int)fast_div((uint32_t)q, h.mdiv), ens = (int)q - site * h.M;
  R b0, b1, b2v;
  if (FUSED) {
    total_field<R, G>(h, HALF == 0 ? e : (const Quad<R> *)p, site, ens, b0, b1, b2v);
  } else {
    const Quad<R> b = beff[q];
    b0 = b.x;
    b1 = b.y;
    b2v = b.z;
  }
  llg_update<R, ALG, HALF>(h, prm, site, ens, q, b0, b1, b2v, e, p, b2, out);
}

// ------------------------------------------------------- LDS-tiled stencil
// For a periodic lattice whose neighbour lists are one translation-invariant
// stencil per basis atom (reference site i = b + NA*(x + N1*(y + N2*z)),
// neighbour j of basis b at basis b' and cell offset (dx, dy, dz) for every
// cell), a workgroup owns a tile of TX*TY*TZ cells (or, in the resident
// launch, a run of adjacent tiles one after the other).  It stages emomM = e*m
// of the tile plus a halo of (RX, RY, RZ) cells for all basis atoms in LDS,
// one plane per component at a fixed stride of PL values (so the
// three reads of a neighbour share one address and differ by an immediate
// offset, and lanes on consecutive cells read consecutive words: no bank
// conflicts along a row of 16 cells); then each of the TX*TY*TZ*NA threads
// takes one site (cell, basis b) and reads neighbour j at its halo position
// plus off[b*W + j] (= b'*HV + dz*HX*HY + dy*HX + dx).  The
// terms are added in list order with the same products as the ELL path, so
// results are identical; the gather traffic moves from L2 to LDS.  M = 1.
struct TileDev {
  int N1, N2, N3, NA;
  int TX, TY, TZ;
  int RX, RY, RZ;
  int HX, HY, HZ, HV;     // halo tile extents, HV = HX*HY*HZ
  // 1/d in FP32 for the small divisors d of the index arithmetic:
  // fdiv(n, 1/d) = n/d exactly for 0 <= n < 2^22
  float rTX, rTXY, rTC, rHX, rntx, rntxy;
  // the same for the halo table: 1/HY, 1/(HY*HZ), 1/N1, 1/N2, 1/N3, and the
  // multiples of N1, N2, N3 at least RX, RY, RZ that keep wrapped
  // coordinates non-negative
  float rHY, rHYZ, rN1, rN2, rN3;
  int ox, oy, oz;
  int rs;                 // halo rows staged per round: blockDim.x / HX
  int per;                // tiles per workgroup (consecutive in the XCD-aware order)
  int ntx, nty, ntz;      // tiles per direction
  int W, Wd;              // exchange / DM stencil widths
  const int *off;         // (NA*W) exchange offsets into the halo tile
  const int *doff;        // (NA*Wd) DM offsets
  // ucls = 1: every site of basis b has site class cls_b[b] (a lattice of one
  // species), so class rows, couplings and offsets are the same for all lanes
  // of a wavefront and are read through the scalar cache
  int ucls;
  int cls_b[8];
};

// n / d for 0 <= n < 2^22 and rd = 1/d rounded to FP32: (n + 1/2) rd is
// within 2^-23 (n + 1/2)/d < 1/(2d) of (n + 1/2)/d, whose fractional part
// lies in [1/(2d), 1 - 1/(2d)], so the truncation is the exact quotient.
__device__ __forceinline__ int fdiv(int n, float rd) {
  return (int)(((float)n + 0.5f) * rd);
}

template <typename R>
__device__ __forceinline__ int dev_slot(const HamDev<R> &h, int i) {
  return h.order_na == 1 ? i : (i % h.order_na) * h.order_nc + i / h.order_na;
}

// Load through the constant address space: the tables it is used for (class
// rows, couplings, DM vectors) are not written by any kernel, and a
// wave-uniform address is then read with a scalar load.
template <typename T>
__device__ __forceinline__ T cload(const T *p) {
  return *(const __attribute__((address_space(4))) T *)p;
}

// Values per component plane of the LDS tile: the launch takes the smallest
// that holds NA*HV.  Each is 8 more than a multiple of 64, so the three
// component reads of a neighbour never pair into one two-address LDS read
// (whose halves would share banks); 2*PL doubles stay within the 16-bit
// immediate offset of a DS instruction.
constexpr int kNumTilePlanes = 4;
constexpr int kTilePlanes[kNumTilePlanes] = {1160, 2312, 3080, 3848};

// Effective field of one site of a tile from the LDS halo.  hbase is the
// halo index of the site's cell.
// UCLS: the site class is the per-basis one of TileDev (wave-uniform).
// FULL: DM, anisotropy and Zeeman terms may be present; without them the
// kernel holds fewer launch arguments live (exchange-only Hamiltonians).
template <typename R, int PL, bool UCLS, bool FULL>
__device__ __forceinline__ void tile_field(const HamDev<R> &h, const TileDev &t, const R *sm,
                                           int hbase, int b, int site, int ens, int64_t q, R &b0,
                                           R &b1, R &b2v) {
  SiteClass sc;
  if (UCLS) {
    // wave-uniform values kept in scalar registers
    const SiteClass *cp = h.classes + __builtin_amdgcn_readfirstlane(t.cls_b[b]);
    sc.ex_row = __builtin_amdgcn_readfirstlane(cload(&cp->ex_row));
    sc.dm_row = __builtin_amdgcn_readfirstlane(cload(&cp->dm_row));
    sc.an_row = __builtin_amdgcn_readfirstlane(cload(&cp->an_row));
    sc.pad = 0;
  } else {
    sc = h.classes[h.cls[site]];
  }
  R s0 = 0, s1 = 0, s2 = 0, qf0 = 0, qf1 = 0, qf2 = 0;
  if (FULL && h.dm_shared) {
    // the DM bonds are the exchange bonds (at most kFusedW): one LDS read of
    // each neighbour serves both sums, the terms added in the order of the
    // separate loops below
    const int len = h.ex_len[sc.ex_row];
    const R *cp = h.cpl + sc.ex_row;
    const int *off = t.off + b * t.W;
    R c0[kFusedW], c1[kFusedW], c2[kFusedW];
#pragma unroll
    for (int j0 = 0; j0 < kFusedW; j0 += 4) {
      R J[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        if (j0 + u < len) {
          const R *v = sm + (hbase + cload(off + j0 + u));
          c0[j0 + u] = v[0];
          c1[j0 + u] = v[PL];
          c2[j0 + u] = v[2 * PL];
          J[u] = UCLS ? cload(cp + (size_t)(j0 + u) * h.ex_rows) : cp[(size_t)(j0 + u) * h.ex_rows];
        }
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        if (j0 + u < len) {
          s0 = fma(J[u], c0[j0 + u], s0);
          s1 = fma(J[u], c1[j0 + u], s1);
          s2 = fma(J[u], c2[j0 + u], s2);
        }
      }
    }
    if (h.Wd > 0 && sc.dm_row >= 0) {
#pragma unroll
      for (int j = 0; j < kFusedW; ++j) {
        if (j < len) {
          const size_t r = (size_t)j * h.dm_rows + sc.dm_row;
          const R D0 = UCLS ? cload(h.dvx + r) : h.dvx[r], D1 = UCLS ? cload(h.dvy + r) : h.dvy[r],
                  D2 = UCLS ? cload(h.dvz + r) : h.dvz[r];
          s0 = s0 + D2 * c1[j] - D1 * c2[j];
          s1 = s1 + D0 * c2[j] - D2 * c0[j];
          s2 = s2 + D1 * c0[j] - D0 * c1[j];
        }
      }
    }
  } else {
    // four neighbours per round: LDS reads first, then the terms in order
    const int len = h.ex_len[sc.ex_row];
    const R *cp = h.cpl + sc.ex_row;
    const int *off = t.off + b * t.W;
    int j = 0;
    for (; j + 4 <= len; j += 4) {
      R a0[4], a1[4], a2[4], J[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const R *v = sm + (hbase + cload(off + j + u));
        a0[u] = v[0];
        a1[u] = v[PL];
        a2[u] = v[2 * PL];
        J[u] = UCLS ? cload(cp + (size_t)(j + u) * h.ex_rows) : cp[(size_t)(j + u) * h.ex_rows];
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        s0 = fma(J[u], a0[u], s0);
        s1 = fma(J[u], a1[u], s1);
        s2 = fma(J[u], a2[u], s2);
      }
    }
    for (; j < len; ++j) {
      const R *v = sm + (hbase + cload(off + j));
      const R J = UCLS ? cload(cp + (size_t)j * h.ex_rows) : cp[(size_t)j * h.ex_rows];
      s0 = fma(J, v[0], s0);
      s1 = fma(J, v[PL], s1);
      s2 = fma(J, v[2 * PL], s2);
    }
  }
  if (FULL && !h.dm_shared && h.Wd > 0 && sc.dm_row >= 0) {
    const int len = h.dm_len[sc.dm_row];
    const int *off = t.doff + b * t.Wd;
    for (int j = 0; j < len; ++j) {
      const R *v = sm + (hbase + cload(off + j));
      const size_t r = (size_t)j * h.dm_rows + sc.dm_row;
      const R D0 = UCLS ? cload(h.dvx + r) : h.dvx[r], D1 = UCLS ? cload(h.dvy + r) : h.dvy[r],
              D2 = UCLS ? cload(h.dvz + r) : h.dvz[r];
      const R m0 = v[0], m1 = v[PL], m2 = v[2 * PL];
      s0 = s0 + D2 * m1 - D1 * m2;
      s1 = s1 + D0 * m2 - D2 * m0;
      s2 = s2 + D1 * m0 - D0 * m1;
    }
  }
  if (FULL && sc.an_row >= 0) {
    const int ty_ = h.an_type[sc.an_row];
    const R *v = sm + (hbase + b * t.HV);
    const R m0 = v[0], m1 = v[PL], m2 = v[2 * PL];
    if (ty_ == 1) {
      uniaxial(h, sc.an_row, m0, m1, m2, s0, s1, s2);
    } else if (ty_ == 2) {
      cubic(h, sc.an_row, m0, m1, m2, s0, s1, s2);
    } else if (ty_ == 7) {
      R c0 = 0, c1 = 0, c2 = 0;
      uniaxial(h, sc.an_row, m0, m1, m2, s0, s1, s2);
      cubic(h, sc.an_row, m0, m1, m2, c0, c1, c2);
      const R sb = h.an_sb[sc.an_row];
      qf0 = qf0 + c0 * sb;
      qf1 = qf1 + c1 * sb;
      qf2 = qf2 + c2 * sb;
    }
  }
  R x0 = 0, x1 = 0, x2 = 0;
  if (FULL && h.ext_mode == 1) {
    x0 = h.ext_u[3 * ens];
    x1 = h.ext_u[3 * ens + 1];
    x2 = h.ext_u[3 * ens + 2];
  } else if (FULL && h.ext_mode == 2) {
    const Quad<R> v = h.ext_q[q];
    x0 = v.x;
    x1 = v.y;
    x2 = v.z;
  }
  b0 = (R)0 + s0 + (qf0 + x0 + (R)0);
  b1 = (R)0 + s1 + (qf1 + x1 + (R)0);
  b2v = (R)0 + s2 + (qf2 + x2 + (R)0);
}

// Halo staging.  The halo of a tile is NR = NA*HZ*HY rows of HX cells; item
// i = row*HX + hx goes to value i of each LDS plane.  Thread t stages column
// t % HX of rows t / HX, t / HX + RS, ... with RS = blockDim.x / HX rows per
// round (the last blockDim.x % HX threads stage nothing).  A table per tile
// holds the x term of every column and the slot of every row's first cell,
// so an item's device slot is one LDS read and an add.
// Items per thread whose loads are held in registers while the previous tile
// is computed (more would spill); the rest are staged after it, in rounds of
// kTileBatch loads issued together.
constexpr int kTilePre = 2;
constexpr int kTileBatch = 4;

// Origin cell of tile `tile` (x fastest, z slowest).
__device__ __forceinline__ void tile_origin(const TileDev &t, int tile, int &X0, int &Y0, int &Z0) {
  const int tz = fdiv(tile, t.rntxy), txy = tile - tz * t.ntx * t.nty;
  const int ty = fdiv(txy, t.rntx), tx = txy - ty * t.ntx;
  X0 = tx * t.TX;
  Y0 = ty * t.TY;
  Z0 = tz * t.TZ;
}

// The staging table of a tile: tab[hx] is the x term of halo column hx,
// tab[HX + row] the device slot of the row's basis atom at its wrapped (y, z)
// and x = 0, so that slot = tab[hx] + tab[HX + row] (periodic images of any
// multiplicity).  Both terms are scaled by M, so that the sum is the quad
// index of the slot's ensemble 0 (N*M < 2^31).
template <typename R>
__device__ __forceinline__ void halo_table(const HamDev<R> &h, const TileDev &t, int X0, int Y0,
                                           int Z0, int ncell, int M, int *tab) {
  const int nr = t.NA * t.HY * t.HZ, hyz = t.HY * t.HZ;
  // wrapped coordinates by the reciprocal divisions (exact: every dividend
  // is non-negative and below 2^22)
  for (int i = threadIdx.x; i < t.HX + nr; i += blockDim.x) {
    int v;
    if (i < t.HX) {
      const int n = X0 + i - t.RX + t.ox;
      const int x = n - t.N1 * fdiv(n, t.rN1);
      v = h.order_na == 1 ? t.NA * x : x;
    } else {
      const int r = i - t.HX;
      const int hb = fdiv(r, t.rHYZ), rz = r - hb * hyz;
      const int hz = fdiv(rz, t.rHY), hy = rz - hz * t.HY;
      const int ny = Y0 + hy - t.RY + t.oy, nz = Z0 + hz - t.RZ + t.oz;
      const int y = ny - t.N2 * fdiv(ny, t.rN2);
      const int z = nz - t.N3 * fdiv(nz, t.rN3);
      const int cyz = t.N1 * (y + t.N2 * z);
      v = h.order_na == 1 ? hb + t.NA * cyz : hb * ncell + cyz;
    }
    tab[i] = v * M;
  }
}

// This thread's staging column and first row.
struct HaloCol {
  int hx, r0;
  bool on;  // the thread stages items
};

__device__ __forceinline__ HaloCol halo_col(const TileDev &t) {
  HaloCol c;
  c.r0 = fdiv(threadIdx.x, t.rHX);
  c.hx = threadIdx.x - c.r0 * t.HX;
  c.on = c.r0 < t.rs;
  return c;
}

template <typename R, int PL>
__device__ __forceinline__ void halo_put(R *sm, int item, const Quad<R> &v) {
  R *d = sm + item;
  d[0] = v.x * v.w;
  d[PL] = v.y * v.w;
  d[2 * PL] = v.z * v.w;
}

// Issue the loads of K rounds of this thread's items, from round k0.  item[u]
// is -1 for rounds past the last row.
template <typename R, int K>
__device__ __forceinline__ void halo_load(const TileDev &t, const int *tab, const HaloCol &c,
                                          int k0, const Quad<R> *src, Quad<R> (&v)[K],
                                          int (&item)[K]) {
  const int nr = t.NA * t.HY * t.HZ;
  const int xterm = tab[c.hx];
#pragma unroll
  for (int u = 0; u < K; ++u) {
    const int r = c.r0 + (k0 + u) * t.rs;
    item[u] = c.on && r < nr ? r * t.HX + c.hx : -1;
  }
#pragma unroll
  for (int u = 0; u < K; ++u)
    if (item[u] >= 0) v[u] = src[tab[t.HX + c.r0 + (k0 + u) * t.rs] + xterm];
}

// Items of this thread from round kTilePre on, in batches of kTileBatch.
template <typename R, int PL>
__device__ __forceinline__ void halo_rest(const TileDev &t, const int *tab, const HaloCol &c,
                                          const Quad<R> *src, R *sm) {
  const int nr = t.NA * t.HY * t.HZ;
  for (int k = kTilePre; c.on && c.r0 + k * t.rs < nr; k += kTileBatch) {
    Quad<R> v[kTileBatch];
    int item[kTileBatch];
    halo_load<R, kTileBatch>(t, tab, c, k, src, v, item);
#pragma unroll
    for (int u = 0; u < kTileBatch; ++u)
      if (item[u] >= 0) halo_put<R, PL>(sm, item[u], v[u]);
  }
}

// Persistent tiles: the work units are (tile, ensemble) pairs, unit u =
// tile*M + ens, so the M ensembles of a tile follow one another (their quads
// share cache lines).  Workgroup w takes the units [w*per, (w+1)*per) of the
// XCD-aware block order.  The halo loads of its next unit are issued before
// the current unit is computed and stored after it, so their latency is
// covered by that computation.  ENS = false is the one-ensemble instance, in
// which every unit term folds away at compile time.
template <typename R, int ALG, int HALF, int PL, bool FULL, bool ENS>
__global__ void __launch_bounds__(1024)
    k_llg_tile(HamDev<R> h, LLGDev<R> prm, TileDev t, const Quad<R> *e, Quad<R> *p,
               Quad<R> *b2, Quad<R> *out) {
  extern __shared__ __attribute__((aligned(16))) unsigned char lds_raw[];
  R *sm = reinterpret_cast<R *>(lds_raw);
  const Quad<R> *src = HALF == 0 ? e : (const Quad<R> *)p;
  const int M = ENS ? h.M : 1;
  const int nunits = t.ntx * t.nty * t.ntz * M;
  const int first = (int)xcd_block() * t.per;
  const int last = min(nunits, first + t.per);
  if (first >= last) return;  // whole workgroups exit before any barrier
  const int ncell = t.N1 * t.N2 * t.N3;
  const int nw = t.HX + t.NA * t.HY * t.HZ;
  int *tabs = reinterpret_cast<int *>(sm + 3 * PL);  // two tables, alternating by unit
  // this thread's site in every tile: cell (lx, ly, lz), basis b.  blockDim.x
  // = tc * NA with tc a multiple of 64, so b is the same for all lanes of a
  // wavefront and the stencil offsets (and, for one species, the couplings)
  // are read with scalar loads.
  const int tc = t.TX * t.TY * t.TZ;
  const int tid = threadIdx.x;
  const int b = __builtin_amdgcn_readfirstlane(fdiv(tid, t.rTC));
  const int lcell = tid - b * tc;
  const int lz = fdiv(lcell, t.rTXY), lxy = lcell - lz * t.TX * t.TY;
  const int ly = fdiv(lxy, t.rTX), lx = lxy - ly * t.TX;
  const int hcell = (lz + t.RZ) * t.HX * t.HY + (ly + t.RY) * t.HX + (lx + t.RX);
  const HaloCol col = halo_col(t);
  // the first unit: table, halo
  int tile = first / M, ens = first - tile * M;
  int X0, Y0, Z0;
  tile_origin(t, tile, X0, Y0, Z0);
  halo_table(h, t, X0, Y0, Z0, ncell, M, tabs);
  __syncthreads();
  Quad<R> v[kTilePre];
  int item[kTilePre];
  halo_load<R, kTilePre>(t, tabs, col, 0, src + ens, v, item);
#pragma unroll
  for (int u = 0; u < kTilePre; ++u)
    if (item[u] >= 0) halo_put<R, PL>(sm, item[u], v[u]);
  halo_rest<R, PL>(t, tabs, col, src + ens, sm);
  for (int unit = first;;) {
    const int k = unit - first;
    const int x = X0 + lx, y = Y0 + ly, z = Z0 + lz;
    const bool valid = x < t.N1 && y < t.N2 && z < t.N3;
    const int cell = x + t.N1 * (y + t.N2 * z);
    const int site = h.order_na == 1 ? b + t.NA * cell : b * ncell + cell;
    const int64_t q = (int64_t)site * M + ens;
    const int next = unit + 1 < last ? unit + 1 : -1;
    const int ntile = ens + 1 < M ? tile : tile + 1, nens = ens + 1 < M ? ens + 1 : 0;
    int *tnext = tabs + ((k + 1) & 1) * nw;
    int nX0 = 0, nY0 = 0, nZ0 = 0;
    if (next >= 0) {
      tile_origin(t, ntile, nX0, nY0, nZ0);
      halo_table(h, t, nX0, nY0, nZ0, ncell, M, tnext);
    }
    __syncthreads();  // this unit's halo and the next unit's table are in place
    if (next >= 0) halo_load<R, kTilePre>(t, tnext, col, 0, src + nens, v, item);
    if (valid) {
      R b0, b1, b2v;
      if (t.ucls)
        tile_field<R, PL, true, FULL>(h, t, sm, hcell, b, site, ens, q, b0, b1, b2v);
      else
        tile_field<R, PL, false, FULL>(h, t, sm, hcell, b, site, ens, q, b0, b1, b2v);
      const SiteIn<R> in = llg_fetch<R, ALG, HALF>(prm, b + t.NA * cell, ens, q, e, p, b2);
      llg_apply<R, ALG, HALF>(prm, site, q, b0, b1, b2v, in, p, b2, out);
    }
    if (next < 0) break;
    __syncthreads();  // every read of this unit's halo is done
#pragma unroll
    for (int u = 0; u < kTilePre; ++u)
      if (item[u] >= 0) halo_put<R, PL>(sm, item[u], v[u]);
    halo_rest<R, PL>(t, tnext, col, src + nens, sm);
    unit = next;
    tile = ntile;
    ens = nens;
    X0 = nX0;
    Y0 = nY0;
    Z0 = nZ0;
  }
}

// ----------------------------------------------------------------- MC
// choose_random_flip (montecarlo_common.f90:25-79)
template <typename R>
__device__ __forceinline__ void trial_move(R e0, R e1, R e2, double delta, double u0, double u1,
                                          double u2, double g0, double g1, double g2, R &n0,
                                          R &n1, R &n2) {
  const int ftype = (int)floor(3 * u0);
  if (ftype == 0) {
    // phi = 2 pi u1, theta = acos(1 - 2 u2) (montecarlo_common.f90:55-59):
    // cos(theta) = z and sin(theta) = sqrt(1 - z^2) >= 0 on [0, pi], and
    // sincospi(2 u1) is sin and cos of phi, all equal to the library forms to
    // rounding without the acos, sin and cos of theta
    const double z = 1 - 2 * u2;
    const double st = sqrt(fmax(0.0, 1 - z * z));
    double sp, cp;
    sincospi(2 * u1, &sp, &cp);
    n0 = (R)(st * cp);
    n1 = (R)(st * sp);
    n2 = (R)z;
  } else if (ftype == 1) {
    R a0 = e0 + (R)(g0 * delta), a1 = e1 + (R)(g1 * delta), a2 = e2 + (R)(g2 * delta);
    // one reciprocal for the three divisions by the length (equal to rounding)
    const R il = (R)1 / fsqrt(a0 * a0 + a1 * a1 + a2 * a2);
    n0 = a0 * il;
    n1 = a1 * il;
    n2 = a2 * il;
  } else {
    // Spin flip, newmom = -emom.  Written as a subtraction: hipcc (clang 22,
    // -O1 and above) drops a bare negation of a loaded value when it merges
    // the three branches of this function, turning every flip into a no-op.
    // 0 - e equals -e except that a +0 component stays +0, which changes no
    // energy or later arithmetic.
    n0 = (R)0 - e0;
    n1 = (R)0 - e1;
    n2 = (R)0 - e2;
  }
}

template <typename R>
__device__ __forceinline__ R sq(R x) {
  return x * x;
}

// calculate_energy (montecarlo_common.f90:431-865), exchange + anisotropy +
// DM + external field, for the U adjacent ensembles e0 .. e0+U-1 of one site:
// de[k] = mub*(E_trial - E_current) of ensemble e0+k, whose current moment is
// me[k], trial direction tr[k] and uniform field x[k].  The site's class, its
// neighbour entries and its couplings are read once for the U ensembles.
template <typename R, int G, int U>
__device__ __forceinline__ void mc_delta_e(const HamDev<R> &h, const Quad<R> *__restrict__ s,
                                           int site, int e0, const Quad<R> (&me)[U],
                                           const R (&tr)[U][3], const R (&x)[U][3],
                                           R (&de)[U]) {
  const SiteClass sc = h.classes[h.cls[site]];
  const int M = h.M, N = h.N;
  // The exchange energy of the current and the trial moment, -sum_j J_j
  // m.m_j (montecarlo_common.f90:512-520), is formed as -m.f with the
  // exchange field f = sum_j J_j m_j of site_field: the same quantity with a
  // third of the arithmetic, equal to the term-by-term sums to rounding.
  R f[U][3];
#pragma unroll
  for (int k = 0; k < U; ++k) f[k][0] = f[k][1] = f[k][2] = (R)0;
  const int row = sc.ex_row;
  if (h.nbr16) {
    if (h.nt)
      mc_exchange_fields<R, G, true, int16_t, U>(h, h.nbr16 + site, s, site, e0, row, f);
    else
      mc_exchange_fields<R, G, false, int16_t, U>(h, h.nbr16 + site, s, site, e0, row, f);
  } else {
    if (h.nt)
      mc_exchange_fields<R, G, true, int, U>(h, h.nbr + site, s, site, e0, row, f);
    else
      mc_exchange_fields<R, G, false, int, U>(h, h.nbr + site, s, site, e0, row, f);
  }
  R c0[U], c1[U], c2[U], t0[U], t1[U], t2[U], e_c[U], e_t[U];
#pragma unroll
  for (int k = 0; k < U; ++k) {
    const R m = me[k].w;
    c0[k] = me[k].x * m;
    c1[k] = me[k].y * m;
    c2[k] = me[k].z * m;
    t0[k] = tr[k][0] * m;
    t1[k] = tr[k][1] * m;
    t2[k] = tr[k][2] * m;
    e_c[k] = (R)0 - (c0[k] * f[k][0] + c1[k] * f[k][1] + c2[k] * f[k][2]);
    e_t[k] = (R)0 - (t0[k] * f[k][0] + t1[k] * f[k][1] + t2[k] * f[k][2]);
  }
  if (sc.an_row >= 0) {
    const int ty = h.an_type[sc.an_row];
    const R k1 = h.an_k1[sc.an_row], k2 = h.an_k2[sc.an_row];
    const R ea0 = h.an_ex[sc.an_row], ea1 = h.an_ey[sc.an_row], ea2 = h.an_ez[sc.an_row];
#pragma unroll
    for (int k = 0; k < U; ++k) {
      if (ty == 1 || ty == 7) {
        R tta = c0[k] * ea0 + c1[k] * ea1 + c2[k] * ea2;
        R ttb = t0[k] * ea0 + t1[k] * ea1 + t2[k] * ea2;
        e_c[k] = e_c[k] + k1 * sq(tta) + k2 * sq(sq(tta));
        e_t[k] = e_t[k] + k1 * sq(ttb) + k2 * sq(sq(ttb));
      } else if (ty == 2) {
        e_c[k] = e_c[k] - k1 * (sq(c0[k]) * sq(c1[k]) + sq(c1[k]) * sq(c2[k]) + sq(c2[k]) * sq(c0[k])) -
                 k2 * (sq(c0[k]) * sq(c1[k]) * sq(c2[k]));
        e_t[k] = e_t[k] - k1 * (sq(t0[k]) * sq(t1[k]) + sq(t1[k]) * sq(t2[k]) + sq(t2[k]) * sq(t0[k])) -
                 k2 * (sq(t0[k]) * sq(t1[k]) * sq(t2[k]));
      }
      if (ty == 7) {
        const R sb = h.an_sb[sc.an_row];
        const R aw1 = k1 * sb, aw2 = k2 * sb;
        e_c[k] = e_c[k] + aw1 * (sq(c0[k]) * sq(c1[k]) + sq(c1[k]) * sq(c2[k]) + sq(c2[k]) * sq(c0[k])) +
                 aw2 * (sq(c0[k]) * sq(c1[k]) * sq(c2[k]));
        e_t[k] = e_t[k] + aw1 * (sq(t0[k]) * sq(t1[k]) + sq(t1[k]) * sq(t2[k]) + sq(t2[k]) * sq(t0[k])) +
                 aw2 * (sq(t0[k]) * sq(t1[k]) * sq(t2[k]));
      }
    }
  }
  if (h.Wd > 0 && sc.dm_row >= 0) {
    const int len = h.dm_len[sc.dm_row];
    for (int j = 0; j < len; ++j) {
      const int nb = h.dnbr16 ? NbDec(h, site)(h.dnbr16[(size_t)j * N + site])
                              : h.dnbr[(size_t)j * N + site];
      const size_t r = (size_t)j * h.dm_rows + sc.dm_row;
      const R D0 = h.dvx[r], D1 = h.dvy[r], D2 = h.dvz[r];
#pragma unroll
      for (int k = 0; k < U; ++k) {
        // unit-vector terms as in montecarlo_common.f90:612,615
        const R u0 = me[k].x, u2 = me[k].z;
        R a0, a1, a2;
        moment(s, (size_t)nb * M + e0 + k, a0, a1, a2);
        e_c[k] = e_c[k] - D0 * (c1[k] * a2 - u2 * a1) - D1 * (c2[k] * a0 - c0[k] * a2) -
                 D2 * (u0 * a1 - c1[k] * a0);
        e_t[k] = e_t[k] - D0 * (t1[k] * a2 - t2[k] * a1) - D1 * (t2[k] * a0 - t0[k] * a2) -
                 D2 * (t0[k] * a1 - t1[k] * a0);
      }
    }
  }
#pragma unroll
  for (int k = 0; k < U; ++k) {
    e_c[k] = e_c[k] - x[k][0] * c0[k] - x[k][1] * c1[k] - x[k][2] * c2[k];
    e_t[k] = e_t[k] - x[k][0] * t0[k] - x[k][1] * t1[k] - x[k][2] * t2[k];
    const R tt = e_t[k] - e_c[k];
    de[k] = (R)kMub * tt;
  }
}

// flip_h (montecarlo_common.f90:371-422)
template <typename R>
__device__ __forceinline__ void heat_bath(double beta, R m, R f0, R f1, R f2, double flipprob,
                                          double qr, R &o0, R &o1, R &o2) {
  double z0 = beta * (double)f0 * kMub * (double)m;
  double z1 = beta * (double)f1 * kMub * (double)m;
  double z2 = beta * (double)f2 * kMub * (double)m;
  double zarg = sqrt(z0 * z0 + z1 * z1 + z2 * z2);
  // the reference's four divisions by zarg and zarg*zst as two reciprocals,
  // its exp(-2 zarg) once, and cos / sin of phi = pi (2 flipprob - 1) by
  // sincospi: equal to the reference's expressions to rounding
  const double rz = 1.0 / zarg;
  double zct = z2 * rz;
  double zst = sqrt(1.0 - zct * zct) + kDblTolerance;
  const double rzs = rz / zst;
  double zcp = z0 * rzs;
  double zsp = z1 * rzs;
  const double ex = exp(-2.0 * zarg);
  double ct = 1.0 + rz * log((1.0 - ex) * qr + ex + kDblTolerance);
  double stv = sqrt(1.0 - ct * ct);
  double sp, cp;
  sincospi(2.0 * flipprob - 1.0, &sp, &cp);
  double s0 = stv * cp, s1 = stv * sp, s2 = ct;
  o0 = (R)(zcp * zct * s0 - zsp * s1 + zcp * zst * s2);
  o1 = (R)(zsp * zct * s0 + zcp * s1 + zsp * zst * s2);
  o2 = (R)(-zst * s0 + zct * s2);
}

struct MCParams {
  const int *sites;     // sites of the colour being updated
  int nsites;
  int site0, sstride;   // sstride > 0: the colour's i-th site is site0 + i*sstride
  const double *delta;  // (M) Gaussian step width per ensemble
  const double *beta;   // (M) Metropolis 1/(k_B (T*temprescale + 1e-15))
  const double *hbeta;  // (M) heat-bath 1/k_B/(temprescale*T)
  const double *ext;    // (3*M) uniform field per ensemble
  uint64_t seed, sweep;      // sweep counter, or offset from *sweep_dev
  const uint64_t *sweep_dev;  // device-held sweep base (graph replays) or nullptr
  int heat_bath;
  FastDiv pdiv;  // division by M / U (the threads of one site)
};

// FP32 engine: the whole Metropolis trial of one (site, ensemble) from one
// Philox block, in single precision.  Each word w_i gives a 24-bit uniform
// v_i from its top bits; the 32 low bits left over pick the move and the
// third Gaussian's angle.  The move (montecarlo_common.f90:25-79's three, with
// equal weights to 2^-16): a uniform direction by z = 1 - 2 v2, sin(theta) =
// sqrt(1 - z^2) and the azimuth 2 pi v1 (the distribution of acos / sin); the
// Gaussian step by Box-Muller, (g0, g1) from (v1, v2) and g2 from radius v0
// and a 16-bit angle; the flip.  acc = v3 is compared with expf.  The
// polynomial ln and sincos are the thermal field's.
__device__ __forceinline__ void trial_f32(const MCParams &mp, uint64_t sweep, uint32_t hs,
                                          int ens, float e0, float e1, float e2, float &n0,
                                          float &n1, float &n2, float &acc) {
  uint32_t w0, w1, w2, w3;
  counter_words(mp.seed, kTagMCU, sweep, hs, (uint32_t)ens, w0, w1, w2, w3);
  const float v0 = u24(w0), v1 = u24(w1), v2 = u24(w2);
  acc = u24(w3);
  const int ftype = (int)((3u * (((w0 & 0xffu) << 8) | (w1 & 0xffu))) >> 16);
  // The lanes of a wave choose their moves independently, so a wave runs
  // every move's arithmetic.  The uniform direction's sincos(2 pi v1) and
  // sqrt(1 - z^2) and the Gaussian pair's sincos(2 pi v2) and sqrt(-2 ln v1)
  // are therefore one sincos and one sqrt of a per-lane argument: each lane
  // computes exactly the values of its own move.
  const bool uni = ftype == 0;
  const float z = 1.0f - 2.0f * v2;
  float sA, cA;
  bm::sincos_q(4.0f * (uni ? v1 : v2), sA, cA);
  const float rA = sqrtf(uni ? fmaxf(0.0f, 1.0f - z * z) : -2.0f * bm::ln(v1));
  if (uni) {
    n0 = rA * cA;
    n1 = rA * sA;
    n2 = z;
  } else if (ftype == 1) {
    const float r2 = sqrtf(-2.0f * bm::ln(v0));
    const float a16 = ((float)(((w2 & 0xffu) << 8) | (w3 & 0xffu)) + 0.5f) * 1.52587890625e-05f;
    float s2, c2;
    bm::sincos_q(4.0f * a16, s2, c2);
    const float dl = (float)mp.delta[ens];
    const float a0 = e0 + (rA * cA) * dl, a1 = e1 + (rA * sA) * dl, a2 = e2 + (r2 * c2) * dl;
    const float il = 1.0f / sqrtf(a0 * a0 + a1 * a1 + a2 * a2);
    n0 = a0 * il;
    n1 = a1 * il;
    n2 = a2 * il;
  } else {
    n0 = 0.0f - e0;
    n1 = 0.0f - e1;
    n2 = 0.0f - e2;
  }
}

// One colour of a sweep.  HB selects heat bath (mp.heat_bath) or Metropolis
// at compile time, so each mode's registers are allocated for its own code
// alone.  A thread updates the U adjacent ensembles e0 .. e0+U-1 of one site
// of the colour (U = 1, or 2 when M is even): the trial moves, acceptances
// and heat-bath draws are those of U = 1, from the same counters, and the
// energy changes or local fields share the site's neighbour entries and
// couplings.
template <typename R, int G, int U, bool HB>
__global__ void __launch_bounds__(256) k_mc_colour(HamDev<R> h, MCParams mp, Quad<R> *s) {
  // (site of the colour, ensemble) pairs number fewer than 2^31
  const uint32_t per = (uint32_t)h.M / U;
  const uint32_t total = (uint32_t)mp.nsites * per;
  const uint32_t t = (uint32_t)xcd_block() * blockDim.x + threadIdx.x;
  if (t >= total) return;
  const int pidx = (int)fast_div(t, U == 1 ? h.mdiv : mp.pdiv);
  const int e0 = ((int)t - pidx * (int)per) * U;
  const int site = mp.sstride ? mp.site0 + pidx * mp.sstride : mp.sites[pidx];
  const size_t q0 = (size_t)site * h.M + e0;
  const uint64_t sweep = mp.sweep_dev ? *mp.sweep_dev + mp.sweep : mp.sweep;
  const uint32_t hs = host_site(h, site);
  Quad<R> me[U];
  R x[U][3];
#pragma unroll
  for (int k = 0; k < U; ++k) {
    me[k] = s[q0 + k];
    x[k][0] = (R)mp.ext[3 * (e0 + k)];
    x[k][1] = (R)mp.ext[3 * (e0 + k) + 1];
    x[k][2] = (R)mp.ext[3 * (e0 + k) + 2];
  }
  if constexpr (HB) {
    // the local field (effective_field_single) of each of the U ensembles
    R sf[U][3], qf[U][3];
    if (U == 1) {
      R y0, y1, y2;
      site_field<R, G>(h, (const Quad<R> *)s, site, e0, sf[0][0], sf[0][1], sf[0][2], qf[0][0],
                       qf[0][1], qf[0][2], y0, y1, y2);
    } else {
      site_field_u<R, G, U>(h, (const Quad<R> *)s, site, e0, sf, qf);
    }
#pragma unroll
    for (int k = 0; k < U; ++k) {
      double u0, u1, u2, u3;
      uniform4(mp.seed, kTagMCU, sweep, hs, (uint32_t)(e0 + k), u0, u1, u2, u3);
      const R f0 = (R)0 + sf[k][0] + (qf[k][0] + x[k][0] + (R)0);
      const R f1 = (R)0 + sf[k][1] + (qf[k][1] + x[k][1] + (R)0);
      const R f2 = (R)0 + sf[k][2] + (qf[k][2] + x[k][2] + (R)0);
      R o0, o1, o2;
      heat_bath<R>(mp.hbeta[e0 + k], me[k].w, f0, f1, f2, u1, u0, o0, o1, o2);
      s[q0 + k] = Quad<R>{o0, o1, o2, me[k].w};
    }
  } else {
    R tr[U][3];
    double acc[U];
#pragma unroll
    for (int k = 0; k < U; ++k) {
      const int ens = e0 + k;
      if (std::is_same<R, float>::value) {
        float n0, n1, n2, a;
        trial_f32(mp, sweep, hs, ens, (float)me[k].x, (float)me[k].y, (float)me[k].z, n0, n1, n2,
                  a);
        tr[k][0] = (R)n0;
        tr[k][1] = (R)n1;
        tr[k][2] = (R)n2;
        acc[k] = a;
      } else {
        double u0, u1, u2, u3;
        uniform4(mp.seed, kTagMCU, sweep, hs, (uint32_t)ens, u0, u1, u2, u3);
        double g0 = 0.0, g1 = 0.0, g2 = 0.0;
        if ((int)floor(3 * u0) == 1)
          normal3_f64(mp.seed, kTagMCG, sweep, hs, (uint32_t)ens, g0, g1, g2);
        trial_move<R>(me[k].x, me[k].y, me[k].z, mp.delta[ens], u0, u1, u2, g0, g1, g2, tr[k][0],
                      tr[k][1], tr[k][2]);
        acc[k] = u3;
      }
    }
    R de[U];
    mc_delta_e<R, G, U>(h, (const Quad<R> *)s, site, e0, me, tr, x, de);
#pragma unroll
    for (int k = 0; k < U; ++k) {
      const int ens = e0 + k;
      bool take;
      if (std::is_same<R, float>::value)
        take = de[k] <= (R)0 || (float)acc[k] < expf(-(float)mp.beta[ens] * (float)de[k]);
      else
        take = de[k] <= (R)0 || acc[k] < exp(-mp.beta[ens] * (double)de[k]);
      if (take) s[q0 + k] = Quad<R>{tr[k][0], tr[k][1], tr[k][2], me[k].w};
    }
  }
}

// ---------------------------------------------------------- FP64 reductions
// Stage 1: block b sums sites [b*chunk, (b+1)*chunk) of ensemble blockIdx.y:
// the three components of emomM = xyz*w (and the energy term when en != null).
template <typename R, int NV>
__global__ void __launch_bounds__(256)
    k_reduce_stage1(int N, int M, int chunk, const Quad<R> *__restrict__ s,
                    const double *__restrict__ en, double *__restrict__ partial) {
  const int ens = blockIdx.y;
  const int b = blockIdx.x;
  const int s0 = b * chunk, s1 = min(N, s0 + chunk);
  double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
  for (int i = s0 + (int)threadIdx.x; i < s1; i += blockDim.x) {
    const size_t q = (size_t)i * M + ens;
    const Quad<R> v = s[q];
    const double m = (double)v.w;
    a0 += (double)v.x * m;
    a1 += (double)v.y * m;
    a2 += (double)v.z * m;
    if (NV > 3) a3 += en[q];
  }
  __shared__ double red[4][256 / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    a0 += __shfl_down(a0, off, 64);
    a1 += __shfl_down(a1, off, 64);
    a2 += __shfl_down(a2, off, 64);
    if (NV > 3) a3 += __shfl_down(a3, off, 64);
  }
  if (lane == 0) {
    red[0][wid] = a0;
    red[1][wid] = a1;
    red[2][wid] = a2;
    red[3][wid] = a3;
  }
  __syncthreads();
  if (threadIdx.x < NV) {
    double x = 0.0;
    for (int w = 0; w < (int)(blockDim.x / 64); ++w) x += red[threadIdx.x][w];
    partial[((size_t)ens * gridDim.x + b) * NV + threadIdx.x] = x;
  }
}

// Stage 2: one thread per (ensemble, value) sums the block partials in order.
template <int NV>
__global__ void k_reduce_stage2(int M, int nblocks, const double *__restrict__ partial,
                                double *__restrict__ out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= M * NV) return;
  const int ens = t / NV, v = t - ens * NV;
  double x = 0.0;
  for (int b = 0; b < nblocks; ++b) x += partial[((size_t)ens * nblocks + b) * NV + v];
  out[t] = x;
}

// ------------------------------------------------- per-replica reductions
// The sites of a batch are split into segments (replicas); each segment is
// cut into chunks of consecutive sites, chunk c = sites [chunks[c].x,
// chunks[c].y).  partial[(c*M + ens)*NV + v] holds the FP64 sum of value v
// (emomM x, y, z and, for NV = 4, the per-q energy) over the chunk.
//
// Wide form, for M dividing the workgroup size: a workgroup reads the chunk's
// q range contiguously and thread t always meets ensemble t % M.
template <typename R, int NV>
__global__ void __launch_bounds__(256)
    k_segsum_wide(int M, const int2 *__restrict__ chunks, const Quad<R> *__restrict__ s,
                  const double *__restrict__ en, double *__restrict__ partial) {
  const int2 ch = chunks[blockIdx.x];
  const size_t q1 = (size_t)ch.y * M;
  double a[NV];
#pragma unroll
  for (int v = 0; v < NV; ++v) a[v] = 0.0;
  for (size_t q = (size_t)ch.x * M + threadIdx.x; q < q1; q += blockDim.x) {
    const Quad<R> x = s[q];
    const double m = (double)x.w;
    a[0] += (double)x.x * m;
    a[1] += (double)x.y * m;
    a[2] += (double)x.z * m;
    if (NV > 3) a[3] += en[q];
  }
  __shared__ double red[256][NV];
#pragma unroll
  for (int v = 0; v < NV; ++v) red[threadIdx.x][v] = a[v];
  __syncthreads();
  // M*NV partials: more than the workgroup's threads once M > blockDim/NV
  for (int t = (int)threadIdx.x; t < M * NV; t += (int)blockDim.x) {
    const int ens = t / NV, v = t - ens * NV;
    double x = 0.0;
    for (int j = ens; j < (int)blockDim.x; j += M) x += red[j][v];
    partial[((size_t)blockIdx.x * M + ens) * NV + v] = x;
  }
}

// General form: grid (chunks, M), one ensemble per workgroup.
template <typename R, int NV>
__global__ void __launch_bounds__(256)
    k_segsum_strided(int M, const int2 *__restrict__ chunks, const Quad<R> *__restrict__ s,
                     const double *__restrict__ en, double *__restrict__ partial) {
  const int2 ch = chunks[blockIdx.x];
  const int ens = blockIdx.y;
  double a[NV];
#pragma unroll
  for (int v = 0; v < NV; ++v) a[v] = 0.0;
  for (int i = ch.x + (int)threadIdx.x; i < ch.y; i += blockDim.x) {
    const size_t q = (size_t)i * M + ens;
    const Quad<R> x = s[q];
    const double m = (double)x.w;
    a[0] += (double)x.x * m;
    a[1] += (double)x.y * m;
    a[2] += (double)x.z * m;
    if (NV > 3) a[3] += en[q];
  }
  __shared__ double red[256][NV];
#pragma unroll
  for (int v = 0; v < NV; ++v) red[threadIdx.x][v] = a[v];
  __syncthreads();
  if ((int)threadIdx.x < NV) {
    double x = 0.0;
    for (int j = 0; j < (int)blockDim.x; ++j) x += red[j][threadIdx.x];
    partial[((size_t)blockIdx.x * M + ens) * NV + threadIdx.x] = x;
  }
}

// One thread per (segment, ensemble, value): the segment's chunk partials
// summed in chunk order into out[(seg*M + ens)*NV + v].
template <int NV>
__global__ void k_segsum_final(int nseg, int M, const int *__restrict__ seg_chunk,
                               const double *__restrict__ partial, double *__restrict__ out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nseg * M * NV) return;
  const int v = t % NV, se = t / NV, ens = se % M, seg = se / M;
  double x = 0.0;
  for (int c = seg_chunk[seg]; c < seg_chunk[seg + 1]; ++c)
    x += partial[((size_t)c * M + ens) * NV + v];
  out[t] = x;
}

// ------------------------------------------------- term-resolved energies
// The site terms of calc_energy (energy.f90:183-324, do_lsf 'N', scalar
// exchange): every field is built from zero and dotted with emomM by
// update_ene (energy.f90:559-571), -factor * m.b, with factor 0.5 for
// exchange, DM and anisotropy (type 7: uniaxial + sb * cubic field,
// energy.f90:311-316) and 1 for the external field.  Grid (chunks, M):
// workgroup (c, ens) sums the chunk's sites in FP64 into
// partial[(c*M + ens)*4 + t], t = exchange, anisotropy, DM, Zeeman;
// k_segsum_final<4> then adds the chunks of each segment.
template <typename R>
__device__ __forceinline__ double update_ene(R m0, R m1, R m2, R b0, R b1, R b2, double f) {
  return -f * ((double)m0 * (double)b0 + (double)m1 * (double)b1 + (double)m2 * (double)b2);
}

template <typename R>
__global__ void __launch_bounds__(256)
    k_energy_terms(HamDev<R> h, const int2 *__restrict__ chunks, const Quad<R> *__restrict__ src,
                   double *__restrict__ partial) {
  constexpr int G = 4;
  const int2 ch = chunks[blockIdx.x];
  const int ens = blockIdx.y, M = h.M;
  double a[4] = {0.0, 0.0, 0.0, 0.0};
  for (int site = ch.x + (int)threadIdx.x; site < ch.y; site += blockDim.x) {
    const size_t q = (size_t)site * M + ens;
    const SiteClass sc = h.classes[h.cls[site]];
    R m0, m1, m2;
    moment(src, q, m0, m1, m2);
    R b0 = (R)0, b1 = (R)0, b2 = (R)0;
    if (h.nbr16)
      exchange_sum<R, G, false>(h, h.nbr16 + site, src, site, ens, sc.ex_row, b0, b1, b2);
    else
      exchange_sum<R, G, false>(h, h.nbr + site, src, site, ens, sc.ex_row, b0, b1, b2);
    a[0] += update_ene(m0, m1, m2, b0, b1, b2, 0.5);
    if (sc.an_row >= 0) {
      const int t = h.an_type[sc.an_row];
      if (t == 1 || t == 2 || t == 7) {
        R u0 = (R)0, u1 = (R)0, u2 = (R)0;
        if (t == 1) {
          uniaxial(h, sc.an_row, m0, m1, m2, u0, u1, u2);
        } else if (t == 2) {
          cubic(h, sc.an_row, m0, m1, m2, u0, u1, u2);
        } else {
          R c0 = (R)0, c1 = (R)0, c2 = (R)0;
          uniaxial(h, sc.an_row, m0, m1, m2, u0, u1, u2);
          cubic(h, sc.an_row, m0, m1, m2, c0, c1, c2);
          const R sb = h.an_sb[sc.an_row];
          u0 = u0 + c0 * sb;
          u1 = u1 + c1 * sb;
          u2 = u2 + c2 * sb;
        }
        a[1] += update_ene(m0, m1, m2, u0, u1, u2, 0.5);
      }
    }
    if (h.Wd > 0 && sc.dm_row >= 0) {
      R d0 = (R)0, d1 = (R)0, d2 = (R)0;
      if (h.dnbr16)
        dm_sum<R, G>(h, h.dnbr16 + site, src, site, ens, sc.dm_row, d0, d1, d2);
      else
        dm_sum<R, G>(h, h.dnbr + site, src, site, ens, sc.dm_row, d0, d1, d2);
      a[2] += update_ene(m0, m1, m2, d0, d1, d2, 0.5);
    }
    if (h.ext_mode == 1) {
      a[3] += update_ene(m0, m1, m2, h.ext_u[3 * ens], h.ext_u[3 * ens + 1],
                         h.ext_u[3 * ens + 2], 1.0);
    } else if (h.ext_mode == 2) {
      const Quad<R> v = h.ext_q[q];
      a[3] += update_ene(m0, m1, m2, v.x, v.y, v.z, 1.0);
    }
  }
  __shared__ double red[256][4];
#pragma unroll
  for (int v = 0; v < 4; ++v) red[threadIdx.x][v] = a[v];
  __syncthreads();
  if ((int)threadIdx.x < 4) {
    double x = 0.0;
    for (int j = 0; j < (int)blockDim.x; ++j) x += red[j][threadIdx.x];
    partial[((size_t)blockIdx.x * M + ens) * 4 + threadIdx.x] = x;
  }
}

// Cumulant accumulators per (segment, ensemble), the running weighted means
// of calc_and_print_cumulant (prn_averages.f90:956-1000): sample n has weight
// n, mean_n = (mean_{n-1} * W_{n-1} + x_n * n) / (W_{n-1} + n).
// acc[8*t + ...] = {W, n, <m>, <m^2>, <m^4>, <e>, <e^2>, 0} with
// m = |sum emomM| / nsites and e the energy per site in mRy.
constexpr int kAccWidth = 8;
template <int W>
__global__ void k_accumulate(int nseg, int M, const int *__restrict__ seg_n,
                             const double *__restrict__ sums, int nv, double *__restrict__ acc) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nseg * M) return;
  const int seg = t / M;
  const double n = (double)seg_n[seg];
  const double *x = sums + (size_t)t * nv;
  double *a = acc + (size_t)t * W;
  const double me = sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]) / n;
  const double m2 = me * me, m4 = m2 * m2;
  const double tot = a[0], w = a[1] + 1.0;
  a[2] = (a[2] * tot + me * w) / (tot + w);
  a[3] = (a[3] * tot + m2 * w) / (tot + w);
  a[4] = (a[4] * tot + m4 * w) / (tot + w);
  if (nv > 3) {
    const double e = x[3] * (kMub / kMry) / n;
    a[5] = (a[5] * tot + e * w) / (tot + w);
    a[6] = (a[6] * tot + e * e * w) / (tot + w);
  }
  a[0] = tot + w;
  a[1] = w;
}

// Graph support: the step / sweep counters live in device memory while a
// captured graph runs; one thread adds or sets them (a vector global store).
template <int I>
__global__ void k_counter(uint64_t *ctr, uint64_t value, int add) {
  if (blockIdx.x == 0 && threadIdx.x == 0) ctr[I] = add ? ctr[I] + value : value;
}

}  // namespace uasd
