/*
 * or_mc.c -- Metropolis / heat-bath Monte Carlo, restated from
 *   source/MonteCarlo/montecarlo.f90        mc_evolve 44-273, choose_random_atom_x 277-301
 *   source/MonteCarlo/montecarlo_common.f90 choose_random_flip 25-79, flip_a 139-267,
 *                                           flip_h 371-422, calculate_energy 431-865
 *   source/Hamiltonian/hamiltonianactions.f90 effective_field_single 254-392
 * or_mc_sweep_ref keeps the reference's visiting order (random permutation,
 * one site at a time) and its MT19937 / ziggurat stream consumption.
 * or_mc_sweep_colored applies the same per-site move, energy and acceptance
 * rules to a colouring of the bond graph (no two bonded sites share a
 * colour) with Philox counter-based random numbers -- the schedule the HIP
 * kernels run, restated for the CPU.
 */
#include "uasd_oracle.h"
#include <math.h>
#include <stddef.h>
#include <stdlib.h>

#define V3(a, c, i, k, N) (a)[(c) + 3 * ((size_t)(i) + (size_t)(N) * (k))]
#define S2(a, i, k, N) (a)[(size_t)(i) + (size_t)(N) * (k)]

/* counter tags shared with the HIP kernels */
#define TAG_MC_U 0x21u
#define TAG_MC_G 0x22u

double or_mc_delta(double temperature)                  /* montecarlo.f90:141 */
{
  return (double)(2.0f / 25.0f) * pow(OR_K_BOLT * temperature / OR_MUB, 0.20);
}

void or_choose_random_flip(const double *e, double delta, const double rn[3],
                           const double gn[3], double newmom[3])
{
  int ftype = (int)floor(3 * rn[0]);
  if (ftype == 0) {
    double phi = rn[1] * 2 * OR_PI;
    double theta = acos(1 - 2 * rn[2]);
    newmom[0] = sin(theta) * cos(phi);
    newmom[1] = sin(theta) * sin(phi);
    newmom[2] = cos(theta);
  } else if (ftype == 1) {
    double g0 = gn[0] * delta, g1 = gn[1] * delta, g2 = gn[2] * delta;
    double a0 = e[0] + g0, a1 = e[1] + g1, a2 = e[2] + g2;
    double l = sqrt(a0 * a0 + a1 * a1 + a2 * a2);
    newmom[0] = (e[0] + g0) / l;
    newmom[1] = (e[1] + g1) / l;
    newmom[2] = (e[2] + g2) / l;
  } else {
    newmom[0] = -e[0];
    newmom[1] = -e[1];
    newmom[2] = -e[2];
  }
}

static inline double sq(double x) { return x * x; }

double or_mc_energy_change(const or_ham *h, int N, const double *emomM,
                           const double *emom, const double *mmom, int iflip, int k,
                           const double newmom[3], const double ext[3])
{
  const int i = iflip;
  double e_c = 0.0, e_t = 0.0;
  double m = S2(mmom, i, k, N);
  double t[3] = {newmom[0] * m, newmom[1] * m, newmom[2] * m};
  double c0 = V3(emomM, 0, i, k, N), c1 = V3(emomM, 1, i, k, N), c2 = V3(emomM, 2, i, k, N);
  const int mnn = h->max_no_neigh;
  for (int j = 0; j < h->nlistsize[i]; ++j) {
    double J = h->ncoup[j + (size_t)mnn * i];
    int nb = h->nlist[j + (size_t)mnn * i] - 1;
    double n0 = V3(emomM, 0, nb, k, N), n1 = V3(emomM, 1, nb, k, N), n2 = V3(emomM, 2, nb, k, N);
    e_c = e_c - J * (c0 * n0 + c1 * n1 + c2 * n2);
    e_t = e_t - J * (t[0] * n0 + t[1] * n1 + t[2] * n2);
  }
  if (h->do_anisotropy) {
    int ty = h->taniso[i];
    const double *ea = &h->eaniso[3 * (size_t)i];
    const double *K = &h->kaniso[2 * (size_t)i];
    if (ty == 1) {
      double tta = c0 * ea[0] + c1 * ea[1] + c2 * ea[2];
      double ttb = t[0] * ea[0] + t[1] * ea[1] + t[2] * ea[2];
      e_c = e_c + K[0] * sq(tta) + K[1] * sq(sq(tta));
      e_t = e_t + K[0] * sq(ttb) + K[1] * sq(sq(ttb));
    } else if (ty == 2) {
      e_c = e_c - K[0] * (sq(c0) * sq(c1) + sq(c1) * sq(c2) + sq(c2) * sq(c0)) -
            K[1] * (sq(c0) * sq(c1) * sq(c2));
      e_t = e_t - K[0] * (sq(t[0]) * sq(t[1]) + sq(t[1]) * sq(t[2]) + sq(t[2]) * sq(t[0])) -
            K[1] * (sq(t[0]) * sq(t[1]) * sq(t[2]));
    }
    if (ty == 7) {
      double tta = (c0 * ea[0] + c1 * ea[1] + c2 * ea[2]);
      double ttb = (t[0] * ea[0] + t[1] * ea[1] + t[2] * ea[2]);
      e_c = e_c + K[0] * sq(tta) + K[1] * sq(sq(tta));
      e_t = e_t + K[0] * sq(ttb) + K[1] * sq(sq(ttb));
      double aw1 = K[0] * h->sb[i], aw2 = K[1] * h->sb[i];
      e_c = e_c + aw1 * (sq(c0) * sq(c1) + sq(c1) * sq(c2) + sq(c2) * sq(c0)) +
            aw2 * (sq(c0) * sq(c1) * sq(c2));
      e_t = e_t + aw1 * (sq(t[0]) * sq(t[1]) + sq(t[1]) * sq(t[2]) + sq(t[2]) * sq(t[0])) +
            aw2 * (sq(t[0]) * sq(t[1]) * sq(t[2]));
    }
  }
  if (h->do_dm) {
    const int mdm = h->max_no_dmneigh;
    double u0 = V3(emom, 0, i, k, N), u2 = V3(emom, 2, i, k, N);
    for (int j = 0; j < h->dmlistsize[i]; ++j) {
      const double *D = &h->dm_vect[3 * (j + (size_t)mdm * i)];
      int nb = h->dmlist[j + (size_t)mdm * i] - 1;
      double n0 = V3(emomM, 0, nb, k, N), n1 = V3(emomM, 1, nb, k, N), n2 = V3(emomM, 2, nb, k, N);
      /* the current-state term mixes emom and emomM (montecarlo_common.f90:612,615) */
      e_c = e_c - D[0] * (c1 * n2 - u2 * n1) - D[1] * (c2 * n0 - c0 * n2) -
            D[2] * (u0 * n1 - c1 * n0);
      e_t = e_t - D[0] * (t[1] * n2 - t[2] * n1) - D[1] * (t[2] * n0 - t[0] * n2) -
            D[2] * (t[0] * n1 - t[1] * n0);
    }
  }
  e_c = e_c - ext[0] * c0 - ext[1] * c1 - ext[2] * c2;
  e_t = e_t - ext[0] * t[0] - ext[1] * t[1] - ext[2] * t[2];
  double tt = e_t - e_c;
  return OR_MUB * tt;
}

void or_flip_h(double *e_out, double mmom, double temperature, double temprescale,
               double flipprob, const double totfield[3], double q)
{
  double beta = 1.0 / OR_K_BOLT / (temprescale * temperature);
  double zfc[3];
  for (int c = 0; c < 3; ++c) zfc[c] = beta * totfield[c] * OR_MUB * mmom;
  double zarg = sqrt(zfc[0] * zfc[0] + zfc[1] * zfc[1] + zfc[2] * zfc[2]);
  double zctheta = zfc[2] / zarg;
  double zstheta = sqrt(1.0 - zctheta * zctheta) + OR_DBL_TOLERANCE;
  double zcphi = zfc[0] / (zarg * zstheta);
  double zsphi = zfc[1] / (zarg * zstheta);
  double ctheta = 1.0 + (1.0 / zarg) *
                  log((1.0 - exp(-2.0 * zarg)) * q + exp(-2.0 * zarg) + OR_DBL_TOLERANCE);
  double stheta = sqrt(1.0 - ctheta * ctheta);
  double phi = OR_PI * (2.0 * flipprob - 1.0);
  double s0 = stheta * cos(phi), s1 = stheta * sin(phi), s2 = ctheta;
  e_out[0] = zcphi * zctheta * s0 - zsphi * s1 + zcphi * zstheta * s2;
  e_out[1] = zsphi * zctheta * s0 + zcphi * s1 + zsphi * zstheta * s2;
  e_out[2] = -zstheta * s0 + zctheta * s2;
}

void or_choose_random_atom_x(or_mt *st, int N, int *iflip_a)
{
  double *dshift = malloc(sizeof(double) * (size_t)N);
  for (int i = 0; i < N; ++i) iflip_a[i] = i + 1;
  for (int i = 0; i < N; ++i) dshift[i] = or_mt_real2(st);
  for (int i = 0; i < N; ++i) {
    int ishift = (int)(dshift[i] * N);
    int itmp = iflip_a[i];
    iflip_a[i] = iflip_a[ishift % N];
    iflip_a[ishift % N] = itmp;
  }
  free(dshift);
}

static inline void metropolis_accept(int N, double *emom, double *emomM, const double *mmom,
                                     int i, int k, const double newmom[3], double de,
                                     double temperature, double temprescale, double flipprob)
{
  double beta = 1.0 / OR_K_BOLT / (temprescale * temperature + 1.0e-15);
  if (de <= 0.0 || flipprob < exp(-beta * de)) {
    double m = S2(mmom, i, k, N);
    for (int c = 0; c < 3; ++c) {
      V3(emom, c, i, k, N) = newmom[c];
      V3(emomM, c, i, k, N) = m * newmom[c];
    }
  }
}

void or_mc_sweep_ref(const or_ham *h, int N, int M, double temperature,
                     double temprescale, char mode, or_mt *st_c, or_zig *zig,
                     const int *iflip_a, double *emomM, double *emom,
                     const double *mmom, const double extfield[3])
{
  size_t n3 = (size_t)3 * N * M, n1 = (size_t)N * M;
  double *fm = malloc(sizeof(double) * n3), *fg = malloc(sizeof(double) * n3);
  double *newm = malloc(sizeof(double) * n3);
  double *mflip = malloc(sizeof(double) * n1), *fa = malloc(sizeof(double) * n1);
  double delta = or_mc_delta(temperature);
  for (size_t x = 0; x < n3; ++x) fm[x] = or_mt_real2(st_c);
  for (size_t x = 0; x < n3; ++x) fg[x] = or_zig_nor(zig);
  for (int i = 0; i < N; ++i)
    for (int k = 0; k < M; ++k) {
      double e[3] = {V3(emom, 0, i, k, N), V3(emom, 1, i, k, N), V3(emom, 2, i, k, N)};
      or_choose_random_flip(e, delta, &V3(fm, 0, i, k, N), &V3(fg, 0, i, k, N),
                            &V3(newm, 0, i, k, N));
    }
  if (mode == 'H')
    for (size_t x = 0; x < n1; ++x) mflip[x] = or_mt_real2(st_c);
  for (size_t x = 0; x < n1; ++x) fa[x] = or_mt_real2(st_c);
  for (int ii = 0; ii < N; ++ii)
    for (int k = 0; k < M; ++k) {
      int i = iflip_a[ii] - 1;
      double loc[3] = {extfield[0] + 0.0 + 0.0, extfield[1] + 0.0 + 0.0, extfield[2] + 0.0 + 0.0};
      if (mode == 'H') {
        double tot[3], e[3];
        or_effective_field_single(h, N, emomM, NULL, i, k, tot);
        double m = S2(mmom, i, k, N);
        or_flip_h(e, m, temperature, temprescale, S2(fa, ii, k, N), tot, S2(mflip, ii, k, N));
        for (int c = 0; c < 3; ++c) {
          V3(emom, c, i, k, N) = e[c];
          V3(emomM, c, i, k, N) = m * e[c];
        }
      } else {
        double de = or_mc_energy_change(h, N, emomM, emom, mmom, i, k, &V3(newm, 0, i, k, N), loc);
        metropolis_accept(N, emom, emomM, mmom, i, k, &V3(newm, 0, i, k, N), de, temperature,
                          temprescale, S2(fa, ii, k, N));
      }
    }
  free(fm); free(fg); free(newm); free(mflip); free(fa);
}

/* The Metropolis trial move of the single-precision MC engine, from one
 * Philox block w (tag TAG_MC_U): choose_random_flip's three moves
 * (montecarlo_common.f90:25-79) drawn in single precision.  v_i = or_u24(w_i)
 * (top 24 bits); the move is (3 b) >> 16 of b = (w0 & 0xff) << 8 | (w1 & 0xff),
 * equal weights to 2^-16; a uniform direction from z = 1 - 2 v2, sin(theta)
 * = sqrt(1 - z^2) and the azimuth 2 pi v1 by the polynomial sincos; the
 * Gaussian step from Box-Muller with the polynomial ln and sincos: (g0, g1)
 * = sqrt(-2 ln v1) (cos, sin)(2 pi v2), g2 = sqrt(-2 ln v0) cos(2 pi a) with
 * a the 16-bit angle ((w2 & 0xff) << 8 | (w3 & 0xff) + 1/2) 2^-16; the flip.
 * Widened to double.  The engine accepts by v3 = or_u24(w3). */
void or_trial_move_f32(const double e[3], double delta, const uint32_t w[4], double out[3])
{
  const float v0 = or_u24(w[0]), v1 = or_u24(w[1]), v2 = or_u24(w[2]);
  const int ftype = (int)((3u * (((w[0] & 0xffu) << 8) | (w[1] & 0xffu))) >> 16);
  if (ftype == 0) {
    const float z = 1.0f - 2.0f * v2;
    const float st = sqrtf(fmaxf(0.0f, 1.0f - z * z));
    float sc[2];
    or_bm_sincos(4.0f * v1, sc);
    out[0] = (double)(st * sc[1]);
    out[1] = (double)(st * sc[0]);
    out[2] = (double)z;
  } else if (ftype == 1) {
    const float r01 = sqrtf(-2.0f * or_bm_ln(v1)), r2 = sqrtf(-2.0f * or_bm_ln(v0));
    const float a16 = ((float)(((w[2] & 0xffu) << 8) | (w[3] & 0xffu)) + 0.5f) *
                      1.52587890625e-05f;
    float sc01[2], sc2[2];
    or_bm_sincos(4.0f * v2, sc01);
    or_bm_sincos(4.0f * a16, sc2);
    const float dl = (float)delta;
    const float a0 = (float)e[0] + (r01 * sc01[1]) * dl, a1 = (float)e[1] + (r01 * sc01[0]) * dl,
                a2 = (float)e[2] + (r2 * sc2[1]) * dl;
    const float il = 1.0f / sqrtf(a0 * a0 + a1 * a1 + a2 * a2);
    out[0] = (double)(a0 * il);
    out[1] = (double)(a1 * il);
    out[2] = (double)(a2 * il);
  } else {
    out[0] = (double)(0.0f - (float)e[0]);
    out[1] = (double)(0.0f - (float)e[1]);
    out[2] = (double)(0.0f - (float)e[2]);
  }
}

void or_mc_sweep_colored(const or_ham *h, int N, int M, const double *temps, char mode,
                         uint64_t seed, uint64_t sweep, int ncolors, const int *color_ptr,
                         const int *color_sites, double *emomM, double *emom,
                         const double *mmom, const double extfield[3])
{
  or_mc_sweep_colored_ex(h, N, M, temps, mode, seed, sweep, ncolors, color_ptr, color_sites,
                         emomM, emom, mmom, extfield, 0);
}

/* f32_moves = 1: Metropolis trial moves as the single-precision engine draws
 * them (or_trial_move_f32); energies and acceptance stay in double. */
void or_mc_sweep_colored_ex(const or_ham *h, int N, int M, const double *temps, char mode,
                            uint64_t seed, uint64_t sweep, int ncolors, const int *color_ptr,
                            const int *color_sites, double *emomM, double *emom,
                            const double *mmom, const double extfield[3], int f32_moves)
{
  for (int col = 0; col < ncolors; ++col)
    for (int p = color_ptr[col]; p < color_ptr[col + 1]; ++p) {
      int i = color_sites[p];
      for (int k = 0; k < M; ++k) {
        double T = temps[k];
        double u[4];
        or_uniform4(seed, TAG_MC_U, sweep, (uint32_t)i, (uint32_t)k, u);
        double m = S2(mmom, i, k, N);
        if (mode == 'H') {
          double tot[3], e[3];
          or_effective_field_single(h, N, emomM, extfield, i, k, tot);
          or_flip_h(e, m, T, 1.0, u[1], tot, u[0]);
          for (int c = 0; c < 3; ++c) {
            V3(emom, c, i, k, N) = e[c];
            V3(emomM, c, i, k, N) = m * e[c];
          }
        } else {
          double e[3] = {V3(emom, 0, i, k, N), V3(emom, 1, i, k, N), V3(emom, 2, i, k, N)};
          double nm[3];
          double uacc = u[3];
          if (f32_moves) {
            uint32_t w[4];
            or_counter_words(seed, TAG_MC_U, sweep, (uint32_t)i, (uint32_t)k, w);
            or_trial_move_f32(e, or_mc_delta(T), w, nm);
            uacc = (double)or_u24(w[3]);
          } else {
            double g[3] = {0.0, 0.0, 0.0};
            if ((int)floor(3 * u[0]) == 1)
              or_noise3_f64(seed, TAG_MC_G, sweep, (uint32_t)i, (uint32_t)k, g);
            double rn[3] = {u[0], u[1], u[2]};
            or_choose_random_flip(e, or_mc_delta(T), rn, g, nm);
          }
          double de = or_mc_energy_change(h, N, emomM, emom, mmom, i, k, nm, extfield);
          metropolis_accept(N, emom, emomM, mmom, i, k, nm, de, T, 1.0, uacc);
        }
      }
    }
}

void or_sum_moments(int N, int M, const double *emomM, double *msum)
{
  for (int k = 0; k < M; ++k) {
    double s[3] = {0.0, 0.0, 0.0};
    for (int i = 0; i < N; ++i)
      for (int c = 0; c < 3; ++c) s[c] = s[c] + V3(emomM, c, i, k, N);
    for (int c = 0; c < 3; ++c) msum[3 * k + c] = s[c];
  }
}

double or_calc_mavrg(int N, int M, const double *emomM)   /* measurements.f90:211-235 */
{
  double tot = 0.0;
  for (int k = 0; k < M; ++k) {
    double s[3] = {0.0, 0.0, 0.0};
    for (int i = 0; i < N; ++i)
      for (int c = 0; c < 3; ++c) s[c] = s[c] + V3(emomM, c, i, k, N);
    tot = tot + pow(s[0] * s[0] + s[1] * s[1] + s[2] * s[2], 0.5);
  }
  return tot / M / N;
}
