/*
 * or_llg.c -- stochastic LLG integrators, restated from
 *   source/Evolution/depondt.f90      depondt_evolve_first 50-193, _second 203-333
 *   source/Evolution/midpoint.f90     smodeulermpt 34-181, modeulermpf 188-323
 *   source/RNG/randomnumbers.f90      rannum scaling 625-771 (llg = 1 branch)
 *   source/Evolution/heun_proper.f90  evolve4p 28-161, heun4p 261-369,
 *                                     evolve4f 168-254, heun4f 376-530
 *   source/Evolution/sphmidpoint.f90  sibt 36-262, sibf 270-477 (SDEalgh 6),
 *                                     sispht 485-713, sisphf 721-923 (SDEalgh 7)
 *   source/Evolution/updatemoments.f90 copym 48-95 (mompar = 0, initexc /= 'I')
 * Spin-transfer torques are off (stt='N', do_she='N', do_sot='N'), so the
 * torque accumulators are the literal 0.0 they start from.
 */
#include "uasd_oracle.h"
#include <math.h>
#include <stddef.h>
#include <stdlib.h>

#define V3(a, c, i, k, N) (a)[(c) + 3 * ((size_t)(i) + (size_t)(N) * (k))]
#define S2(a, i, k, N) (a)[(size_t)(i) + (size_t)(N) * (k)]

static inline void rodrigues(const double bd[3], double dt, double lam, const double e[3],
                             double out[3])
{
  double lldamp = 1.0 / (1.0 + lam * lam);
  double Bnorm = bd[0] * bd[0] + bd[1] * bd[1] + bd[2] * bd[2];
  Bnorm = sqrt(Bnorm) + 1.0e-15;
  double hx = bd[0] / Bnorm, hy = bd[1] / Bnorm, hz = bd[2] / Bnorm;
  double v = Bnorm * dt * OR_GAMA * lldamp;
  double cosv = cos(v), sinv = sin(v), u = 1.0 - cosv;
  out[0] = hx * hx * u * e[0] + cosv * e[0] + hx * hy * u * e[1] - hz * sinv * e[1] +
           hx * hz * u * e[2] + hy * sinv * e[2];
  out[1] = hy * hx * u * e[0] + hz * sinv * e[0] + hy * hy * u * e[1] + cosv * e[1] +
           hy * hz * u * e[2] - hx * sinv * e[2];
  out[2] = hx * hz * u * e[0] - hy * sinv * e[0] + hz * hy * u * e[1] + hx * sinv * e[1] +
           hz * hz * u * e[2] + cosv * e[2];
}

/* btherm holds N(0,1) deviates on entry (rng_gaussianP, depondt.f90:133) and
 * the scaled thermal field on exit (depondt.f90:143-145). */
void or_depondt_first(int N, int M, const double *lambda, const double *temp,
                      double temprescale, double dt, const double *beff,
                      double *btherm, const double *mmom, double *emom,
                      double *emom2, double *emomM, double *b2eff,
                      double *thermal_field)
{
#pragma omp parallel for collapse(2) schedule(static) if (OR_PAR(N, M))
  for (int k = 0; k < M; ++k)
    for (int i = 0; i < N; ++i) {
      double lam = lambda[i];
      double Dp = (2.0 * lam * OR_K_BOLT) / (dt * OR_GAMA * OR_MUB);
      double sigma = sqrt(Dp * temprescale * temp[i] / S2(mmom, i, k, N));
      double bl[3], bd[3], e[3], mr[3];
      for (int c = 0; c < 3; ++c) {
        V3(btherm, c, i, k, N) = V3(btherm, c, i, k, N) * sigma;
        bl[c] = V3(beff, c, i, k, N) + V3(btherm, c, i, k, N);
        if (thermal_field) V3(thermal_field, c, i, k, N) = V3(btherm, c, i, k, N);
        e[c] = V3(emom, c, i, k, N);
      }
      bd[0] = 0.0 + bl[0] + lam * e[1] * bl[2] - lam * e[2] * bl[1];
      bd[1] = 0.0 + bl[1] + lam * e[2] * bl[0] - lam * e[0] * bl[2];
      bd[2] = 0.0 + bl[2] + lam * e[0] * bl[1] - lam * e[1] * bl[0];
      rodrigues(bd, dt, lam, e, mr);
      double m = S2(mmom, i, k, N);
      for (int c = 0; c < 3; ++c) {
        V3(emom2, c, i, k, N) = e[c];
        V3(emomM, c, i, k, N) = mr[c] * m;
        V3(emom, c, i, k, N) = mr[c];
        V3(b2eff, c, i, k, N) = bd[c];
      }
    }
}

void or_depondt_second(int N, int M, const double *lambda, double dt,
                       const double *beff, const double *btherm,
                       const double *b2eff, double *emom, double *emom2)
{
#pragma omp parallel for collapse(2) schedule(static) if (OR_PAR(N, M))
  for (int k = 0; k < M; ++k)
    for (int i = 0; i < N; ++i) {
      double lam = lambda[i];
      double bl[3], bd[3], e[3], mr[3];
      for (int c = 0; c < 3; ++c) {
        bl[c] = V3(beff, c, i, k, N) + V3(btherm, c, i, k, N);
        e[c] = V3(emom, c, i, k, N);
      }
      bd[0] = 0.0 + bl[0] + lam * e[1] * bl[2] - lam * e[2] * bl[1];
      bd[1] = 0.0 + bl[1] + lam * e[2] * bl[0] - lam * e[0] * bl[2];
      bd[2] = 0.0 + bl[2] + lam * e[0] * bl[1] - lam * e[1] * bl[0];
      for (int c = 0; c < 3; ++c) bd[c] = 0.5 * bd[c] + 0.5 * V3(b2eff, c, i, k, N);
      for (int c = 0; c < 3; ++c) {
        e[c] = V3(emom2, c, i, k, N);
        V3(emom, c, i, k, N) = e[c];
      }
      rodrigues(bd, dt, lam, e, mr);
      for (int c = 0; c < 3; ++c) V3(emom2, c, i, k, N) = mr[c];
    }
}

/* rannum, llg = 1: Dk = (lambda/(1+lambda**2)*k_bolt/gama/mub)*(gama/bn),
 * D = Dk*mmomi*T*temprescale, sigma = sqrt(2D). ranv holds N(0,1) on entry. */
void or_rannum_scale(int N, int M, int llg, const double *lambda, double bn,
                     const double *mmomi, const double *temp, double temprescale,
                     double *ranv)
{
#pragma omp parallel for collapse(2) schedule(static) if (OR_PAR(N, M))
  for (int k = 0; k < M; ++k)
    for (int i = 0; i < N; ++i) {
      double lam = lambda[i], Dk;
      if (llg == 1)
        Dk = (lam / (1.0 + lam * lam) * OR_K_BOLT / OR_GAMA / (OR_MUB)) * (OR_GAMA / bn);
      else
        Dk = (lam * OR_K_BOLT / OR_GAMA / (OR_MUB)) * (OR_GAMA / bn);
      double D = Dk * S2(mmomi, i, k, N) * temp[i] * temprescale;
      double sigma = sqrt(2.0 * D);
      for (int c = 0; c < 3; ++c) V3(ranv, c, i, k, N) = V3(ranv, c, i, k, N) * sigma;
    }
}

static inline void midpoint_solve(const double A[3], const double e[3], double et[3],
                                  double *detAi)
{
  *detAi = 1.0 / (1.0 + (A[0] * A[0] + A[1] * A[1] + A[2] * A[2]));
  double a2[3];
  a2[0] = e[0] + e[1] * A[2] - e[2] * A[1];
  a2[1] = e[1] + e[2] * A[0] - e[0] * A[2];
  a2[2] = e[2] + e[0] * A[1] - e[1] * A[0];
  et[0] = a2[0] * (1 + A[0] * A[0]) + a2[1] * (A[0] * A[1] + A[2]) + a2[2] * (A[0] * A[2] - A[1]);
  et[1] = a2[0] * (A[1] * A[0] - A[2]) + a2[1] * (1 + A[1] * A[1]) + a2[2] * (A[1] * A[2] + A[0]);
  et[2] = a2[0] * (A[2] * A[0] + A[1]) + a2[1] * (A[2] * A[1] - A[0]) + a2[2] * (1 + A[2] * A[2]);
}

static inline void midpoint_A(const double *landeg, double bn, const double *lambda,
                              double deltat, int i, const double b[3], const double r[3],
                              const double ep[3], double A[3], double s1out[3])
{
  double lam = lambda[i];
  double lldamp = 1.0 / (1.0 + lam * lam);
  double dt = deltat * bn * OR_GAMA * lldamp;
  double sqrtdt = sqrt(dt);
  double dtg = dt * landeg[i];
  double sqrtdtg = sqrtdt * landeg[i];
  double a1[3], s1[3];
  a1[0] = -0.0 - b[0] - lam * (ep[1] * b[2] - ep[2] * b[1]);
  a1[1] = -0.0 - b[1] - lam * (ep[2] * b[0] - ep[0] * b[2]);
  a1[2] = -0.0 - b[2] - lam * (ep[0] * b[1] - ep[1] * b[0]);
  s1[0] = -r[0] - lam * (ep[1] * r[2] - ep[2] * r[1]);
  s1[1] = -r[1] - lam * (ep[2] * r[0] - ep[0] * r[2]);
  s1[2] = -r[2] - lam * (ep[0] * r[1] - ep[1] * r[0]);
  for (int c = 0; c < 3; ++c) A[c] = 0.5 * dtg * a1[c] + 0.5 * sqrtdtg * s1[c];
  if (s1out)
    for (int c = 0; c < 3; ++c) s1out[c] = s1[c];
}

void or_midpoint_first(int N, int M, const double *landeg, double bn,
                       const double *lambda, const double *beff, const double *emom,
                       double *emom2, double *emomM, const double *mmom,
                       double deltat, const double *ranv, double *thermal_field)
{
#pragma omp parallel for collapse(2) schedule(static) if (OR_PAR(N, M))
  for (int i = 0; i < N; ++i)
    for (int k = 0; k < M; ++k) {
      double b[3], r[3], e[3], A[3], s1[3], et[3], detAi;
      for (int c = 0; c < 3; ++c) {
        b[c] = V3(beff, c, i, k, N);
        r[c] = V3(ranv, c, i, k, N);
        e[c] = V3(emom, c, i, k, N);
      }
      midpoint_A(landeg, bn, lambda, deltat, i, b, r, e, A, s1);
      if (thermal_field)
        for (int c = 0; c < 3; ++c) V3(thermal_field, c, i, k, N) = s1[c];
      midpoint_solve(A, e, et, &detAi);
      for (int c = 0; c < 3; ++c) et[c] = 0.5 * (e[c] + et[c] * detAi);
      double m = S2(mmom, i, k, N);
      for (int c = 0; c < 3; ++c) {
        V3(emom2, c, i, k, N) = et[c];
        V3(emomM, c, i, k, N) = et[c] * m;
      }
    }
}

void or_midpoint_second(int N, int M, const double *landeg, double bn,
                        const double *lambda, const double *beff, const double *emom,
                        double *emom2, double deltat, const double *ranv)
{
#pragma omp parallel for collapse(2) schedule(static) if (OR_PAR(N, M))
  for (int i = 0; i < N; ++i)
    for (int k = 0; k < M; ++k) {
      double b[3], r[3], e[3], ep[3], A[3], et[3], detAi;
      for (int c = 0; c < 3; ++c) {
        b[c] = V3(beff, c, i, k, N);
        r[c] = V3(ranv, c, i, k, N);
        e[c] = V3(emom, c, i, k, N);
        ep[c] = V3(emom2, c, i, k, N);
      }
      midpoint_A(landeg, bn, lambda, deltat, i, b, r, ep, A, NULL);
      midpoint_solve(A, e, et, &detAi);
      for (int c = 0; c < 3; ++c) V3(emom2, c, i, k, N) = et[c] * detAi;
    }
}

void or_moment_update(int N, int M, const double *mmom, double *emom,
                      const double *emom2, double *emomM, double *mmomi)
{
#pragma omp parallel for collapse(2) schedule(static) if (OR_PAR(N, M))
  for (int k = 0; k < M; ++k)
    for (int i = 0; i < N; ++i) {
      double m = S2(mmom, i, k, N);
      for (int c = 0; c < 3; ++c) {
        V3(emom, c, i, k, N) = V3(emom2, c, i, k, N);
        V3(emomM, c, i, k, N) = V3(emom2, c, i, k, N) * m;
      }
      if (mmomi) S2(mmomi, i, k, N) = 1.0 / m;
    }
}

/* Heun predictor-corrector (SDEalgh 4).  evolve4p (heun_proper.f90:28-161):
 * ranv (N(0,1) on entry) is scaled in place by sigma = sqrt(dt*2*D),
 * D = lam/(1+lam^2) kB/muB / mmom * T * temprescale, dt = delta_t*gama
 * (compensate_drift 0); heun4p (261-369) starts from emom2 (the current
 * state), keeps b2eff = beff and writes the Euler predictor into emom and
 * emomM = emom*mmom.  The noise enters through dW = ranv*Landeg. */
void or_heun_first(int N, int M, const double *landeg, const double *lambda, const double *temp,
                   double temprescale, double delta_t, const double *beff, double *ranv,
                   const double *mmom, double *emom, const double *emom2, double *emomM,
                   double *b2eff)
{
  const double dt = delta_t * OR_GAMA;
  for (int k = 0; k < M; ++k)
    for (int i = 0; i < N; ++i) {
      double Dk = lambda[i] / (1 + lambda[i] * lambda[i]) * OR_K_BOLT / OR_MUB;
      double D = Dk * (1.0 / S2(mmom, i, k, N)) * temp[i] * temprescale;
      double sigma = sqrt(dt * 2 * D);
      for (int c = 0; c < 3; ++c) V3(ranv, c, i, k, N) = V3(ranv, c, i, k, N) * sigma;
    }
  for (int i = 0; i < N; ++i)
    for (int k = 0; k < M; ++k) {
      const double lam = lambda[i], lg = landeg[i];
      for (int c = 0; c < 3; ++c) V3(b2eff, c, i, k, N) = V3(beff, c, i, k, N);
      double e1x = V3(emom2, 0, i, k, N), e1y = V3(emom2, 1, i, k, N), e1z = V3(emom2, 2, i, k, N);
      double lldamp = 1.0 / (1.0 + lam * lam);
      double dtg = dt * lg * lldamp;
      double dwx = V3(ranv, 0, i, k, N) * lg, dwy = V3(ranv, 1, i, k, N) * lg,
             dwz = V3(ranv, 2, i, k, N) * lg;
      double hesx = V3(beff, 0, i, k, N), hesy = V3(beff, 1, i, k, N), hesz = V3(beff, 2, i, k, N);
      double hlpx = lam * hesx, hlpy = lam * hesy, hlpz = lam * hesz;
      double prot = e1x * hlpx + e1y * hlpy + e1z * hlpz;
      /* the spin-torque field is 0: (hes + 0.0) */
      double a1x = hlpx + e1z * (hesy + 0.0) - e1y * (hesz + 0.0) - e1x * prot;
      double a1y = hlpy - e1z * (hesx + 0.0) + e1x * (hesz + 0.0) - e1y * prot;
      double a1z = hlpz + e1y * (hesx + 0.0) - e1x * (hesy + 0.0) - e1z * prot;
      double b1xx = lam - e1x * e1x * lam, b1xy = e1z - e1x * e1y * lam,
             b1xz = -e1y - e1x * e1z * lam;
      double b1yx = -e1z - e1y * e1x * lam, b1yy = lam - e1y * e1y * lam,
             b1yz = e1x - e1y * e1z * lam;
      double b1zx = e1y - e1z * e1x * lam, b1zy = -e1x - e1z * e1y * lam,
             b1zz = lam - e1z * e1z * lam;
      double etx = e1x + a1x * dtg + b1xx * dwx + b1xy * dwy + b1xz * dwz;
      double ety = e1y + a1y * dtg + b1yx * dwx + b1yy * dwy + b1yz * dwz;
      double etz = e1z + a1z * dtg + b1zx * dwx + b1zy * dwy + b1zz * dwz;
      double m = S2(mmom, i, k, N);
      V3(emom, 0, i, k, N) = etx;
      V3(emom, 1, i, k, N) = ety;
      V3(emom, 2, i, k, N) = etz;
      V3(emomM, 0, i, k, N) = etx * m;
      V3(emomM, 1, i, k, N) = ety * m;
      V3(emomM, 2, i, k, N) = etz * m;
    }
}

/* heun4f (heun_proper.f90:376-530) and the renormalisation of evolve4f
 * (236-245): emom = predictor, emom2 = state at step start, beff the field
 * of the predictor, b2eff that of the start; the drift and diffusion of
 * both points are averaged.  The reference's b1zz reads
 * lam - e1z*e1z - lam (heun_proper.f90:480), kept as written: it multiplies
 * the noise only. */
void or_heun_second(int N, int M, const double *landeg, const double *lambda, double delta_t,
                    const double *beff, const double *b2eff, const double *ranv,
                    const double *emom, double *emom2)
{
  const double dt = delta_t * OR_GAMA;
  for (int i = 0; i < N; ++i)
    for (int k = 0; k < M; ++k) {
      const double lam = lambda[i], lg = landeg[i];
      double lldamp = 1.0 / (1.0 + lam * lam);
      double dtg = dt * lg * lldamp;
      double dwx = V3(ranv, 0, i, k, N) * lg, dwy = V3(ranv, 1, i, k, N) * lg,
             dwz = V3(ranv, 2, i, k, N) * lg;
      double e1x = V3(emom, 0, i, k, N), e1y = V3(emom, 1, i, k, N), e1z = V3(emom, 2, i, k, N);
      double e2x = V3(emom2, 0, i, k, N), e2y = V3(emom2, 1, i, k, N), e2z = V3(emom2, 2, i, k, N);
      double h1esx = V3(beff, 0, i, k, N), h1esy = V3(beff, 1, i, k, N), h1esz = V3(beff, 2, i, k, N);
      double h1px = lam * h1esx, h1py = lam * h1esy, h1pz = lam * h1esz;
      double h2esx = V3(b2eff, 0, i, k, N), h2esy = V3(b2eff, 1, i, k, N),
             h2esz = V3(b2eff, 2, i, k, N);
      double h2px = lam * h2esx, h2py = lam * h2esy, h2pz = lam * h2esz;
      double prot = e1x * h1px + e1y * h1py + e1z * h1pz;
      double a1x = h1px + e1z * (h1esy + 0.0) - e1y * (h1esz + 0.0) - e1x * prot;
      double a1y = h1py - e1z * (h1esx + 0.0) + e1x * (h1esz + 0.0) - e1y * prot;
      double a1z = h1pz + e1y * (h1esx + 0.0) - e1x * (h1esy + 0.0) - e1z * prot;
      double b1xx = lam - e1x * e1x * lam, b1xy = e1z - e1x * e1y * lam,
             b1xz = -e1y - e1x * e1z * lam;
      double b1yx = -e1z - e1y * e1x * lam, b1yy = lam - e1y * e1y * lam,
             b1yz = e1x - e1y * e1z * lam;
      double b1zx = e1y - e1z * e1x * lam, b1zy = -e1x - e1z * e1y * lam,
             b1zz = lam - e1z * e1z - lam;
      prot = e2x * h2px + e2y * h2py + e2z * h2pz;
      double a2x = h2px + e2z * h2esy - e2y * h2esz - e2x * prot;
      double a2y = h2py - e2z * h2esx + e2x * h2esz - e2y * prot;
      double a2z = h2pz + e2y * h2esx - e2x * h2esy - e2z * prot;
      double b2xx = lam - e2x * e2x * lam, b2xy = e2z - e2x * e2y * lam,
             b2xz = -e2y - e2x * e2z * lam;
      double b2yx = -e2z - e2y * e2x * lam, b2yy = lam - e2y * e2y * lam,
             b2yz = e2x - e2y * e2z * lam;
      double b2zx = e2y - e2z * e2x * lam, b2zy = -e2x - e2z * e2y * lam,
             b2zz = lam - e2z * e2z * lam;
      double ax = a1x + a2x, ay = a1y + a2y, az = a1z + a2z;
      double bxx = b1xx + b2xx, bxy = b1xy + b2xy, bxz = b1xz + b2xz;
      double byx = b1yx + b2yx, byy = b1yy + b2yy, byz = b1yz + b2yz;
      double bzx = b1zx + b2zx, bzy = b1zy + b2zy, bzz = b1zz + b2zz;
      double etx = ax * dtg + bxx * dwx + bxy * dwy + bxz * dwz;
      double ety = ay * dtg + byx * dwx + byy * dwy + byz * dwz;
      double etz = az * dtg + bzx * dwx + bzy * dwy + bzz * dwz;
      V3(emom2, 0, i, k, N) = e2x + 0.5 * etx;
      V3(emom2, 1, i, k, N) = e2y + 0.5 * ety;
      V3(emom2, 2, i, k, N) = e2z + 0.5 * etz;
    }
  for (int i = 0; i < N; ++i)
    for (int k = 0; k < M; ++k) {
      double x = V3(emom2, 0, i, k, N), y = V3(emom2, 1, i, k, N), z = V3(emom2, 2, i, k, N);
      double etot = 1.0 / sqrt(x * x + y * y + z * z);
      V3(emom2, 0, i, k, N) = x * etot;
      V3(emom2, 1, i, k, N) = y * etot;
      V3(emom2, 2, i, k, N) = z * etot;
    }
}

/* Semi-implicit midpoint solvers with fixed-point iteration (SDEalgh 6,
 * sibt/sibf) and its spherical variant (SDEalgh 7, sispht/sisphf).  Both
 * halves solve e = e1 + dtg u x a1 + sqrtdtg u x s1 by at most 100
 * fixed-point iterations from e = e1, u = (e1 + e)/2 (SIB) or
 * (e1 + e)/|e1 + e| (spherical), stopping when every component moved by
 * less than epsilona = 1e-12, a default-real literal (so float(1e-12)).
 * a1 = -b - lam v x b and s1 = -f - lam v x f with f the rannum-scaled
 * noise and v = e1 (predictor) or the predicted midpoint (corrector).
 * Predictor: SIB stores (e1 + e)/2, spherical (e1 + e0)/|e1 + e0| with e0
 * the iterate before the last (the loop exits before e0 = e); emomM =
 * emom2*mmom.  Corrector: emom2 = the last iterate. */
static void sib_solve(int sph, const double e1[3], const double a1[3], const double s1[3],
                      double dtg, double sqrtdtg, double out[3], double prev[3])
{
  const double epsilona = (double)1e-12f;
  double e0[3] = {e1[0], e1[1], e1[2]}, ef[3] = {e1[0], e1[1], e1[2]};
  for (int k = 1; k <= 100; ++k) {
    double u[3];
    if (sph) {
      double unorm = sqrt((e1[0] + e0[0]) * (e1[0] + e0[0]) + (e1[1] + e0[1]) * (e1[1] + e0[1]) +
                          (e1[2] + e0[2]) * (e1[2] + e0[2]));
      for (int c = 0; c < 3; ++c) u[c] = (e1[c] + e0[c]) / unorm;
    } else {
      for (int c = 0; c < 3; ++c) u[c] = 0.5 * (e1[c] + e0[c]);
    }
    double d[3] = {dtg * u[0], dtg * u[1], dtg * u[2]};
    double q[3] = {sqrtdtg * u[0], sqrtdtg * u[1], sqrtdtg * u[2]};
    ef[0] = e1[0] + d[1] * a1[2] - d[2] * a1[1] + q[1] * s1[2] - q[2] * s1[1];
    ef[1] = e1[1] + d[2] * a1[0] - d[0] * a1[2] + q[2] * s1[0] - q[0] * s1[2];
    ef[2] = e1[2] + d[0] * a1[1] - d[1] * a1[0] + q[0] * s1[1] - q[1] * s1[0];
    double dx = sqrt((e0[0] - ef[0]) * (e0[0] - ef[0]));
    double dy = sqrt((e0[1] - ef[1]) * (e0[1] - ef[1]));
    double dz = sqrt((e0[2] - ef[2]) * (e0[2] - ef[2]));
    if (dx < epsilona && dy < epsilona && dz < epsilona) break;
    for (int c = 0; c < 3; ++c) e0[c] = ef[c];
  }
  for (int c = 0; c < 3; ++c) {
    out[c] = ef[c];
    prev[c] = e0[c];
  }
}

static void sib_drift(double lam, const double v[3], const double b[3], const double f[3],
                      double a1[3], double s1[3])
{
  a1[0] = -b[0] - lam * (v[1] * b[2] - v[2] * b[1]);
  a1[1] = -b[1] - lam * (v[2] * b[0] - v[0] * b[2]);
  a1[2] = -b[2] - lam * (v[0] * b[1] - v[1] * b[0]);
  s1[0] = -f[0] - lam * (v[1] * f[2] - v[2] * f[1]);
  s1[1] = -f[1] - lam * (v[2] * f[0] - v[0] * f[2]);
  s1[2] = -f[2] - lam * (v[0] * f[1] - v[1] * f[0]);
}

void or_sib_first(int alg, int N, int M, const double *landeg, double bn, const double *lambda,
                  const double *beff, const double *emom, double *emom2, double *emomM,
                  const double *mmom, double deltat, const double *ranv)
{
  for (int k = 0; k < M; ++k)
    for (int i = 0; i < N; ++i) {
      double lldamp = 1.0 / (1.0 + lambda[i] * lambda[i]);
      double dt = deltat * bn * OR_GAMA * lldamp;
      double sqrtdt = sqrt(dt);
      double dtg = dt * landeg[i], sqrtdtg = sqrtdt * landeg[i];
      double e1[3], b[3], f[3], a1[3], s1[3], ef[3], e0[3];
      for (int c = 0; c < 3; ++c) {
        e1[c] = V3(emom, c, i, k, N);
        b[c] = V3(beff, c, i, k, N);
        f[c] = V3(ranv, c, i, k, N);
      }
      sib_drift(lambda[i], e1, b, f, a1, s1);
      sib_solve(alg == 7, e1, a1, s1, dtg, sqrtdtg, ef, e0);
      double o[3];
      if (alg == 7) {
        double unorm = sqrt((e1[0] + e0[0]) * (e1[0] + e0[0]) + (e1[1] + e0[1]) * (e1[1] + e0[1]) +
                            (e1[2] + e0[2]) * (e1[2] + e0[2]));
        for (int c = 0; c < 3; ++c) o[c] = (e1[c] + e0[c]) / unorm;
      } else {
        for (int c = 0; c < 3; ++c) o[c] = 0.5 * (e1[c] + ef[c]);
      }
      double m = S2(mmom, i, k, N);
      for (int c = 0; c < 3; ++c) {
        V3(emom2, c, i, k, N) = o[c];
        V3(emomM, c, i, k, N) = o[c] * m;
      }
    }
}

void or_sib_second(int alg, int N, int M, const double *landeg, double bn, const double *lambda,
                   const double *beff, const double *emom, double *emom2, double deltat,
                   const double *ranv)
{
  for (int k = 0; k < M; ++k)
    for (int i = 0; i < N; ++i) {
      double lldamp = 1.0 / (1.0 + lambda[i] * lambda[i]);
      double dt = deltat * bn * OR_GAMA * lldamp;
      double sqrtdt = sqrt(dt);
      double dtg = dt * landeg[i], sqrtdtg = sqrtdt * landeg[i];
      double e1[3], etp[3], b[3], f[3], a1[3], s1[3], ef[3], e0[3];
      for (int c = 0; c < 3; ++c) {
        e1[c] = V3(emom, c, i, k, N);
        etp[c] = V3(emom2, c, i, k, N);
        b[c] = V3(beff, c, i, k, N);
        f[c] = V3(ranv, c, i, k, N);
      }
      sib_drift(lambda[i], etp, b, f, a1, s1);
      sib_solve(alg == 7, e1, a1, s1, dtg, sqrtdtg, ef, e0);
      for (int c = 0; c < 3; ++c) V3(emom2, c, i, k, N) = ef[c];
    }
}

/* The sd_mphase step body (sd_driver.f90:668-765) for nsteps, with the
 * thermal field drawn from the Philox counter stream the HIP kernels use --
 * the FP32 engine's single-precision deviates (noise = 1) or the FP64
 * engine's double-precision ones (noise = 2) -- or absent (noise = 0).  Used
 * as the CPU baseline and for long trajectory checks. */
void or_sd_run(const or_ham *h, int N, int M, int alg, long nsteps, uint64_t step0, double dt,
               const double *lam, const double *temp, const double *landeg,
               const double *mmom, double *emom, double *emomM, uint64_t seed, int noise,
               const double *ext)
{
  size_t n3 = (size_t)3 * N * M, n1 = (size_t)N * M;
  double *beff = malloc(sizeof(double) * n3), *e2 = malloc(sizeof(double) * n3);
  double *b2 = malloc(sizeof(double) * n3), *rv = malloc(sizeof(double) * n3);
  double *mmomi = malloc(sizeof(double) * n1);
  for (size_t x = 0; x < n1; ++x) mmomi[x] = 1.0 / mmom[x];
  for (long s = 0; s < nsteps; ++s) {
    uint64_t step = step0 + (uint64_t)s;
#pragma omp parallel for collapse(2) schedule(static) if (OR_PAR(N, M))
    for (int k = 0; k < M; ++k)
      for (int i = 0; i < N; ++i) {
        float z[3] = {0.0f, 0.0f, 0.0f};
        double zd[3] = {0.0, 0.0, 0.0};
        if (noise == 1) {
          or_thermal3(seed, 0x11u, step, (uint32_t)i, (uint32_t)k, z);
          for (int c = 0; c < 3; ++c) zd[c] = (double)z[c];
        } else if (noise == 2) {
          or_thermal3_f64(seed, step, (uint32_t)i, (uint32_t)k, zd);
        }
        for (int c = 0; c < 3; ++c) V3(rv, c, i, k, N) = zd[c];
      }
    or_effective_field(h, N, M, emomM, ext, beff, NULL, NULL, NULL);
    if (alg == 1) {
      or_rannum_scale(N, M, 1, lam, 1.0, mmomi, temp, 1.0, rv);
      or_midpoint_first(N, M, landeg, 1.0, lam, beff, emom, e2, emomM, mmom, dt, rv, NULL);
      or_effective_field(h, N, M, emomM, ext, beff, NULL, NULL, NULL);
      or_midpoint_second(N, M, landeg, 1.0, lam, beff, emom, e2, dt, rv);
    } else if (alg == 6 || alg == 7) {
      or_rannum_scale(N, M, 1, lam, 1.0, mmomi, temp, 1.0, rv);
      or_sib_first(alg, N, M, landeg, 1.0, lam, beff, emom, e2, emomM, mmom, dt, rv);
      or_effective_field(h, N, M, emomM, ext, beff, NULL, NULL, NULL);
      or_sib_second(alg, N, M, landeg, 1.0, lam, beff, emom, e2, dt, rv);
    } else if (alg == 4) {
      /* at step start emom2 equals emom (moment_update copies it) */
      for (size_t x = 0; x < n3; ++x) e2[x] = emom[x];
      or_heun_first(N, M, landeg, lam, temp, 1.0, dt, beff, rv, mmom, emom, e2, emomM, b2);
      or_effective_field(h, N, M, emomM, ext, beff, NULL, NULL, NULL);
      or_heun_second(N, M, landeg, lam, dt, beff, b2, rv, emom, e2);
    } else {
      or_depondt_first(N, M, lam, temp, 1.0, dt, beff, rv, mmom, emom, e2, emomM, b2, NULL);
      or_effective_field(h, N, M, emomM, ext, beff, NULL, NULL, NULL);
      or_depondt_second(N, M, lam, dt, beff, rv, b2, emom, e2);
    }
    or_moment_update(N, M, mmom, emom, e2, emomM, mmomi);
  }
  free(beff); free(e2); free(b2); free(rv); free(mmomi);
}
