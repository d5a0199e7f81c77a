/*
 * or_init.c -- initial spin configuration, restated from
 *   source/System/magnetizationinit.f90 magninit, initmag = 1 (141-175).
 * The reference draws three uniforms per site from state_c; a draw outside
 * the unit ball is redrawn with half the spread (x = 1*(rn-0.5)), which is
 * kept as written because it fixes the stream consumption.
 * Also the thread control of the oracle's OpenMP site loops.
 */
#include "uasd_oracle.h"
#include <math.h>
#include <stddef.h>

void or_magninit_random(or_mt *st, int NA, int N1, int N2, int N3, int M,
                        const double *mmom, double *emom, double *emomM)
{
  const size_t N = (size_t)NA * N1 * N2 * N3;
  for (int i3 = 0; i3 < N3; ++i3)
    for (int i2 = 0; i2 < N2; ++i2)
      for (int i1 = 0; i1 < N1; ++i1)
        for (int i0 = 0; i0 < NA; ++i0) {
          size_t i = (size_t)i0 + (size_t)i1 * NA + (size_t)i2 * N1 * NA + (size_t)i3 * N2 * N1 * NA;
          double rn[3];
          for (int c = 0; c < 3; ++c) rn[c] = or_mt_real2(st);
          double x = 2.0 * (rn[0] - 0.5), y = 2.0 * (rn[1] - 0.5), z = 2.0 * (rn[2] - 0.5);
          while (x * x + y * y + z * z > 1) {
            for (int c = 0; c < 3; ++c) rn[c] = or_mt_real2(st);
            x = 1.0 * (rn[0] - 0.5);
            y = 1.0 * (rn[1] - 0.5);
            z = 1.0 * (rn[2] - 0.5);
          }
          double u[3];
          u[0] = x / sqrt(x * x + y * y + z * z);
          u[1] = y / sqrt(x * x + y * y + z * z);
          u[2] = z / sqrt(x * x + y * y + z * z);
          for (int j = 0; j < M; ++j) {
            double m = mmom[(size_t)i0 + N * j];   /* mmom(anumb(I0), j) */
            for (int c = 0; c < 3; ++c) {
              emom[c + 3 * (i + N * j)] = u[c];
              emomM[c + 3 * (i + N * j)] = u[c] * m;
            }
          }
        }
}

#ifdef _OPENMP
#include <omp.h>
#endif

void or_set_threads(int n)
{
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

int or_threads(void)
{
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
