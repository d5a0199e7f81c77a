/*
 * or_pt.c -- the replica-exchange step of heat-bath parallel tempering,
 * restated from source/pt_driver.f90 (pt_iphase):
 *   exchange attempts   205-232 (two passes over neighbouring pairs)
 *   replica selection   246-250 (first replica of lowest energy)
 * Test infrastructure only: the checker of uppasd_amd/tempering.py.
 */
#include "uasd_oracle.h"
#include <math.h>

/* temps (nrep) are swapped in place between neighbouring replicas;
 * repene (nrep) the field-routine energies in mRy (summed over the
 * replica's ensembles); flipprob (nrep) uniform draws; hit/trial counters
 * are incremented. */
void or_pt_exchange(int nrep, double *temps, const double *repene, int natom,
                    const double *flipprob, double temprescale, double *n_hit, double *n_trial)
{
  for (int ishift = 0; ishift <= 1; ++ishift)
    for (int irep = 1 + ishift; irep <= nrep - 1; irep += 2) {
      *n_trial = *n_trial + 1.0;
      double temp_1 = temps[irep - 1], temp_2 = temps[irep];
      double beta_1 = 1.0 / OR_K_BOLT / (temprescale * temp_1 + 1.0e-15);
      double beta_2 = 1.0 / OR_K_BOLT / (temprescale * temp_2 + 1.0e-15);
      double ene_1 = repene[irep - 1] * OR_MRY / natom;
      double ene_2 = repene[irep] * OR_MRY / natom;
      double try_pow = (ene_2 - ene_1) * (beta_2 - beta_1);
      double trial_prob = fmin(1.0, exp(try_pow));
      if (flipprob[irep - 1] <= trial_prob) {
        double t = temps[irep - 1];
        temps[irep - 1] = temps[irep];
        temps[irep] = t;
        *n_hit = *n_hit + 1.0;
      }
    }
}

/* pt_driver.f90:246-250: the first replica whose energy is the minimum
 * (0-based). */
int or_pt_select(int nrep, const double *repene)
{
  double mn = repene[0];
  for (int r = 1; r < nrep; ++r) mn = repene[r] < mn ? repene[r] : mn;
  int irep = 0;
  while (repene[irep] > mn) ++irep;
  return irep;
}
