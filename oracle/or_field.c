/*
 * or_field.c -- effective field and field-routine energy, restated from
 * source/Hamiltonian/hamiltonianactions.f90:
 *   effective_field_full        108-252 (loop body 183-248, energy 245-250)
 *   effective_field_single      254-392 (heat-bath single-site variant)
 *   heisenberg_field            431-466
 *   dzyaloshinskii_moriya_field 546-580
 *   uniaxial_anisotropy_field   842-879
 *   cubic_anisotropy_field      883-919
 * and the term-resolved energies of source/Hamiltonian/energy.f90:
 *   calc_energy site loop       183-324 (do_lsf 'N', scalar exchange)
 *   division per atom           353-366
 *   update_ene                  559-571
 * Only the terms on the accelerated path are restated (exchange, DM,
 * uniaxial / cubic / type-7 anisotropy, static external field); the
 * time-dependent field is taken as zero as in sd_mphase without pulses.
 */
#include "uasd_oracle.h"
#include <math.h>
#include <stddef.h>

#define V3(a, c, i, k, N) (a)[(c) + 3 * ((size_t)(i) + (size_t)(N) * (k))]

/* The exchange sum accumulates by fused multiply-add, as the reference's
 * release build (-Ofast -march=native, cmake/SetFortranFlags.cmake:77-157)
 * contracts it; the HIP kernels do the same, so both round alike. */
static void heisenberg_field(const or_ham *h, int i, int k, int N, const double *emomM,
                             double f[3])
{
  const int mnn = h->max_no_neigh;
  for (int j = 0; j < h->nlistsize[i]; ++j) {
    double J = h->ncoup[j + (size_t)mnn * i];
    int nb = h->nlist[j + (size_t)mnn * i] - 1;
    f[0] = fma(J, V3(emomM, 0, nb, k, N), f[0]);
    f[1] = fma(J, V3(emomM, 1, nb, k, N), f[1]);
    f[2] = fma(J, V3(emomM, 2, nb, k, N), f[2]);
  }
}

static void dm_field(const or_ham *h, int i, int k, int N, const double *emomM, double f[3])
{
  const int mdm = h->max_no_dmneigh;
  for (int j = 0; j < h->dmlistsize[i]; ++j) {
    const double *D = &h->dm_vect[3 * (j + (size_t)mdm * i)];
    int nb = h->dmlist[j + (size_t)mdm * i] - 1;
    double m1 = V3(emomM, 0, nb, k, N), m2 = V3(emomM, 1, nb, k, N), m3 = V3(emomM, 2, nb, k, N);
    f[0] = f[0] + D[2] * m2 - D[1] * m3;
    f[1] = f[1] + D[0] * m3 - D[2] * m1;
    f[2] = f[2] + D[1] * m1 - D[0] * m2;
  }
}

static void uniaxial_field(const or_ham *h, int i, int k, int N, const double *emomM, double f[3])
{
  const double *e = &h->eaniso[3 * (size_t)i];
  const double *K = &h->kaniso[2 * (size_t)i];
  double m1 = V3(emomM, 0, i, k, N), m2 = V3(emomM, 1, i, k, N), m3 = V3(emomM, 2, i, k, N);
  double tt1 = m1 * e[0] + m2 * e[1] + m3 * e[2];
  double tt2 = K[0] + 2.0 * K[1] * (1.0 - tt1 * tt1);
  double tt3 = 2.0 * tt1 * tt2;
  f[0] = f[0] - tt3 * e[0];
  f[1] = f[1] - tt3 * e[1];
  f[2] = f[2] - tt3 * e[2];
}

static void cubic_field(const or_ham *h, int i, int k, int N, const double *emomM, double f[3])
{
  const double *K = &h->kaniso[2 * (size_t)i];
  double m1 = V3(emomM, 0, i, k, N), m2 = V3(emomM, 1, i, k, N), m3 = V3(emomM, 2, i, k, N);
  double s1 = m1 * m1, s2 = m2 * m2, s3 = m3 * m3;
  f[0] = f[0] + 2.0 * K[0] * m1 * (s2 + s3) + 2.0 * K[1] * m1 * (s2 * s3);
  f[1] = f[1] + 2.0 * K[0] * m2 * (s3 + s1) + 2.0 * K[1] * m2 * (s3 * s1);
  f[2] = f[2] + 2.0 * K[0] * m3 * (s1 + s2) + 2.0 * K[1] * m3 * (s1 * s2);
}

static void site_field(const or_ham *h, int N, const double *emomM, const double *ext,
                       int i, int k, double bs[3], double bq[3], double bext[3])
{
  bs[0] = bs[1] = bs[2] = 0.0;
  bq[0] = bq[1] = bq[2] = 0.0;
  heisenberg_field(h, i, k, N, emomM, bs);
  if (h->do_dm) dm_field(h, i, k, N, emomM, bs);
  if (h->do_anisotropy) {
    int t = h->taniso[i];
    if (t == 1) {
      uniaxial_field(h, i, k, N, emomM, bs);
    } else if (t == 2) {
      cubic_field(h, i, k, N, emomM, bs);
    } else if (t == 7) {
      double tf[3] = {0.0, 0.0, 0.0};
      uniaxial_field(h, i, k, N, emomM, bs);
      cubic_field(h, i, k, N, emomM, tf);
      double sb = h->sb[i];
      bq[0] = bq[0] + tf[0] * sb;
      bq[1] = bq[1] + tf[1] * sb;
      bq[2] = bq[2] + tf[2] * sb;
    }
  }
  for (int c = 0; c < 3; ++c) bext[c] = ext ? V3(ext, c, i, k, N) : 0.0;
}

void or_effective_field(const or_ham *h, int N, int M, const double *emomM,
                        const double *ext, double *beff, double *beff1,
                        double *beff2, double *energy)
{
  double en = 0.0;
  /* sites are independent; the energy sum keeps its serial order */
#pragma omp parallel for collapse(2) schedule(static) if (!energy && OR_PAR(N, M))
  for (int k = 0; k < M; ++k)
    for (int i = 0; i < N; ++i) {
      double bs[3], bq[3], be[3];
      site_field(h, N, emomM, ext, i, k, bs, bq, be);
      double b2[3];
      for (int c = 0; c < 3; ++c) {
        b2[c] = bq[c] + be[c] + 0.0;                 /* + time_external_field (zero) */
        if (beff1) V3(beff1, c, i, k, N) = bs[c];
        if (beff2) V3(beff2, c, i, k, N) = b2[c];
        V3(beff, c, i, k, N) = 0.0 + bs[c] + b2[c];
      }
      double tf[3];
      for (int c = 0; c < 3; ++c) tf[c] = 0.5 * (bs[c] + 2.0 * bq[c] + 2.0 * be[c] + 0.0);
      if (energy)
        en = en - V3(emomM, 0, i, k, N) * tf[0] - V3(emomM, 1, i, k, N) * tf[1] -
             V3(emomM, 2, i, k, N) * tf[2];
    }
  if (energy) *energy = en * OR_MUB / OR_MRY;
}

void or_effective_field_single(const or_ham *h, int N, const double *emomM,
                               const double *ext_uniform, int i, int k, double out[3])
{
  double bs[3], bq[3], be[3];
  site_field(h, N, emomM, NULL, i, k, bs, bq, be);
  if (ext_uniform)
    for (int c = 0; c < 3; ++c) be[c] = ext_uniform[c];
  for (int c = 0; c < 3; ++c) {
    double b2 = bq[c] + be[c] + 0.0;
    out[c] = 0.0 + bs[c] + b2;
  }
}

/* update_ene (energy.f90:559-571): -factor * sum(m * b). */
static double update_ene(const double m[3], const double b[3], double factor)
{
  return -factor * (m[0] * b[0] + m[1] * b[1] + m[2] * b[2]);
}

/* calc_energy for the terms on the accelerated path (energy.f90:183-366):
 * per ensemble k, each field built from zero at every site and dotted with
 * emomM, summed in site order and divided by Natom.  ene[4*k + t] in field
 * units (the caller scales by mub/mry), t = 0 exchange, 1 anisotropy, 2 DM,
 * 3 external field.  Anisotropy type 7 is uniaxial + sb * cubic
 * (energy.f90:311-316); the external field is external_field +
 * time_external_field with the latter zero. */
void or_energy_terms(const or_ham *h, int N, int M, const double *emomM, const double *ext,
                     double *ene)
{
  for (int k = 0; k < M; ++k) {
    double exc = 0.0, eani = 0.0, edm = 0.0, eext = 0.0;
    for (int i = 0; i < N; ++i) {
      double m[3] = {V3(emomM, 0, i, k, N), V3(emomM, 1, i, k, N), V3(emomM, 2, i, k, N)};
      double bxc[3] = {0.0, 0.0, 0.0};
      heisenberg_field(h, i, k, N, emomM, bxc);
      exc = exc + update_ene(m, bxc, 0.5);
      if (h->do_dm) {
        double bdm[3] = {0.0, 0.0, 0.0};
        dm_field(h, i, k, N, emomM, bdm);
        edm = edm + update_ene(m, bdm, 0.5);
      }
      if (h->do_anisotropy) {
        int t = h->taniso[i];
        double ba[3] = {0.0, 0.0, 0.0};
        if (t == 1) {
          uniaxial_field(h, i, k, N, emomM, ba);
        } else if (t == 2) {
          cubic_field(h, i, k, N, emomM, ba);
        } else if (t == 7) {
          double bc[3] = {0.0, 0.0, 0.0};
          uniaxial_field(h, i, k, N, emomM, ba);
          cubic_field(h, i, k, N, emomM, bc);
          for (int c = 0; c < 3; ++c) ba[c] = ba[c] + bc[c] * h->sb[i];
        }
        if (t == 1 || t == 2 || t == 7) eani = eani + update_ene(m, ba, 0.5);
      }
      double be[3];
      for (int c = 0; c < 3; ++c) be[c] = 0.0 + (ext ? V3(ext, c, i, k, N) : 0.0);
      eext = eext + update_ene(m, be, 1.0);
    }
    ene[4 * k + 0] = exc / N;
    ene[4 * k + 1] = eani / N;
    ene[4 * k + 2] = edm / N;
    ene[4 * k + 3] = eext / N;
  }
}
