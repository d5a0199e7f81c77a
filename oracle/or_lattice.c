/*
 * or_lattice.c -- neighbour map and pair Hamiltonian set-up, restated from
 *   source/Input/inputhandler_ext.f90   read_exchange (444-618), read_dmdata (1095-1251)
 *   source/Hamiltonian/neighbourmap.f90 setup_nm (32-359), get_symops (361-497),
 *                                       get_fullnnlist (508-640)
 *   source/Hamiltonian/hamiltonianinit.f90 setup_anisotropies (857-976),
 *                                       setup_neighbour_hamiltonian (985-1157)
 * Loops, tolerances and floating-point operation order follow the Fortran
 * so that the lists come out bit-identical.  Atom numbering follows
 * setup_type_and_numb (System/geometry.f90:144-184) with block size 1:
 *   i = I0 + I1*NA + I2*N1*NA + I3*N2*N1*NA  (1-based).
 */
#include "uasd_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define F3(a, i, j, k, n1, n2) (a)[((i) - 1) + (n1) * (((j) - 1) + (n2) * ((k) - 1))]

/* read_exchange / read_dmdata: dedupe bond vectors per type (tol 1e-5) */
int or_read_exchange_lines(const or_lattice *lat, int do_ralloy, int nchmax, int hdim,
                           int nlines, const int *isite, const int *jsite,
                           const int *ichem, const int *jchem, const double *r,
                           const double *val, or_shells *out)
{
  const double tol = 1.0e-5;
  int NT = lat->NT;
  int msh = nlines > 0 ? nlines : 1;   /* upper bound, reduced below */
  int *nn = calloc(NT, sizeof(int));
  double *red = calloc((size_t)NT * msh * 3, sizeof(double));
  int *nntype = calloc((size_t)NT * msh, sizeof(int));
  int nchj = do_ralloy ? nchmax : (nchmax > lat->NT ? nchmax : lat->NT);
  double *jc = calloc((size_t)hdim * NT * msh * nchmax * nchj, sizeof(double));
  if (!nn || !red || !nntype || !jc) return -1;
#define RED(t, s, c) red[((t) - 1) + NT * (((s) - 1) + msh * ((c) - 1))]
#define JC(h, t, s, ci, cj) jc[(h) + hdim * (((t) - 1) + NT * (((s) - 1) + msh * (((ci) - 1) + nchmax * ((cj) - 1))))]
  for (int l = 0; l < nlines; ++l) {
    int itype = lat->atype_inp[isite[l] - 1];
    int ic = do_ralloy ? ichem[l] : 1;
    int jcx = do_ralloy ? jchem[l] : 1;
    double rr[3] = {r[3 * l], r[3 * l + 1], r[3 * l + 2]};  /* posfiletype C, maptype 1 */
    int unique = 1;
    for (int s = 1; s <= nn[itype - 1]; ++s) {
      double d0 = rr[0] - RED(itype, s, 1), d1 = rr[1] - RED(itype, s, 2),
             d2 = rr[2] - RED(itype, s, 3);
      double norm = d0 * d0 + d1 * d1 + d2 * d2;
      if (norm < tol) {
        unique = 0;
        for (int h = 0; h < hdim; ++h) JC(h, itype, s, ic, jcx) = val[hdim * l + h];
      }
    }
    if (unique) {
      int s = ++nn[itype - 1];
      nntype[(itype - 1) + NT * (s - 1)] = lat->atype_inp[jsite[l] - 1];
      for (int c = 1; c <= 3; ++c) RED(itype, s, c) = rr[c - 1];
      for (int h = 0; h < hdim; ++h) JC(h, itype, s, ic, jcx) = val[hdim * l + h];
    }
  }
  /* reduce to max_no_shells (read_exchange_reduce* helpers) */
  int ms = 0;
  for (int t = 0; t < NT; ++t) ms = nn[t] > ms ? nn[t] : ms;
  if (ms < 1) ms = 1;
  out->NT = NT; out->max_no_shells = ms; out->nchmax = nchmax; out->hdim = hdim;
  out->nn = nn;
  out->redcoord = calloc((size_t)NT * ms * 3, sizeof(double));
  out->nntype = calloc((size_t)NT * ms, sizeof(int));
  out->jc = calloc((size_t)hdim * NT * ms * nchmax * nchmax, sizeof(double));
  for (int t = 1; t <= NT; ++t)
    for (int s = 1; s <= ms; ++s) {
      for (int c = 1; c <= 3; ++c)
        out->redcoord[(t - 1) + NT * ((s - 1) + ms * (c - 1))] = RED(t, s, c);
      out->nntype[(t - 1) + NT * (s - 1)] = nntype[(t - 1) + NT * (s - 1)];
      for (int ci = 1; ci <= nchmax; ++ci)
        for (int cj = 1; cj <= nchmax; ++cj)
          for (int h = 0; h < hdim; ++h)
            out->jc[h + hdim * ((t - 1) + NT * ((s - 1) + ms * ((ci - 1) + nchmax * (cj - 1))))] =
                JC(h, t, s, ci, cj);
    }
#undef RED
#undef JC
  free(red); free(nntype); free(jc);
  return 0;
}

void or_shells_free(or_shells *s)
{
  free(s->nn); free(s->redcoord); free(s->nntype); free(s->jc);
  memset(s, 0, sizeof(*s));
}

/* get_symops (neighbourmap.f90:361-497) for sym 0..3; returns count, fills
 * mats[(isym)*9 + (i-1) + 3*(j-1)] = sym_mats(i,j,isym). */
static int get_symops(int isym, double *mats)
{
  int cnt = 0;
  memset(mats, 0, sizeof(double) * 9 * 48);
#define SM(i, j, s) mats[(s) * 9 + ((i) - 1) + 3 * ((j) - 1)]
  if (isym == 0) {
    for (int i = 1; i <= 3; ++i) SM(i, i, 0) = 1.0;
    cnt = 1;
  } else if (isym == 1) {
    for (int i = 1; i <= 3; ++i)
      for (int j = 0; j <= 1; ++j) {
        int js = (j == 0) ? 1 : -1;
        for (int x = 0; x <= 1; ++x)
          for (int y = 0; y <= 1; ++y)
            for (int z = 0; z <= 1; ++z) {
              SM(1, (i - js) % 3 + 1, cnt) = x ? -1.0 : 1.0;
              SM(2, i % 3 + 1, cnt) = y ? -1.0 : 1.0;
              SM(3, (i + js) % 3 + 1, cnt) = z ? -1.0 : 1.0;
              cnt++;
            }
      }
  } else if (isym == 2) {
    for (int j = 0; j <= 1; ++j)
      for (int x = 0; x <= 1; ++x)
        for (int y = 0; y <= 1; ++y) {
          SM(1, j % 2 + 1, cnt) = x ? -1.0 : 1.0;
          SM(2, (j + 1) % 2 + 1, cnt) = y ? -1.0 : 1.0;
          SM(3, 3, cnt) = 1.0;
          cnt++;
        }
  } else if (isym == 3) {
    const double half = 0.5, roothalf = sqrt(3.0) * 0.5;
    for (int x = 0; x <= 1; ++x)
      for (int y = 0; y <= 1; ++y)
        for (int z = 0; z <= 1; ++z) {
          SM(1, 1, cnt) = x ? -1.0 : 1.0;
          SM(2, 2, cnt) = y ? -1.0 : 1.0;
          SM(3, 3, cnt) = z ? -1.0 : 1.0;
          cnt++;
        }
    for (int x1 = 0; x1 <= 1; ++x1)
      for (int x2 = 0; x2 <= 1; ++x2)
        for (int y1 = 0; y1 <= 1; ++y1)
          for (int y2 = 0; y2 <= 1; ++y2) {
            double sx1 = x1 ? -1.0 : 1.0, sx2 = x2 ? -1.0 : 1.0;
            double sy1 = y1 ? -1.0 : 1.0, sy2 = y2 ? -1.0 : 1.0;
            if (sx1 * sx2 * sy1 * sy2 < 0.0) {
              for (int z = 0; z <= 1; ++z) {
                SM(1, 1, cnt) = sx1 * half;
                SM(2, 1, cnt) = sx2 * roothalf;
                SM(1, 2, cnt) = sy1 * roothalf;
                SM(2, 2, cnt) = sy2 * half;
                SM(3, 3, cnt) = z ? -1.0 : 1.0;
                cnt++;
              }
            }
          }
  }
#undef SM
  return cnt;
}

int or_setup_nm(const or_lattice *lat, int sym, const or_shells *sh, int use_nntype,
                or_nmap *out)
{
  const double tol = 0.01;
  const int NA = lat->NA, NT = lat->NT, N1 = lat->N1, N2 = lat->N2, N3 = lat->N3;
  const int Natom = NA * N1 * N2 * N3;
  const int msh = sh->max_no_shells;
  const double *C1 = lat->C1, *C2 = lat->C2, *C3 = lat->C3;
  double mats[9 * 48];
  int nsym = get_symops(sym, mats);
  int max_no_equiv = nsym;
  double inv[3][3] = {{0}};                       /* inv[i-1][j-1] = invmatrix(i,j) */
  double det = C1[0] * C2[1] * C3[2] - C1[0] * C2[2] * C3[1] +
               C1[1] * C2[2] * C3[0] - C1[1] * C2[0] * C3[2] +
               C1[2] * C2[0] * C3[1] - C1[2] * C2[1] * C3[0];
  if (fabs(det) > OR_DBL_TOLERANCE) {
    inv[0][0] = (C2[1] * C3[2] - C3[1] * C2[2]) / det;
    inv[0][1] = (C1[2] * C3[1] - C3[2] * C1[1]) / det;
    inv[0][2] = (C1[1] * C2[2] - C2[1] * C1[2]) / det;
    inv[1][0] = (C2[2] * C3[0] - C3[2] * C2[0]) / det;
    inv[1][1] = (C1[0] * C3[2] - C3[0] * C1[2]) / det;
    inv[1][2] = (C1[2] * C2[0] - C2[2] * C1[0]) / det;
    inv[2][0] = (C2[0] * C3[1] - C3[0] * C2[1]) / det;
    inv[2][1] = (C1[1] * C3[0] - C3[1] * C1[0]) / det;
    inv[2][2] = (C1[0] * C2[1] - C2[0] * C1[1]) / det;
  }
  /* get_fullnnlist with nelem = 1 */
  double *nncoord = calloc((size_t)3 * max_no_equiv * msh * NT, sizeof(double));
  int *nmdimt = calloc((size_t)msh * NT, sizeof(int));
#define NNC(c, e, s, t) nncoord[(c) + 3 * ((e) + max_no_equiv * ((s) + msh * (t)))]
#define REDC(t, s, c) sh->redcoord[(t) + NT * ((s) + msh * (c))]
  for (int t = 0; t < NT; ++t)
    for (int s = 0; s < sh->nn[t]; ++s) {
      if (nsym == 1) {
        for (int c = 0; c < 3; ++c) NNC(c, 0, s, t) = REDC(t, s, c);
        nmdimt[s + msh * t] = 1;
      } else {
        int counter = 0;
        for (int is = 0; is < nsym; ++is) {
          double tv[3] = {0.0, 0.0, 0.0};
          for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
              tv[i] = tv[i] + REDC(t, s, j) * mats[is * 9 + i + 3 * j];
          int unique = 1;
          for (int k = 0; k < counter; ++k) {
            double d0 = tv[0] - NNC(0, k, s, t), d1 = tv[1] - NNC(1, k, s, t),
                   d2 = tv[2] - NNC(2, k, s, t);
            if (d0 * d0 + d1 * d1 + d2 * d2 < tol) unique = 0;
          }
          if (unique) {
            for (int i = 0; i < 3; ++i) NNC(i, counter, s, t) = tv[i];
            counter++;
          }
        }
        nmdimt[s + msh * t] = counter;
      }
    }
  /* nm_cell / nm_trunk for the first cell */
  int *nm_cell = calloc((size_t)max_no_equiv * msh * NA, sizeof(int));
  int *nm_trunk = calloc((size_t)3 * max_no_equiv * msh * NA, sizeof(int));
  int *nnm_cell = calloc((size_t)msh * NA, sizeof(int));
  const int ncell = N1 * N2 * N3;
  for (int i0 = 0; i0 < NA; ++i0) {
    int itype = lat->atype_inp[i0] - 1;
    for (int s = 0; s < sh->nn[itype]; ++s) {
      int counter = 0;
      for (int e = 0; e < nmdimt[s + msh * itype]; ++e) {
        double cvec[3], icvec[3], bsf[3], rvec[3];
        for (int c = 0; c < 3; ++c) cvec[c] = NNC(c, e, s, itype) + lat->bas[3 * i0 + c];
        for (int c = 0; c < 3; ++c)
          icvec[c] = cvec[0] * inv[0][c] + cvec[1] * inv[1][c] + cvec[2] * inv[2][c];
        for (int c = 0; c < 3; ++c) bsf[c] = floor(icvec[c] + 5.0e-5);
        rvec[0] = cvec[0] - bsf[0] * C1[0] - bsf[1] * C2[0] - bsf[2] * C3[0];
        rvec[1] = cvec[1] - bsf[0] * C1[1] - bsf[1] * C2[1] - bsf[2] * C3[1];
        rvec[2] = cvec[2] - bsf[0] * C1[2] - bsf[1] * C2[2] - bsf[2] * C3[2];
        for (int ia = 0; ia < NA; ++ia) {
          int jtype = lat->atype_inp[ia];
          int typematch = 1;
          if (use_nntype) typematch = (jtype == sh->nntype[itype + NT * s]);
          double d0 = rvec[0] - lat->bas[3 * ia], d1 = rvec[1] - lat->bas[3 * ia + 1],
                 d2 = rvec[2] - lat->bas[3 * ia + 2];
          int tolmatch = (d0 * d0 + d1 * d1 + d2 * d2) < tol;
          if (tolmatch && typematch) {
            counter++;
            if (max_no_equiv >= counter) {
              int e2 = counter - 1;
              nm_cell[e2 + max_no_equiv * (s + msh * i0)] = ia + 1;
              if (ncell > 1)
                for (int c = 0; c < 3; ++c)
                  nm_trunk[c + 3 * (e2 + max_no_equiv * (s + msh * i0))] = (int)lrint(bsf[c]);
            }
          }
        }
      }
      nnm_cell[s + msh * i0] = counter;
    }
  }
  free(nncoord);
  /* full map */
  int *nm = calloc((size_t)Natom * msh * max_no_equiv, sizeof(int));
  int *nmdim = calloc((size_t)msh * Natom, sizeof(int));
  for (int iz = 0; iz < N3; ++iz)
    for (int iy = 0; iy < N2; ++iy)
      for (int ix = 0; ix < N1; ++ix)
        for (int i0 = 0; i0 < NA; ++i0) {
          int itype = lat->atype_inp[i0] - 1;
          int iat = i0 + ix * NA + iy * N1 * NA + iz * N2 * N1 * NA;   /* 0-based */
          for (int s = 0; s < sh->nn[itype]; ++s) {
            nmdim[s + msh * iat] = nmdimt[s + msh * itype];
            for (int e = 0; e < nnm_cell[s + msh * i0]; ++e) {
              int xh = 0, yh = 0, zh = 0;
              if (ncell > 1) {
                const int *tr = &nm_trunk[3 * (e + max_no_equiv * (s + msh * i0))];
                xh = tr[0]; yh = tr[1]; zh = tr[2];
              }
              int j0 = nm_cell[e + max_no_equiv * (s + msh * i0)];
              int jx = xh + ix, jy = yh + iy, jz = zh + iz;
              if (lat->BC[0] == 'P') jx = (jx + 1000 * N1) % N1; else if (ncell <= 1) jx = 0;
              if (lat->BC[1] == 'P') jy = (jy + 1000 * N2) % N2; else if (ncell <= 1) jy = 0;
              if (lat->BC[2] == 'P') jz = (jz + 1000 * N3) % N3; else if (ncell <= 1) jz = 0;
              if (jx >= 0 && jx < N1 && jy >= 0 && jy < N2 && jz >= 0 && jz < N3) {
                int jat = j0 + jx * NA + jy * N1 * NA + jz * N2 * N1 * NA;  /* 1-based */
                nm[iat + Natom * (s + msh * e)] = jat;
              }
            }
          }
        }
  int mnn = 1;
  for (int i0 = 0; i0 < NA; ++i0) {
    int itype = lat->atype_inp[i0] - 1, counter = 0;
    for (int s = 0; s < sh->nn[itype]; ++s) counter += nnm_cell[s + msh * i0];
    mnn = counter > mnn ? counter : mnn;
  }
  free(nm_cell); free(nm_trunk); free(nnm_cell); free(nmdimt);
#undef NNC
#undef REDC
  out->Natom = Natom; out->max_no_shells = msh; out->max_no_equiv = max_no_equiv;
  out->max_no_neigh = mnn; out->nm = nm; out->nmdim = nmdim;
  return 0;
}

void or_nmap_free(or_nmap *m)
{
  free(m->nm); free(m->nmdim);
  memset(m, 0, sizeof(*m));
}

int or_setup_neighbour_hamiltonian(const or_lattice *lat, const or_shells *sh,
                                   const or_nmap *nm, const double *ammom,
                                   const int *achem, int lexp, int map_multiple,
                                   or_pairlist *out)
{
  const int NA = lat->NA, NT = lat->NT, Natom = nm->Natom, msh = nm->max_no_shells;
  const int mnn = nm->max_no_neigh, meq = nm->max_no_equiv, hdim = sh->hdim;
  const int nch = sh->nchmax;
  const double fc2 = 2.0 * OR_MRY / OR_MUB;
  int *nlist = calloc((size_t)mnn * Natom, sizeof(int));
  int *nlistsize = calloc((size_t)Natom, sizeof(int));
  double *ncoup = calloc((size_t)hdim * mnn * Natom, sizeof(double));
  if (!nlist || !nlistsize || !ncoup) return -1;
  for (int i = 0; i < Natom; ++i) {
    int ncount = 1;
    int anumb_i = i % NA;                     /* 0-based basis index */
    int itype = lat->atype_inp[anumb_i];      /* 1-based */
    for (int k = 0; k < sh->nn[itype - 1]; ++k) {
      for (int j = 0; j < nm->nmdim[k + msh * i]; ++j) {
        int nij = nm->nm[i + Natom * (k + msh * j)];
        if (nij > 0) {
          int exis = 0;
          for (int l = 1; l <= ncount - 1; ++l)
            if (nlist[(l - 1) + mnn * i] == nij) exis = 1;
          if (!exis || map_multiple) {
            nlist[(ncount - 1) + mnn * i] = nij;
            int anumb_j = (nij - 1) % NA;
            int ci = achem ? achem[i] : 1;
            int cj = achem ? achem[nij - 1] : 1;
            double mi = ammom[anumb_i + NA * (ci - 1)];
            double mj = ammom[anumb_j + NA * (cj - 1)];
            int jslot = achem ? cj : 1;
            for (int h = 0; h < hdim; ++h) {
              double v;
              if (fabs(mi * mj) < 1e-6f) {   /* 1e-6 is a default-real literal */
                v = 0.0;
              } else {
                double x = sh->jc[h + hdim * ((itype - 1) + NT * (k + msh * ((ci - 1) + nch * (jslot - 1))))];
                double a = mi, b = mj;
                if (lexp == 1) {
                  v = x * fc2 / a / b;
                } else {
                  v = x * fc2 / pow(a, lexp) / pow(b, lexp);
                }
              }
              ncoup[h + hdim * ((ncount - 1) + mnn * i)] = v;
            }
            ncount++;
          }
        }
      }
    }
    nlistsize[i] = ncount - 1;
  }
  (void)meq;
  out->Natom = Natom; out->hdim = hdim; out->max_no_neigh = mnn;
  out->nlist = nlist; out->nlistsize = nlistsize; out->ncoup = ncoup;
  return 0;
}

void or_pairlist_free(or_pairlist *p)
{
  free(p->nlist); free(p->nlistsize); free(p->ncoup);
  memset(p, 0, sizeof(*p));
}

void or_setup_anisotropies(const or_lattice *lat, const int *type_inp,
                           const double *aniso_inp, const double *ammom,
                           int *taniso, double *eaniso, double *kaniso, double *sb)
{
  const int NA = lat->NA, Natom = NA * lat->N1 * lat->N2 * lat->N3;
  const double fc = OR_MRY / OR_MUB;
#define AN(a, c) aniso_inp[(a) + NA * (c)]
  for (int i = 0; i < Natom; ++i) {
    int a = i % NA;
    taniso[i] = type_inp[a];
    double anorm = AN(a, 2) * AN(a, 2) + AN(a, 3) * AN(a, 3) + AN(a, 4) * AN(a, 4);
    double s = sqrt(anorm + 1.0e-15);
    for (int c = 0; c < 3; ++c) eaniso[3 * i + c] = AN(a, 2 + c) / s;
    /* integer powers as flang expands them: x**4 = (x*x)*(x*x), x**6 = x**2 * x**4 */
    double m = ammom[a], m2 = m * m, m4 = m2 * m2, m6 = m2 * m4;
    if (taniso[i] == 2) {
      kaniso[2 * i] = fc * AN(a, 0) / m4;
      kaniso[2 * i + 1] = fc * AN(a, 1) / m6;
    } else {
      kaniso[2 * i] = fc * AN(a, 0) / m2;
      kaniso[2 * i + 1] = fc * AN(a, 1) / m4;
    }
    sb[i] = AN(a, 5);
  }
#undef AN
}
