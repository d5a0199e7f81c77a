// uasd_setup.cpp -- host-side Hamiltonian set-up of libuasd_hip.so: the
// reference neighbour map and pair couplings, reproduced bit for bit.
//
// Follows source/Hamiltonian/neighbourmap.f90 (setup_nm 32-359, get_symops
// 361-497, get_fullnnlist 508-640) and hamiltonianinit.f90
// (setup_neighbour_hamiltonian 985-1157) for a Bravais lattice with basis.
// The per-basis-atom part (symmetry-unique bond vectors, cell hops) is small
// and done once; the per-site part runs in parallel over sites and writes the
// list rows directly, without the reference's (Natom,shell,equiv) map.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/uasd_hip.h"

namespace {

constexpr double kMry = 2.179872325e-21;
constexpr double kMub = 9.274009994e-24;
constexpr double kDblTolerance = (double)1e-14f;

struct Mat3 {
  double a[3][3];  // a[i][j] = sym_mats(i+1,j+1)
};

std::vector<Mat3> symops(int isym) {
  std::vector<Mat3> ops;
  auto blank = [] {
    Mat3 m;
    std::memset(&m, 0, sizeof(m));
    return m;
  };
  if (isym == 0) {
    Mat3 m = blank();
    for (int i = 0; i < 3; ++i) m.a[i][i] = 1.0;
    ops.push_back(m);
  } else if (isym == 1) {  // 48 cubic operations, generation order of get_symops
    for (int i = 1; i <= 3; ++i)
      for (int j = 0; j <= 1; ++j) {
        const int js = j == 0 ? 1 : -1;
        for (int x = 0; x <= 1; ++x)
          for (int y = 0; y <= 1; ++y)
            for (int z = 0; z <= 1; ++z) {
              Mat3 m = blank();
              m.a[0][(i - js) % 3] = x ? -1.0 : 1.0;
              m.a[1][i % 3] = y ? -1.0 : 1.0;
              m.a[2][(i + js) % 3] = z ? -1.0 : 1.0;
              ops.push_back(m);
            }
      }
  } else if (isym == 2) {  // 8 operations of the square lattice in the xy-plane
    for (int j = 0; j <= 1; ++j)
      for (int x = 0; x <= 1; ++x)
        for (int y = 0; y <= 1; ++y) {
          Mat3 m = blank();
          m.a[0][j % 2] = x ? -1.0 : 1.0;
          m.a[1][(j + 1) % 2] = y ? -1.0 : 1.0;
          m.a[2][2] = 1.0;
          ops.push_back(m);
        }
  } else if (isym == 3) {  // 24 hexagonal operations
    const double half = 0.5, roothalf = std::sqrt(3.0) * 0.5;
    for (int x = 0; x <= 1; ++x)
      for (int y = 0; y <= 1; ++y)
        for (int z = 0; z <= 1; ++z) {
          Mat3 m = blank();
          m.a[0][0] = x ? -1.0 : 1.0;
          m.a[1][1] = y ? -1.0 : 1.0;
          m.a[2][2] = z ? -1.0 : 1.0;
          ops.push_back(m);
        }
    for (int x1 = 0; x1 <= 1; ++x1)
      for (int x2 = 0; x2 <= 1; ++x2)
        for (int y1 = 0; y1 <= 1; ++y1)
          for (int y2 = 0; y2 <= 1; ++y2) {
            const double sx1 = x1 ? -1.0 : 1.0, sx2 = x2 ? -1.0 : 1.0;
            const double sy1 = y1 ? -1.0 : 1.0, sy2 = y2 ? -1.0 : 1.0;
            if (sx1 * sx2 * sy1 * sy2 < 0.0)
              for (int z = 0; z <= 1; ++z) {
                Mat3 m = blank();
                m.a[0][0] = sx1 * half;
                m.a[1][0] = sx2 * roothalf;
                m.a[0][1] = sy1 * roothalf;
                m.a[1][1] = sy2 * half;
                m.a[2][2] = z ? -1.0 : 1.0;
                ops.push_back(m);
              }
          }
  }
  return ops;
}

struct Hop {
  int j0;          // basis atom of the neighbour, 0-based
  int dx, dy, dz;  // cell translation
};

}  // namespace

extern "C" int uasd_build_pair_list(int NA, int NT, int N1, int N2, int N3, const double *C1,
                                    const double *C2, const double *C3, const double *bas,
                                    const int *atype, const char *bc, int sym, int msh,
                                    const int *nn, const double *redcoord, const int *nntype,
                                    int hdim, int nchmax, const double *jc, const double *ammom,
                                    const int *achem, int *max_no_neigh, int *nlist,
                                    int *nlistsize, double *ncoup) {
  if (NA <= 0 || NT <= 0 || N1 <= 0 || N2 <= 0 || N3 <= 0 || !C1 || !C2 || !C3 || !bas ||
      !atype || !bc || !nn || !redcoord || msh <= 0 || !max_no_neigh)
    return UASD_ERR_ARG;
  if (sym < 0 || sym > 3) return UASD_ERR_UNSUPPORTED;
  if (nlist && (!nlistsize || !ncoup || !jc || !ammom || hdim <= 0 || nchmax <= 0))
    return UASD_ERR_ARG;
  const std::vector<Mat3> ops = symops(sym);
  const int nsym = (int)ops.size();
  const double tol = 0.01;
  // inverse of the cell matrix, setup_nm lines 109-126
  double inv[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  const double det = C1[0] * C2[1] * C3[2] - C1[0] * C2[2] * C3[1] + C1[1] * C2[2] * C3[0] -
                     C1[1] * C2[0] * C3[2] + C1[2] * C2[0] * C3[1] - C1[2] * C2[1] * C3[0];
  if (std::fabs(det) > kDblTolerance) {
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
  auto red = [&](int t, int s, int c) { return redcoord[t + NT * (s + msh * c)]; };
  // symmetry-unique bond vectors per (type, shell): get_fullnnlist
  std::vector<std::vector<std::vector<std::array<double, 3>>>> vecs(NT);
  for (int t = 0; t < NT; ++t) {
    vecs[t].resize(nn[t]);
    for (int s = 0; s < nn[t]; ++s) {
      auto &out = vecs[t][s];
      if (nsym == 1) {
        out.push_back({red(t, s, 0), red(t, s, 1), red(t, s, 2)});
        continue;
      }
      for (int is = 0; is < nsym; ++is) {
        std::array<double, 3> tv = {0.0, 0.0, 0.0};
        for (int i = 0; i < 3; ++i)
          for (int j = 0; j < 3; ++j) tv[i] = tv[i] + red(t, s, j) * ops[is].a[i][j];
        bool unique = true;
        for (const auto &o : out) {
          const double d0 = tv[0] - o[0], d1 = tv[1] - o[1], d2 = tv[2] - o[2];
          if (d0 * d0 + d1 * d1 + d2 * d2 < tol) unique = false;
        }
        if (unique) out.push_back(tv);
      }
    }
  }
  // hops of each basis atom, in (shell, equivalent vector) order
  const int ncell = N1 * N2 * N3;
  std::vector<std::vector<std::vector<Hop>>> hops(NA);
  int mnn = 1;
  for (int i0 = 0; i0 < NA; ++i0) {
    const int it = atype[i0] - 1;
    hops[i0].resize(nn[it]);
    int total = 0;
    for (int s = 0; s < nn[it]; ++s) {
      for (const auto &v : vecs[it][s]) {
        double cv[3], icv[3], bsf[3], rv[3];
        for (int c = 0; c < 3; ++c) cv[c] = v[c] + bas[3 * i0 + c];
        for (int c = 0; c < 3; ++c) icv[c] = cv[0] * inv[0][c] + cv[1] * inv[1][c] + cv[2] * inv[2][c];
        for (int c = 0; c < 3; ++c) bsf[c] = std::floor(icv[c] + 5.0e-5);
        rv[0] = cv[0] - bsf[0] * C1[0] - bsf[1] * C2[0] - bsf[2] * C3[0];
        rv[1] = cv[1] - bsf[0] * C1[1] - bsf[1] * C2[1] - bsf[2] * C3[1];
        rv[2] = cv[2] - bsf[0] * C1[2] - bsf[1] * C2[2] - bsf[2] * C3[2];
        for (int ia = 0; ia < NA; ++ia) {
          const bool typematch = nntype ? (atype[ia] == nntype[it + NT * s]) : true;
          const double d0 = rv[0] - bas[3 * ia], d1 = rv[1] - bas[3 * ia + 1],
                       d2 = rv[2] - bas[3 * ia + 2];
          if (d0 * d0 + d1 * d1 + d2 * d2 < tol && typematch) {
            Hop h{ia, 0, 0, 0};
            if (ncell > 1) {
              h.dx = (int)std::lrint(bsf[0]);
              h.dy = (int)std::lrint(bsf[1]);
              h.dz = (int)std::lrint(bsf[2]);
            }
            if ((int)hops[i0][s].size() < nsym) hops[i0][s].push_back(h);
            ++total;
          }
        }
      }
    }
    mnn = std::max(mnn, total);
  }
  *max_no_neigh = mnn;
  if (!nlist) return UASD_OK;

  const int Natom = NA * ncell;
  const double fc2 = 2.0 * kMry / kMub;
  const bool pbc[3] = {bc[0] == 'P', bc[1] == 'P', bc[2] == 'P'};
  const int nd[3] = {N1, N2, N3};
  int status = UASD_OK;
#pragma omp parallel for schedule(static)
  for (int iat = 0; iat < Natom; ++iat) {
    const int i0 = iat % NA;
    const int cellid = iat / NA;
    const int ix = cellid % N1, iy = (cellid / N1) % N2, iz = cellid / (N1 * N2);
    const int it = atype[i0] - 1;
    const int ci = achem ? achem[iat] : 1;
    const double mi = ammom[i0 + (size_t)NA * (ci - 1)];
    int *row = nlist + (size_t)mnn * iat;
    double *cp = ncoup + (size_t)hdim * mnn * iat;
    int count = 0;
    for (int s = 0; s < nn[it]; ++s) {
      // nmdim(shell) = number of symmetry-unique vectors bounds the walk
      const size_t nh = std::min(hops[i0][s].size(), vecs[it][s].size());
      for (size_t e = 0; e < nh; ++e) {
        const Hop &h = hops[i0][s][e];
        int j[3] = {h.dx + ix, h.dy + iy, h.dz + iz};
        bool inside = true;
        for (int c = 0; c < 3; ++c) {
          if (pbc[c]) j[c] = (j[c] + 1000 * nd[c]) % nd[c];
          else if (ncell <= 1) j[c] = 0;
          if (j[c] < 0 || j[c] >= nd[c]) inside = false;
        }
        if (!inside) continue;
        const int jat = h.j0 + 1 + j[0] * NA + j[1] * N1 * NA + j[2] * N2 * N1 * NA;  // 1-based
        bool exists = false;
        for (int l = 0; l < count; ++l)
          if (row[l] == jat) exists = true;
        if (exists) continue;
        row[count] = jat;
        const int cj = achem ? achem[jat - 1] : 1;
        const double mj = ammom[h.j0 + (size_t)NA * (cj - 1)];
        const int jslot = achem ? cj : 1;
        for (int d = 0; d < hdim; ++d) {
          double v = 0.0;
          if (!(std::fabs(mi * mj) < (double)1e-6f)) {
            const double x =
                jc[d + (size_t)hdim * (it + (size_t)NT * (s + (size_t)msh * ((ci - 1) + (size_t)nchmax * (jslot - 1))))];
            v = x * fc2 / mi / mj;
          }
          cp[(size_t)hdim * count + d] = v;
        }
        ++count;
      }
    }
    for (int l = count; l < mnn; ++l) {
      row[l] = 0;
      for (int d = 0; d < hdim; ++d) cp[(size_t)hdim * l + d] = 0.0;
    }
    nlistsize[iat] = count;
  }
  return status;
}
