// san_driver.cpp -- exercises the host set-up of the engine (uasd_setup.cpp:
// uasd_build_pair_list) and the oracle's restatements under AddressSanitizer
// and UndefinedBehaviorSanitizer (built by tests/test_sanitizers.py with g++,
// CPU only).  Each lattice is built by both implementations and the lists
// must agree bit for bit; then a few LLG steps (every solver) and MC sweeps
// (both modes, both trial-move forms) run on the oracle.  Prints "ok".
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

extern "C" {
#include "../../oracle/uasd_oracle.h"
int uasd_build_pair_list(int NA, int NT, int N1, int N2, int N3, const double *C1,
                         const double *C2, const double *C3, const double *bas,
                         const int *atype, const char *bc, int sym, int max_no_shells,
                         const int *nn, const double *redcoord, const int *nntype, int hdim,
                         int nchmax, const double *jc, const double *ammom, const int *achem,
                         int *max_no_neigh, int *nlist, int *nlistsize, double *ncoup);
}

struct Line {
  int i, j;
  double r[3], v[3];
};

static int fails = 0;
#define CHECK(c)                                                   \
  do {                                                             \
    if (!(c)) {                                                    \
      std::fprintf(stderr, "check failed line %d: %s\n", __LINE__, #c); \
      ++fails;                                                     \
    }                                                              \
  } while (0)

// Builds a lattice's pair list both ways; returns the oracle's list.
static or_pairlist build_both(const or_lattice &lat, const std::vector<Line> &lines, int sym,
                              int hdim, int use_nntype) {
  const int nl = (int)lines.size();
  std::vector<int> isite(nl), jsite(nl), one(nl, 1);
  std::vector<double> r(3 * nl), v(hdim * nl);
  for (int k = 0; k < nl; ++k) {
    isite[k] = lines[k].i;
    jsite[k] = lines[k].j;
    for (int c = 0; c < 3; ++c) r[3 * k + c] = lines[k].r[c];
    for (int c = 0; c < hdim; ++c) v[hdim * k + c] = lines[k].v[c];
  }
  or_shells sh;
  std::memset(&sh, 0, sizeof(sh));
  CHECK(or_read_exchange_lines(&lat, 0, 1, hdim, nl, isite.data(), jsite.data(), one.data(),
                               one.data(), r.data(), v.data(), &sh) == 0);
  or_nmap nm;
  std::memset(&nm, 0, sizeof(nm));
  CHECK(or_setup_nm(&lat, sym, &sh, use_nntype, &nm) == 0);
  std::vector<double> ammom(lat.NA, 2.2);
  or_pairlist pl;
  std::memset(&pl, 0, sizeof(pl));
  CHECK(or_setup_neighbour_hamiltonian(&lat, &sh, &nm, ammom.data(), nullptr, 1, 0, &pl) == 0);
  const int N = pl.Natom;
  int mnn = 0;
  const char bc[4] = {lat.BC[0], lat.BC[1], lat.BC[2], 0};
  CHECK(uasd_build_pair_list(lat.NA, lat.NT, lat.N1, lat.N2, lat.N3, lat.C1, lat.C2, lat.C3,
                             lat.bas, lat.atype_inp, bc, sym, sh.max_no_shells, sh.nn,
                             sh.redcoord, use_nntype ? sh.nntype : nullptr, hdim, 1, sh.jc,
                             ammom.data(), nullptr, &mnn, nullptr, nullptr, nullptr) == 0);
  std::vector<int> nlist((size_t)mnn * N), nsize(N);
  std::vector<double> nc((size_t)hdim * mnn * N);
  CHECK(uasd_build_pair_list(lat.NA, lat.NT, lat.N1, lat.N2, lat.N3, lat.C1, lat.C2, lat.C3,
                             lat.bas, lat.atype_inp, bc, sym, sh.max_no_shells, sh.nn,
                             sh.redcoord, use_nntype ? sh.nntype : nullptr, hdim, 1, sh.jc,
                             ammom.data(), nullptr, &mnn, nlist.data(), nsize.data(),
                             nc.data()) == 0);
  for (int i = 0; i < N; ++i) {
    CHECK(nsize[i] == pl.nlistsize[i]);
    for (int j = 0; j < nsize[i] && j < pl.max_no_neigh; ++j) {
      CHECK(nlist[(size_t)j + (size_t)mnn * i] == pl.nlist[(size_t)j + (size_t)pl.max_no_neigh * i]);
      for (int c = 0; c < hdim; ++c)
        CHECK(std::memcmp(&nc[c + hdim * ((size_t)j + (size_t)mnn * i)],
                          &pl.ncoup[c + hdim * ((size_t)j + (size_t)pl.max_no_neigh * i)],
                          sizeof(double)) == 0);
    }
  }
  or_nmap_free(&nm);
  or_shells_free(&sh);
  return pl;
}

static or_lattice make_lattice(const double *cell, const double *bas, const int *atype, int NA,
                               int n1, int n2, int n3, const char *bc) {
  or_lattice lat;
  std::memset(&lat, 0, sizeof(lat));
  lat.NA = NA;
  lat.NT = 1;
  for (int a = 0; a < NA; ++a) lat.NT = atype[a] > lat.NT ? atype[a] : lat.NT;
  lat.N1 = n1;
  lat.N2 = n2;
  lat.N3 = n3;
  for (int c = 0; c < 3; ++c) {
    lat.C1[c] = cell[c];
    lat.C2[c] = cell[3 + c];
    lat.C3[c] = cell[6 + c];
  }
  lat.bas = bas;
  lat.atype_inp = atype;
  for (int c = 0; c < 3; ++c) lat.BC[c] = bc[c];
  return lat;
}

int main() {
  // bcc, five shells, sym 1 (cubic), periodic
  const double cub[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
  const double bbas[6] = {0, 0, 0, 0.5, 0.5, 0.5};
  const int bty[2] = {1, 1};
  or_lattice bcc = make_lattice(cub, bbas, bty, 2, 5, 5, 5, "PPP");
  const double sh[5][3] = {{.5, .5, .5}, {1, 0, 0}, {1, 1, 0}, {1.5, .5, .5}, {1, 1, 1}};
  std::vector<Line> bl;
  for (int k = 0; k < 5; ++k) bl.push_back({1, 1, {sh[k][0], sh[k][1], sh[k][2]}, {1.34 / (k + 1), 0, 0}});
  or_pairlist pb = build_both(bcc, bl, 1, 1, 1);
  // triangular with DM vectors (hdim 3, sym 0) and vacuum in z
  const double tri[9] = {1, 0, 0, -0.5, std::sqrt(3.0) / 2, 0, 0, 0, 1};
  const double tbas[3] = {0, 0, 0};
  const int tty[1] = {1};
  or_lattice tl = make_lattice(tri, tbas, tty, 1, 8, 8, 1, "PP0");
  std::vector<Line> dl;
  const double b6[6][2] = {{1, 0}, {-1, 0}, {.5, std::sqrt(3.0) / 2}, {-.5, -std::sqrt(3.0) / 2},
                           {-.5, std::sqrt(3.0) / 2}, {.5, -std::sqrt(3.0) / 2}};
  for (auto &b : b6) dl.push_back({1, 1, {b[0], b[1], 0}, {-0.2 * b[1], 0.2 * b[0], 0}});
  or_pairlist pd = build_both(tl, dl, 0, 3, 0);
  // stripe with vacuum in x and y (the HeisStripe geometry)
  or_lattice st = make_lattice(cub, tbas, tty, 1, 6, 2, 12, "00P");
  std::vector<Line> sl = {{1, 1, {1, 0, 0}, {1, 0, 0}}};
  or_pairlist ps = build_both(st, sl, 1, 1, 1);
  or_pairlist_free(&ps);

  // LLG steps on the bcc lattice, every solver, FP64 and FP32 deviates
  const int N = pb.Natom, M = 2;
  or_ham h;
  std::memset(&h, 0, sizeof(h));
  h.Natom = N;
  h.max_no_neigh = pb.max_no_neigh;
  h.nlist = pb.nlist;
  h.nlistsize = pb.nlistsize;
  h.ncoup = pb.ncoup;
  std::vector<double> lam(N, 0.1), T(N, 300.0), lg(N, 1.0), mm((size_t)N * M, 2.2);
  for (int alg : {1, 4, 5, 6, 7})
    for (int noise : {1, 2}) {
      std::vector<double> e((size_t)3 * N * M), eM((size_t)3 * N * M);
      for (size_t x = 0; x < (size_t)N * M; ++x) {
        const double th = 0.1 * (double)(x % 31), ph = 0.37 * (double)x;
        e[3 * x] = std::sin(th) * std::cos(ph);
        e[3 * x + 1] = std::sin(th) * std::sin(ph);
        e[3 * x + 2] = std::cos(th);
        for (int c = 0; c < 3; ++c) eM[3 * x + c] = e[3 * x + c] * mm[x];
      }
      or_sd_run(&h, N, M, alg, 3, 0, 1e-16, lam.data(), T.data(), lg.data(), mm.data(),
                e.data(), eM.data(), 99, noise, nullptr);
      for (size_t x = 0; x < (size_t)N * M; ++x) {
        const double n = std::sqrt(e[3 * x] * e[3 * x] + e[3 * x + 1] * e[3 * x + 1] +
                                   e[3 * x + 2] * e[3 * x + 2]);
        CHECK(std::fabs(n - 1.0) < 1e-9);
      }
    }
  // MC on the triangular lattice with DM and anisotropy: a proper colouring
  const int Nt = pd.Natom;
  or_ham ht;
  std::memset(&ht, 0, sizeof(ht));
  ht.Natom = Nt;
  ht.max_no_neigh = pd.max_no_neigh;
  std::vector<double> jc((size_t)pd.max_no_neigh * Nt, 1.0);
  ht.nlist = pd.nlist;
  ht.nlistsize = pd.nlistsize;
  ht.ncoup = jc.data();
  ht.do_dm = 1;
  ht.max_no_dmneigh = pd.max_no_neigh;
  ht.dmlist = pd.nlist;
  ht.dmlistsize = pd.nlistsize;
  ht.dm_vect = pd.ncoup;
  std::vector<int> taniso(Nt);
  std::vector<double> eaniso(3 * (size_t)Nt), kaniso(2 * (size_t)Nt), sb(Nt, 0.5);
  for (int i = 0; i < Nt; ++i) {
    taniso[i] = i % 3 == 0 ? 1 : (i % 3 == 1 ? 2 : 7);
    eaniso[3 * i + 2] = 1.0;
    kaniso[2 * i] = 0.3;
    kaniso[2 * i + 1] = -0.1;
  }
  ht.do_anisotropy = 1;
  ht.taniso = taniso.data();
  ht.eaniso = eaniso.data();
  ht.kaniso = kaniso.data();
  ht.sb = sb.data();
  std::vector<int> col(Nt, -1);
  int ncol = 0;
  for (int i = 0; i < Nt; ++i) {
    std::vector<char> used(16, 0);
    for (int j = 0; j < pd.nlistsize[i]; ++j) {
      const int n = pd.nlist[(size_t)j + (size_t)pd.max_no_neigh * i] - 1;
      if (col[n] >= 0) used[col[n]] = 1;
    }
    int k = 0;
    while (used[k]) ++k;
    col[i] = k;
    ncol = k + 1 > ncol ? k + 1 : ncol;
  }
  std::vector<int> ptr(ncol + 1, 0), sites;
  for (int k = 0; k < ncol; ++k) {
    for (int i = 0; i < Nt; ++i)
      if (col[i] == k) sites.push_back(i);
    ptr[k + 1] = (int)sites.size();
  }
  const double temps[2] = {100.0, 900.0}, ext[3] = {0.1, 0.0, 0.5};
  std::vector<double> mt((size_t)Nt * M, 1.7);
  for (char mode : {'M', 'H'})
    for (int f32 : {0, 1}) {
      std::vector<double> e((size_t)3 * Nt * M, 0.0), eM((size_t)3 * Nt * M, 0.0);
      for (size_t x = 0; x < (size_t)Nt * M; ++x) {
        e[3 * x + 2] = 1.0;
        eM[3 * x + 2] = 1.7;
      }
      for (int s = 0; s < 2; ++s)
        or_mc_sweep_colored_ex(&ht, Nt, M, temps, mode, 77, s, ncol, ptr.data(), sites.data(),
                               eM.data(), e.data(), mt.data(), ext, f32);
      double et[8];
      or_energy_terms(&ht, Nt, M, eM.data(), nullptr, et);
      CHECK(std::isfinite(et[0]) && std::isfinite(et[4]));
    }
  or_pairlist_free(&pb);
  or_pairlist_free(&pd);
  if (fails) {
    std::fprintf(stderr, "%d checks failed\n", fails);
    return 1;
  }
  std::printf("ok\n");
  return 0;
}
