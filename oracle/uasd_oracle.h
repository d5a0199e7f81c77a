/*
 * uasd_oracle.h -- CPU restatement (FP64, plain C) of UppASD's spin-update
 * hot path.  TEST INFRASTRUCTURE ONLY: this library is the checker for the
 * HIP path.  Only tests/, __graft_entry__.smoke() and the cpu_baseline leg
 * of bench.py may load it.  Nothing in uppasd_amd/ links or calls it.
 *
 * Every function follows a reference routine term for term (operation order
 * kept so that results are reproducible to the last bit on an IEEE double
 * machine built with -ffp-contract=off); the citation is given at each
 * function in the .c files.  Arrays use the reference's Fortran layout:
 *   vec3 fields  (3,Natom,Mensemble) column-major -> v[c + 3*(i + N*k)]
 *   per-site     (Natom,Mensemble)                  -> a[i + N*k]
 *   neighbour lists (mnn,Natom) with 1-based atom indices, as in ham%nlist.
 */
#ifndef UASD_ORACLE_H
#define UASD_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------- threads ----------------------------------------------- */
/* The per-site loops of the field and LLG routines run on OpenMP threads
 * when a call covers at least 4096 (site, ensemble) pairs: every site is
 * computed by the same operations whatever the thread count, so results do
 * not depend on it (the energy sums stay serial).  or_set_threads(n) sets
 * the thread count (n <= 0 leaves the OpenMP default). */
void or_set_threads(int n);
int  or_threads(void);
#define OR_PAR(N, M) ((long)(N) * (long)(M) >= 4096L)

/* ---------------- physical constants (source/Parameters/constants.f90) --- */
extern const double OR_GAMA, OR_K_BOLT, OR_MUB, OR_MRY, OR_PI;
extern const double OR_DBL_TOLERANCE;   /* real(1e-14,4) widened, parameters.f90:12 */

/* ---------------- RNG: MT19937 as written in source/RNG/mtprng.f90 ------- */
typedef struct {
  int64_t  mti;
  uint64_t mt[624];
} or_mt;
void     or_mt_init(or_mt *s, int32_t seed);
uint64_t or_mt_rand64(or_mt *s);
int32_t  or_mt_rand(or_mt *s);
double   or_mt_real1(or_mt *s);
double   or_mt_real2(or_mt *s);

/* ---------------- ziggurat normal (source/RNG/randomnumbers.f90) --------- */
typedef struct {
  int32_t kn[128];
  double  fn[128], wn[128];
  or_mt   st;             /* state_z */
} or_zig;
void   or_zig_setup(or_zig *z, int32_t seed);
double or_zig_nor(or_zig *z);

/* ---------------- counter-based RNG used by the HIP kernels -------------- */
void   or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
void   or_counter_words(uint64_t seed, uint32_t tag, uint64_t step, uint32_t site,
                        uint32_t ens, uint32_t out[4]);
float  or_u24(uint32_t w);
double or_u32_to_open01(uint32_t x);                       /* (x+0.5)*2^-32 */
void   or_noise3_f64(uint64_t seed, uint32_t tag, uint64_t step, uint32_t site,
                     uint32_t ens, double out[3]);
void   or_thermal3(uint64_t seed, uint32_t tag, uint64_t step, uint32_t site,
                   uint32_t ens, float out[3]);        /* LLG thermal deviates */
void   or_thermal3_f64(uint64_t seed, uint64_t step, uint32_t site, uint32_t ens,
                       double out[3]);                 /* FP64 engine's deviates */
float  or_bm_ln(float u);
void   or_bm_sincos(float v, float out[2]);           /* sin, cos of (pi/2) v */
void   or_uniform4(uint64_t seed, uint32_t tag, uint64_t step, uint32_t site,
                   uint32_t ens, double out[4]);

/* ---------------- lattice + neighbour map ------------------------------- */
typedef struct {
  int    NA, NT, N1, N2, N3;
  double C1[3], C2[3], C3[3];
  const double *bas;        /* (3,NA) Cartesian basis, units of alat */
  const int    *atype_inp;  /* (NA) 1-based type of each basis atom   */
  char   BC[3];             /* 'P' periodic, '0' vacuum              */
} or_lattice;

/* Output of read_exchange (inputhandler_ext.f90:444-618). */
typedef struct {
  int     NT, max_no_shells, nchmax;
  int    *nn;          /* (NT)                                        */
  double *redcoord;    /* (NT,max_no_shells,3) Fortran order          */
  int    *nntype;      /* (NT,max_no_shells)                          */
  double *jc;          /* (hdim,NT,max_no_shells,nchmax,nchmax) Fortran order */
  int     hdim;
} or_shells;

int  or_read_exchange_lines(const or_lattice *lat, int do_ralloy, int nchmax, int hdim,
                            int nlines, const int *isite, const int *jsite,
                            const int *ichem, const int *jchem, const double *r,
                            const double *val, or_shells *out);
void or_shells_free(or_shells *s);

typedef struct {
  int  Natom, max_no_shells, max_no_equiv, max_no_neigh;
  int *nm;      /* (Natom,max_no_shells,max_no_equiv), 0 = none */
  int *nmdim;   /* (max_no_shells,Natom) */
} or_nmap;

int  or_setup_nm(const or_lattice *lat, int sym, const or_shells *sh, int use_nntype,
                 or_nmap *out);
void or_nmap_free(or_nmap *m);

typedef struct {
  int     Natom, hdim, max_no_neigh;
  int    *nlist;      /* (max_no_neigh,Natom) 1-based */
  int    *nlistsize;  /* (Natom) */
  double *ncoup;      /* (hdim,max_no_neigh,Natom) */
} or_pairlist;

/* setup_neighbour_hamiltonian (hamiltonianinit.f90:985-1157).  ammom is
 * (NA,nchmax); achem is (Natom) 1-based chemistry or NULL (not an alloy). */
int  or_setup_neighbour_hamiltonian(const or_lattice *lat, const or_shells *sh,
                                    const or_nmap *nm, const double *ammom,
                                    const int *achem, int lexp, int map_multiple,
                                    or_pairlist *out);
void or_pairlist_free(or_pairlist *p);

/* setup_anisotropies (hamiltonianinit.f90:857-976), non-alloy branch.
 * aniso_inp is (NA,6): K1,K2,ex,ey,ez,ratio; type_inp is (NA). */
void or_setup_anisotropies(const or_lattice *lat, const int *type_inp,
                           const double *aniso_inp, const double *ammom,
                           int *taniso, double *eaniso, double *kaniso, double *sb);

/* ---------------- Hamiltonian view used by field / MC routines ---------- */
typedef struct {
  int Natom;
  /* exchange */
  int max_no_neigh; const int *nlist; const int *nlistsize; const double *ncoup;
  /* DMI */
  int do_dm; int max_no_dmneigh; const int *dmlist; const int *dmlistsize;
  const double *dm_vect;
  /* single-ion anisotropy */
  int do_anisotropy; const int *taniso; const double *eaniso; const double *kaniso;
  const double *sb;
} or_ham;

/* effective_field_full (hamiltonianactions.f90:108-252) with exchange, DM,
 * uniaxial/cubic/type-7 anisotropy and external field.  ext may be NULL. */
void or_effective_field(const or_ham *h, int N, int M, const double *emomM,
                        const double *ext, double *beff, double *beff1,
                        double *beff2, double *energy);

/* ---------------- LLG solvers ------------------------------------------ */
void or_depondt_first(int N, int M, const double *lambda, const double *temp,
                      double temprescale, double dt, const double *beff,
                      double *btherm, const double *mmom, double *emom,
                      double *emom2, double *emomM, double *b2eff,
                      double *thermal_field);
void or_depondt_second(int N, int M, const double *lambda, double dt,
                       const double *beff, const double *btherm,
                       const double *b2eff, double *emom, double *emom2);
void or_rannum_scale(int N, int M, int llg, const double *lambda, double bn,
                     const double *mmomi, const double *temp, double temprescale,
                     double *ranv);
void or_midpoint_first(int N, int M, const double *landeg, double bn,
                       const double *lambda, const double *beff, const double *emom,
                       double *emom2, double *emomM, const double *mmom,
                       double dt, const double *ranv, double *thermal_field);
void or_midpoint_second(int N, int M, const double *landeg, double bn,
                        const double *lambda, const double *beff, const double *emom,
                        double *emom2, double dt, const double *ranv);
/* Heun (SDEalgh 4): heun_proper.f90 evolve4p/heun4p and heun4f/evolve4f */
void or_heun_first(int N, int M, const double *landeg, const double *lambda, const double *temp,
                   double temprescale, double delta_t, const double *beff, double *ranv,
                   const double *mmom, double *emom, const double *emom2, double *emomM,
                   double *b2eff);
void or_heun_second(int N, int M, const double *landeg, const double *lambda, double delta_t,
                    const double *beff, const double *b2eff, const double *ranv,
                    const double *emom, double *emom2);
/* SIB (alg 6) and spherical midpoint (alg 7): sphmidpoint.f90 */
void or_sib_first(int alg, int N, int M, const double *landeg, double bn, const double *lambda,
                  const double *beff, const double *emom, double *emom2, double *emomM,
                  const double *mmom, double deltat, const double *ranv);
void or_sib_second(int alg, int N, int M, const double *landeg, double bn, const double *lambda,
                   const double *beff, const double *emom, double *emom2, double deltat,
                   const double *ranv);
void or_moment_update(int N, int M, const double *mmom, double *emom,
                      const double *emom2, double *emomM, double *mmomi);
/* nsteps of the sd_mphase body; noise 1 = Philox stream of the HIP kernels */
void or_sd_run(const or_ham *h, int N, int M, int alg, long nsteps, uint64_t step0, double dt,
               const double *lam, const double *temp, const double *landeg,
               const double *mmom, double *emom, double *emomM, uint64_t seed, int noise,
               const double *ext);

/* magninit, initmag = 1 (magnetizationinit.f90:141-175); mmom is (N,M) */
void or_magninit_random(or_mt *st, int NA, int N1, int N2, int N3, int M,
                        const double *mmom, double *emom, double *emomM);

/* ---------------- Monte Carlo ------------------------------------------- */
double or_mc_delta(double temperature);
void   or_choose_random_flip(const double *e, double delta, const double rn[3],
                             const double gn[3], double newmom[3]);
double or_mc_energy_change(const or_ham *h, int N, const double *emomM,
                           const double *emom, const double *mmom, int iflip, int k,
                           const double newmom[3], const double extfield[3]);
/* ext_uniform: one field vector applied to every site, or NULL */
void   or_effective_field_single(const or_ham *h, int N, const double *emomM,
                                 const double *ext_uniform, int i, int k, double out[3]);
/* calc_energy terms (energy.f90:183-366): ene (M,4) per atom in field units,
 * columns exchange, anisotropy, DM, external field. */
void   or_energy_terms(const or_ham *h, int N, int M, const double *emomM,
                       const double *ext, double *ene);
void   or_flip_h(double *e_out, double mmom, double temperature, double temprescale,
                 double flipprob, const double totfield[3], double q);
void   or_choose_random_atom_x(or_mt *st, int N, int *iflip_a);
void   or_mc_sweep_ref(const or_ham *h, int N, int M, double temperature,
                       double temprescale, char mode, or_mt *st_c, or_zig *zig,
                       const int *iflip_a, double *emomM, double *emom,
                       const double *mmom, const double extfield[3]);
void   or_mc_sweep_colored(const or_ham *h, int N, int M, const double *temps,
                           char mode, uint64_t seed, uint64_t sweep, int ncolors,
                           const int *color_ptr, const int *color_sites,
                           double *emomM, double *emom, const double *mmom,
                           const double extfield[3]);
void   or_mc_sweep_colored_ex(const or_ham *h, int N, int M, const double *temps, char mode,
                              uint64_t seed, uint64_t sweep, int ncolors, const int *color_ptr,
                              const int *color_sites, double *emomM, double *emom,
                              const double *mmom, const double extfield[3], int f32_moves);
void   or_trial_move_f32(const double e[3], double delta, const uint32_t w[4], double out[3]);

/* bulk fills in the reference's stream order (rng_uniformP / fill_rngarray) */
void   or_mt_fill_real2(or_mt *s, double *out, long n);
void   or_zig_fill(or_zig *z, double *out, long n);

/* ---------------- observables ------------------------------------------ */
void   or_sum_moments(int N, int M, const double *emomM, double *msum /*(3,M)*/);
double or_calc_mavrg(int N, int M, const double *emomM);

/* parallel tempering exchange step (pt_driver.f90:205-232) and the final
 * replica choice (246-250) */
void or_pt_exchange(int nrep, double *temps, const double *repene, int natom,
                    const double *flipprob, double temprescale, double *n_hit, double *n_trial);
int  or_pt_select(int nrep, const double *repene);

#ifdef __cplusplus
}
#endif
#endif
