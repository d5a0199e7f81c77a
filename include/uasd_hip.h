/*
 * uasd_hip.h -- C-ABI of libuasd_hip.so, the MI355X (gfx950) spin-update
 * engine that stands in for UppASD's hot path.
 *
 * Conventions
 *  - Every entry point returns an int status (UASD_OK = 0, negative on
 *    failure); uasd_last_error() gives the message of the calling thread's
 *    last failure.  No entry point exits the process.
 *  - Host arrays use the reference's Fortran layout and 1-based atom
 *    indices, so a bind(C) interface can pass module arrays unchanged:
 *      vec3 per site    (3,Natom,Mensemble)  e.g. emom, emomM, beff
 *      per site/ens     (Natom,Mensemble)    e.g. mmom
 *      neighbour list   (max_no_neigh,Natom) ham%nlist
 *      couplings        (max_no_neigh,nHam)  ham%ncoup(:,:,1)
 *  - A context owns device copies of the Hamiltonian and of the moments;
 *    the *_host entry points copy host arrays in and out on every call
 *    (drop-in semantics), the resident entry points (uasd_sd_steps,
 *    uasd_mc_sweeps, uasd_measure) keep everything in HBM.
 *  - Calls on one context are not thread safe; different contexts are.
 */
#ifndef UASD_HIP_H
#define UASD_HIP_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define UASD_API_VERSION 1

enum {
  UASD_OK = 0,
  UASD_ERR_ARG = -1,          /* invalid argument / shape                      */
  UASD_ERR_HIP = -2,          /* HIP runtime failure                           */
  UASD_ERR_NOMEM = -3,        /* device allocation failed                      */
  UASD_ERR_UNSUPPORTED = -4,  /* option outside the accelerated path           */
  UASD_ERR_STATE = -5,        /* call order (e.g. stepping before set-up)      */
  UASD_ERR_NODEVICE = -6      /* no usable gfx950 device                       */
};

enum { UASD_FP64 = 0, UASD_FP32 = 1 };

/* Thermal-noise source for the LLG solvers. */
enum {
  UASD_NOISE_NONE = 0,        /* no noise drawn (the reference's T = 0 result) */
  UASD_NOISE_PHILOX = 1,      /* counter-based Philox4x32-10, drawn in-kernel   */
  UASD_NOISE_INJECTED = 2     /* caller-supplied N(0,1) array, scaled in-kernel */
};

/* Solvers (the reference's SDEalgh values, evolution.f90:134-217). */
enum {
  UASD_SDE_MIDPOINT = 1,
  UASD_SDE_HEUN = 4,
  UASD_SDE_DEPONDT = 5,
  UASD_SDE_SIB = 6,     /* sibt/sibf, sphmidpoint.f90:36-477 */
  UASD_SDE_SPHMID = 7   /* sispht/sisphf, sphmidpoint.f90:485-923 */
};

typedef struct uasd_ctx uasd_ctx;

const char *uasd_last_error(void);
int uasd_version(void);
int uasd_device_count(int *count);

/* Context: Natom sites, Mensemble ensembles, FP64 or FP32 arithmetic.
 * Replaces the state set-up of FortranData_Initiate / CudaMdSimulation::
 * initiateMatrices (chelper.f90:166-188, gpu_files/cudaMdSimulation.cu:133). */
int uasd_ctx_create(uasd_ctx **ctx, int natom, int mensemble, int precision);
int uasd_ctx_destroy(uasd_ctx *ctx);
/* Run on a caller stream (hipStream_t); NULL restores the context's own. */
int uasd_ctx_set_stream(uasd_ctx *ctx, void *hip_stream);
void *uasd_ctx_stream(uasd_ctx *ctx);
int uasd_ctx_sync(uasd_ctx *ctx);

/* Device site order (no reference counterpart; results do not depend on
 * it): na > 1 stores the sites of each of the na basis atoms contiguously
 * (reference site i = cell*na + basis goes to slot basis*(natom/na) + cell),
 * so the lanes of a wavefront gather one sublattice from contiguous memory.
 * Host arrays keep the reference order in every call.  Must precede the
 * uploads; not combined with several segments.  Default 1 (reference order). */
int uasd_set_site_order(uasd_ctx *ctx, int na);

/* Lattice dimensions (the NA, N1, N2, N3 arguments of effective_field_full,
 * hamiltonianactions.f90:108-112): reference site i = b + NA*(x + N1*(y +
 * N2*z)).  When the exchange (and DM) lists are one translation-invariant
 * stencil per basis atom over the periodic lattice, the fused LLG kernels can
 * gather from an LDS-staged tile of cells (a tuning choice, same results).
 * All zero forgets the lattice.  has_stencil reports whether the lists are
 * such a stencil and its reach in cells (x, y, z). */
int uasd_set_lattice_dims(uasd_ctx *ctx, int na, int n1, int n2, int n3);
int uasd_has_stencil(uasd_ctx *ctx, int *has, int *radius3);

/* ---- Hamiltonian: the ham_t arrays (hamiltoniandatatype.f90:14-100) ---- */
/* aham (Natom) maps site -> Hamiltonian row (ham%aHam); NULL means row = site.
 * Rows with identical couplings are stored once on the device. */
int uasd_set_exchange(uasd_ctx *ctx, int max_no_neigh, int nham, const int *nlist,
                      const int *nlistsize, const double *ncoup, const int *aham);
/* dm_vect is (3,max_no_dmneigh,nHam) (ham%dm_vect). */
int uasd_set_dm(uasd_ctx *ctx, int max_no_dmneigh, int nham, const int *dmlist,
                const int *dmlistsize, const double *dm_vect, const int *aham);
/* ham%taniso (Natom), eaniso (3,Natom), kaniso (2,Natom), sb (Natom). */
int uasd_set_anisotropy(uasd_ctx *ctx, const int *taniso, const double *eaniso,
                        const double *kaniso, const double *sb);
/* external_field (3,Natom,Mensemble), or NULL to clear (calculatefields.f90:23-82). */
int uasd_set_external_field(uasd_ctx *ctx, const double *field);

/* ---- LLG parameters: lambda1_array, Temp_array, Landeg (Natom each) ---- */
int uasd_set_llg(uasd_ctx *ctx, const double *lambda1, const double *temp,
                 const double *landeg, double delta_t, double temprescale, double bn);
/* Temperature rescaling of the quantum heat bath (temprescale), used by the
 * LLG noise and the MC acceptance; uasd_set_llg also sets it. */
int uasd_set_temprescale(uasd_ctx *ctx, double temprescale);
/* Noise source; normals is (3,Natom,Mensemble) N(0,1) for UASD_NOISE_INJECTED
 * and is used by the next evolve call (both halves of the step). */
int uasd_set_noise(uasd_ctx *ctx, int mode, uint64_t seed, const double *normals);

/* ---- moments ---- */
/* emom (3,N,M) unit vectors and mmom (N,M); emomM = emom*mmom on the device. */
int uasd_set_moments(uasd_ctx *ctx, const double *emom, const double *mmom);
int uasd_get_moments(uasd_ctx *ctx, double *emom, double *emomM);

/* ---- drop-in entry points, host arrays in and out on every call ---- */
/* effective_field_full (hamiltonianactions.f90:108-252): beff from emomM,
 * and (each may be NULL) beff1 = pair + anisotropy field, beff2 = cubic*sb +
 * external field, energy in mRy as the reference's field routine reports it. */
int uasd_effective_field_host(uasd_ctx *ctx, const double *emomM, double *beff, double *beff1,
                              double *beff2, double *energy);
/* evolve_first (evolution.f90:38-273) for SDEalgh 1, 4, 5, 6, 7: reads beff,
 * emom, mmom; writes emom2, emomM (and, for Depondt, b2eff and the updated
 * emom).  Heun (evolve4p, heun_proper.f90:28-369) starts from emom2 and
 * writes the predictor into emom and emomM and b2eff = beff. */
int uasd_evolve_first_host(uasd_ctx *ctx, int sdealgh, const double *beff, double *emom,
                           double *emom2, double *emomM, double *b2eff);
/* evolve_second (evolution.f90:280-406) for SDEalgh 1, 4, 5, 6, 7; Heun
 * (evolve4f, heun_proper.f90:168-254) reads the predictor in emom and the
 * start in emom2 and writes the renormalised result into emom2. */
int uasd_evolve_second_host(uasd_ctx *ctx, int sdealgh, const double *beff,
                            const double *b2eff, double *emom, double *emom2);

/* ---- device-resident hot path ---- */
/* nsteps of the sd_mphase body (sd_driver.f90:668-765): field -> evolve_first
 * -> field -> evolve_second -> moment_update, fused into two kernels/step. */
int uasd_sd_steps(uasd_ctx *ctx, int sdealgh, int64_t nsteps);
/* Graph-captured stepping (no reference counterpart: the reference has no
 * launches to capture): with steps_per_graph = K > 0, uasd_sd_steps and
 * uasd_mc_sweeps replay a hipGraph of K steps (K sweeps) for every full run of
 * K, and launch the rest one by one; the numbers are the same either way.
 * The Philox counters then advance on the device.  0 (the default) disables. */
int uasd_set_graph_steps(uasd_ctx *ctx, int steps_per_graph);
/* Field of the resident moments into a host array (NULL skips the copy). */
int uasd_effective_field(uasd_ctx *ctx, double *beff, double *energy);
/* nsweeps of Metropolis ('M') or heat-bath ('H') updates (mc_evolve,
 * montecarlo.f90:44-273) over a colouring of the bond graph; temps (M) is the
 * temperature of each ensemble. */
int uasd_mc_sweeps(uasd_ctx *ctx, char mode, const double *temps, int64_t nsweeps,
                   uint64_t seed);
/* FP64 reductions: msum (3,M) = sum_i emomM(:,i,k); energy (M) or NULL. */
int uasd_measure(uasd_ctx *ctx, double *msum, double *energy);

/* ---- batches: replicas of uneven size as segments of one lattice ----
 * A batch concatenates R systems (their neighbour lists shifted by the site
 * offset of each) into one context; one launch then advances every replica.
 * offsets (nseg+1): segment s = sites [offsets[s], offsets[s+1]).  Without a
 * call the whole lattice is one segment.  Resets the accumulators. */
int uasd_set_segments(uasd_ctx *ctx, int nseg, const int *offsets);
/* FP64 per (segment, ensemble) sums: msum (3,M,nseg) = sum emomM over the
 * segment, energy (M,nseg) the field-routine energy in mRy, or NULL. */
int uasd_measure_segments(uasd_ctx *ctx, double *msum, double *energy);
/* Term-resolved energies of calc_energy (energy.f90:60-366 with do_lsf 'N'
 * and scalar exchange; the terms of the accelerated Hamiltonian):
 * ene (4,M,nseg), ene[4*(seg*M + k) + t] = energy per site in mRy of term t
 * (0 exchange "Exc", 1 anisotropy "Ani", 2 DM, 3 external field "Zeeman") in
 * ensemble k of segment seg (one segment = the whole lattice), i.e. the
 * reference's ene%ene_xc(k) ... times mub/mry.  Anisotropy type 7 counts
 * 0.5*(uniaxial + sb*cubic), as calc_energy does. */
int uasd_energy_terms(uasd_ctx *ctx, double *ene);
/* uasd_energy_terms without any device allocation (no reference counterpart:
 * for callers that cannot allocate, e.g. tight measurement loops or code
 * that must not call hipMalloc).  Same kernels, same results.  The partial
 * sums live in a fixed device array of 2^19 doubles (4 MiB) per device, held
 * until the process ends: (ceil(n_s / 4096) summed over segments + nseg) *
 * Mensemble * 4 must not exceed 2^19 (UASD_ERR_ARG otherwise).  The site-class
 * tables must already be on the device: after a Hamiltonian upload that
 * changes them, call uasd_energy_terms (or a field, step or sweep entry point)
 * once first (UASD_ERR_STATE otherwise).  Every call on one device shares the
 * array: calls that overlap on one device, from any context or thread, race. */
int uasd_energy_terms_noalloc(uasd_ctx *ctx, double *ene);
/* One cumulant sample per (segment, ensemble), kept on the device in FP64:
 * m = |sum emomM|/n and e = energy/n (mRy per site), as running means with
 * the linearly increasing weights of calc_and_print_cumulant
 * (prn_averages.f90:956-1000), each (segment, ensemble) its own sequence.
 * Asynchronous: no host copy, no synchronisation. */
int uasd_accumulate(uasd_ctx *ctx, int with_energy);
/* out (8,M,nseg): W, n, <m>, <m^2>, <m^4>, <e>, <e^2>, 0 per (segment,
 * ensemble); U4 = 1 - <m^4>/(3<m^2>^2).  nsamples may be NULL. */
int uasd_get_accumulators(uasd_ctx *ctx, double *out, int64_t *nsamples);
int uasd_reset_accumulators(uasd_ctx *ctx);

/* ---- colouring used by the MC sweeps ---- */
int uasd_get_coloring(uasd_ctx *ctx, int *ncolors, int *color_of_site);

/* ---- per-kernel timing with HIP events on the context stream ----
 * enable = n > 0 records a pair of events around one launch in n of each
 * kind (1: every launch) and around every uasd_sd_steps / uasd_mc_sweeps
 * call; -1 times the calls only, with no events between kernels; 0 switches
 * timing off.  Events between kernels cost time of their own, so a sampled
 * run keeps the stream close to an untimed one. */
int uasd_set_kernel_timing(uasd_ctx *ctx, int enable);
/* kind: 0 predictor, 1 corrector, 2 field, 3 MC colour pass, 4 LLG graph
 * replay (K steps), 5 MC graph replay (K sweeps), 6 a whole uasd_sd_steps or
 * uasd_mc_sweeps call */
int uasd_get_kernel_timing(uasd_ctx *ctx, int kind, int64_t *launches, double *total_ms);

/* Bytes per neighbour-list entry the LLG kernels read under the current
 * tuning (no reference counterpart): 2 when the list has packed int16
 * entries (every cell difference fits; periodic lattices numbered cell by
 * cell) and the tuning uses them, else 4. */
int uasd_index_bytes(uasd_ctx *ctx, int *exchange, int *dm);
/* The kernel the fused LLG step launches under the current tuning (no
 * reference counterpart): *tile = the LDS-tiled stencil shape k (1..15), or 0
 * for the list-gather kernels (also when a requested shape does not fit the
 * context: no stencil, LDS, thread count). */
int uasd_llg_tile(uasd_ctx *ctx, int *tile);

/* ---- launch shapes of the gathering kernels (no reference counterpart:
 * the reference's OpenMP loops have no launch shape) ----
 * kind: 0 fused LLG step, 1 field, 2 MC colour pass.  gather = neighbours
 * whose loads are issued together (2, 4, 8, 16); block = workgroup size
 * (64, 128, 256), plus 0x10000 to read the neighbour lists with
 * non-temporal loads, 0x20000 to read int32 entries where packed int16
 * entries exist, and (k << 18) for the LDS-tiled stencil kernel with tile
 * shape k = 1..15 (kind 0 only; needs uasd_set_lattice_dims and a stencil):
 * cells (x, y, z) = (8,8,4) (16,8,2) (32,4,2) (16,4,4) (16,16,1) (32,8,1)
 * (64,4,1) (8,4,4) (4,4,4) (8,8,2) (16,8,1) (8,8,1) (16,8,4) (8,16,4) (8,8,8),
 * one thread per site (cells x basis atoms <= 1024), work units (tile,
 * ensemble); plus 0x400000 for resident workgroups that each loop over a run
 * of adjacent units, with the next unit's halo loads in flight during the
 * current unit's computation;
 * plus 0x800000 (any kind, for the whole context) to gather DM neighbours
 * separately even where every site's DM bonds are its exchange bonds (at
 * most 8), which the list kernels otherwise gather once for both sums;
 * plus 0x1000000 (kind 2 only) for two adjacent ensembles of a site per
 * thread, sharing its neighbour entries and couplings (M even; one per
 * thread otherwise).  Results do not depend on any of them. */
int uasd_set_tuning(uasd_ctx *ctx, int kind, int gather, int block);
int uasd_get_tuning(uasd_ctx *ctx, int kind, int *gather, int *block);
/* Autotuner: times every shape on the context's own lattice and state
 * (steps LLG steps / sweeps MC sweeps per candidate), keeps the fastest and
 * restores the spins and step counters, so a run is unchanged by tuning.
 * best_ms (may be NULL) receives the fastest candidate's time. */
int uasd_autotune_llg(uasd_ctx *ctx, int sdealgh, int steps, double *best_ms);
int uasd_autotune_mc(uasd_ctx *ctx, char mode, const double *temps, int sweeps,
                     double *best_ms);
/* Autotuner cache (no reference counterpart).  enable = 1 keeps every
 * autotuner winner for the process, keyed by the workload (kernel family,
 * solver or mode, precision, sizes, list widths, terms, site order, lattice
 * dimensions), and later uasd_autotune_* calls on the same workload take the
 * kept shape without timing (best_ms is then the kept time).  path (may be
 * NULL) names a text file whose entries are loaded now and to which new
 * winners are appended.  enable = 0 switches the cache off and forgets it
 * (the default: every call times). */
int uasd_set_tune_cache(int enable, const char *path);

/* ---- host set-up: the reference neighbour map, bit-exact ----
 * setup_nm + setup_neighbour_hamiltonian (neighbourmap.f90:32-359,
 * hamiltonianinit.f90:985-1157) for a Bravais lattice with basis.
 * Shells as read_exchange leaves them: nn (NT), redcoord (NT,max_no_shells,3),
 * nntype (NT,max_no_shells) or NULL, jc (hdim,NT,max_no_shells,nchmax,nchmax).
 * Pass nlist == NULL to query max_no_neigh only.  ammom (NA,nchmax); achem
 * (Natom) 1-based chemistry or NULL. */
int uasd_build_pair_list(int NA, int NT, int N1, int N2, int N3, const double *C1,
                         const double *C2, const double *C3, const double *bas,
                         const int *atype, const char *bc, int sym, int max_no_shells,
                         const int *nn, const double *redcoord, const int *nntype,
                         int hdim, int nchmax, const double *jc, const double *ammom,
                         const int *achem, int *max_no_neigh, int *nlist, int *nlistsize,
                         double *ncoup);

#ifdef __cplusplus
}
#endif
#endif
