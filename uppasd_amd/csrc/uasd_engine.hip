// uasd_engine.hip -- context, uploads and the C-ABI of libuasd_hip.so.
//
// A context mirrors the state the reference keeps in module arrays (ham_t,
// MomentData, FieldData, the LLG inputs) on one device, in the layout
// described in uasd_kernels.hpp.  Uploads convert the Fortran layout
// (3,Natom,Mensemble) to quads {x, y, z, w} indexed site*M+ens (w = mmom for
// spins) and compress the per-site coupling rows of ham%ncoup / ham%dm_vect /
// the anisotropy arrays to their distinct rows.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/uasd_hip.h"
#include "uasd_launch.hpp"

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg) {
  g_err = msg;
  return code;
}

#define HIP_TRY(expr)                                                                  \
  do {                                                                                 \
    hipError_t _e = (expr);                                                            \
    if (_e != hipSuccess)                                                              \
      return fail(_e == hipErrorOutOfMemory ? UASD_ERR_NOMEM : UASD_ERR_HIP,          \
                  std::string(#expr) + ": " + hipGetErrorString(_e));                  \
  } while (0)

// Owned device allocation.
struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  ~DevBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);  // teardown: no caller to report to
    p = nullptr;
    bytes = 0;
  }
  hipError_t ensure(size_t n) {
    if (n <= bytes && p) return hipSuccess;
    release();
    if (n == 0) return hipSuccess;
    hipError_t e = hipMalloc(&p, n);
    if (e == hipSuccess) bytes = n;
    else p = nullptr;
    return e;
  }
  template <typename T>
  T *as() const { return static_cast<T *>(p); }
};

// Hash of a byte string (FNV-1a) for deduplicating rows.
struct RowKey {
  std::string bytes;
  bool operator==(const RowKey &o) const { return bytes == o.bytes; }
};
struct RowKeyHash {
  size_t operator()(const RowKey &k) const {
    uint64_t h = 1469598103934665603ULL;
    for (unsigned char c : k.bytes) h = (h ^ c) * 1099511628211ULL;
    return (size_t)h;
  }
};

struct Timing {
  int64_t launches = 0;
  double ms = 0.0;
};

// Fixed device scratch of uasd_energy_terms_noalloc: partial sums per
// (chunk, ensemble) and the sums per (segment, ensemble), 4 doubles each.
constexpr size_t kScratchDoubles = (size_t)1 << 19;  // 4 MiB per device
constexpr int kMaxDevices = 64;

}  // namespace

// One copy per device, held until the process ends.  On the host the name is
// only a stand-in for the symbol: kernels and copies take the device address
// that scratch_init looks up, never `scratch` itself.
static __device__ __align__(256) double scratch[kScratchDoubles];

namespace {

std::mutex g_scratch_mu;
double *g_scratch_ptr[kMaxDevices] = {};

// Device address of `device`'s copy of the scratch array, looked up once per
// device; the caller has made `device` current.
int scratch_init(int device) {
  if (device < 0 || device >= kMaxDevices)
    return fail(UASD_ERR_UNSUPPORTED, "device index beyond the scratch table (64 devices)");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[device]) return UASD_OK;
  void *p = nullptr;
  HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  if (!p) return fail(UASD_ERR_HIP, "hipGetSymbolAddress returned a null scratch address");
  g_scratch_ptr[device] = static_cast<double *>(p);
  return UASD_OK;
}

double *scratch_of(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  return g_scratch_ptr[device];
}

}  // namespace

// ----------------------------------------------------------------- context
struct uasd_ctx {
  int N = 0, M = 0, prec = UASD_FP64, device = 0;
  // device site order: slot d holds reference site site_of(d) (see HamDev)
  int order_na = 1;
  int order_nc = 0;  // sites per basis atom (N when order_na = 1)
  hipStream_t own_stream = nullptr, stream = nullptr;
  size_t rsize = 8;  // sizeof(R)

  // host-side descriptors kept for class building / colouring
  int W = 0, ex_rows = 0;
  std::vector<int> ex_row_of_site;  // (N)
  std::vector<int> nbr_host;        // (W*N) slot-major, 0-based
  std::vector<int> ex_len_host;     // (ex_rows)
  bool ex_full = false;             // every exchange row holds W neighbours
  int Wd = 0, dm_rows = 0;
  std::vector<int> dm_row_of_site, dnbr_host, dm_len_host;
  std::vector<int> an_row_of_site;
  int an_rows = 0;
  int ext_mode = 0;
  std::vector<double> ext_u_host;  // (3*M)

  // device buffers
  DevBuf nbr, ex_len, cpl;
  DevBuf dnbr, dm_len, dv;
  DevBuf nbr16, dnbr16;  // int16 index differences when they all fit (else empty)
  bool use16 = false, use16_dm = false;
  DevBuf an_type, an_vals;
  DevBuf cls, classes;
  DevBuf ext_u, ext_full;
  DevBuf spins;   // quad arrays: e, p, b2
  DevBuf field;   // beff quads for the unfused path
  DevBuf scratch; // emomM quads (w = 1) for host field calls
  DevBuf noise;   // injected N(0,1) quads
  DevBuf nsave;   // fused steps: deviates drawn by the predictor for the corrector
  DevBuf llgarr;  // lambda, temp, lande (double, N each)
  DevBuf energy;  // per q energy terms
  DevBuf partial, redout;
  DevBuf mcbuf;   // per-ensemble delta/beta/hbeta/ext + colour site lists

  bool have_ex = false, have_state = false, have_llg = false;
  bool classes_dirty = true;
  bool dm_shared = false;  // DM bonds = exchange bonds at every site (fused gather)
  bool tune_no_fuse = false;  // keep the separate DM gather (a tuning choice)

  // LLG parameters
  double dt = 1e-16, temprescale = 1.0, bn = 1.0;
  double lam0 = 0, temp0 = 0, lande0 = 1;
  bool lam_uniform = true, temp_uniform = true, lande_uniform = true;
  bool m_uniform = false;  // every moment magnitude equal (m0), as uploaded
  double m0 = 0;
  int noise_mode = UASD_NOISE_NONE;
  uint64_t seed = 0;
  uint64_t step = 0;      // LLG step counter (Philox counter)
  uint64_t mc_sweep = 0;  // MC sweep counter

  // launch shapes of the gathering kernels: 0 fused LLG, 1 field, 2 MC
  uasd::Tune tune[3];

  // lattice dimensions (reference site i = b + NA*(x + N1*(y + N2*z))) and
  // the translation-invariant stencil of the lists, when there is one
  int lat_na = 0, lat_n1 = 0, lat_n2 = 0, lat_n3 = 0;
  int64_t lds_max = 64 * 1024;  // LDS one workgroup may use (opt-in maximum)
  struct Stencil {
    bool ok = false;
    int W = 0, rx = 0, ry = 0, rz = 0;
    std::vector<int> bp, dx, dy, dz;  // (NA*W)
  } st_ex, st_dm;
  bool stencil_ok = false, stencil_dirty = true;
  DevBuf tile_off;
  int tile_off_shape = 0;
  std::vector<int> cls_host;       // site class per device slot
  int64_t cls_version = 0;         // bumped by every class rebuild
  int64_t ucls_version = -1;       // cls_version the per-basis check below is for
  int ucls = 0, cls_b[8] = {0};

  // segments (replicas of a batch): site ranges, their chunks, accumulators
  int nseg = 0;
  std::vector<int> seg_off;  // (nseg+1)
  int nchunks = 0;
  DevBuf seg_dev;            // int2 chunks (nchunks), int seg_chunk (nseg+1), int seg_n (nseg)
  DevBuf seg_partial, seg_out, acc;
  int64_t acc_samples = 0;

  // captured step graphs: replayed for runs of graph_steps LLG steps / MC
  // sweeps; the Philox counters then live in ctr (0 LLG step, 1 MC sweep).
  int graph_steps = 0;
  uint64_t gen = 0;           // bumped whenever a captured graph may be stale
  uint64_t graphs_gen = 0;
  DevBuf ctr;
  bool ctr_valid[2] = {false, false};
  uint64_t ctr_val[2] = {0, 0};
  bool capturing = false;
  struct Graph {
    hipGraph_t g = nullptr;
    hipGraphExec_t exec = nullptr;
  };
  std::map<std::tuple<int, int, int, uint64_t>, Graph> graphs;  // kind, alg/mode, K, seed

  // colouring
  int ncolors = 0;
  std::vector<int> color_of_site, color_ptr, color_sites;
  // per colour: its slots are color_start + i * color_stride (stride 0: not
  // an arithmetic progression, the kernels read color_sites)
  std::vector<int> color_start, color_stride;
  bool coloring_dirty = true;

  // Per-kernel timing: an event pair is recorded around each launch without
  // a host wait; pairs are resolved when the totals are read.
  bool timing = false;
  int timing_every = 1;       // time one launch in timing_every of each kind
  bool timing_region = false; // time every uasd_sd_steps / uasd_mc_sweeps call (kind 6)
  int64_t timing_tick[7] = {};
  Timing timers[7];  // 0 predictor, 1 corrector, 2 field, 3 MC, 4 LLG graph, 5 MC graph, 6 call
  hipEvent_t region_cur = nullptr;
  std::vector<hipEvent_t> ev_pool;  // free events
  struct Pending {
    hipEvent_t a, b;
    int kind;
  };
  std::vector<Pending> pending;
  hipEvent_t cur = nullptr;
};

void build_stencil(uasd_ctx *c);
bool tile_fits(const uasd_ctx *c, int k);
int make_tile(uasd_ctx *c, int k, uasd::TileDev &t);

namespace {

void destroy_graphs(uasd_ctx *c);
int sd_steps_run(uasd_ctx *c, int alg, int64_t nsteps);
int mc_sweeps_run(uasd_ctx *c, char mode, const uasd::MCParams &mp, int64_t nsweeps,
                  uint64_t seed);

// Reference site held by device slot d, and the slot of reference site i.
inline int site_of(const uasd_ctx *c, int d) {
  return c->order_na == 1 ? d : (d % c->order_nc) * c->order_na + d / c->order_nc;
}
inline int slot_of(const uasd_ctx *c, int i) {
  return c->order_na == 1 ? i : (i % c->order_na) * c->order_nc + i / c->order_na;
}

// Bits that hold the basis shift in an int16 list entry.
int idx_kb(const uasd_ctx *c) {
  int kb = 0;
  while ((1 << kb) < c->order_na) ++kb;
  return kb;
}


// Quad buffers: one Quad<R> = {x, y, z, w} per q = site*M + ens.
size_t quad_bytes(const uasd_ctx *c) { return (size_t)c->N * c->M * 4 * c->rsize; }

// spins holds three quad arrays: 0 e (emom, w = mmom), 1 p (emom2, w = mmom),
// 2 b2 (b2eff, w unused).
template <typename R>
uasd::Quad<R> *quads(const uasd_ctx *c, const DevBuf &b, int k = 0) {
  return reinterpret_cast<uasd::Quad<R> *>(static_cast<char *>(b.p) + (size_t)k * quad_bytes(c));
}

template <typename R>
uasd::HamDev<R> ham_view(const uasd_ctx *c) {
  uasd::HamDev<R> h{};
  h.N = c->N;
  h.M = c->M;
  h.mdiv = uasd::make_fast_div((uint32_t)c->M);
  h.W = c->W;
  h.ex_rows = c->ex_rows;
  h.ex_full = c->ex_full ? 1 : 0;
  h.nbr = c->nbr.as<int>();
  h.nbr16 = c->use16 ? c->nbr16.as<int16_t>() : nullptr;
  h.ex_len = c->ex_len.as<int>();
  h.cpl = c->cpl.as<R>();
  h.Wd = c->Wd;
  h.dm_rows = c->dm_rows;
  h.dnbr = c->dnbr.as<int>();
  h.dnbr16 = c->use16_dm ? c->dnbr16.as<int16_t>() : nullptr;
  h.dm_len = c->dm_len.as<int>();
  h.dm_shared = c->dm_shared && !c->tune_no_fuse ? 1 : 0;
  if (c->Wd > 0) {
    size_t n = (size_t)c->Wd * c->dm_rows;
    h.dvx = c->dv.as<R>();
    h.dvy = h.dvx + n;
    h.dvz = h.dvy + n;
  }
  if (c->an_rows > 0) {
    h.an_type = c->an_type.as<int>();
    const R *v = c->an_vals.as<R>();
    size_t n = (size_t)c->an_rows;
    h.an_ex = v;
    h.an_ey = v + n;
    h.an_ez = v + 2 * n;
    h.an_k1 = v + 3 * n;
    h.an_k2 = v + 4 * n;
    h.an_sb = v + 5 * n;
  }
  h.cls = c->cls.as<int>();
  h.classes = c->classes.as<uasd::SiteClass>();
  h.ext_mode = c->ext_mode;
  h.order_na = c->order_na;
  h.idx_kb = idx_kb(c);
  h.order_nc = c->order_na == 1 ? c->N : c->order_nc;
  h.ext_u = c->ext_u.as<R>();
  if (c->ext_mode == 2) h.ext_q = quads<R>(c, c->ext_full);
  return h;
}

template <typename R>
int upload_converted(hipStream_t s, DevBuf &dst, const std::vector<double> &src) {
  std::vector<R> tmp(src.begin(), src.end());
  HIP_TRY(dst.ensure(std::max<size_t>(1, tmp.size()) * sizeof(R)));
  if (!tmp.empty())
    HIP_TRY(hipMemcpyAsync(dst.p, tmp.data(), tmp.size() * sizeof(R), hipMemcpyHostToDevice, s));
  HIP_TRY(hipStreamSynchronize(s));
  return UASD_OK;
}

int upload_real(uasd_ctx *c, DevBuf &dst, const std::vector<double> &src) {
  return c->prec == UASD_FP64 ? upload_converted<double>(c->stream, dst, src)
                              : upload_converted<float>(c->stream, dst, src);
}

template <typename T>
int upload_raw(uasd_ctx *c, DevBuf &dst, const std::vector<T> &src) {
  HIP_TRY(dst.ensure(std::max<size_t>(1, src.size()) * sizeof(T)));
  if (!src.empty())
    HIP_TRY(hipMemcpyAsync(dst.p, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice,
                           c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return UASD_OK;
}

// (3,N,M) Fortran array (+ w per (N,M) Fortran array, or a constant) -> quads
// in device site order
std::vector<double> to_quads(const uasd_ctx *c, const double *a, const double *w, double wconst) {
  const int N = c->N, M = c->M;
  std::vector<double> out((size_t)4 * N * M);
  for (int k = 0; k < M; ++k)
    for (int i = 0; i < N; ++i) {
      const double *v = a + 3 * ((size_t)i + (size_t)N * k);
      const size_t q = (size_t)slot_of(c, i) * M + k;
      out[4 * q] = v[0];
      out[4 * q + 1] = v[1];
      out[4 * q + 2] = v[2];
      out[4 * q + 3] = w ? w[(size_t)i + (size_t)N * k] : wconst;
    }
  return out;
}

// upload host quads (double) into quad array k of buffer b (converted to R)
int upload_quads(uasd_ctx *c, DevBuf &b, int k, const std::vector<double> &hq) {
  HIP_TRY(b.ensure(std::max(b.bytes, (size_t)(k + 1) * quad_bytes(c))));
  if (c->prec == UASD_FP64) {
    HIP_TRY(hipMemcpyAsync(quads<double>(c, b, k), hq.data(), hq.size() * 8,
                           hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
  } else {
    std::vector<float> f(hq.begin(), hq.end());
    HIP_TRY(hipMemcpyAsync(quads<float>(c, b, k), f.data(), f.size() * 4, hipMemcpyHostToDevice,
                           c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
  }
  return UASD_OK;
}

int download_quads(uasd_ctx *c, const DevBuf &b, int k, std::vector<double> &hq) {
  const size_t n = (size_t)4 * c->N * c->M;
  if (c->prec == UASD_FP64) {
    hq.resize(n);
    HIP_TRY(hipMemcpyAsync(hq.data(), quads<double>(c, b, k), n * 8, hipMemcpyDeviceToHost,
                           c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
  } else {
    std::vector<float> f(n);
    HIP_TRY(hipMemcpyAsync(f.data(), quads<float>(c, b, k), n * 4, hipMemcpyDeviceToHost,
                           c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    hq.assign(f.begin(), f.end());
  }
  return UASD_OK;
}

// quads (device site order) -> Fortran (3,N,M); times_w gives emomM = xyz*w
void from_quads(const uasd_ctx *c, const std::vector<double> &hq, double *out, bool times_w) {
  const int N = c->N, M = c->M;
  for (int k = 0; k < M; ++k)
    for (int i = 0; i < N; ++i) {
      const size_t q = (size_t)slot_of(c, i) * M + k;
      double *v = out + 3 * ((size_t)i + (size_t)N * k);
      const double w = hq[4 * q + 3];
      for (int d = 0; d < 3; ++d) v[d] = times_w ? hq[4 * q + d] * w : hq[4 * q + d];
    }
}

// Whether every site's DM neighbours are its exchange neighbours, in the same
// order and number, at most kFusedW of them (or there is no DM term): the
// field kernels then gather each neighbour once for both sums.
bool dm_lists_shared(const uasd_ctx *c) {
  if (c->Wd == 0 || c->dm_row_of_site.empty()) return false;
  if (c->W > uasd::kFusedW) return false;
  const int N = c->N;
  for (int i = 0; i < N; ++i) {
    const int dr = c->dm_row_of_site[i];
    if (dr < 0) continue;
    const int er = c->ex_row_of_site.empty() ? 0 : c->ex_row_of_site[i];
    const int L = c->ex_len_host[er];
    if (c->dm_len_host[dr] != L) return false;
    for (int j = 0; j < L; ++j)
      if (c->dnbr_host[(size_t)j * N + i] != c->nbr_host[(size_t)j * N + i]) return false;
  }
  return true;
}

// may_alloc = false: fail rather than grow the device tables (the no-alloc
// entry point), i.e. when the class tables outgrow the buffers of the last
// build.
int rebuild_classes(uasd_ctx *c, bool may_alloc = true) {
  if (!c->classes_dirty) return UASD_OK;
  std::map<std::tuple<int, int, int>, int> ids;
  std::vector<int> cls(c->N);
  std::vector<uasd::SiteClass> table;
  for (int i = 0; i < c->N; ++i) {
    int er = c->ex_row_of_site.empty() ? 0 : c->ex_row_of_site[i];
    int dr = c->dm_row_of_site.empty() ? -1 : c->dm_row_of_site[i];
    int ar = c->an_row_of_site.empty() ? -1 : c->an_row_of_site[i];
    auto key = std::make_tuple(er, dr, ar);
    auto it = ids.find(key);
    if (it == ids.end()) {
      it = ids.emplace(key, (int)table.size()).first;
      table.push_back(uasd::SiteClass{er, dr, ar, 0});
    }
    cls[i] = it->second;
  }
  c->dm_shared = dm_lists_shared(c);
  if (!may_alloc && (!c->cls.p || !c->classes.p || c->cls.bytes < cls.size() * sizeof(int) ||
                     c->classes.bytes < table.size() * sizeof(uasd::SiteClass)))
    return fail(UASD_ERR_STATE,
                "the site-class tables are not on the device yet; call uasd_energy_terms (or "
                "any field, step or sweep entry point) once after the last Hamiltonian upload");
  int r = upload_raw(c, c->cls, cls);
  if (r) return r;
  r = upload_raw(c, c->classes, table);
  if (r) return r;
  c->cls_host = cls;
  c->cls_version += 1;
  c->classes_dirty = false;
  c->gen += 1;
  return UASD_OK;
}

// Distinct rows of a per-site table.  row_ref[i] is the Hamiltonian row of
// site i; `len` and `vals(width*hdim)` describe that row.
struct RowSet {
  std::vector<int> row_of_site;
  std::vector<int> len;
  std::vector<std::vector<double>> vals;
};

RowSet dedupe_rows(int N, const std::vector<int> &row_ref, const int *lenarr, const double *vals,
                   int width, int hdim) {
  RowSet rs;
  rs.row_of_site.resize(N);
  std::unordered_map<int, int> by_ref;  // Hamiltonian row -> distinct row
  std::unordered_map<RowKey, int, RowKeyHash> by_content;
  for (int i = 0; i < N; ++i) {
    int ref = row_ref[i];
    auto f = by_ref.find(ref);
    if (f != by_ref.end()) {
      rs.row_of_site[i] = f->second;
      continue;
    }
    int L = lenarr[ref];
    RowKey key;
    key.bytes.assign(reinterpret_cast<const char *>(&L), sizeof(int));
    const double *v = vals + (size_t)ref * width * hdim;
    key.bytes.append(reinterpret_cast<const char *>(v), sizeof(double) * (size_t)L * hdim);
    auto g = by_content.find(key);
    int id;
    if (g == by_content.end()) {
      id = (int)rs.len.size();
      by_content.emplace(std::move(key), id);
      rs.len.push_back(L);
      rs.vals.emplace_back(v, v + (size_t)L * hdim);
    } else {
      id = g->second;
    }
    by_ref.emplace(ref, id);
    rs.row_of_site[i] = id;
  }
  return rs;
}

int check_ctx(uasd_ctx *c) {
  if (!c) return fail(UASD_ERR_ARG, "null context");
  return UASD_OK;
}

hipEvent_t take_event(uasd_ctx *c) {
  if (!c->ev_pool.empty()) {
    hipEvent_t e = c->ev_pool.back();
    c->ev_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

void tic(uasd_ctx *c, int kind) {
  if (!c->timing || c->capturing) return;
  if (c->timing_tick[kind]++ % c->timing_every != 0) return;
  c->cur = take_event(c);
  // a start event that could not be recorded times nothing: back to the pool
  if (c->cur && hipEventRecord(c->cur, c->stream) != hipSuccess) {
    c->ev_pool.push_back(c->cur);
    c->cur = nullptr;
  }
}

int toc(uasd_ctx *c, int kind) {
  if (!c->timing || c->capturing || !c->cur) return UASD_OK;
  hipEvent_t b = take_event(c);
  if (!b) return fail(UASD_ERR_HIP, "hipEventCreate failed");
  HIP_TRY(hipEventRecord(b, c->stream));
  c->pending.push_back({c->cur, b, kind});
  c->cur = nullptr;
  return UASD_OK;
}

// Event pair around a whole stepping call (kind 6), on the context stream.
int region_tic(uasd_ctx *c) {
  if (!c->timing_region || c->capturing) return UASD_OK;
  c->region_cur = take_event(c);
  if (!c->region_cur) return fail(UASD_ERR_HIP, "hipEventCreate failed");
  const hipError_t err = hipEventRecord(c->region_cur, c->stream);
  if (err != hipSuccess) {  // no region opened: the event goes back to the pool
    c->ev_pool.push_back(c->region_cur);
    c->region_cur = nullptr;
  }
  HIP_TRY(err);
  return UASD_OK;
}

int region_toc(uasd_ctx *c) {
  if (!c->region_cur) return UASD_OK;
  hipEvent_t b = take_event(c);
  if (!b) return fail(UASD_ERR_HIP, "hipEventCreate failed");
  HIP_TRY(hipEventRecord(b, c->stream));
  c->pending.push_back({c->region_cur, b, 6});
  c->region_cur = nullptr;
  return UASD_OK;
}

int resolve_timing(uasd_ctx *c) {
  for (auto &p : c->pending) {
    HIP_TRY(hipEventSynchronize(p.b));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, p.a, p.b));
    c->timers[p.kind].launches += 1;
    c->timers[p.kind].ms += ms;
    c->ev_pool.push_back(p.a);
    c->ev_pool.push_back(p.b);
  }
  c->pending.clear();
  return UASD_OK;
}

template <typename R>
uasd::LLGDev<R> llg_view(uasd_ctx *c) {
  uasd::LLGDev<R> p{};
  const double *arr = c->llgarr.as<double>();
  p.lam = c->lam_uniform ? nullptr : arr;
  p.temp = c->temp_uniform ? nullptr : arr + c->N;
  p.lande = c->lande_uniform ? nullptr : arr + 2 * (size_t)c->N;
  p.lam0 = c->lam0;
  p.temp0 = c->temp0;
  p.lande0 = c->lande0;
  p.dt = c->dt;
  p.temprescale = c->temprescale;
  p.bn = c->bn;
  p.noise_mode = c->noise_mode;
  // a uniform T = 0 scales every deviate by zero: draw none
  if (c->temp_uniform && c->temp0 == 0.0 && p.noise_mode == UASD_NOISE_PHILOX)
    p.noise_mode = UASD_NOISE_NONE;
  p.seed = c->seed;
  p.step = c->step;
  if (c->noise_mode == UASD_NOISE_INJECTED) p.noise = quads<R>(c, c->noise);
  p.lc0 = uasd::lam_const<R>(c->lam0, c->dt, c->bn);
  if (c->lam_uniform && c->temp_uniform && c->m_uniform) {
    const double m = (double)(R)c->m0;  // as the kernels read it from the quads
    p.sig_uniform = 1;
    p.sigk0 = uasd::midpoint_sigma<R>(p.lc0.Dk, m, c->temp0, c->temprescale);
    p.sigp0 = uasd::depondt_sigma<R>(p.lc0.Dp, m, c->temp0, c->temprescale);
    p.sigh0 = uasd::heun_sigma<R>(p.lc0.Dh, m, c->temp0, c->temprescale, c->dt);
  }
  return p;
}

int ready_for_llg(uasd_ctx *c) {
  if (!c->have_ex) return fail(UASD_ERR_STATE, "exchange not set (uasd_set_exchange)");
  if (!c->have_state) return fail(UASD_ERR_STATE, "moments not set (uasd_set_moments)");
  if (!c->have_llg) return fail(UASD_ERR_STATE, "LLG parameters not set (uasd_set_llg)");
  if (c->noise_mode == UASD_NOISE_INJECTED && !c->noise.p)
    return fail(UASD_ERR_STATE, "injected noise selected but no normals uploaded");
  if (c->noise_mode == UASD_NOISE_PHILOX) {
    void *old = c->nsave.p;
    HIP_TRY(c->nsave.ensure((size_t)c->N * c->M * 4 * c->rsize));
    if (c->nsave.p != old) c->gen += 1;  // captured graphs hold the old address
  }
  return rebuild_classes(c);
}

// Quad arrays in c->spins: 0 e (emom), 1 p (emom2), 2 b2 (b2eff).
// step_off >= 0: inside a graph capture, the step is *ctr[0] + step_off.
bool sde_supported(int alg) {
  return alg == UASD_SDE_MIDPOINT || alg == UASD_SDE_HEUN || alg == UASD_SDE_DEPONDT ||
         alg == UASD_SDE_SIB || alg == UASD_SDE_SPHMID;
}

// solvers that keep b2eff between the halves
bool sde_b2eff(int alg) { return alg == UASD_SDE_HEUN || alg == UASD_SDE_DEPONDT; }

template <typename R>
int run_llg_half(uasd_ctx *c, int alg, int half, bool fused, bool write_e,
                 int64_t step_off = -1) {
  uasd::HamDev<R> h = ham_view<R>(c);
  uasd::LLGDev<R> prm = llg_view<R>(c);
  if (step_off >= 0) {
    prm.step = (uint64_t)step_off;
    prm.step_dev = c->ctr.as<uint64_t>();
  }
  // a fused predictor/corrector pair draws the deviates once: the predictor
  // stores them, the corrector of the same step reads them back
  if (fused && prm.noise_mode == UASD_NOISE_PHILOX) prm.nsave = c->nsave.as<uasd::Quad<R>>();
  uasd::Quad<R> *e = quads<R>(c, c->spins, 0), *p = quads<R>(c, c->spins, 1),
                *b2 = quads<R>(c, c->spins, 2);
  const uasd::Quad<R> *beff = fused ? nullptr : quads<R>(c, c->field);
  uasd::Quad<R> *out = write_e ? e : p;
  if (fused && c->tune[0].tile > 0) {
    build_stencil(c);
    if (tile_fits(c, c->tune[0].tile)) {
      uasd::TileDev t;
      if (int r = make_tile(c, c->tune[0].tile, t)) return r;
      tic(c, half);
      HIP_TRY(uasd::launch_llg_tile<R>(c->stream, t, c->tune[0].pipe, alg, half, h, prm, e, p, b2,
                                       out));
      return toc(c, half);
    }
  }
  tic(c, half);
  HIP_TRY(uasd::launch_llg<R>(c->stream, c->tune[0], alg, half, fused, h, prm, e, p, b2, beff,
                              out));
  return toc(c, half);
}

// parts: also write beff1 and beff2 into quad arrays 1 and 2 of c->field.
template <typename R>
int run_field(uasd_ctx *c, bool from_scratch, bool want_energy, bool parts = false) {
  uasd::HamDev<R> h = ham_view<R>(c);
  const uasd::Quad<R> *src = from_scratch ? quads<R>(c, c->scratch) : quads<R>(c, c->spins, 0);
  HIP_TRY(c->field.ensure(std::max(c->field.bytes, (parts ? 3 : 1) * quad_bytes(c))));
  double *en = nullptr;
  if (want_energy) {
    HIP_TRY(c->energy.ensure((size_t)c->N * c->M * sizeof(double)));
    en = c->energy.as<double>();
  }
  tic(c, 2);
  HIP_TRY(uasd::launch_field<R>(c->stream, c->tune[1], h, src, quads<R>(c, c->field), en,
                                parts ? quads<R>(c, c->field, 1) : nullptr));
  return toc(c, 2);
}

// FP64 sums per ensemble of emomM (and of the per-q energy terms written by
// the field kernel when with_energy).
template <typename R>
int reduce_moments(uasd_ctx *c, bool with_energy, bool from_scratch, std::vector<double> &out) {
  const int nv = with_energy ? 4 : 3;
  const int nblocks = std::max(1, std::min(1024, (c->N + 4095) / 4096));
  HIP_TRY(c->partial.ensure((size_t)c->M * nblocks * nv * sizeof(double)));
  HIP_TRY(c->redout.ensure((size_t)c->M * nv * sizeof(double)));
  const uasd::Quad<R> *s = from_scratch ? quads<R>(c, c->scratch) : quads<R>(c, c->spins, 0);
  HIP_TRY(uasd::launch_reduce<R>(c->stream, c->N, c->M, s,
                                 with_energy ? c->energy.as<double>() : nullptr, nv,
                                 c->partial.as<double>(), nblocks, c->redout.as<double>()));
  out.resize((size_t)c->M * nv);
  HIP_TRY(hipMemcpyAsync(out.data(), c->redout.p, out.size() * sizeof(double),
                         hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return UASD_OK;
}

int build_coloring(uasd_ctx *c) {
  if (!c->coloring_dirty) return UASD_OK;
  const int N = c->N;
  // undirected bond graph from exchange and DM lists
  std::vector<std::vector<int>> adj(N);
  auto add_list = [&](const std::vector<int> &row_of_site, const std::vector<int> &len,
                      const std::vector<int> &nb, int W) {
    if (W == 0 || row_of_site.empty()) return;
    for (int i = 0; i < N; ++i) {
      int L = len[row_of_site[i]];
      for (int j = 0; j < L; ++j) {
        int n = nb[(size_t)j * N + i];
        if (n == i) continue;
        adj[i].push_back(n);
        adj[n].push_back(i);
      }
    }
  };
  add_list(c->ex_row_of_site, c->ex_len_host, c->nbr_host, c->W);
  add_list(c->dm_row_of_site, c->dm_len_host, c->dnbr_host, c->Wd);
  std::vector<int> col(N, -1);
  int ncol = 0;
  std::vector<int> mark;
  // greedy colouring visiting the sites in reference order, so the colour
  // classes (and hence MC chains) do not depend on the device site order
  for (int h = 0; h < N; ++h) {
    const int i = slot_of(c, h);
    for (int n : adj[i])
      if (col[n] >= 0) {
        if ((int)mark.size() <= col[n]) mark.resize(col[n] + 1, -1);
        mark[col[n]] = h;
      }
    int k = 0;
    while (k < (int)mark.size() && mark[k] == h) ++k;
    col[i] = k;
    ncol = std::max(ncol, k + 1);
  }
  for (int i = 0; i < N; ++i)
    for (int n : adj[i])
      if (col[n] == col[i]) return fail(UASD_ERR_ARG, "colouring failed (self-bond?)");
  c->ncolors = ncol;
  c->color_of_site = col;
  c->color_ptr.assign(ncol + 1, 0);
  for (int i = 0; i < N; ++i) c->color_ptr[col[i] + 1]++;
  for (int k = 0; k < ncol; ++k) c->color_ptr[k + 1] += c->color_ptr[k];
  c->color_sites.assign(N, 0);
  std::vector<int> fillp(c->color_ptr.begin(), c->color_ptr.end() - 1);
  for (int i = 0; i < N; ++i) c->color_sites[fillp[col[i]]++] = i;
  c->color_start.assign(ncol, 0);
  c->color_stride.assign(ncol, 0);
  for (int k = 0; k < ncol; ++k) {
    const int p0 = c->color_ptr[k], n = c->color_ptr[k + 1] - p0;
    if (n == 0) continue;
    const int a0 = c->color_sites[p0], d = n > 1 ? c->color_sites[p0 + 1] - a0 : 1;
    bool ap = d > 0;
    for (int i = 2; ap && i < n; ++i)
      ap = c->color_sites[p0 + i] - c->color_sites[p0 + i - 1] == d;
    c->color_start[k] = a0;
    c->color_stride[k] = ap ? d : 0;
  }
  c->coloring_dirty = false;
  c->gen += 1;
  return UASD_OK;
}

}  // namespace

// Slot-major int16 entries of a slot-major list of device slots (see NbDec):
// v = delta * 2^kb + beta with beta = (b' - b) mod na and delta = (c' - c)
// mod nc taken in [-nc/2, nc/2).  False when any delta does not fit.
bool diff16(const uasd_ctx *c, const std::vector<int> &nb, int N, int W,
            std::vector<int16_t> &out) {
  const int na = c->order_na, nc = na == 1 ? N : c->order_nc, kb = idx_kb(c);
  const long lim = 1L << (15 - kb);
  out.resize((size_t)W * N);
  for (int j = 0; j < W; ++j)
    for (int d = 0; d < N; ++d) {
      const int n = nb[(size_t)j * N + d];
      const int b = d / nc, cc = d - b * nc, b2 = n / nc, c2 = n - b2 * nc;
      const int beta = ((b2 - b) % na + na) % na;
      long delta = (long)c2 - cc;
      if (2 * delta >= nc) delta -= nc;
      if (2 * delta < -(long)nc) delta += nc;
      if (delta < -lim || delta >= lim) return false;
      out[(size_t)j * N + d] = (int16_t)(delta * (1L << kb) + beta);
    }
  return true;
}

// ------------------------------------------------------------ stencil
// True when the list (device order) is one translation of a per-basis
// stencil over the periodic lattice; fills st.
bool detect_stencil(const uasd_ctx *c, const std::vector<int> &nb, const std::vector<int> &row,
                    const std::vector<int> &len, int W, uasd_ctx::Stencil &st) {
  const int NA = c->lat_na, N1 = c->lat_n1, N2 = c->lat_n2, N3 = c->lat_n3, N = c->N;
  st = uasd_ctx::Stencil{};
  if (NA <= 0 || W <= 0 || row.empty()) return false;
  st.W = W;
  st.bp.assign((size_t)NA * W, 0);
  st.dx = st.dy = st.dz = st.bp;
  auto wrap = [](int d, int n) { return 2 * d > n ? d - n : d; };
  std::vector<int> L(NA);
  for (int b = 0; b < NA; ++b) {
    const int d0 = slot_of(c, b);
    L[b] = len[row[d0]];
    for (int j = 0; j < L[b]; ++j) {
      const int n = site_of(c, nb[(size_t)j * N + d0]);
      const int bb = n % NA, cell = n / NA;
      const int x = cell % N1, y = (cell / N1) % N2, z = cell / (N1 * N2);
      st.bp[b * W + j] = bb;
      st.dx[b * W + j] = wrap(x, N1);
      st.dy[b * W + j] = wrap(y, N2);
      st.dz[b * W + j] = wrap(z, N3);
    }
  }
  for (int i = 0; i < N; ++i) {
    const int b = i % NA, cell = i / NA;
    const int x = cell % N1, y = (cell / N1) % N2, z = cell / (N1 * N2);
    const int d = slot_of(c, i);
    if (len[row[d]] != L[b]) return false;
    for (int j = 0; j < L[b]; ++j) {
      const int k = b * W + j;
      const int xx = ((x + st.dx[k]) % N1 + N1) % N1, yy = ((y + st.dy[k]) % N2 + N2) % N2,
                zz = ((z + st.dz[k]) % N3 + N3) % N3;
      const int want = st.bp[k] + NA * (xx + N1 * (yy + N2 * zz));
      if (site_of(c, nb[(size_t)j * N + d]) != want) return false;
    }
  }
  for (size_t k = 0; k < st.bp.size(); ++k) {
    st.rx = std::max(st.rx, std::abs(st.dx[k]));
    st.ry = std::max(st.ry, std::abs(st.dy[k]));
    st.rz = std::max(st.rz, std::abs(st.dz[k]));
  }
  st.ok = true;
  return true;
}

void build_stencil(uasd_ctx *c) {
  if (!c->stencil_dirty) return;
  c->stencil_dirty = false;
  c->stencil_ok = false;
  if (c->lat_na <= 0 || !c->have_ex) return;
  if (!detect_stencil(c, c->nbr_host, c->ex_row_of_site, c->ex_len_host, c->W, c->st_ex)) return;
  if (c->Wd > 0 && !detect_stencil(c, c->dnbr_host, c->dm_row_of_site, c->dm_len_host, c->Wd,
                                   c->st_dm))
    return;
  c->stencil_ok = true;
}

// Tile shape k (1-based) usable for this lattice and precision?
bool tile_fits(const uasd_ctx *c, int k) {
  if (!c->stencil_ok || k < 1 || k > uasd::kNumTileShapes) return false;
  // quad indices of the staging tables and the (tile, ensemble) units in int
  if ((int64_t)c->N * c->M >= (int64_t)1 << 31) return false;
  if (c->order_na != 1 && c->order_na != c->lat_na) return false;  // tile kernel's slot formula
  const int *sh = uasd::kTileShapes[k - 1];
  if ((c->lat_n3 > 1) != (sh[2] > 1)) return false;
  const int rx = std::max(c->st_ex.rx, c->st_dm.rx), ry = std::max(c->st_ex.ry, c->st_dm.ry),
            rz = std::max(c->st_ex.rz, c->st_dm.rz);
  const int hv = (sh[0] + 2 * rx) * (sh[1] + 2 * ry) * (sh[2] + 2 * rz);
  // one thread per site; whole wavefronts per basis atom
  if (sh[0] * sh[1] * sh[2] * c->lat_na > 1024 || (sh[0] * sh[1] * sh[2]) % 64 != 0) return false;
  const int plane = uasd::tile_plane(hv * c->lat_na);
  if (plane == 0) return false;
  // halo coordinates in FP32 reciprocal divisions: dividends below 2^22
  if (c->lat_n1 >= (1 << 20) || c->lat_n2 >= (1 << 20) || c->lat_n3 >= (1 << 20)) return false;
  // tile index arithmetic in FP32 reciprocals: fewer than 2^22 tiles
  if ((int64_t)((c->lat_n1 + sh[0] - 1) / sh[0]) * ((c->lat_n2 + sh[1] - 1) / sh[1]) *
          ((c->lat_n3 + sh[2] - 1) / sh[2]) >= (1 << 22))
    return false;
  const int ntab = sh[0] + 2 * rx + c->lat_na * (sh[1] + 2 * ry) * (sh[2] + 2 * rz);
  if ((sh[0] + 2 * rx) > sh[0] * sh[1] * sh[2] * c->lat_na) return false;  // a row per round
  return (int64_t)uasd::tile_lds_bytes(c->rsize, plane, ntab) <= c->lds_max;
}

// TileDev for shape k, uploading its halo offsets when the shape changes.
int make_tile(uasd_ctx *c, int k, uasd::TileDev &t) {
  const int *sh = uasd::kTileShapes[k - 1];
  t = uasd::TileDev{};
  t.N1 = c->lat_n1;
  t.N2 = c->lat_n2;
  t.N3 = c->lat_n3;
  t.NA = c->lat_na;
  t.TX = sh[0];
  t.TY = sh[1];
  t.TZ = sh[2];
  t.RX = std::max(c->st_ex.rx, c->st_dm.rx);
  t.RY = std::max(c->st_ex.ry, c->st_dm.ry);
  t.RZ = std::max(c->st_ex.rz, c->st_dm.rz);
  t.HX = t.TX + 2 * t.RX;
  t.HY = t.TY + 2 * t.RY;
  t.HZ = t.TZ + 2 * t.RZ;
  t.HV = t.HX * t.HY * t.HZ;
  t.rTX = 1.0f / (float)t.TX;
  t.rTXY = 1.0f / (float)(t.TX * t.TY);
  t.rTC = 1.0f / (float)(t.TX * t.TY * t.TZ);
  t.rHX = 1.0f / (float)t.HX;
  t.rs = t.TX * t.TY * t.TZ * t.NA / t.HX;  // halo rows staged per round
  t.ntx = (t.N1 + t.TX - 1) / t.TX;
  t.nty = (t.N2 + t.TY - 1) / t.TY;
  t.ntz = (t.N3 + t.TZ - 1) / t.TZ;
  t.rntx = 1.0f / (float)t.ntx;
  t.rntxy = 1.0f / (float)(t.ntx * t.nty);
  t.rHY = 1.0f / (float)t.HY;
  t.rHYZ = 1.0f / (float)(t.HY * t.HZ);
  t.rN1 = 1.0f / (float)t.N1;
  t.rN2 = 1.0f / (float)t.N2;
  t.rN3 = 1.0f / (float)t.N3;
  t.ox = t.N1 * ((t.RX + t.N1 - 1) / t.N1);
  t.oy = t.N2 * ((t.RY + t.N2 - 1) / t.N2);
  t.oz = t.N3 * ((t.RZ + t.N3 - 1) / t.N3);
  t.W = c->st_ex.W;
  t.Wd = c->Wd > 0 ? c->st_dm.W : 0;
  if (c->tile_off_shape != k) {
    std::vector<int> off((size_t)t.NA * (t.W + t.Wd), 0);
    auto fill = [&](const uasd_ctx::Stencil &st, int W, int base) {
      for (int b = 0; b < t.NA; ++b)
        for (int j = 0; j < W; ++j) {
          const int q = b * W + j;
          off[base + q] = st.bp[q] * t.HV + st.dz[q] * t.HX * t.HY + st.dy[q] * t.HX + st.dx[q];
        }
    };
    fill(c->st_ex, t.W, 0);
    if (t.Wd > 0) fill(c->st_dm, t.Wd, t.NA * t.W);
    if (int r = upload_raw(c, c->tile_off, off)) return r;
    c->tile_off_shape = k;
  }
  t.off = c->tile_off.as<int>();
  t.doff = t.off + (size_t)t.NA * t.W;
  // one site class per basis atom? (checked once per class rebuild)
  if (c->ucls_version != c->cls_version) {
    c->ucls = 0;
    if (t.NA <= 8 && (int)c->cls_host.size() == c->N) {
      c->ucls = 1;
      for (int b = 0; b < t.NA; ++b) c->cls_b[b] = c->cls_host[slot_of(c, b)];
      for (int i = 0; i < c->N && c->ucls; ++i)
        if (c->cls_host[slot_of(c, i)] != c->cls_b[i % t.NA]) c->ucls = 0;
    }
    c->ucls_version = c->cls_version;
  }
  t.ucls = c->ucls;
  for (int b = 0; b < 8; ++b) t.cls_b[b] = c->cls_b[b];
  return UASD_OK;
}

// ------------------------------------------------------------ segments
constexpr int kSegChunk = 4096;  // sites per reduction chunk

int set_segments(uasd_ctx *c, int nseg, const int *off) {
  std::vector<int> chunks, seg_chunk(1, 0), seg_n;
  for (int s = 0; s < nseg; ++s) {
    for (int a = off[s]; a < off[s + 1]; a += kSegChunk) {
      chunks.push_back(a);
      chunks.push_back(std::min(off[s + 1], a + kSegChunk));
    }
    seg_chunk.push_back((int)chunks.size() / 2);
    seg_n.push_back(off[s + 1] - off[s]);
  }
  std::vector<int> all(chunks);
  all.insert(all.end(), seg_chunk.begin(), seg_chunk.end());
  all.insert(all.end(), seg_n.begin(), seg_n.end());
  if (int r = upload_raw(c, c->seg_dev, all)) return r;
  c->nseg = nseg;
  c->seg_off.assign(off, off + nseg + 1);
  c->nchunks = (int)chunks.size() / 2;
  const size_t nacc = (size_t)nseg * c->M * uasd::kAccWidth;
  HIP_TRY(c->acc.ensure(nacc * sizeof(double)));
  std::vector<double> zero(nacc, 0.0);  // uploaded, not memset: the accumulators
  if (int r = upload_raw(c, c->acc, zero)) return r;  // may later sit under a graph
  c->acc_samples = 0;
  return UASD_OK;
}

int ensure_segments(uasd_ctx *c) {
  if (c->nseg > 0) return UASD_OK;
  const int off[2] = {0, c->N};
  return set_segments(c, 1, off);
}

// Per-(segment, ensemble) FP64 sums into c->seg_out; energy terms from the
// last field launch when with_energy.  Asynchronous.
template <typename R>
int segment_sums(uasd_ctx *c, bool with_energy) {
  const int nv = with_energy ? 4 : 3;
  HIP_TRY(c->seg_partial.ensure((size_t)c->nchunks * c->M * nv * sizeof(double)));
  HIP_TRY(c->seg_out.ensure((size_t)c->nseg * c->M * 4 * sizeof(double)));
  const int2 *chunks = c->seg_dev.as<int2>();
  const int *seg_chunk = reinterpret_cast<const int *>(chunks + c->nchunks);
  HIP_TRY(uasd::launch_segsum<R>(c->stream, c->M, c->nseg, c->nchunks, chunks, seg_chunk,
                                 quads<R>(c, c->spins, 0),
                                 with_energy ? c->energy.as<double>() : nullptr, nv,
                                 c->seg_partial.as<double>(), c->seg_out.as<double>()));
  return UASD_OK;
}

// =================================================================== C-ABI
extern "C" {

const char *uasd_last_error(void) { return g_err.c_str(); }

int uasd_version(void) { return UASD_API_VERSION; }

int uasd_device_count(int *count) {
  if (!count) return fail(UASD_ERR_ARG, "count is null");
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess) n = 0;
  *count = n;
  return UASD_OK;
}

int uasd_ctx_create(uasd_ctx **out, int natom, int mensemble, int precision) {
  if (!out) return fail(UASD_ERR_ARG, "out is null");
  *out = nullptr;
  if (natom <= 0 || mensemble <= 0) return fail(UASD_ERR_ARG, "natom and mensemble must be > 0");
  if ((int64_t)natom * mensemble > (int64_t)1 << 31)
    return fail(UASD_ERR_ARG, "natom*mensemble exceeds 2^31");
  if (precision != UASD_FP64 && precision != UASD_FP32)
    return fail(UASD_ERR_ARG, "precision must be UASD_FP64 or UASD_FP32");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
    return fail(UASD_ERR_NODEVICE, "no HIP device available");
  auto c = std::make_unique<uasd_ctx>();
  c->N = natom;
  c->M = mensemble;
  c->prec = precision;
  c->rsize = precision == UASD_FP64 ? 8 : 4;
  HIP_TRY(hipGetDevice(&c->device));
  hipDeviceProp_t prop;
  HIP_TRY(hipGetDeviceProperties(&prop, c->device));
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    return fail(UASD_ERR_NODEVICE, std::string("device is ") + prop.gcnArchName +
                                       ", the library is built for gfx950");
  c->lds_max = (int64_t)std::max(prop.sharedMemPerBlock, prop.sharedMemPerBlockOptin);
  if (int r = scratch_init(c->device)) return r;
  HIP_TRY(hipStreamCreateWithFlags(&c->own_stream, hipStreamNonBlocking));
  c->stream = c->own_stream;
  HIP_TRY(c->spins.ensure(3 * quad_bytes(c.get())));
  HIP_TRY(hipMemsetAsync(c->spins.p, 0, 3 * quad_bytes(c.get()), c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  // the whole lattice as one segment until uasd_set_segments says otherwise
  if (int r = ensure_segments(c.get())) return r;
  *out = c.release();
  return UASD_OK;
}

int uasd_ctx_destroy(uasd_ctx *c) {
  if (!c) return UASD_OK;
  // the context is released whatever happens; a failure of the work still
  // queued on its stream is reported, the release calls' own are not
  const hipError_t serr = c->stream ? hipStreamSynchronize(c->stream) : hipSuccess;
  for (const auto &pe : c->pending) {  // timings not yet read are dropped
    (void)hipEventDestroy(pe.a);
    (void)hipEventDestroy(pe.b);
  }
  destroy_graphs(c);
  if (c->cur) (void)hipEventDestroy(c->cur);
  if (c->region_cur) (void)hipEventDestroy(c->region_cur);
  for (hipEvent_t e : c->ev_pool) (void)hipEventDestroy(e);
  if (c->own_stream) (void)hipStreamDestroy(c->own_stream);
  delete c;
  if (serr != hipSuccess)
    return fail(UASD_ERR_HIP, std::string("uasd_ctx_destroy: queued work failed: ") +
                                  hipGetErrorString(serr));
  return UASD_OK;
}

int uasd_ctx_set_stream(uasd_ctx *c, void *s) {
  if (int r = check_ctx(c)) return r;
  HIP_TRY(hipStreamSynchronize(c->stream));
  c->stream = s ? static_cast<hipStream_t>(s) : c->own_stream;
  return UASD_OK;
}

void *uasd_ctx_stream(uasd_ctx *c) { return c ? (void *)c->stream : nullptr; }

int uasd_ctx_sync(uasd_ctx *c) {
  if (int r = check_ctx(c)) return r;
  HIP_TRY(hipStreamSynchronize(c->stream));
  return UASD_OK;
}

int uasd_set_exchange(uasd_ctx *c, int mnn, int nham, const int *nlist, const int *nlistsize,
                      const double *ncoup, const int *aham) {
  if (int r = check_ctx(c)) return r;
  c->gen += 1;
  const int N = c->N;
  if (mnn < 0 || nham <= 0 || !nlist || !nlistsize || (mnn > 0 && !ncoup))
    return fail(UASD_ERR_ARG, "uasd_set_exchange: bad arguments");
  std::vector<int> ref(N);
  for (int i = 0; i < N; ++i) {
    int r = aham ? aham[i] - 1 : i;
    if (r < 0 || r >= nham) return fail(UASD_ERR_ARG, "aham entry out of range");
    if (nlistsize[r] < 0 || nlistsize[r] > mnn) return fail(UASD_ERR_ARG, "nlistsize > max_no_neigh");
    ref[i] = r;
  }
  RowSet rs = dedupe_rows(N, ref, nlistsize, ncoup, mnn, 1);
  int W = 0;
  for (int L : rs.len) W = std::max(W, L);
  // device order: slot d = slot_of(i) holds site i; list values are slots
  std::vector<int> nb((size_t)std::max(W, 1) * N), row_dev(N);
  for (int i = 0; i < N; ++i) {
    const int d = slot_of(c, i);
    int L = rs.len[rs.row_of_site[i]];
    row_dev[d] = rs.row_of_site[i];
    for (int j = 0; j < W; ++j) {
      int v = d;
      if (j < L) {
        v = nlist[(size_t)j + (size_t)mnn * i] - 1;
        if (v < 0 || v >= N) return fail(UASD_ERR_ARG, "nlist entry out of range");
        v = slot_of(c, v);
      }
      nb[(size_t)j * N + d] = v;
    }
  }
  const int rows = (int)rs.len.size();
  std::vector<double> cp((size_t)std::max(W, 1) * rows, 0.0);
  for (int r = 0; r < rows; ++r)
    for (int j = 0; j < rs.len[r]; ++j) cp[(size_t)j * rows + r] = rs.vals[r][j];
  c->W = W;
  c->ex_rows = rows;
  c->ex_row_of_site = row_dev;
  c->ex_len_host = rs.len;
  c->ex_full = W > 0 && std::all_of(rs.len.begin(), rs.len.end(), [W](int l) { return l == W; });
  c->nbr_host = nb;
  if (int r = upload_raw(c, c->nbr, nb)) return r;
  std::vector<int16_t> d16;
  c->use16 = W > 0 && diff16(c, nb, N, W, d16);
  if (c->use16) {
    if (int r = upload_raw(c, c->nbr16, d16)) return r;
  } else {
    c->nbr16.release();
  }
  if (int r = upload_raw(c, c->ex_len, rs.len)) return r;
  if (int r = upload_real(c, c->cpl, cp)) return r;
  c->have_ex = true;
  c->classes_dirty = true;
  c->coloring_dirty = true;
  c->stencil_dirty = true;
  c->tile_off_shape = 0;
  return UASD_OK;
}

int uasd_set_dm(uasd_ctx *c, int mdm, int nham, const int *dmlist, const int *dmlistsize,
                const double *dm_vect, const int *aham) {
  if (int r = check_ctx(c)) return r;
  c->gen += 1;
  const int N = c->N;
  if (!dmlist || mdm <= 0) {  // clear
    c->Wd = 0;
    c->dm_rows = 0;
    c->dm_row_of_site.clear();
    c->classes_dirty = c->coloring_dirty = true;
  c->stencil_dirty = true;
  c->tile_off_shape = 0;
    return UASD_OK;
  }
  if (nham <= 0 || !dmlistsize || !dm_vect) return fail(UASD_ERR_ARG, "uasd_set_dm: bad arguments");
  std::vector<int> ref(N);
  for (int i = 0; i < N; ++i) {
    int r = aham ? aham[i] - 1 : i;
    if (r < 0 || r >= nham) return fail(UASD_ERR_ARG, "aham entry out of range");
    if (dmlistsize[r] < 0 || dmlistsize[r] > mdm) return fail(UASD_ERR_ARG, "dmlistsize > max");
    ref[i] = r;
  }
  RowSet rs = dedupe_rows(N, ref, dmlistsize, dm_vect, mdm, 3);
  int W = 0;
  for (int L : rs.len) W = std::max(W, L);
  std::vector<int> nb((size_t)std::max(W, 1) * N), row_dev(N);
  for (int i = 0; i < N; ++i) {
    const int d = slot_of(c, i);
    int L = rs.len[rs.row_of_site[i]];
    row_dev[d] = rs.row_of_site[i];
    for (int j = 0; j < W; ++j) {
      int v = d;
      if (j < L) {
        v = dmlist[(size_t)j + (size_t)mdm * i] - 1;
        if (v < 0 || v >= N) return fail(UASD_ERR_ARG, "dmlist entry out of range");
        v = slot_of(c, v);
      }
      nb[(size_t)j * N + d] = v;
    }
  }
  const int rows = (int)rs.len.size();
  const size_t n = (size_t)std::max(W, 1) * rows;
  std::vector<double> dv(3 * n, 0.0);
  for (int r = 0; r < rows; ++r)
    for (int j = 0; j < rs.len[r]; ++j)
      for (int d = 0; d < 3; ++d) dv[d * n + (size_t)j * rows + r] = rs.vals[r][3 * j + d];
  c->Wd = W;
  c->dm_rows = rows;
  c->dm_row_of_site = row_dev;
  c->dm_len_host = rs.len;
  c->dnbr_host = nb;
  if (int r = upload_raw(c, c->dnbr, nb)) return r;
  std::vector<int16_t> d16;
  c->use16_dm = W > 0 && diff16(c, nb, N, W, d16);
  if (c->use16_dm) {
    if (int r = upload_raw(c, c->dnbr16, d16)) return r;
  } else {
    c->dnbr16.release();
  }
  if (int r = upload_raw(c, c->dm_len, rs.len)) return r;
  if (int r = upload_real(c, c->dv, dv)) return r;
  c->classes_dirty = c->coloring_dirty = true;
  c->stencil_dirty = true;
  c->tile_off_shape = 0;
  return UASD_OK;
}

int uasd_set_anisotropy(uasd_ctx *c, const int *taniso, const double *eaniso,
                        const double *kaniso, const double *sb) {
  if (int r = check_ctx(c)) return r;
  c->gen += 1;
  const int N = c->N;
  if (!taniso) {
    c->an_rows = 0;
    c->an_row_of_site.clear();
    c->classes_dirty = true;
    return UASD_OK;
  }
  if (!eaniso || !kaniso || !sb) return fail(UASD_ERR_ARG, "uasd_set_anisotropy: null array");
  std::map<std::vector<double>, int> rows;
  std::vector<int> row_of(N, -1);
  std::vector<std::vector<double>> vals;
  std::vector<int> types;
  for (int i = 0; i < N; ++i) {
    int t = taniso[i];
    if (t != 0 && t != 1 && t != 2 && t != 7)
      return fail(UASD_ERR_UNSUPPORTED, "anisotropy type must be 0, 1, 2 or 7");
    if (t == 0) continue;
    std::vector<double> key = {(double)t,     eaniso[3 * (size_t)i], eaniso[3 * (size_t)i + 1],
                               eaniso[3 * (size_t)i + 2], kaniso[2 * (size_t)i],
                               kaniso[2 * (size_t)i + 1], sb[i]};
    auto it = rows.find(key);
    if (it == rows.end()) {
      it = rows.emplace(key, (int)vals.size()).first;
      vals.push_back(key);
      types.push_back(t);
    }
    row_of[i] = it->second;
  }
  const int nr = (int)vals.size();
  c->an_rows = nr;
  c->an_row_of_site.assign(N, -1);
  for (int i = 0; i < N; ++i) c->an_row_of_site[slot_of(c, i)] = row_of[i];
  if (nr > 0) {
    std::vector<double> v(6 * (size_t)nr);
    for (int r = 0; r < nr; ++r)
      for (int k = 0; k < 6; ++k) v[(size_t)k * nr + r] = vals[r][1 + k];
    if (int r = upload_raw(c, c->an_type, types)) return r;
    if (int r = upload_real(c, c->an_vals, v)) return r;
  }
  c->classes_dirty = true;
  return UASD_OK;
}

int uasd_set_external_field(uasd_ctx *c, const double *field) {
  if (int r = check_ctx(c)) return r;
  c->gen += 1;
  const int N = c->N, M = c->M;
  if (!field) {
    c->ext_mode = 0;
    c->ext_u_host.assign(3 * (size_t)M, 0.0);
    return upload_real(c, c->ext_u, c->ext_u_host);
  }
  bool uniform = true, zero = true;
  for (int k = 0; k < M && uniform; ++k) {
    const double *f0 = field + 3 * ((size_t)N * k);
    for (int i = 0; i < N && uniform; ++i) {
      const double *f = field + 3 * ((size_t)i + (size_t)N * k);
      for (int d = 0; d < 3; ++d) {
        if (std::memcmp(&f[d], &f0[d], sizeof(double)) != 0) uniform = false;
        if (f[d] != 0.0) zero = false;
      }
    }
  }
  c->ext_u_host.assign(3 * (size_t)M, 0.0);
  if (uniform) {
    for (int k = 0; k < M; ++k)
      for (int d = 0; d < 3; ++d) c->ext_u_host[3 * k + d] = field[3 * ((size_t)N * k) + d];
    c->ext_mode = zero ? 0 : 1;
    return upload_real(c, c->ext_u, c->ext_u_host);
  }
  c->ext_mode = 2;
  if (int r = upload_real(c, c->ext_u, c->ext_u_host)) return r;
  return upload_quads(c, c->ext_full, 0, to_quads(c, field, nullptr, 0.0));
}

int uasd_set_llg(uasd_ctx *c, const double *lambda1, const double *temp, const double *landeg,
                 double delta_t, double temprescale, double bn) {
  if (int r = check_ctx(c)) return r;
  c->gen += 1;
  const int N = c->N;
  if (!lambda1 || !temp) return fail(UASD_ERR_ARG, "uasd_set_llg: lambda1 and temp required");
  if (!(delta_t > 0.0)) return fail(UASD_ERR_ARG, "delta_t must be > 0");
  auto is_uniform = [N](const double *a) {
    for (int i = 1; i < N; ++i)
      if (std::memcmp(&a[i], &a[0], sizeof(double)) != 0) return false;
    return true;
  };
  std::vector<double> arr(3 * (size_t)N);
  for (int i = 0; i < N; ++i) {
    const int d = slot_of(c, i);
    arr[d] = lambda1[i];
    arr[N + d] = temp[i];
    arr[2 * (size_t)N + d] = landeg ? landeg[i] : 1.0;
  }
  c->lam_uniform = is_uniform(lambda1);
  c->temp_uniform = is_uniform(temp);
  c->lande_uniform = landeg ? is_uniform(landeg) : true;
  c->lam0 = lambda1[0];
  c->temp0 = temp[0];
  c->lande0 = landeg ? landeg[0] : 1.0;
  c->dt = delta_t;
  c->temprescale = temprescale;
  c->bn = bn;
  if (int r = upload_raw(c, c->llgarr, arr)) return r;
  c->have_llg = true;
  return UASD_OK;
}

int uasd_set_temprescale(uasd_ctx *c, double temprescale) {
  if (int r = check_ctx(c)) return r;
  c->gen += 1;
  if (!(temprescale > 0.0)) return fail(UASD_ERR_ARG, "temprescale must be > 0");
  c->temprescale = temprescale;
  return UASD_OK;
}

int uasd_set_noise(uasd_ctx *c, int mode, uint64_t seed, const double *normals) {
  if (int r = check_ctx(c)) return r;
  c->gen += 1;
  if (mode != UASD_NOISE_NONE && mode != UASD_NOISE_PHILOX && mode != UASD_NOISE_INJECTED)
    return fail(UASD_ERR_ARG, "unknown noise mode");
  if (mode == UASD_NOISE_INJECTED) {
    if (!normals) return fail(UASD_ERR_ARG, "injected noise needs the normals array");
    if (int r = upload_quads(c, c->noise, 0, to_quads(c, normals, nullptr, 0.0))) return r;
  }
  if (mode == UASD_NOISE_PHILOX && (seed != c->seed || c->noise_mode != mode)) c->step = 0;
  c->noise_mode = mode;
  c->seed = seed;
  return UASD_OK;
}

int uasd_set_moments(uasd_ctx *c, const double *emom, const double *mmom) {
  if (int r = check_ctx(c)) return r;
  if (!emom || !mmom) return fail(UASD_ERR_ARG, "uasd_set_moments: null array");
  std::vector<double> hq = to_quads(c, emom, mmom, 0.0);
  if (int r = upload_quads(c, c->spins, 0, hq)) return r;
  if (int r = upload_quads(c, c->spins, 1, hq)) return r;  // emom2 starts as emom
  c->have_state = true;
  const size_t nm = (size_t)c->N * c->M;
  c->m_uniform = std::all_of(mmom, mmom + nm, [&](double v) { return v == mmom[0]; });
  c->m0 = mmom[0];
  c->gen += 1;  // captured steps hold the noise amplitudes
  return UASD_OK;
}

int uasd_get_moments(uasd_ctx *c, double *emom, double *emomM) {
  if (int r = check_ctx(c)) return r;
  if (!c->have_state) return fail(UASD_ERR_STATE, "moments not set");
  std::vector<double> hq;
  if (int r = download_quads(c, c->spins, 0, hq)) return r;
  if (emom) from_quads(c, hq, emom, false);
  if (emomM) from_quads(c, hq, emomM, true);
  return UASD_OK;
}

// ---------------------------------------------------------- drop-in (host)
// Host vec3 arrays are uploaded as quads; spin quads keep the resident mmom
// in w, the others carry a constant.
// dst[q] = (src[q].xyz, wsrc[q].w): uploaded directions merged with the
// moment magnitudes already on the device
extern "C++" {
template <typename R>
__global__ void __launch_bounds__(256)
    k_merge_xyz(const uasd::Quad<R> *__restrict__ src, const uasd::Quad<R> *wsrc,
                uasd::Quad<R> *dst, size_t n) {  // wsrc and dst may be the same array
  const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  const uasd::Quad<R> v = src[q];
  dst[q] = uasd::Quad<R>{v.x, v.y, v.z, wsrc[q].w};
}
}

// quad array k of the spins <- the (3,N,M) array a, with w = mmom of the
// state (quad array 0): a goes to the scratch quads, the merge runs on the
// device, so the magnitudes are not read back to the host
static int upload_vec_keep_w(uasd_ctx *c, int k, const double *a) {
  if (int r = upload_quads(c, c->scratch, 0, to_quads(c, a, nullptr, 0.0))) return r;
  const size_t n = (size_t)c->N * c->M;
  const dim3 grid((unsigned)((n + 255) / 256)), block(256);
  if (c->prec == UASD_FP64)
    hipLaunchKernelGGL(k_merge_xyz<double>, grid, block, 0, c->stream,
                       quads<double>(c, c->scratch), quads<double>(c, c->spins, 0),
                       quads<double>(c, c->spins, k), n);
  else
    hipLaunchKernelGGL(k_merge_xyz<float>, grid, block, 0, c->stream,
                       quads<float>(c, c->scratch), quads<float>(c, c->spins, 0),
                       quads<float>(c, c->spins, k), n);
  HIP_TRY(hipGetLastError());
  return UASD_OK;
}

int uasd_effective_field_host(uasd_ctx *c, const double *emomM, double *beff, double *beff1,
                              double *beff2, double *energy) {
  if (int r = check_ctx(c)) return r;
  if (!c->have_ex) return fail(UASD_ERR_STATE, "exchange not set");
  if (!emomM || !beff) return fail(UASD_ERR_ARG, "null array");
  if (int r = rebuild_classes(c)) return r;
  if (int r = upload_quads(c, c->scratch, 0, to_quads(c, emomM, nullptr, 1.0))) return r;
  const bool parts = beff1 || beff2;
  int r = c->prec == UASD_FP64 ? run_field<double>(c, true, energy != nullptr, parts)
                               : run_field<float>(c, true, energy != nullptr, parts);
  if (r) return r;
  std::vector<double> hq;
  if ((r = download_quads(c, c->field, 0, hq))) return r;
  from_quads(c, hq, beff, false);
  if (beff1) {
    if ((r = download_quads(c, c->field, 1, hq))) return r;
    from_quads(c, hq, beff1, false);
  }
  if (beff2) {
    if ((r = download_quads(c, c->field, 2, hq))) return r;
    from_quads(c, hq, beff2, false);
  }
  if (energy) {
    std::vector<double> red;
    r = c->prec == UASD_FP64 ? reduce_moments<double>(c, true, true, red)
                             : reduce_moments<float>(c, true, true, red);
    if (r) return r;
    double e = 0.0;
    for (int k = 0; k < c->M; ++k) e += red[4 * (size_t)k + 3];
    *energy = e * uasd::kMub / uasd::kMry;
  }
  return UASD_OK;
}

int uasd_evolve_first_host(uasd_ctx *c, int alg, const double *beff, double *emom, double *emom2,
                           double *emomM, double *b2eff) {
  if (int r = check_ctx(c)) return r;
  if (!sde_supported(alg))
    return fail(UASD_ERR_UNSUPPORTED, "SDEalgh must be 1 (midpoint), 4 (Heun), 5 (Depondt), 6 (SIB) or 7 (spherical midpoint)");
  if (!beff || !emom || !emom2 || !emomM || (sde_b2eff(alg) && !b2eff))
    return fail(UASD_ERR_ARG, "null array");
  if (int r = ready_for_llg(c)) return r;
  if (int r = upload_quads(c, c->field, 0, to_quads(c, beff, nullptr, 0.0))) return r;
  // heun4p starts from emom2 (heun_proper.f90:309-311), the others from emom
  if (int r = upload_vec_keep_w(c, 0, alg == UASD_SDE_HEUN ? emom2 : emom)) return r;
  int r = c->prec == UASD_FP64 ? run_llg_half<double>(c, alg, 0, false, false)
                               : run_llg_half<float>(c, alg, 0, false, false);
  if (r) return r;
  std::vector<double> pq;
  if ((r = download_quads(c, c->spins, 1, pq))) return r;
  if (alg == UASD_SDE_DEPONDT) {
    // depondt.f90:182-187: emom2 = m(t), emom = rotated, emomM = emom*mmom, b2eff = bdup
    std::memcpy(emom2, emom, sizeof(double) * 3 * (size_t)c->N * c->M);
    from_quads(c, pq, emom, false);
    from_quads(c, pq, emomM, true);
    std::vector<double> bq;
    if ((r = download_quads(c, c->spins, 2, bq))) return r;
    from_quads(c, bq, b2eff, false);
  } else if (alg == UASD_SDE_HEUN) {
    // heun_proper.f90:303-365: b2eff = beff, emom = predictor, emomM =
    // emom*mmom, emom2 unchanged
    from_quads(c, pq, emom, false);
    from_quads(c, pq, emomM, true);
    std::vector<double> bq;
    if ((r = download_quads(c, c->spins, 2, bq))) return r;
    from_quads(c, bq, b2eff, false);
  } else {
    // midpoint.f90:167-175 (and sibt / sispht, sphmidpoint.f90:247-258,
    // 698-710): emom2 = et', emomM = et'*mmom, emom unchanged
    from_quads(c, pq, emom2, false);
    from_quads(c, pq, emomM, true);
  }
  return UASD_OK;
}

int uasd_evolve_second_host(uasd_ctx *c, int alg, const double *beff, const double *b2eff,
                            double *emom, double *emom2) {
  if (int r = check_ctx(c)) return r;
  if (!sde_supported(alg))
    return fail(UASD_ERR_UNSUPPORTED, "SDEalgh must be 1 (midpoint), 4 (Heun), 5 (Depondt), 6 (SIB) or 7 (spherical midpoint)");
  if (!beff || !emom || !emom2 || (sde_b2eff(alg) && !b2eff))
    return fail(UASD_ERR_ARG, "null array");
  if (int r = ready_for_llg(c)) return r;
  if (int r = upload_quads(c, c->field, 0, to_quads(c, beff, nullptr, 0.0))) return r;
  if (alg == UASD_SDE_DEPONDT || alg == UASD_SDE_HEUN) {
    // reference on entry: emom = predicted, emom2 = m(t)
    if (int r = upload_vec_keep_w(c, 1, emom)) return r;
    if (int r = upload_vec_keep_w(c, 0, emom2)) return r;
    if (int r = upload_quads(c, c->spins, 2, to_quads(c, b2eff, nullptr, 0.0))) return r;
  } else {
    if (int r = upload_vec_keep_w(c, 0, emom)) return r;
    if (int r = upload_vec_keep_w(c, 1, emom2)) return r;
  }
  int r = c->prec == UASD_FP64 ? run_llg_half<double>(c, alg, 1, false, false)
                               : run_llg_half<float>(c, alg, 1, false, false);
  if (r) return r;
  if (alg == UASD_SDE_DEPONDT) std::memcpy(emom, emom2, sizeof(double) * 3 * (size_t)c->N * c->M);
  std::vector<double> pq;
  if ((r = download_quads(c, c->spins, 1, pq))) return r;
  from_quads(c, pq, emom2, false);
  c->step += 1;
  return UASD_OK;
}

// ----------------------------------------------------------- resident path
}  // extern "C"

namespace {

void destroy_graphs(uasd_ctx *c) {
  for (auto &kv : c->graphs) {
    if (kv.second.exec) (void)hipGraphExecDestroy(kv.second.exec);
    if (kv.second.g) (void)hipGraphDestroy(kv.second.g);
  }
  c->graphs.clear();
}

// Records what body() enqueues on the context stream and instantiates it.
template <typename F>
int capture_graph(uasd_ctx *c, F &&body, uasd_ctx::Graph &out) {
  HIP_TRY(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  c->capturing = true;
  const int r = body();
  c->capturing = false;
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(c->stream, &g);
  if (r) {
    if (g) (void)hipGraphDestroy(g);
    return r;
  }
  if (e != hipSuccess)
    return fail(UASD_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
  hipGraphExec_t x = nullptr;
  const hipError_t e2 = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
  if (e2 != hipSuccess) {
    (void)hipGraphDestroy(g);
    return fail(UASD_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e2));
  }
  out.g = g;
  out.exec = x;
  return UASD_OK;
}

// Replays graph `key` (capturing it on first use), which advances device
// counter idx by K.  The device counter is first set to host_ctr when they
// differ (after eager steps), with a kernel, so no memset or copy node is
// ever recorded.
template <typename F>
int replay_graph(uasd_ctx *c, const std::tuple<int, int, int, uint64_t> &key, int idx,
                 uint64_t &host_ctr, int K, int timer, F &&body) {
  if (c->graphs_gen != c->gen || c->graphs.size() > 16) {
    destroy_graphs(c);
    c->graphs_gen = c->gen;
  }
  if (!c->ctr.p) {
    HIP_TRY(c->ctr.ensure(2 * sizeof(uint64_t)));
    c->ctr_valid[0] = c->ctr_valid[1] = false;
  }
  auto it = c->graphs.find(key);
  if (it == c->graphs.end()) {
    uasd_ctx::Graph gr;
    if (int r = capture_graph(c, body, gr)) return r;
    it = c->graphs.emplace(key, gr).first;
  }
  if (!c->ctr_valid[idx] || c->ctr_val[idx] != host_ctr) {
    HIP_TRY(uasd::launch_counter(c->stream, idx, c->ctr.as<uint64_t>(), host_ctr, 0));
    c->ctr_valid[idx] = true;
  }
  tic(c, timer);
  HIP_TRY(hipGraphLaunch(it->second.exec, c->stream));
  if (int r = toc(c, timer)) return r;
  host_ctr += (uint64_t)K;
  c->ctr_val[idx] = host_ctr;
  return UASD_OK;
}

}  // namespace

extern "C" {

int uasd_set_graph_steps(uasd_ctx *c, int steps_per_graph) {
  if (int r = check_ctx(c)) return r;
  if (steps_per_graph < 0 || steps_per_graph > 4096)
    return fail(UASD_ERR_ARG, "steps_per_graph must be 0 (off) .. 4096");
  c->graph_steps = steps_per_graph;
  return UASD_OK;
}

int uasd_sd_steps(uasd_ctx *c, int alg, int64_t nsteps) {
  if (int r = check_ctx(c)) return r;
  if (!sde_supported(alg))
    return fail(UASD_ERR_UNSUPPORTED, "SDEalgh must be 1 (midpoint), 4 (Heun), 5 (Depondt), 6 (SIB) or 7 (spherical midpoint)");
  if (nsteps < 0) return fail(UASD_ERR_ARG, "nsteps < 0");
  if (int r = ready_for_llg(c)) return r;
  if (int r = region_tic(c)) return r;
  if (int r = sd_steps_run(c, alg, nsteps)) return r;
  return region_toc(c);
}

}  // extern "C"

namespace {

int sd_steps_run(uasd_ctx *c, int alg, int64_t nsteps) {
  int64_t s = 0;
  const int K = c->graph_steps;
  if (K > 0) {
    auto body = [&]() -> int {
      for (int j = 0; j < K; ++j) {
        int r = c->prec == UASD_FP64 ? run_llg_half<double>(c, alg, 0, true, false, j)
                                     : run_llg_half<float>(c, alg, 0, true, false, j);
        if (r) return r;
        r = c->prec == UASD_FP64 ? run_llg_half<double>(c, alg, 1, true, true, j)
                                 : run_llg_half<float>(c, alg, 1, true, true, j);
        if (r) return r;
      }
      HIP_TRY(uasd::launch_counter(c->stream, 0, c->ctr.as<uint64_t>(), (uint64_t)K, 1));
      return UASD_OK;
    };
    for (; nsteps - s >= K; s += K)
      if (int r = replay_graph(c, std::make_tuple(0, alg, K, (uint64_t)0), 0, c->step, K, 4, body))
        return r;
  }
  for (; s < nsteps; ++s) {
    int r = c->prec == UASD_FP64 ? run_llg_half<double>(c, alg, 0, true, false)
                                 : run_llg_half<float>(c, alg, 0, true, false);
    if (r) return r;
    r = c->prec == UASD_FP64 ? run_llg_half<double>(c, alg, 1, true, true)
                             : run_llg_half<float>(c, alg, 1, true, true);
    if (r) return r;
    c->step += 1;
  }
  return UASD_OK;
}

}  // namespace

extern "C" {

int uasd_effective_field(uasd_ctx *c, double *beff, double *energy) {
  if (int r = check_ctx(c)) return r;
  if (!c->have_ex || !c->have_state) return fail(UASD_ERR_STATE, "exchange or moments not set");
  if (int r = rebuild_classes(c)) return r;
  int r = c->prec == UASD_FP64 ? run_field<double>(c, false, energy != nullptr)
                               : run_field<float>(c, false, energy != nullptr);
  if (r) return r;
  if (beff) {
    std::vector<double> hq;
    if ((r = download_quads(c, c->field, 0, hq))) return r;
    from_quads(c, hq, beff, false);
  }
  if (energy) {
    std::vector<double> red;
    r = c->prec == UASD_FP64 ? reduce_moments<double>(c, true, false, red)
                             : reduce_moments<float>(c, true, false, red);
    if (r) return r;
    double e = 0.0;
    for (int k = 0; k < c->M; ++k) e += red[4 * (size_t)k + 3];
    *energy = e * uasd::kMub / uasd::kMry;
  }
  HIP_TRY(hipStreamSynchronize(c->stream));
  return UASD_OK;
}

// Uploads the per-ensemble MC constants and the colour lists and fills mp.
static int mc_prepare(uasd_ctx *c, char mode, const double *temps, uint64_t seed,
                      uasd::MCParams &mp) {
  if (mode != 'M' && mode != 'H') return fail(UASD_ERR_UNSUPPORTED, "mode must be 'M' or 'H'");
  if (!temps) return fail(UASD_ERR_ARG, "temps is null");
  if (!c->have_ex || !c->have_state) return fail(UASD_ERR_STATE, "exchange or moments not set");
  if (c->ext_mode == 2)
    return fail(UASD_ERR_UNSUPPORTED, "MC takes a uniform external field (as mc_evolve's extfield)");
  if (int r = rebuild_classes(c)) return r;
  if (int r = build_coloring(c)) return r;
  const int M = c->M;
  for (int k = 0; k < M; ++k)
    if (!(temps[k] > 0.0)) return fail(UASD_ERR_ARG, "MC temperatures must be > 0");
  // per-ensemble constants (montecarlo.f90:141, montecarlo_common.f90:190, 397)
  std::vector<double> host(6 * (size_t)M);
  for (int k = 0; k < M; ++k) {
    host[k] = (double)(2.0f / 25.0f) * std::pow(uasd::kKBolt * temps[k] / uasd::kMub, 0.20);
    host[M + k] = 1.0 / uasd::kKBolt / (c->temprescale * temps[k] + 1.0e-15);
    host[2 * (size_t)M + k] = 1.0 / uasd::kKBolt / (c->temprescale * temps[k]);
  }
  for (int k = 0; k < M; ++k)
    for (int d = 0; d < 3; ++d)
      host[3 * (size_t)M + 3 * k + d] = c->ext_u_host.empty() ? 0.0 : c->ext_u_host[3 * k + d];
  const size_t cbytes = host.size() * sizeof(double);
  HIP_TRY(c->mcbuf.ensure(cbytes + c->color_sites.size() * sizeof(int)));
  HIP_TRY(hipMemcpyAsync(c->mcbuf.p, host.data(), cbytes, hipMemcpyHostToDevice, c->stream));
  int *sites_dev = reinterpret_cast<int *>(static_cast<char *>(c->mcbuf.p) + cbytes);
  HIP_TRY(hipMemcpyAsync(sites_dev, c->color_sites.data(), c->color_sites.size() * sizeof(int),
                         hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  mp = uasd::MCParams{};
  const double *base = c->mcbuf.as<double>();
  mp.sites = sites_dev;
  mp.delta = base;
  mp.beta = base + M;
  mp.hbeta = base + 2 * (size_t)M;
  mp.ext = base + 3 * (size_t)M;
  mp.seed = seed;
  mp.heat_bath = mode == 'H';
  return UASD_OK;
}

// nsweeps colour-ordered sweeps; every launch is asynchronous on the stream.
static int mc_run(uasd_ctx *c, uasd::MCParams mp, int64_t nsweeps) {
  const int *sites_dev = mp.sites;
  for (int64_t s = 0; s < nsweeps; ++s) {
    mp.sweep = c->mc_sweep;
    for (int col = 0; col < c->ncolors; ++col) {
      mp.sites = sites_dev + c->color_ptr[col];
      mp.nsites = c->color_ptr[col + 1] - c->color_ptr[col];
      mp.site0 = c->color_start[col];
      mp.sstride = c->color_stride[col];
      tic(c, 3);
      if (c->prec == UASD_FP64) {
        HIP_TRY(uasd::launch_mc<double>(c->stream, c->tune[2], ham_view<double>(c), mp,
                                        quads<double>(c, c->spins, 0)));
      } else {
        HIP_TRY(uasd::launch_mc<float>(c->stream, c->tune[2], ham_view<float>(c), mp,
                                       quads<float>(c, c->spins, 0)));
      }
      if (int r = toc(c, 3)) return r;
    }
    c->mc_sweep += 1;
  }
  return UASD_OK;
}

int uasd_mc_sweeps(uasd_ctx *c, char mode, const double *temps, int64_t nsweeps, uint64_t seed) {
  if (int r = check_ctx(c)) return r;
  if (nsweeps < 0) return fail(UASD_ERR_ARG, "nsweeps < 0");
  uasd::MCParams mp;
  if (int r = mc_prepare(c, mode, temps, seed, mp)) return r;
  if (int r = region_tic(c)) return r;
  if (int r = mc_sweeps_run(c, mode, mp, nsweeps, seed)) return r;
  return region_toc(c);
}

}  // extern "C"

namespace {

int mc_sweeps_run(uasd_ctx *c, char mode, const uasd::MCParams &mp, int64_t nsweeps,
                  uint64_t seed) {
  int64_t s = 0;
  const int K = c->graph_steps;
  if (K > 0) {
    auto body = [&]() -> int {
      uasd::MCParams g = mp;
      g.sweep_dev = c->ctr.as<uint64_t>() + 1;
      for (int j = 0; j < K; ++j) {
        g.sweep = (uint64_t)j;
        for (int col = 0; col < c->ncolors; ++col) {
          g.sites = mp.sites + c->color_ptr[col];
          g.nsites = c->color_ptr[col + 1] - c->color_ptr[col];
          g.site0 = c->color_start[col];
          g.sstride = c->color_stride[col];
          if (c->prec == UASD_FP64) {
            HIP_TRY(uasd::launch_mc<double>(c->stream, c->tune[2], ham_view<double>(c), g,
                                            quads<double>(c, c->spins, 0)));
          } else {
            HIP_TRY(uasd::launch_mc<float>(c->stream, c->tune[2], ham_view<float>(c), g,
                                           quads<float>(c, c->spins, 0)));
          }
        }
      }
      HIP_TRY(uasd::launch_counter(c->stream, 1, c->ctr.as<uint64_t>(), (uint64_t)K, 1));
      return UASD_OK;
    };
    for (; nsweeps - s >= K; s += K)
      if (int r = replay_graph(c, std::make_tuple(1, (int)mode, K, seed), 1, c->mc_sweep, K, 5,
                               body))
        return r;
  }
  return mc_run(c, mp, nsweeps - s);
}

}  // namespace

extern "C" {

// ---------------------------------------------------------------- tuning
int uasd_index_bytes(uasd_ctx *c, int *exchange, int *dm) {
  if (int r = check_ctx(c)) return r;
  const bool i32 = c->tune[0].idx32 != 0;  // as the LLG kernels are tuned
  if (exchange) *exchange = c->use16 && !i32 ? 2 : 4;
  if (dm) *dm = c->use16_dm && !i32 ? 2 : 4;
  return UASD_OK;
}

int uasd_llg_tile(uasd_ctx *c, int *tile) {
  if (int r = check_ctx(c)) return r;
  if (!tile) return fail(UASD_ERR_ARG, "tile must not be NULL");
  *tile = 0;
  if (c->tune[0].tile > 0) {
    build_stencil(c);
    if (tile_fits(c, c->tune[0].tile)) *tile = c->tune[0].tile;
  }
  return UASD_OK;
}

int uasd_set_tuning(uasd_ctx *c, int kind, int gather, int block) {
  if (int r = check_ctx(c)) return r;
  c->gen += 1;
  if (kind < 0 || kind > 2) return fail(UASD_ERR_ARG, "kind must be 0 (LLG), 1 (field) or 2 (MC)");
  // block may carry flags: 0x10000 non-temporal list loads, 0x20000 int32
  // entries, (k << 18) LDS-tiled stencil shape k and 0x400000 resident
  // workgroups looping over tiles (LLG only), 0x800000 a separate DM gather
  // where the DM bonds are the exchange bonds (context-wide), 0x1000000 two
  // adjacent ensembles per thread (MC only)
  uasd::Tune t{gather, block & 0xffff, (block >> 16) & 1, (block >> 17) & 1, (block >> 18) & 0xf,
               (block >> 22) & 1, (block >> 24) & 1};
  if (!uasd::tune_valid(t) || (block >> 25) != 0 || (kind != 0 && t.tile != 0) ||
      (kind != 2 && t.pair != 0))
    return fail(UASD_ERR_ARG, "gather must be 2, 4, 8 or 16 and block 64, 128 or 256");
  c->tune[kind] = t;
  c->tune_no_fuse = ((block >> 23) & 1) != 0;
  return UASD_OK;
}

int uasd_get_tuning(uasd_ctx *c, int kind, int *gather, int *block) {
  if (int r = check_ctx(c)) return r;
  if (kind < 0 || kind > 2) return fail(UASD_ERR_ARG, "kind must be 0 (LLG), 1 (field) or 2 (MC)");
  if (gather) *gather = c->tune[kind].gather;
  if (block)
    *block = c->tune[kind].block | (c->tune[kind].nt << 16) | (c->tune[kind].idx32 << 17) |
             (c->tune[kind].tile << 18) | (c->tune[kind].pipe << 22) |
             ((c->tune_no_fuse ? 1 : 0) << 23) | (c->tune[kind].pair << 24);
  return UASD_OK;
}

}  // extern "C"

namespace {

// Copies of the spin quads and the counters, restored after a tuning run so
// that tuning leaves the trajectory untouched.
struct StateGuard {
  uasd_ctx *c;
  DevBuf save;
  uint64_t step, sweep;
  bool timing;
  int err = UASD_OK;
  explicit StateGuard(uasd_ctx *cc) : c(cc), step(cc->step), sweep(cc->mc_sweep), timing(cc->timing) {
    const size_t n = 3 * quad_bytes(c);
    hipError_t e = save.ensure(n);
    if (e == hipSuccess) e = hipMemcpyAsync(save.p, c->spins.p, n, hipMemcpyDeviceToDevice, c->stream);
    if (e != hipSuccess) err = fail(UASD_ERR_HIP, std::string("tuning state copy: ") + hipGetErrorString(e));
    c->timing = false;
  }
  int restore() {
    const size_t n = 3 * quad_bytes(c);
    HIP_TRY(hipMemcpyAsync(c->spins.p, save.p, n, hipMemcpyDeviceToDevice, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    c->step = step;
    c->mc_sweep = sweep;
    c->timing = timing;
    return UASD_OK;
  }
};

struct EventPair {
  hipEvent_t a = nullptr, b = nullptr;
  ~EventPair() {
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
  }
};

// Times `run` (asynchronous launches on c->stream) for every candidate shape,
// once untimed and once timed, and leaves the fastest in *slot.
// ------------------------------------------------------ autotuner cache
// Winners of uasd_autotune_* per workload, kept for the process once
// uasd_set_tune_cache has switched the cache on, and read from / appended to
// its text file when it names one: one line per workload, "key gather block
// nt idx32 tile pipe best_ms".  The key holds what the launch shapes' speed
// depends on: the kernel family, solver or mode, precision, sizes, list
// widths, terms, site order, lattice dimensions, list entry width and noise
// source (none at a uniform T = 0).
std::mutex g_tune_mu;
bool g_tune_on = false;
std::string g_tune_path;
std::map<std::string, std::pair<uasd::Tune, double>> g_tune;

std::string tune_key(const uasd_ctx *c, const char *kind, int alg) {
  char buf[320];
  std::snprintf(buf, sizeof buf,
                "v2/gfx950/%s/a%d/%s/N%d/M%d/seg%d/W%d/Wd%d/an%d/ext%d/order%d/lat%d,%d,%d,%d/i16%d%d/"
                "noise%d",
                kind, alg, c->prec == UASD_FP64 ? "fp64" : "fp32", c->N, c->M, c->nseg, c->W,
                c->Wd, c->an_rows > 0 ? 1 : 0, c->ext_mode, c->order_na, c->lat_na, c->lat_n1,
                c->lat_n2, c->lat_n3, c->use16 ? 1 : 0, c->use16_dm ? 1 : 0,
                c->temp_uniform && c->temp0 == 0.0 ? 0 : c->noise_mode);
  return buf;
}

bool tune_lookup(const std::string &key, uasd::Tune &t, double &ms) {
  std::lock_guard<std::mutex> lock(g_tune_mu);
  if (!g_tune_on) return false;
  auto it = g_tune.find(key);
  if (it == g_tune.end()) return false;
  t = it->second.first;
  ms = it->second.second;
  return true;
}

void tune_store(const std::string &key, const uasd::Tune &t, double ms) {
  std::lock_guard<std::mutex> lock(g_tune_mu);
  if (!g_tune_on) return;
  g_tune[key] = {t, ms};
  if (g_tune_path.empty()) return;
  if (FILE *f = std::fopen(g_tune_path.c_str(), "a")) {
    std::fprintf(f, "%s %d %d %d %d %d %d %d %.6f\n", key.c_str(), t.gather, t.block, t.nt,
                 t.idx32, t.tile, t.pipe, t.pair, ms);
    std::fclose(f);
  }
}

template <typename Run>
// pair: also time two ensembles per thread (the MC colour pass, M even)
int tune_search(uasd_ctx *c, uasd::Tune *slot, Run &&run, double *best_ms, bool pair = false) {
  EventPair ev;
  HIP_TRY(hipEventCreate(&ev.a));
  HIP_TRY(hipEventCreate(&ev.b));
  const int gathers[] = {2, 4, 8, 16};
  const int blocks[] = {64, 128, 256};
  uasd::Tune best = *slot;
  float best_t = 1e30f;
  std::vector<uasd::Tune> cand;
  const int nidx = c->use16 || c->use16_dm ? 2 : 1;
  for (int ix = 0; ix < nidx; ++ix)
    for (int nt = 0; nt < 2; ++nt)
      for (int g : gathers)
        for (int bl : blocks) {
          cand.push_back(uasd::Tune{g, bl, nt, ix, 0});
          if (pair && c->M % 2 == 0) cand.push_back(uasd::Tune{g, bl, nt, ix, 0, 0, 1});
        }
  if (slot == &c->tune[0]) {
    build_stencil(c);
    for (int k = 1; k <= uasd::kNumTileShapes; ++k)
      if (tile_fits(c, k))
        for (int pipe = 0; pipe < 2; ++pipe) cand.push_back(uasd::Tune{4, 256, 0, 0, k, pipe});
  }
  for (const uasd::Tune &cd : cand) {
    {
      *slot = cd;
      if (int r = run()) return r;
      HIP_TRY(hipEventRecord(ev.a, c->stream));
      if (int r = run()) return r;
      HIP_TRY(hipEventRecord(ev.b, c->stream));
      HIP_TRY(hipEventSynchronize(ev.b));
      float ms = 0.f;
      HIP_TRY(hipEventElapsedTime(&ms, ev.a, ev.b));
      if (ms < best_t) {
        best_t = ms;
        best = *slot;
      }
    }
  }
  *slot = best;
  if (best_ms) *best_ms = best_t;
  return UASD_OK;
}

}  // namespace

extern "C" {

int uasd_autotune_llg(uasd_ctx *c, int alg, int steps, double *best_ms) {
  if (int r = check_ctx(c)) return r;
  if (!sde_supported(alg))
    return fail(UASD_ERR_UNSUPPORTED, "SDEalgh must be 1 (midpoint), 4 (Heun), 5 (Depondt), 6 (SIB) or 7 (spherical midpoint)");
  if (steps <= 0) return fail(UASD_ERR_ARG, "steps must be > 0");
  if (int r = ready_for_llg(c)) return r;
  build_stencil(c);
  const std::string key = tune_key(c, "llg", alg);
  uasd::Tune cached;
  double cached_ms = 0.0;
  if (tune_lookup(key, cached, cached_ms) && uasd::tune_valid(cached) && cached.pair == 0 &&
      (cached.tile == 0 || tile_fits(c, cached.tile))) {
    c->tune[0] = c->tune[1] = cached;
    c->tune[1].tile = c->tune[1].pipe = 0;
    if (best_ms) *best_ms = cached_ms;
    c->gen += 1;
    return UASD_OK;
  }
  StateGuard g(c);
  if (g.err) return g.err;
  double ms_found = 0.0;
  int r = tune_search(c, &c->tune[0], [&]() -> int {
    for (int s = 0; s < steps; ++s) {
      int e = c->prec == UASD_FP64 ? run_llg_half<double>(c, alg, 0, true, false)
                                   : run_llg_half<float>(c, alg, 0, true, false);
      if (e) return e;
      e = c->prec == UASD_FP64 ? run_llg_half<double>(c, alg, 1, true, true)
                               : run_llg_half<float>(c, alg, 1, true, true);
      if (e) return e;
      c->step += 1;
    }
    return UASD_OK;
  }, &ms_found);
  if (r) return r;
  if (best_ms) *best_ms = ms_found;
  tune_store(key, c->tune[0], ms_found);
  c->tune[1] = c->tune[0];  // the field kernel gathers the same way
  c->gen += 1;
  return g.restore();
}

int uasd_autotune_mc(uasd_ctx *c, char mode, const double *temps, int sweeps, double *best_ms) {
  if (int r = check_ctx(c)) return r;
  if (sweeps <= 0) return fail(UASD_ERR_ARG, "sweeps must be > 0");
  uasd::MCParams mp;
  if (int r = mc_prepare(c, mode, temps, 0x7u, mp)) return r;
  const std::string key = tune_key(c, "mc", (int)mode);
  uasd::Tune cached;
  double cached_ms = 0.0;
  if (tune_lookup(key, cached, cached_ms) && uasd::tune_valid(cached) && cached.tile == 0) {
    c->tune[2] = cached;
    if (best_ms) *best_ms = cached_ms;
    c->gen += 1;
    return UASD_OK;
  }
  StateGuard g(c);
  if (g.err) return g.err;
  double ms_found = 0.0;
  int r = tune_search(c, &c->tune[2], [&]() -> int { return mc_run(c, mp, sweeps); }, &ms_found,
                      true);
  if (r) return r;
  if (best_ms) *best_ms = ms_found;
  tune_store(key, c->tune[2], ms_found);
  c->gen += 1;
  return g.restore();
}

int uasd_set_tune_cache(int enable, const char *path) {
  std::lock_guard<std::mutex> lock(g_tune_mu);
  g_tune_on = enable != 0;
  g_tune_path = (enable && path) ? path : "";
  if (!g_tune_on) {
    g_tune.clear();
    return UASD_OK;
  }
  if (g_tune_path.empty()) return UASD_OK;
  FILE *f = std::fopen(g_tune_path.c_str(), "r");
  if (!f) return UASD_OK;  // a new file: created by the first result
  char key[320];
  uasd::Tune t;
  double ms = 0.0;
  int loaded = 0;
  while (std::fscanf(f, "%319s %d %d %d %d %d %d %d %lf", key, &t.gather, &t.block, &t.nt,
                     &t.idx32, &t.tile, &t.pipe, &t.pair, &ms) == 9) {
    if (uasd::tune_valid(t)) {
      g_tune[key] = {t, ms};
      ++loaded;
    }
  }
  std::fclose(f);
  (void)loaded;
  return UASD_OK;
}

int uasd_measure(uasd_ctx *c, double *msum, double *energy) {
  if (int r = check_ctx(c)) return r;
  if (!c->have_state) return fail(UASD_ERR_STATE, "moments not set");
  if (!msum) return fail(UASD_ERR_ARG, "msum is null");
  if (energy) {
    if (!c->have_ex) return fail(UASD_ERR_STATE, "exchange not set");
    if (int r = rebuild_classes(c)) return r;
    int r = c->prec == UASD_FP64 ? run_field<double>(c, false, true)
                                 : run_field<float>(c, false, true);
    if (r) return r;
  }
  std::vector<double> red;
  int r = c->prec == UASD_FP64 ? reduce_moments<double>(c, energy != nullptr, false, red)
                               : reduce_moments<float>(c, energy != nullptr, false, red);
  if (r) return r;
  const int nv = energy ? 4 : 3;
  for (int k = 0; k < c->M; ++k) {
    for (int d = 0; d < 3; ++d) msum[3 * k + d] = red[(size_t)nv * k + d];
    if (energy) energy[k] = red[(size_t)nv * k + 3] * uasd::kMub / uasd::kMry;
  }
  return UASD_OK;
}

int uasd_set_site_order(uasd_ctx *c, int na) {
  if (int r = check_ctx(c)) return r;
  if (na < 1 || c->N % na != 0) return fail(UASD_ERR_ARG, "na must be >= 1 and divide natom");
  if (c->have_ex || c->have_state || c->have_llg || c->Wd > 0 || c->an_rows > 0 || c->ext_mode == 2)
    return fail(UASD_ERR_STATE, "set the site order before any Hamiltonian, moment or LLG upload");
  if (c->nseg > 1 && na > 1) return fail(UASD_ERR_UNSUPPORTED, "site order with several segments");
  c->order_na = na;
  c->order_nc = c->N / na;
  c->gen += 1;
  return UASD_OK;
}

int uasd_set_lattice_dims(uasd_ctx *c, int na, int n1, int n2, int n3) {
  if (int r = check_ctx(c)) return r;
  if (na == 0 && n1 == 0 && n2 == 0 && n3 == 0) {  // forget the lattice
    c->lat_na = c->lat_n1 = c->lat_n2 = c->lat_n3 = 0;
  } else {
    if (na < 1 || n1 < 1 || n2 < 1 || n3 < 1 || (int64_t)na * n1 * n2 * n3 != c->N)
      return fail(UASD_ERR_ARG, "NA*N1*N2*N3 must equal natom");
    c->lat_na = na;
    c->lat_n1 = n1;
    c->lat_n2 = n2;
    c->lat_n3 = n3;
  }
  c->stencil_dirty = true;
  c->tile_off_shape = 0;
  c->gen += 1;
  return UASD_OK;
}

int uasd_has_stencil(uasd_ctx *c, int *has, int *radius) {
  if (int r = check_ctx(c)) return r;
  build_stencil(c);
  if (has) *has = c->stencil_ok ? 1 : 0;
  if (radius) {
    radius[0] = std::max(c->st_ex.rx, c->st_dm.rx);
    radius[1] = std::max(c->st_ex.ry, c->st_dm.ry);
    radius[2] = std::max(c->st_ex.rz, c->st_dm.rz);
  }
  return UASD_OK;
}

int uasd_set_segments(uasd_ctx *c, int nseg, const int *offsets) {
  if (int r = check_ctx(c)) return r;
  if (nseg > 1 && c->order_na > 1)
    return fail(UASD_ERR_UNSUPPORTED, "segments need the reference site order");
  if (nseg <= 0 || !offsets) return fail(UASD_ERR_ARG, "nseg must be > 0 with offsets (nseg+1)");
  if (offsets[0] != 0 || offsets[nseg] != c->N)
    return fail(UASD_ERR_ARG, "segment offsets must start at 0 and end at natom");
  for (int s = 0; s < nseg; ++s)
    if (offsets[s + 1] <= offsets[s]) return fail(UASD_ERR_ARG, "segments must be non-empty and increasing");
  return set_segments(c, nseg, offsets);
}

int uasd_measure_segments(uasd_ctx *c, double *msum, double *energy) {
  if (int r = check_ctx(c)) return r;
  if (!c->have_state) return fail(UASD_ERR_STATE, "moments not set");
  if (!msum) return fail(UASD_ERR_ARG, "msum is null");
  if (int r = ensure_segments(c)) return r;
  if (energy) {
    if (!c->have_ex) return fail(UASD_ERR_STATE, "exchange not set");
    if (int r = rebuild_classes(c)) return r;
    int r = c->prec == UASD_FP64 ? run_field<double>(c, false, true)
                                 : run_field<float>(c, false, true);
    if (r) return r;
  }
  int r = c->prec == UASD_FP64 ? segment_sums<double>(c, energy != nullptr)
                               : segment_sums<float>(c, energy != nullptr);
  if (r) return r;
  const int nv = energy ? 4 : 3;
  std::vector<double> out((size_t)c->nseg * c->M * nv);
  HIP_TRY(hipMemcpyAsync(out.data(), c->seg_out.p, out.size() * sizeof(double),
                         hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  for (size_t t = 0; t < (size_t)c->nseg * c->M; ++t) {
    for (int d = 0; d < 3; ++d) msum[3 * t + d] = out[nv * t + d];
    if (energy) energy[t] = out[nv * t + 3] * uasd::kMub / uasd::kMry;
  }
  return UASD_OK;
}

}  // extern "C"

namespace {

// calc_energy's four terms per (segment, ensemble) into ene (4,M,nseg), with
// the chunk partials in `partial` (nchunks*M*4 doubles) and the sums in `sums`
// (nseg*M*4 doubles), both device memory.  The sums are copied straight into
// ene and scaled there, so no host memory is taken either.
int energy_terms_into(uasd_ctx *c, double *partial, double *sums, double *ene) {
  const int2 *chunks = c->seg_dev.as<int2>();
  const int *seg_chunk = reinterpret_cast<const int *>(chunks + c->nchunks);
  if (c->prec == UASD_FP64) {
    HIP_TRY(uasd::launch_energy_terms<double>(c->stream, ham_view<double>(c), c->nseg,
                                              c->nchunks, chunks, seg_chunk,
                                              quads<double>(c, c->spins, 0), partial, sums));
  } else {
    HIP_TRY(uasd::launch_energy_terms<float>(c->stream, ham_view<float>(c), c->nseg,
                                             c->nchunks, chunks, seg_chunk,
                                             quads<float>(c, c->spins, 0), partial, sums));
  }
  const size_t n = (size_t)c->nseg * c->M * 4;
  HIP_TRY(hipMemcpyAsync(ene, sums, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  // per site in mRy: energy.f90:353-366 divides by the number of atoms,
  // calculate_mean_and_deviation then scales by mub/mry
  for (int sg = 0; sg < c->nseg; ++sg) {
    const double na = (double)(c->seg_off[sg + 1] - c->seg_off[sg]);
    for (int k = 0; k < c->M; ++k)
      for (int t = 0; t < 4; ++t) {
        const size_t i = ((size_t)sg * c->M + k) * 4 + t;
        ene[i] = ene[i] / na * (uasd::kMub / uasd::kMry);
      }
  }
  return UASD_OK;
}

int energy_terms_check(uasd_ctx *c, double *ene) {
  if (int r = check_ctx(c)) return r;
  if (!c->have_state) return fail(UASD_ERR_STATE, "moments not set");
  if (!c->have_ex) return fail(UASD_ERR_STATE, "exchange not set");
  if (!ene) return fail(UASD_ERR_ARG, "ene is null");
  if (c->M > 65535) return fail(UASD_ERR_UNSUPPORTED, "more than 65535 ensembles");
  return UASD_OK;
}

}  // namespace

extern "C" {

int uasd_energy_terms(uasd_ctx *c, double *ene) {
  if (int r = energy_terms_check(c, ene)) return r;
  if (int r = ensure_segments(c)) return r;
  if (int r = rebuild_classes(c)) return r;
  HIP_TRY(c->seg_partial.ensure((size_t)c->nchunks * c->M * 4 * sizeof(double)));
  HIP_TRY(c->seg_out.ensure((size_t)c->nseg * c->M * 4 * sizeof(double)));
  return energy_terms_into(c, c->seg_partial.as<double>(), c->seg_out.as<double>(), ene);
}

int uasd_energy_terms_noalloc(uasd_ctx *c, double *ene) {
  if (int r = energy_terms_check(c, ene)) return r;
  if (c->nseg == 0) return fail(UASD_ERR_STATE, "the context has no segments");
  const size_t need = ((size_t)c->nchunks + (size_t)c->nseg) * (size_t)c->M * 4;
  if (need > kScratchDoubles)
    return fail(UASD_ERR_ARG,
                "uasd_energy_terms_noalloc needs " + std::to_string(need) +
                    " doubles of scratch, more than its fixed " + std::to_string(kScratchDoubles) +
                    ": (chunks of 4096 sites + segments) * ensembles * 4 must not exceed it; "
                    "use uasd_energy_terms");
  double *base = scratch_of(c->device);
  if (!base) return fail(UASD_ERR_STATE, "no scratch address for this device");
  if (int r = rebuild_classes(c, false)) return r;
  return energy_terms_into(c, base, base + (size_t)c->nchunks * c->M * 4, ene);
}

int uasd_accumulate(uasd_ctx *c, int with_energy) {
  if (int r = check_ctx(c)) return r;
  if (!c->have_state) return fail(UASD_ERR_STATE, "moments not set");
  if (int r = ensure_segments(c)) return r;
  if (with_energy) {
    if (!c->have_ex) return fail(UASD_ERR_STATE, "exchange not set");
    if (int r = rebuild_classes(c)) return r;
    int r = c->prec == UASD_FP64 ? run_field<double>(c, false, true)
                                 : run_field<float>(c, false, true);
    if (r) return r;
  }
  int r = c->prec == UASD_FP64 ? segment_sums<double>(c, with_energy != 0)
                               : segment_sums<float>(c, with_energy != 0);
  if (r) return r;
  const int *seg_n = reinterpret_cast<const int *>(c->seg_dev.as<int2>() + c->nchunks) + c->nseg + 1;
  HIP_TRY(uasd::launch_accumulate(c->stream, c->nseg, c->M, seg_n, c->seg_out.as<double>(),
                                  with_energy ? 4 : 3, c->acc.as<double>()));
  c->acc_samples += 1;
  return UASD_OK;
}

int uasd_get_accumulators(uasd_ctx *c, double *out, int64_t *nsamples) {
  if (int r = check_ctx(c)) return r;
  if (int r = ensure_segments(c)) return r;
  if (out) {
    HIP_TRY(hipMemcpyAsync(out, c->acc.p, (size_t)c->nseg * c->M * uasd::kAccWidth * sizeof(double),
                           hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
  }
  if (nsamples) *nsamples = c->acc_samples;
  return UASD_OK;
}

int uasd_reset_accumulators(uasd_ctx *c) {
  if (int r = check_ctx(c)) return r;
  if (c->nseg == 0) return ensure_segments(c);
  std::vector<int> off(c->seg_off);
  return set_segments(c, c->nseg, off.data());
}

int uasd_get_coloring(uasd_ctx *c, int *ncolors, int *color_of_site) {
  if (int r = check_ctx(c)) return r;
  if (!c->have_ex) return fail(UASD_ERR_STATE, "exchange not set");
  if (int r = build_coloring(c)) return r;
  if (ncolors) *ncolors = c->ncolors;
  if (color_of_site)
    for (int i = 0; i < c->N; ++i) color_of_site[i] = c->color_of_site[slot_of(c, i)];
  return UASD_OK;
}

int uasd_set_kernel_timing(uasd_ctx *c, int enable) {
  if (int r = check_ctx(c)) return r;
  if (int r = resolve_timing(c)) return r;
  if (enable < -1) return fail(UASD_ERR_ARG, "enable must be >= -1");
  c->timing = enable > 0;
  c->timing_every = enable > 0 ? enable : 1;
  c->timing_region = enable != 0;
  for (auto &t : c->timers) t = Timing{};
  for (auto &k : c->timing_tick) k = 0;
  return UASD_OK;
}

int uasd_get_kernel_timing(uasd_ctx *c, int kind, int64_t *launches, double *total_ms) {
  if (int r = check_ctx(c)) return r;
  if (kind < 0 || kind > 6) return fail(UASD_ERR_ARG, "kind must be 0..6");
  if (int r = resolve_timing(c)) return r;
  if (launches) *launches = c->timers[kind].launches;
  if (total_ms) *total_ms = c->timers[kind].ms;
  return UASD_OK;
}

}  // extern "C"
