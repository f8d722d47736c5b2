// stepper.hip — the C ABI of include/tdglhip.h: device state, the order of one step
// (gtdgl/src/step.jl:13-35) and of the band update (bands.jl:330-521), the separable solve
// and the device-resident PCG (implicit.jl:62-104, 279-446).
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "kernels.h"

using namespace tdgl;

namespace {

thread_local std::string g_err;

int fail(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define HIPCHK(expr)                                                                         \
  do {                                                                                       \
    hipError_t e_ = (expr);                                                                  \
    if (e_ != hipSuccess) return fail(TDGL_E_DEVICE, "%s: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

// Dense 1D bases of the separable solve: V (nn x nn) on nodes (cosine, DCT-II), U on the nn-1
// edges (sine, DST-I), the singular values s of the forward difference in ascending order
// (the SVD of implicit.jl:15-22, whose signs cancel in the solve).  Stored for the transform
// passes as A[n*L + m] = M(n, m) (forward) and M(m, n) (inverse).
struct Bases {
  int nn;
  std::vector<double> V, U, s;
};
Bases make_bases(int nn) {
  Bases b;
  b.nn = nn;
  b.V.assign((size_t)nn * nn, 0.0);
  b.U.assign((size_t)(nn - 1) * (nn - 1), 0.0);
  b.s.assign(nn, 0.0);
  const double pi = 3.14159265358979323846;
  for (int k = 0; k < nn; ++k) {
    b.s[k] = k == 0 ? 0.0 : 2.0 * std::sin(pi * k / (2.0 * nn));
    for (int i = 0; i < nn; ++i)
      b.V[i + (size_t)nn * k] = k == 0 ? 1.0 / std::sqrt((double)nn) : std::sqrt(2.0 / nn) * std::cos(pi * k * (i + 0.5) / nn);
  }
  for (int k = 1; k < nn; ++k)
    for (int e = 0; e < nn - 1; ++e) b.U[e + (size_t)(nn - 1) * (k - 1)] = -std::sqrt(2.0 / nn) * std::sin(pi * k * (e + 1.0) / nn);
  return b;
}

// kernel classes of the per-kernel profile (tdgl_profile_get)
enum { K_STAGE = 0, K_QP, K_FIELD, K_CURL, K_GHOSTS, K_SCATTER, K_TX, K_TX32, K_MODES, K_APPLYA, K_DOT, K_FINISH, K_VEC, K_METER,
       K_FOLD, K_COUNT };
const char *const K_NAMES[K_COUNT] = {"band_stage", "normal_fluids", "current_field", "links_curl_bc", "neumann_ghosts",
                                      "solve_scatter", "sep_transform_f64", "sep_transform_f32", "sep_modes", "pcg_apply_A",
                                      "pcg_dot", "pcg_finish", "pcg_vector", "cycle_meter", "sep_fold"};
static_assert(K_COUNT == TDGL_NKERNELS, "kernel classes");

// a device or pinned host block of the process's cache
struct Block {
  void *p;
  size_t cap;
  int dev;  // -1: pinned host
};

struct PendingGet {
  void *host;
  int n1, n2, n3;
  size_t es;
};

struct HostSnap {
  cplx *G[TDGL_MAX_BANDS], *F[TDGL_MAX_BANDS];
  double *th[3], *ph[3];
  cplx *U[3], *V[3];
  int parity;
  bool warm, have2;
};
// A captured run of steps is valid for the host-side pointer set it was captured with (the parity
// of the advance's swaps), the PCG's warm-start flags and its iteration budget.
struct GraphKey {
  int n, parity, warm, have2, budget, cls;
  bool operator==(const GraphKey &o) const {
    return n == o.n && parity == o.parity && warm == o.warm && have2 == o.have2 && budget == o.budget && cls == o.cls;
  }
};
struct CachedGraph {
  bool live = false;
  GraphKey key{};
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  HostSnap end{};        // the host-side pointers and flags after its steps
  long long used = 0;    // the last use (the least recently used entry is replaced)
};
constexpr int GRAPH_CACHE = 16;  // the one-step graphs of the step classes take two parities each
constexpr int GRAPH_CACHE_CALL = 4;  // graphs of whole calls (up to GRAPH_MAX_STEPS steps each)
}  // namespace

struct tdgl_handle {
  tdgl_desc d;
  Geo g;
  Ranges rg;
  int dev = 0;
  hipStream_t s = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  std::vector<Block> allocs;   // device blocks, given back to the process's cache by tdgl_destroy
  int64_t bytes = 0;
  // state (state.jl:51-60); advance! swaps G/F, θ/φ, V/U
  cplx *G[TDGL_MAX_BANDS] = {}, *F[TDGL_MAX_BANDS] = {};
  double *th[3] = {}, *ph[3] = {};
  cplx *U[3] = {}, *V[3] = {};
  double *J[3] = {}, *h[3] = {};
  // workspace
  cplx *W = nullptr, *W2 = nullptr;
  double *Om = nullptr;
  // masks and tables
  double *f = nullptr, *tau[3] = {}, *D[3] = {};
  double *bar[6] = {};
  // separable solve: per axis, forward/inverse A for V and U, double and float; s
  double *Ad[3][2][2] = {};  // [axis][0=V,1=U][0=fwd,1=inv]
  float *Af[3][2][2] = {};
  double *sd[3] = {};
  float *sf[3] = {};
  double *T1 = nullptr, *T2 = nullptr;
  float *T32a = nullptr, *T32b = nullptr;
  ProfPair ppair{};                       // the events of the launch being profiled
  // F of the last upload (host layout): outside the block no step writes F, so it is what the
  // reference's F holds there; `stepped`: a step ran since (then F equals G on the block)
  std::vector<double> Fup[TDGL_MAX_BANDS];
  bool stepped = false;
  // the ladder's per-cycle measurements (tdgl_meter_begin): its own allocations
  std::vector<Block> mallocs;
  MeterArgs ma{};
  double *mrec = nullptr, *mrung = nullptr;
  bool meter = false;
  bool plane32 = false, plane64 = false;  // the fused transform launches fit (per precision)
  // the parity-folded Float32 preconditioner: quarters, their two buffers, the half-size bases
  // [axis][0=V,1=U][parity][0=fwd,1=inv]
  bool fold = false;
  Quarters q{};
  float *qa = nullptr, *qb = nullptr;
  float *Hq[2][2][2][2] = {};
  float *Hp[2][2][2][2] = {};  // Hq zero-padded to whole 16 x 16 tiles
  int HpL[2][2] = {};          // their unpadded order
  int qplane = 1;              // TDGL_HIP_QPLANE=0: the quarter x-y transforms on k_plane_mf1 / k_plane_mf
  int qparts = 1;              // k_qplane's blocks a quarter plane (its m tiles split; TDGL_HIP_QPARTS)
  int zq_waves = 8;        // the quarter z pass: 4 or 8 waves a block on the vector ALUs, 0 on the matrix cores (k_zmodes_mf)
  int stxcd = 1;           // the stencil kernels' blocks in XCD runs (TDGL_HIP_STENCILXCD=0: the hardware's order)
  float *zqb[4] = {};      // its padded z bases (ZQ_BASES)
  float *zqmf = nullptr;   // the same as k_zmodes_mf's image (zmodes_mf_image floats)
  // the parity-folded FP64 separable solve (field_mode SEPARABLE): the same quarters, in T1 and T2,
  // the FP64 half-size bases [axis][0=V,1=U][parity][0=fwd,1=inv] and padded z bases
  bool fold64 = false;
  size_t t12cap = 0;           // the elements of T1 and T2
  double *Hq64[2][2][2][2] = {};
  double *Hp64[2][2][2][2] = {};  // Hq64 zero-padded to whole 16 x 16 tiles (k_qplane64)
  double *zqb64[4] = {};
  bool tx_only = false;                   // the separable solve as single-axis passes (a mixed batch)
  int box_kc = 1;                         // levels of k per block of the PCG box kernels
  int plane1 = 1;                         // PlanePass::one of the Float32 launches
  // normal fluids: the sources / solved modes (region-shaped, one per mode), work, the bases
  double *qS[TDGL_MAX_BANDS] = {};
  double *qT1 = nullptr, *qT2 = nullptr;
  double *qV[3] = {}, *qW[3] = {}, *qe[3] = {};
  int qn = 0;                  // unknowns of the region
  // open boundary: the clock, h_δ of the last two recomputations (h-shaped), δJ and θ_δ per link
  // lattice (compact boxes), the padded transform (M0 x M1 x n2 complex) and Ĝ (real) per lattice
  int *ofl = nullptr;
  double *ohold[3] = {}, *ohnew[3] = {};
  double *odj[3] = {}, *oth[3] = {}, *oghat[3] = {};
  cplx *obuf[3] = {};
  cplx *otw[3][3] = {};
  int on[3][3] = {}, oM[3][3] = {};
  FftAxis oax[3][3] = {};
  double *olgf = nullptr, *oJbar = nullptr;
  // right-hand side of the implicit solve and the PCG vectors: node-shaped, zero outside the ranges
  Vec3 B{}, x{}, r{}, z{}, P{}, P2{}, Qv{}, xprev{}, zold{};
  // the refinement (d.pcg_refine): its Float32 vectors and the rounds a captured solve holds
  Pcg32 f32v{};
  float *D32[3] = {};
  bool refine = false;
  int rounds = 3;
  PcgDev pd{};
  bool warm = false, have2 = false;
  int last_iters = 0;
  int *flh = nullptr;      // pinned host copy of the flags
  // (H, HJ) of the step: device copy read by the field boundary, pinned host staging
  double *hp = nullptr, *hph = nullptr;
  // the captured ladder cycle
  bool graphs = true, capturing = false;
  int parity = 0;          // swaps of advance! modulo 2
  int budget = 0;          // PCG iterations per solve in the graph (0: not known yet)
  bool budget_fixed = false;  // set by TDGL_HIP_BUDGET: the budget only grows (roll backs)
  int shrink_run = 0, shrink_max = 0;  // cycles in a row far under the budget, their longest solve
  int shrink_hold = 0;     // cycles after a roll back in which the budget does not shrink
  // budgets per step class (PCG_NCLS): with the open boundary, a call's steps replay one-step graphs,
  // each with the budget of its class (the steps since the last recomputation of h_δ, which the host
  // follows on its copy of the device's clock); TDGL_HIP_STEPWISE=0: one budget, one graph a call
  bool stepwise = false;
  int cur_cls = 0;         // the class of the step being launched or captured
  int oclock = -1;         // the host's copy of the clock's count (ofl[OF_COUNT]), read at each call
  bool clock_known = false;  // inside tdgl_step, after the read: cur_cls is the step's own class
  int *oclk_h = nullptr;   // pinned: the clock read back
  int cbudget[PCG_NCLS] = {};
  int cshrink_run[PCG_NCLS] = {}, cshrink_max[PCG_NCLS] = {}, cshrink_hold[PCG_NCLS] = {};
  bool cseen[PCG_NCLS] = {};  // a step of the class has run (the others' budgets are never used)
  CachedGraph gc[GRAPH_CACHE];
  long long gtick = 0;
  long long captures = 0;  // graphs captured and instantiated
  int zcols = 0;           // TDGL_HIP_ZCOL=1: the z pass of whole columns per block (k_zmodes)
  int paptile = 1;         // TDGL_HIP_PAPTILE=0: the PCG operator with a gather per stencil value (k_pcg_pap)
  int boxxcd = 1;          // TDGL_HIP_BOXXCD=0: the box kernels' blocks in the hardware's order
  int zswap = 1;           // TDGL_HIP_ZSWAP=0: the folded PCG copies z into zold instead of alternating them
  uint32_t *Dcode = nullptr;   // D as a palette for the PCG operator (TDGL_HIP_DPAL=0: off)
  double *Dpal = nullptr;
  float *Dpal32 = nullptr;
  double *bk[48] = {};     // device copy of a cycle's start (roll back)
  // thermal noise: the stream's step counter on the device ([0] steps taken, [1] the step of the
  // first of the caller's draws) and its host copy; the caller's draws; SK-ROCK's δθ and source;
  // the frame rate between its two low-pass sweeps
  bool noise = false;
  uint64_t *nctr = nullptr;
  double *ndraw = nullptr;
  Block ndraw_b{};
  long long nstride = 0, nslots = 0;
  uint64_t nctr_h = 0, nbase_h = 0;
  cplx *Sp = nullptr;
  double *Qth = nullptr, *Om2 = nullptr;
  int64_t rollbacks = 0, graph_cycles = 0;
  // counters
  int64_t steps = 0, calls = 0, pcg_iters = 0;
  double device_ms = 0;
  // host staging
  std::vector<unsigned char> stage;
  std::vector<Block> stages;   // pinned staging of the copies queued by put_async / get_async
  std::vector<PendingGet> gets;                    // where get_flush writes the queued downloads
  std::vector<Block> pins;                         // pinned blocks of flh and hph
  // per-kernel profile (HIP events around each launch while enabled)
  bool prof = false;
  std::vector<hipEvent_t> evs;
  std::vector<int> evcls;
  size_t evn = 0;
  int64_t launches[K_COUNT] = {};
  double kms[K_COUNT] = {};
};

namespace {

#define CHK(expr)          \
  do {                     \
    int rc_ = (expr);      \
    if (rc_) return rc_;   \
  } while (0)

// Copies and fills of a handle go on its own stream, never on the legacy null stream: that one
// waits for the work of every blocking stream, and another handle's stream may be capturing a
// graph in another host thread at that moment.  (None of these runs inside a capture.)
int hcopy(tdgl_handle *h, void *dst, const void *src, size_t bytes, hipMemcpyKind kind) {
  HIPCHK(hipMemcpyAsync(dst, src, bytes, kind, h->s));
  HIPCHK(hipStreamSynchronize(h->s));
  return 0;
}
int hfill0(tdgl_handle *h, void *dst, size_t bytes) {
  HIPCHK(hipMemsetAsync(dst, 0, bytes, h->s));
  HIPCHK(hipStreamSynchronize(h->s));
  return 0;
}

// Device and pinned host blocks are kept by the process once a handle is done with them and
// handed to the next handle that asks for at most as much (and at least 2/3 of it): a free, of
// either kind, waits for the whole device and gives the memory back to the system, and mapping it
// again costs as much, so creating and destroying one run would stall the runs that other host
// threads drive meanwhile.  A handle waits for its own stream before it gives its blocks back,
// so a block is idle when another handle takes it.  Above CACHE_MAX idle bytes of one kind, the
// largest idle blocks are freed.
constexpr size_t CACHE_MAX = (size_t)64 << 30;
std::mutex g_cache_mu;
std::vector<Block> g_cache;
size_t g_cached[2] = {0, 0};  // idle bytes: device, pinned host

size_t round_cap(size_t b) { return (std::max<size_t>(b, 1) + 255) / 256 * 256; }
bool cache_take(int dev, size_t bytes, Block *out) {
  std::lock_guard<std::mutex> lk(g_cache_mu);
  int best = -1;
  for (size_t q = 0; q < g_cache.size(); ++q) {
    const Block &b = g_cache[q];
    if (b.dev != dev || b.cap < bytes || 2 * b.cap > 3 * bytes + 768) continue;
    if (best < 0 || b.cap < g_cache[best].cap) best = (int)q;
  }
  if (best < 0) return false;
  *out = g_cache[best];
  g_cache[best] = g_cache.back();
  g_cache.pop_back();
  g_cached[dev < 0] -= out->cap;
  return true;
}
void cache_give(const Block &b) {
  std::vector<Block> drop;
  {
    std::lock_guard<std::mutex> lk(g_cache_mu);
    g_cache.push_back(b);
    const int k = b.dev < 0;
    g_cached[k] += b.cap;
    while (g_cached[k] > CACHE_MAX) {
      int big = -1;
      for (size_t q = 0; q < g_cache.size(); ++q)
        if ((g_cache[q].dev < 0) == (k == 1) && (big < 0 || g_cache[q].cap > g_cache[big].cap)) big = (int)q;
      drop.push_back(g_cache[big]);
      g_cached[k] -= g_cache[big].cap;
      g_cache[big] = g_cache.back();
      g_cache.pop_back();
    }
  }
  for (const Block &x : drop) {
    if (x.dev < 0) {
      (void)hipHostFree(x.p);
    } else {
      int cur = 0;
      (void)hipGetDevice(&cur);
      (void)hipSetDevice(x.dev);
      (void)hipFree(x.p);
      (void)hipSetDevice(cur);
    }
  }
}
// a device block of at least `bytes` on device `dev` (the current device)
hipError_t dblock_on(int dev, size_t bytes, Block *out) {
  const size_t cap = round_cap(bytes);
  if (cache_take(dev, cap, out)) return hipSuccess;
  out->cap = cap;
  out->dev = dev;
  return hipMalloc(&out->p, cap);
}
hipError_t dblock(tdgl_handle *h, size_t bytes, Block *out) { return dblock_on(h->dev, bytes, out); }
// a pinned host block of at least `bytes`
hipError_t hblock(size_t bytes, Block *out) {
  const size_t cap = round_cap(bytes);
  if (cache_take(-1, cap, out)) return hipSuccess;
  out->cap = cap;
  out->dev = -1;
  return hipHostMalloc(&out->p, cap, hipHostMallocDefault);
}

template <class T>
int dalloc(tdgl_handle *h, T **p, size_t n) {
  if (n == 0) n = 1;
  Block b{};
  if (dblock(h, n * sizeof(T), &b) != hipSuccess) return fail(TDGL_E_NOMEM, "hipMalloc of %zu bytes failed", n * sizeof(T));
  h->allocs.push_back(b);
  void *q = b.p;
  // zeroed on the handle's stream; tdgl_create waits for the fills once, at its end
  if (hipMemsetAsync(q, 0, n * sizeof(T), h->s) != hipSuccess) return fail(TDGL_E_DEVICE, "hipMemsetAsync failed");
  h->bytes += (int64_t)(n * sizeof(T));
  *p = (T *)q;
  return 0;
}

#define ALLOC(ptr, n)                        \
  do {                                       \
    int rc_ = dalloc(h, &(ptr), (size_t)(n)); \
    if (rc_) return rc_;                     \
  } while (0)

// a host vector to the device: staged in a pinned block of the handle's and queued on its stream
// (tdgl_create waits once, at its end, for every copy and fill it queued, then gives the staging
// back); a synchronous pageable copy each would wait on the device's copy queue, shared with the
// runs other host threads drive
template <class T>
int upload_vec(tdgl_handle *h, T *dst, const std::vector<T> &v) {
  const size_t bytes = v.size() * sizeof(T);
  if (bytes == 0) return 0;
  Block b{};
  HIPCHK(hblock(bytes, &b));
  h->stages.push_back(b);
  std::memcpy(b.p, v.data(), bytes);
  HIPCHK(hipMemcpyAsync(dst, b.p, bytes, hipMemcpyHostToDevice, h->s));
  return 0;
}

// Julia-shaped (n1, n2, n3) host array of `es`-byte elements -> node-shaped device array, the
// copy queued on the handle's stream from a staging buffer of its own; `flush` waits for the
// queued copies and drops their staging
void stages_back(tdgl_handle *h) {
  for (const Block &b : h->stages) cache_give(b);
  h->stages.clear();
}
int put_async(tdgl_handle *h, void *dev, const void *host, int n1, int n2, int n3, size_t es) {
  const Geo &g = h->g;
  const size_t bytes = (size_t)g.n * es;
  Block b{};
  HIPCHK(hblock(bytes, &b));
  h->stages.push_back(b);
  unsigned char *st = (unsigned char *)b.p;
  const unsigned char *src = (const unsigned char *)host;
  const size_t row = (size_t)n1 * es, prow = (size_t)g.px * es;
  for (int k = 0; k < g.Nz + 1; ++k)
    for (int j = 0; j < g.py; ++j) {
      unsigned char *dst = st + ((size_t)g.px * (j + (size_t)g.py * k)) * es;
      if (j < n2 && k < n3) {
        std::memcpy(dst, src + ((size_t)n1 * (j + (size_t)n2 * k)) * es, row);
        std::memset(dst + row, 0, prow - row);
      } else {
        std::memset(dst, 0, prow);
      }
    }
  HIPCHK(hipMemcpyAsync(dev, st, bytes, hipMemcpyHostToDevice, h->s));
  return 0;
}
int flush(tdgl_handle *h) {
  const hipError_t e = hipStreamSynchronize(h->s);
  stages_back(h);
  if (e != hipSuccess) return fail(TDGL_E_DEVICE, "the queued copies: %s", hipGetErrorString(e));
  return 0;
}
int put(tdgl_handle *h, void *dev, const void *host, int n1, int n2, int n3, size_t es) {
  CHK(put_async(h, dev, host, n1, n2, n3, es));
  return flush(h);
}
// node-shaped device array -> Julia-shaped host array: the copy queued into a staging buffer of
// its own, the host array written by get_flush after one wait for all queued copies
int get_async(tdgl_handle *h, void *host, const void *dev, int n1, int n2, int n3, size_t es) {
  const Geo &g = h->g;
  Block b{};
  HIPCHK(hblock((size_t)g.n * es, &b));
  h->stages.push_back(b);
  h->gets.push_back(PendingGet{host, n1, n2, n3, es});
  HIPCHK(hipMemcpyAsync(b.p, dev, (size_t)g.n * es, hipMemcpyDeviceToHost, h->s));
  return 0;
}
int get_flush(tdgl_handle *h) {
  const Geo &g = h->g;
  const hipError_t e = hipStreamSynchronize(h->s);
  if (e == hipSuccess)
    for (size_t q = 0; q < h->gets.size(); ++q) {
      const PendingGet &pg = h->gets[q];
      const unsigned char *st = (const unsigned char *)h->stages[q].p;
      unsigned char *dst = (unsigned char *)pg.host;
      for (int k = 0; k < pg.n3; ++k)
        for (int j = 0; j < pg.n2; ++j)
          std::memcpy(dst + ((size_t)pg.n1 * (j + (size_t)pg.n2 * k)) * pg.es,
                      &st[((size_t)g.px * (j + (size_t)g.py * k)) * pg.es], (size_t)pg.n1 * pg.es);
    }
  h->gets.clear();
  stages_back(h);
  if (e != hipSuccess) return fail(TDGL_E_DEVICE, "the queued copies: %s", hipGetErrorString(e));
  return 0;
}
int get(tdgl_handle *h, void *host, const void *dev, int n1, int n2, int n3, size_t es) {
  const Geo &g = h->g;
  h->stage.resize((size_t)g.n * es);
  CHK(hcopy(h, h->stage.data(), dev, (size_t)g.n * es, hipMemcpyDeviceToHost));
  unsigned char *dst = (unsigned char *)host;
  for (int k = 0; k < n3; ++k)
    for (int j = 0; j < n2; ++j)
      std::memcpy(dst + ((size_t)n1 * (j + (size_t)n2 * k)) * es, &h->stage[((size_t)g.px * (j + (size_t)g.py * k)) * es],
                  (size_t)n1 * es);
  return 0;
}



int prof_flush(tdgl_handle *h) {
  if (h->evn == 0) return 0;
  HIPCHK(hipStreamSynchronize(h->s));
  for (size_t q = 0; q < h->evn; q += 2) {
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, h->evs[q], h->evs[q + 1]));
    h->kms[h->evcls[q / 2]] += ms;
    h->launches[h->evcls[q / 2]] += 1;
  }
  h->evn = 0;
  return 0;
}
int prof_mark(tdgl_handle *h, int cls, bool begin) {
  if (begin && h->evn + 2 > 16384) CHK(prof_flush(h));
  if (h->evn + 1 >= h->evs.size()) {
    hipEvent_t e;
    HIPCHK(hipEventCreate(&e));
    h->evs.push_back(e);
    HIPCHK(hipEventCreate(&e));
    h->evs.push_back(e);
  }
  if (begin) {
    if (h->evcls.size() <= h->evn / 2) h->evcls.resize(h->evn / 2 + 1);
    h->evcls[h->evn / 2] = cls;
    h->ppair = ProfPair{h->evs[h->evn], h->evs[h->evn + 1], false};
    prof_pair = &h->ppair;  // the launch records the pair at the kernel's start and end
  } else {
    prof_pair = nullptr;
    if (!h->ppair.used) {  // no kernel was launched: an empty interval
      HIPCHK(hipEventRecord(h->evs[h->evn], h->s));
      HIPCHK(hipEventRecord(h->evs[h->evn + 1], h->s));
    }
    h->evn += 2;
  }
  return 0;
}
// a launch, with its events when the profile is on
#define PL(cls, expr)                                                                                 \
  do {                                                                                                \
    if (h->prof) CHK(prof_mark(h, (cls), true));                                                      \
    hipError_t e_ = (expr);                                                                           \
    if (h->prof) prof_pair = nullptr;                                                                 \
    if (e_ != hipSuccess) return fail(TDGL_E_DEVICE, "launch %s: %s", #expr, hipGetErrorString(e_));  \
    if (h->prof) CHK(prof_mark(h, (cls), false));                                                     \
  } while (0)
#define LCHK(expr)                                                                                   \
  do {                                                                                               \
    hipError_t e_ = (expr);                                                                          \
    if (e_ != hipSuccess) return fail(TDGL_E_DEVICE, "launch %s: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

// ---------------------------------------------------------------------------- band update
NoiseArgs noise_args(const tdgl_handle *h) {
  NoiseArgs z{};
  z.ctr = h->nctr;
  z.hd = h->ndraw;
  z.hstride = h->nstride;
  z.nslots = h->nslots;
  z.seed = h->d.noise_seed;
  z.host = h->d.noise_host;
  z.nb = h->d.nbands;
  z.Tdt = h->d.noise_T * h->d.dt;
  z.dV = h->d.dx * h->d.dy * h->d.dz;
  return z;
}

StageArgs stage_base(tdgl_handle *h, int mu, const tdgl_bandcoef &cf) {
  StageArgs a{};
  a.xcd = h->stxcd;
  for (int nu = 0; nu < h->d.nbands; ++nu) a.G[nu] = h->G[nu];
  for (int c = 0; c < 3; ++c) {
    a.V[c] = h->V[c];
    a.U[c] = h->U[c];
  }
  a.f = h->f;
  a.Om = h->Om;
  a.kx = cf.kx;
  a.ky = cf.ky;
  a.kz = cf.kz;
  a.c = cf.c;
  a.eps = cf.eps;
  a.beta = cf.beta;
  a.gamma2 = cf.gamma2;
  for (int nu = 0; nu < TDGL_MAX_BANDS; ++nu) a.cpl[nu] = cf.cpl[nu];
  a.mu = mu;
  a.nb = h->d.nbands;
  a.wmax = 0.5 / h->d.dt;  // _ROT_MAXDT / dt (bands.jl:209, 216)
  a.dt = h->d.dt;
  a.hard = h->d.hard;
  a.lsrc = LS_DIRECT;
  if (h->noise) {
    a.kwt_rot = 1;
    a.nz = noise_args(h);
    a.Om2 = h->Om2;
    a.Qth = h->Qth;
    a.nGm = h->d.noise_Gamma[mu];
    a.ng2 = h->d.bcs[mu].gamma2;
    a.nsk = h->d.noise_sk[mu];
  }
  return a;
}

// the frame mode of a band's first (sub)stage and of the later ones: with noise the rate is
// low-passed and the terms are rotations (bands.jl:210-211)
int frame_first(const tdgl_handle *h) { return h->d.rotframe ? (h->noise ? FRAME_RAW : FRAME_RATE) : FRAME_NONE; }
int frame_later(const tdgl_handle *h) { return h->d.rotframe ? (h->noise ? FRAME_TURNTERM : FRAME_TERM) : FRAME_NONE; }

// a stage launch, and with FRAME_RAW the low-pass sweeps that finish it
int stage_launch(tdgl_handle *h, const StageArgs &a) {
  PL(K_STAGE, launch_stage(h->g, a, h->s));
  if (a.frame == FRAME_RAW) PL(K_STAGE, launch_frame_smooth(h->g, a, h->s));
  return 0;
}

// the turn of band µ by its normal-fluid potential, on the launch that writes its new values
void stage_qp(tdgl_handle *h, StageArgs &a, int mu) {
  const tdgl_desc &d = h->d;
  if (d.qp_mode == TDGL_QP_NONE) return;
  a.qp = d.qp_mode;
  a.qpK = d.qp_nmodes;
  for (int k = 0; k < d.qp_nmodes; ++k) a.qpM[k] = h->qS[k];
  a.qn1 = d.Nx - 1;
  a.qnj = d.qp_nj;
  a.qj0 = d.qp_j0;
  a.qk0 = d.qp_k0;
  if (d.qp_mode == TDGL_QP_TWO)
    a.qpc[0] = d.qp_turn[mu];
  else
    for (int k = 0; k < d.qp_nmodes; ++k) a.qpc[k] = d.qp_P[mu][k];
}

// _band_rkl! (bands.jl:464-521): stage j writes buf[(j - ns) mod 3] of (F, W, W2) so that the
// last stage lands in F and no stage writes a buffer it reads.  The links of stage j >= 2 are
// V (V Ū)^c_j, or V _cis(R c_j) in the rotating frame, evaluated inside the stage kernel.
int band_rkl(tdgl_handle *h, int mu, int ns) {
  const tdgl_desc &d = h->d;
  const bool kc = d.rkc[mu] != 0;
  const double q = kc ? d.divisor[mu] : (double)(ns * (ns + 1) / 2);
  const double hh = d.dt / q;
  const bool rot = d.rotframe, ex = d.extrap;
  cplx *buf[3] = {h->F[mu], h->W, h->W2};
  const cplx *A = h->G[mu], *B = h->G[mu];
  const bool lg = h->noise && d.bcs[mu].gamma2 > 0;   // _logcomb (bands.jl:455)
  const bool skq = h->noise && d.noise_sk[mu];         // _skrock (bands.jl:446)
  for (int j = 1; j <= ns; ++j) {
    cplx *C = buf[(((j - ns) % 3) + 3) % 3];
    StageArgs a = stage_base(h, mu, d.bcs[mu]);
    a.D = C;
    a.S = B;
    a.A = A;
    if (ex && j >= 2) {
      // the links of the previous stage's fraction c_{j-1} (set after stage j-1, bands.jl:502-513)
      a.lsrc = rot ? LS_ROTLINKS : LS_ROTFRAC;
      a.lc = kc ? d.rkc_c[mu][j - 2] : (double)((j - 1) * j) / (double)(ns * (ns + 1));
    }
    a.ghost_fly = j >= 2;
    if (j == 1 && skq) {
      // SK-ROCK's first stage: K₁ = E(Ψⁿ e^{iν₁δθ}) e^{i(κ₁−ν₁)δθ} (bands.jl:482-487)
      SkArgs k{};
      k.G = h->G[mu];
      k.f = h->f;
      k.Qth = h->Qth;
      k.Sp = h->Sp;
      k.mu = mu;
      k.hard = d.hard;
      k.c = std::sqrt(d.noise_T * d.dt / (d.dx * d.dy * d.dz));
      k.g2 = d.bcs[mu].gamma2;
      k.Gm = d.noise_Gamma[mu];
      k.nu1 = d.noise_sk_w[mu][0];
      k.nz = noise_args(h);
      PL(K_STAGE, launch_sk_src(h->g, k, h->s));
      a.S = h->Sp;
      a.ghost_fly = 1;
      a.sk_on = 1;
      a.sk_post = d.noise_sk_w[mu][1] - d.noise_sk_w[mu][0];
    }
    a.frame = j == 1 ? frame_first(h) : frame_later(h);
    a.h = hh;
    if (j >= 2) {
      a.combine = 1;
      a.logcomb = lg;
      a.cmu = kc ? d.rkc_mu[mu][j - 1] : (2.0 * j - 1) / j;
      a.cnu = kc ? d.rkc_nu[mu][j - 1] : (1.0 - j) / j;
    }
    a.turn = rot && j == ns;
    if (j == ns) {
      stage_qp(h, a, mu);
      a.noise = h->noise;
    }
    CHK(stage_launch(h, a));
    A = B;
    B = C;
  }
  return 0;
}

// _band_substeps! (bands.jl:378-406): substep m >= 2 uses V Q^(m-1) (times the frame's turn of
// each link per substep), evaluated inside the stage kernel.
int band_substeps(tdgl_handle *h, int mu, int n) {
  const tdgl_desc &d = h->d;
  const bool rot = d.rotframe, ex = d.extrap;
  const double hh = d.dt / n;
  const cplx *src = h->G[mu];
  for (int m = 1; m <= n; ++m) {
    cplx *dst = ((n - m) % 2 == 0) ? h->F[mu] : h->W;
    StageArgs a = stage_base(h, mu, d.bcs[mu]);
    a.D = dst;
    a.S = src;
    if (ex && m >= 2) {
      a.lsrc = LS_SUBSTEP;
      a.msub = m;
      a.nsub = n;
      a.lframe = rot;
    }
    a.ghost_fly = m >= 2;
    a.frame = m == 1 ? frame_first(h) : frame_later(h);
    a.h = hh;
    a.turn = rot && m == n;
    if (m == n) {
      stage_qp(h, a, mu);
      a.noise = h->noise;
    }
    CHK(stage_launch(h, a));
    src = dst;
  }
  return 0;
}

// update_bands! (bands.jl:330-346)
int update_bands(tdgl_handle *h) {
  const tdgl_desc &d = h->d;
  for (int mu = 0; mu < d.nbands; ++mu) {
    if (d.stages[mu] > 0) {
      CHK(band_rkl(h, mu, d.stages[mu]));
    } else if (d.nsub[mu] == 1) {
      StageArgs a = stage_base(h, mu, d.bc[mu]);
      a.D = h->F[mu];
      a.S = h->G[mu];
      a.frame = frame_first(h);
      a.h = d.dt;
      a.turn = d.rotframe;
      stage_qp(h, a, mu);
      a.noise = h->noise;
      CHK(stage_launch(h, a));
    } else {
      CHK(band_substeps(h, mu, d.nsub[mu]));
    }
  }
  return 0;
}

// ---------------------------------------------------------------------------- separable solve
View node_view(tdgl_handle *h, const double *base) {
  return View{(void *)(base + h->g.at(2, 2, 2)), h->g.px, h->g.px * h->g.py};
}
View node_view(tdgl_handle *h, const float *base) {
  return View{(void *)(base + h->g.at(2, 2, 2)), h->g.px, h->g.px * h->g.py};
}
View compact_view(tdgl_handle *h, const void *base, int c, size_t es) {
  return View{(void *)((const unsigned char *)base + (size_t)h->rg.off[c] * es), h->rg.e[c][0],
              h->rg.e[c][0] * h->rg.e[c][1]};
}

TxPass tx_pass(tdgl_handle *h, int axis, bool inv, bool f32) {
  TxPass p{};
  p.axis = axis;
  p.d = h->pd;
  for (int c = 0; c < 3; ++c) {
    for (int q = 0; q < 3; ++q) p.e[c][q] = h->rg.e[c][q];
    const int which = (axis == c) ? 1 : 0;  // U on the component's own axis, V elsewhere
    p.A[c] = f32 ? (const void *)h->Af[axis][which][inv] : (const void *)h->Ad[axis][which][inv];
  }
  return p;
}

// Forward (x then y) or inverse (x then y) transforms of the plane launch: U on a component's own
// axis, V on the others
PlanePass plane_pass(tdgl_handle *h, bool inv, bool f32) {
  PlanePass p{};
  p.d = h->pd;
  p.one = h->plane1;
  for (int c = 0; c < 3; ++c) {
    for (int q = 0; q < 3; ++q) p.e[c][q] = h->rg.e[c][q];
    const int wx = c == 0 ? 1 : 0, wy = c == 1 ? 1 : 0;
    p.X[c] = f32 ? (const void *)h->Af[0][wx][inv] : (const void *)h->Ad[0][wx][inv];
    p.Y[c] = f32 ? (const void *)h->Af[1][wy][inv] : (const void *)h->Ad[1][wy][inv];
    // the inverse basis is the forward one transposed
    p.XT[c] = f32 ? (const void *)h->Af[0][wx][!inv] : (const void *)h->Ad[0][wx][!inv];
    p.YT[c] = f32 ? (const void *)h->Af[1][wy][!inv] : (const void *)h->Ad[1][wy][!inv];
  }
  return p;
}

ZModes zmodes(tdgl_handle *h, const void *in, void *out, bool f32) {
  ZModes m{};
  m.in = in;
  m.out = out;
  for (int w = 0; w < 2; ++w) {
    m.Az[w] = f32 ? (const void *)h->Af[2][w][0] : (const void *)h->Ad[2][w][0];
    m.Ai[w] = f32 ? (const void *)h->Af[2][w][1] : (const void *)h->Ad[2][w][1];
  }
  m.sx = f32 ? (const void *)h->sf[0] : (const void *)h->sd[0];
  m.sy = f32 ? (const void *)h->sf[1] : (const void *)h->sd[1];
  m.sz = f32 ? (const void *)h->sf[2] : (const void *)h->sd[2];
  for (int c = 0; c < 3; ++c) m.off[c] = h->rg.off[c];
  m.nx = h->g.Nx - 1;
  m.ny = h->g.Ny - 1;
  m.nz = h->g.Nz - 1;
  m.a = h->d.a;
  m.columns = h->zcols;
  return m;
}

// What the last inverse pass does with the solution
struct SepOut {
  Vec3 *out;          // a node-shaped vector
  bool addth;         // or φ = θ + value
  int dot;            // with the partials of r·z (1) and r·zold (2)
  const Vec3 *zold = nullptr;  // the zold of the dot (the handle's when null)
  // Float32 (the refinement's inner solve; the folded path only): out, its r and zold
  Vec3f *out32 = nullptr;
  const Vec3f *r32 = nullptr, *zold32 = nullptr;
};

// The quarters of the parity folds of the solve region (Quarters): four per component, in the
// order (component, 2 px + py); qe[4 c + k] their extents
Quarters quarters_of(const Ranges &rg, int qe[12][3]) {
  Quarters q{};
  int off = 0;
  for (int c = 0; c < 3; ++c) {
    q.L[c][0] = rg.e[c][0];
    q.L[c][1] = rg.e[c][1];
    q.h[c][0] = (rg.e[c][0] + 1) / 2;
    q.h[c][1] = (rg.e[c][1] + 1) / 2;
    q.e2[c] = rg.e[c][2];
    for (int k = 0; k < 4; ++k) {
      q.off[c][k] = off;
      off += q.h[c][0] * q.h[c][1] * q.e2[c];
      qe[4 * c + k][0] = q.h[c][0];
      qe[4 * c + k][1] = q.h[c][1];
      qe[4 * c + k][2] = q.e2[c];
    }
  }
  q.total = off;
  return q;
}

// (I + aK)^{-1} of the node-shaped vector `in` (_sep_solve!, implicit.jl:62-83, or in Float32
// _precond32!, :353-373): forward passes along x and y, the pass along z fused with the mode
// projector, the three inverse passes.  `conv`: the convergence check of the first pass.
// The x-y passes of the folded preconditioner over the twelve quarters (three components, four
// parities each): the quarter (c, px, py) takes the half-size bases of parity px along x and py
// along y of the component's own bases (U on its own axis, V elsewhere).
PlanePass quarter_pass(tdgl_handle *h, bool inv, float *in, float *out) {
  PlanePass p{};
  p.d = h->pd;
  p.one = h->plane1;
  p.ncomp = 12;
  const Quarters &q = h->q;
  for (int c = 0; c < 3; ++c)
    for (int px = 0; px < 2; ++px)
      for (int py = 0; py < 2; ++py) {
        const int k = 4 * c + 2 * px + py, qo = q.off[c][2 * px + py];
        p.e[k][0] = q.h[c][0];
        p.e[k][1] = q.h[c][1];
        p.e[k][2] = q.e2[c];
        p.in[k] = View{(void *)(in + qo), q.h[c][0], q.h[c][0] * q.h[c][1]};
        p.out[k] = View{(void *)(out + qo), q.h[c][0], q.h[c][0] * q.h[c][1]};
        const int wx = c == 0 ? 1 : 0, wy = c == 1 ? 1 : 0;
        p.X[k] = h->Hq[0][wx][px][inv];
        p.Y[k] = h->Hq[1][wy][py][inv];
        p.XT[k] = h->Hq[0][wx][px][!inv];
        p.YT[k] = h->Hq[1][wy][py][!inv];
      }
  return p;
}

// the same passes for k_qplane (the padded bases)
QPass qplane_pass(tdgl_handle *h, bool inv, float *in, float *out) {
  const PlanePass p = quarter_pass(h, inv, in, out);
  QPass qp{};
  qp.d = h->pd;
  qp.parts = h->qparts;
  for (int c = 0; c < 3; ++c)
    for (int px = 0; px < 2; ++px)
      for (int py = 0; py < 2; ++py) {
        const int k = 4 * c + 2 * px + py;
        qp.in[k] = p.in[k];
        qp.out[k] = p.out[k];
        for (int e = 0; e < 3; ++e) qp.e[k][e] = p.e[k][e];
        const int wx = c == 0 ? 1 : 0, wy = c == 1 ? 1 : 0;
        qp.XT[k] = h->Hp[0][wx][px][!inv];
        qp.YT[k] = h->Hp[1][wy][py][!inv];
        // the padded bases are 16 ceil(e/16) square: a quarter of another extent stays off k_qplane
        if (h->HpL[0][wx] != p.e[k][0] || h->HpL[1][wy] != p.e[k][1]) qp.e[k][2] = 0;
      }
  return qp;
}

// one x-y launch of the quarters: k_qplane where it takes them, else the plane kernels
int quarter_xy(tdgl_handle *h, bool inv, float *in, float *out, const int *gate, Conv conv) {
  if (h->qplane) {
    QPass qp = qplane_pass(h, inv, in, out);
    qp.conv = conv;
    if (qplane_fits(qp)) {
      PL(K_TX32, launch_qplane(qp, gate, h->s));
      return 0;
    }
  }
  PlanePass p = quarter_pass(h, inv, in, out);
  p.conv = conv;
  PL(K_TX32, (launch_plane<float, float, float>(p, gate, h->s)));
  return 0;
}

// (I + aK)^{-1} in Float32 with the folds (_precond32!, implicit.jl:353-373): fold x and y, the x-y
// passes of the quarters, the z pass with the projector on the quarters, the inverse x-y passes,
// unfold into the node-shaped output with the partials of its dots.
FoldPass fold_pass(tdgl_handle *h, const Vec3 &in) {
  FoldPass f{};
  for (int c = 0; c < 3; ++c) f.in[c] = node_view(h, in.c[c]);
  f.out = h->qa;
  f.q = h->q;
  return f;
}
FoldPass fold_pass(tdgl_handle *h, const Vec3f &in) {
  FoldPass f{};
  for (int c = 0; c < 3; ++c) f.in[c] = node_view(h, in.c[c]);
  f.out = h->qa;
  f.q = h->q;
  f.f32 = 1;
  return f;
}

// what follows the fold (h->qa holds the folded input): the forward x-y launch, which takes the
// convergence check of a first pass, the z pass, the inverse x-y launch, the unfold
int sep_folded_tail(tdgl_handle *h, const SepOut &o, const int *gate, Conv conv) {
  CHK(quarter_xy(h, false, h->qa, h->qb, gate, conv));
  const int *g2 = conv.mode != CONV_NONE ? conv.out : gate;
  ZModesQ zq{};
  zq.modes = h->qb;
  zq.q = h->q;
  for (int k = 0; k < 4; ++k) zq.zb[k] = h->zqb[k];
  zq.zmf = h->zqmf;
  zq.sx = h->sf[0];
  zq.sy = h->sf[1];
  zq.sz = h->sf[2];
  zq.nx = h->g.Nx - 1;
  zq.ny = h->g.Ny - 1;
  zq.nz = h->g.Nz - 1;
  zq.a = h->d.a;
  zq.waves = h->zq_waves;
  PL(K_MODES, launch_zmodes_q(zq, g2, h->s));
  CHK(quarter_xy(h, true, h->qb, h->qa, g2, Conv{}));
  UnfoldPass u{};
  u.in = h->qa;
  u.q = h->q;
  for (int c = 0; c < 3; ++c) {
    if (o.out32) {
      u.f32 = 1;
      u.out[c] = node_view(h, o.out32->c[c]);
      if (o.dot) {
        u.r[c] = node_view(h, o.r32->c[c]);
        if (o.zold32) u.zold[c] = node_view(h, o.zold32->c[c]);
      }
      continue;
    }
    u.out[c] = node_view(h, o.out->c[c]);
    if (o.dot) {
      u.r[c] = node_view(h, h->r.c[c]);
      u.zold[c] = node_view(h, (o.zold ? o.zold : &h->zold)->c[c]);
    }
  }
  u.dot = o.dot;
  u.d = h->pd;
  PL(K_FOLD, launch_unfold(u, g2, h->s));
  return 0;
}

int sep_solve_folded(tdgl_handle *h, const Vec3 &in, const SepOut &o, const int *gate, Conv conv) {
  if (o.addth) return fail(TDGL_E_ARG, "the folded preconditioner writes z, not φ");
  // the fold runs where the solve does (the incoming flag of a first pass); the forward x-y launch
  // decides on convergence
  PL(K_FOLD, launch_fold(fold_pass(h, in), conv.mode != CONV_NONE ? conv.in : gate, h->s));
  return sep_folded_tail(h, o, gate, conv);
}

// The FP64 separable solve with the parity folds (_sep_solve!, implicit.jl:62-83): fold x and y
// of the three components into the quarters (T1), their x-y transforms with the FP64 half-size
// bases (T1 -> T2), the z pass with the projector on the quarters (in T2), the inverse x-y
// transforms (T2 -> T1), the unfold into the output (φ = θ + the solve with addth).
PlanePass quarter_pass64(tdgl_handle *h, bool inv, double *in, double *out) {
  PlanePass p{};
  p.d = h->pd;
  p.one = h->plane1;
  p.ncomp = 12;
  const Quarters &q = h->q;
  for (int c = 0; c < 3; ++c)
    for (int px = 0; px < 2; ++px)
      for (int py = 0; py < 2; ++py) {
        const int k = 4 * c + 2 * px + py, qo = q.off[c][2 * px + py];
        p.e[k][0] = q.h[c][0];
        p.e[k][1] = q.h[c][1];
        p.e[k][2] = q.e2[c];
        p.in[k] = View{(void *)(in + qo), q.h[c][0], q.h[c][0] * q.h[c][1]};
        p.out[k] = View{(void *)(out + qo), q.h[c][0], q.h[c][0] * q.h[c][1]};
        const int wx = c == 0 ? 1 : 0, wy = c == 1 ? 1 : 0;
        p.X[k] = h->Hq64[0][wx][px][inv];
        p.Y[k] = h->Hq64[1][wy][py][inv];
        p.XT[k] = h->Hq64[0][wx][px][!inv];
        p.YT[k] = h->Hq64[1][wy][py][!inv];
      }
  return p;
}
// one FP64 x-y launch of the quarters: k_qplane64 where it takes them (TDGL_HIP_QPLANE=0: never),
// else the plane kernels
int quarter_xy64(tdgl_handle *h, bool inv, double *in, double *out) {
  const PlanePass p = quarter_pass64(h, inv, in, out);
  if (h->qplane) {
    QPass64 qp{};
    for (int k = 0; k < 12; ++k) {
      qp.in[k] = p.in[k];
      qp.out[k] = p.out[k];
      for (int e = 0; e < 3; ++e) qp.e[k][e] = p.e[k][e];
      const int c = k / 4, px = (k / 2) % 2, py = k % 2;
      const int wx = c == 0 ? 1 : 0, wy = c == 1 ? 1 : 0;
      qp.XT[k] = h->Hp64[0][wx][px][!inv];
      qp.YT[k] = h->Hp64[1][wy][py][!inv];
      // the padded bases are 16 ceil(e/16) square: a quarter of another extent stays off k_qplane64
      if (h->HpL[0][wx] != p.e[k][0] || h->HpL[1][wy] != p.e[k][1]) qp.e[k][2] = 0;
    }
    if (qplane64_fits(qp)) {
      PL(K_TX, launch_qplane64(qp, h->s));
      return 0;
    }
  }
  PL(K_TX, (launch_plane<double, double, double>(p, nullptr, h->s)));
  return 0;
}
int sep_solve_folded64(tdgl_handle *h, const Vec3 &in, const SepOut &o) {
  FoldPass f{};
  for (int c = 0; c < 3; ++c) f.in[c] = node_view(h, in.c[c]);
  f.out64 = h->T1;
  f.q = h->q;
  PL(K_FOLD, launch_fold(f, nullptr, h->s));
  CHK(quarter_xy64(h, false, h->T1, h->T2));
  ZModesQ64 zq{};
  zq.modes = h->T2;
  zq.q = h->q;
  for (int k = 0; k < 4; ++k) zq.zb[k] = h->zqb64[k];
  zq.sx = h->sd[0];
  zq.sy = h->sd[1];
  zq.sz = h->sd[2];
  zq.nx = h->g.Nx - 1;
  zq.ny = h->g.Ny - 1;
  zq.nz = h->g.Nz - 1;
  zq.a = h->d.a;
  zq.waves = h->zq_waves == 4 ? 4 : 8;
  PL(K_MODES, launch_zmodes_q64(zq, h->s));
  CHK(quarter_xy64(h, true, h->T2, h->T1));
  UnfoldPass u{};
  u.in64 = h->T1;
  u.q = h->q;
  u.addth = o.addth;
  for (int c = 0; c < 3; ++c) {
    if (o.addth) {
      u.out[c] = node_view(h, h->ph[c]);
      u.th[c] = node_view(h, h->th[c]);
    } else {
      u.out[c] = node_view(h, o.out->c[c]);
    }
  }
  u.d = h->pd;
  PL(K_FOLD, launch_unfold(u, nullptr, h->s));
  return 0;
}

int sep_solve(tdgl_handle *h, const Vec3 &in, const SepOut &o, const int *gate, bool f32, Conv conv) {
  if (f32 && h->fold) return sep_solve_folded(h, in, o, gate, conv);
  if (!f32 && h->fold64 && !h->tx_only && !gate && conv.mode == CONV_NONE && !o.dot && !o.out32)
    return sep_solve_folded64(h, in, o);
  void *Ta = f32 ? (void *)h->T32a : (void *)h->T1;
  void *Tb = f32 ? (void *)h->T32b : (void *)h->T2;
  const size_t es = f32 ? 4 : 8;
  if (!h->tx_only && (f32 ? h->plane32 : h->plane64)) {
    PlanePass f = plane_pass(h, false, f32);
    for (int c = 0; c < 3; ++c) {
      f.in[c] = node_view(h, in.c[c]);
      f.out[c] = compact_view(h, Ta, c, es);
    }
    f.conv = conv;
    if (f32)
      PL(K_TX32, (launch_plane<double, float, float>(f, gate, h->s)));
    else
      PL(K_TX, (launch_plane<double, double, double>(f, gate, h->s)));
    const int *g2 = conv.mode != CONV_NONE ? conv.out : gate;
    ZModes m = zmodes(h, Ta, Tb, f32);
    if (f32)
      PL(K_MODES, launch_zmodes<float>(m, g2, h->s));
    else
      PL(K_MODES, launch_zmodes<double>(m, g2, h->s));
    PlanePass b = plane_pass(h, true, f32);
    for (int c = 0; c < 3; ++c) {
      b.in[c] = compact_view(h, Tb, c, es);
      if (o.addth) {
        b.out[c] = node_view(h, h->ph[c]);
        b.th[c] = node_view(h, h->th[c]);
      } else {
        b.out[c] = node_view(h, o.out->c[c]);
      }
      if (o.dot) {
        b.r[c] = node_view(h, h->r.c[c]);
        b.zold[c] = node_view(h, h->zold.c[c]);
      }
    }
    b.addth = o.addth;
    b.dot = o.dot;
    if (f32)
      PL(K_TX32, (launch_plane<float, float, double>(b, g2, h->s)));
    else
      PL(K_TX, (launch_plane<double, double, double>(b, g2, h->s)));
    return 0;
  }
  {
    TxPass p = tx_pass(h, 0, false, f32);
    for (int c = 0; c < 3; ++c) {
      p.in[c] = node_view(h, in.c[c]);
      p.out[c] = compact_view(h, Ta, c, es);
    }
    p.conv = conv;
    if (f32)
      PL(K_TX32, (launch_tx<double, float, float>(p, gate, h->s)));
    else
      PL(K_TX, (launch_tx<double, double, double>(p, gate, h->s)));
  }
  const int *g2 = conv.mode != CONV_NONE ? conv.out : gate;
  {
    TxPass p = tx_pass(h, 1, false, f32);
    for (int c = 0; c < 3; ++c) {
      p.in[c] = compact_view(h, Ta, c, es);
      p.out[c] = compact_view(h, Tb, c, es);
    }
    if (f32)
      PL(K_TX32, (launch_tx<float, float, float>(p, g2, h->s)));
    else
      PL(K_TX, (launch_tx<double, double, double>(p, g2, h->s)));
  }
  {
    Tx2Modes m{};
    m.in = Tb;
    m.out = Ta;
    m.Az[0] = f32 ? (const void *)h->Af[2][0][0] : (const void *)h->Ad[2][0][0];
    m.Az[1] = f32 ? (const void *)h->Af[2][1][0] : (const void *)h->Ad[2][1][0];
    m.sx = f32 ? (const void *)h->sf[0] : (const void *)h->sd[0];
    m.sy = f32 ? (const void *)h->sf[1] : (const void *)h->sd[1];
    m.sz = f32 ? (const void *)h->sf[2] : (const void *)h->sd[2];
    for (int c = 0; c < 3; ++c) m.off[c] = h->rg.off[c];
    m.nx = h->g.Nx - 1;
    m.ny = h->g.Ny - 1;
    m.nz = h->g.Nz - 1;
    m.a = h->d.a;
    if (f32)
      PL(K_MODES, launch_tx2_modes<float>(m, g2, h->s));
    else
      PL(K_MODES, launch_tx2_modes<double>(m, g2, h->s));
  }
  for (int ax = 0; ax < 2; ++ax) {
    TxPass p = tx_pass(h, ax, true, f32);
    void *src = ax == 0 ? Ta : Tb, *dst = ax == 0 ? Tb : Ta;
    for (int c = 0; c < 3; ++c) {
      p.in[c] = compact_view(h, src, c, es);
      p.out[c] = compact_view(h, dst, c, es);
    }
    if (f32)
      PL(K_TX32, (launch_tx<float, float, float>(p, g2, h->s)));
    else
      PL(K_TX, (launch_tx<double, double, double>(p, g2, h->s)));
  }
  {
    TxPass p = tx_pass(h, 2, true, f32);
    for (int c = 0; c < 3; ++c) {
      p.in[c] = compact_view(h, Ta, c, es);
      if (o.addth) {
        p.out[c] = node_view(h, h->ph[c]);
        p.th[c] = node_view(h, h->th[c]);
      } else {
        p.out[c] = node_view(h, o.out->c[c]);
      }
      if (o.dot) {
        p.r[c] = node_view(h, h->r.c[c]);
        p.zold[c] = node_view(h, h->zold.c[c]);
      }
    }
    p.addth = o.addth;
    p.dot = o.dot;
    if (f32)
      PL(K_TX32, (launch_tx<float, float, double>(p, g2, h->s)));
    else
      PL(K_TX, (launch_tx<double, double, double>(p, g2, h->s)));
  }
  return 0;
}

// the partials of r·z written by the last launch of a preconditioner solve
int sep_last_parts(tdgl_handle *h, bool f32) {
  if (f32 && h->fold) return fold_blocks(h->q);
  if (!h->tx_only && (f32 ? h->plane32 : h->plane64)) return f32 ? plane_parts<float>(plane_pass(h, true, true)) : plane_parts<double>(plane_pass(h, true, false));
  return tx_parts(tx_pass(h, 2, true, false));
}

// ---------------------------------------------------------------------------- PCG
PcgArgs pcg_args(tdgl_handle *h) {
  const tdgl_desc &d = h->d;
  PcgArgs a{};
  a.cls = h->cur_cls;
  a.g = h->g;
  a.rg = h->rg;
  a.d = h->pd;
  a.kc = h->box_kc;
  a.tile = h->paptile;
  a.zcopy = 1;
  a.xcd = h->boxxcd;
  a.f = h->f32v;
  a.Dcode = h->Dcode;
  a.Dpal = h->Dpal;
  a.Dpal32 = h->Dpal32;
  for (int c = 0; c < 3; ++c) {
    a.D[c] = h->D[c];
    a.th[c] = h->th[c];
    a.ph[c] = h->ph[c];
  }
  a.sxy = d.sxy;
  a.sxz = d.sxz;
  a.syx = d.syx;
  a.syz = d.syz;
  a.szx = d.szx;
  a.szy = d.szy;
  a.dydz = d.dy * d.dz;
  a.dzdx = d.dz * d.dx;
  a.dxdy = d.dx * d.dy;
  a.x = h->x;
  a.r = h->r;
  a.b = h->B;
  a.z = h->z;
  a.Q = h->Qv;
  a.zold = h->zold;
  a.xprev = h->xprev;
  return a;
}

// Iteration q of the loop of _implicit_curl_fast! / implicit_curl! (implicit.jl:416-437,
// 309-325): the direction and A P, the update of x and r, the preconditioner (whose first pass
// decides on convergence, whose last pass gives r·z).  Gated by ACT[q % 2].
int pcg_iteration(tdgl_handle *h, int q, bool flexible) {
  PcgArgs a = pcg_args(h);
  a.q = q;
  a.flexible = flexible;
  a.Pnew = (q % 2 == 0) ? h->P : h->P2;
  a.Pold = (q % 2 == 0) ? h->P2 : h->P;
  a.nparts_in = sep_last_parts(h, flexible);
  if (flexible && h->fold) {
    // z and zold alternate between the two buffers instead of a copy: iteration q's z is the
    // preconditioner's output of iteration q - 1 (of the start for q = 0, in h->z), and iteration
    // q's preconditioner writes the other buffer, its r·zold read from this one
    const bool swap = h->zswap;
    Vec3 *zq = (swap && (q & 1)) ? &h->zold : &h->z, *zn = (swap && !(q & 1)) ? &h->zold : &h->z;
    a.z = *zq;
    a.zcopy = swap ? 0 : 1;
    PL(K_APPLYA, launch_pcg_pap(a, h->s));
    a.nparts_in = pcg_box_blocks(h->g, h->box_kc);
    // the update of x and r and the fold of r in one pass; its partials of r·r feed the forward
    // x-y launch's convergence check
    PL(K_VEC, launch_pcg_xr_fold(a, fold_pass(h, h->r), h->s));
    Conv cv{CONV_RR, h->pd.fl + FL_ACT0 + (q & 1), h->pd.fl + FL_ACT0 + ((q + 1) & 1), fold_blocks(h->q)};
    SepOut o{zn, false, 2, swap ? zq : nullptr};
    return sep_folded_tail(h, o, nullptr, cv);
  }
  PL(K_APPLYA, launch_pcg_pap(a, h->s));
  a.nparts_in = pcg_box_blocks(h->g, h->box_kc);
  SepOut o{&h->z, false, flexible ? 2 : 1};
  PL(K_VEC, launch_pcg_xr(a, h->s));
  Conv cv{CONV_RR, h->pd.fl + FL_ACT0 + (q & 1), h->pd.fl + FL_ACT0 + ((q + 1) & 1), pcg_box_blocks(h->g, h->box_kc)};
  CHK(sep_solve(h, h->r, o, nullptr, flexible, cv));
  return 0;
}

// Iteration q of a Float32 inner solve of the refinement: the iteration of pcg_iteration's folded
// flexible path on the Float32 vectors (x the round's increment d), its convergence against
// SC_TOLI x SC_NBI (the round's starting residual).  Gated by ACT[q % 2].
int pcg32_iteration(tdgl_handle *h, int q) {
  PcgArgs a = pcg_args(h);
  a.q = q;
  a.flexible = 1;
  a.f.Pnew = (q % 2 == 0) ? h->f32v.Pnew : h->f32v.Pold;
  a.f.Pold = (q % 2 == 0) ? h->f32v.Pold : h->f32v.Pnew;
  Vec3f *zq = (q & 1) ? &h->f32v.zold : &h->f32v.z, *zn = (q & 1) ? &h->f32v.z : &h->f32v.zold;
  a.f.z = *zq;
  a.nparts_in = sep_last_parts(h, true);
  PL(K_APPLYA, launch_pcg32_pap(a, h->s));
  a.nparts_in = pcg_box_blocks(h->g, h->box_kc);
  PL(K_VEC, launch_pcg32_xr_fold(a, fold_pass(h, h->f32v.r), h->s));
  Conv cv{CONV_RR, h->pd.fl + FL_ACT0 + (q & 1), h->pd.fl + FL_ACT0 + ((q + 1) & 1), fold_blocks(h->q), 1};
  SepOut o{nullptr, false, 2, nullptr, zn, &h->f32v.r, zq};
  return sep_folded_tail(h, o, nullptr, cv);
}

// One round of the refinement: r32 = b − A x (FP64), the round's check and the fold of r32, the
// Float32 solve A d = r32 from d = 0 (its first preconditioner, then `budget` gated iterations, or
// eagerly while active), x += d.  Round 0 is gated by NZ, the others by OACT.
int refine_round(tdgl_handle *h, int round) {
  const tdgl_desc &d = h->d;
  PcgArgs a = pcg_args(h);
  a.gate_fl = round == 0 ? FL_NZ : FL_OACT;
  PL(K_APPLYA, launch_pcg_resid32(a, h->s));
  RefFold rf{};
  rf.f = fold_pass(h, h->f32v.r);
  rf.gate = h->pd.fl + a.gate_fl;
  rf.nparts = pcg_box_blocks(h->g, h->box_kc);
  rf.fold = 1;
  rf.d = h->pd;
  PL(K_FOLD, launch_ref_fold(rf, h->s));
  const int *oact = h->pd.fl + FL_OACT;
  {
    SepOut o{nullptr, false, 1, nullptr, &h->f32v.z, &h->f32v.r, nullptr};
    CHK(sep_folded_tail(h, o, oact, Conv{}));
  }
  int q = 0;
  if (h->capturing) {
    for (; q < h->budget; ++q) CHK(pcg32_iteration(h, q));
  } else {
    int batch = 4;
    for (;;) {
      HIPCHK(hipMemcpyAsync(h->flh, h->pd.fl, FL_COUNT * sizeof(int), hipMemcpyDeviceToHost, h->s));
      HIPCHK(hipStreamSynchronize(h->s));
      if (!h->flh[FL_ACT0 + (q & 1)] || h->flh[FL_IT] >= d.pcg_maxit) break;
      for (int t = 0; t < batch; ++t, ++q) CHK(pcg32_iteration(h, q));
      batch = 2;
    }
  }
  a.gate_fl = FL_OACT;
  PL(K_VEC, launch_pcg_xupdate(a, h->s));
  return 0;
}

// The refinement (d.pcg_refine): the FP64 operator, residual and solution around Float32 inner
// PCG solves with the folded Float32 preconditioner.  Each round solves A d = b − A x to SC_TOLI of
// its starting residual; the rounds stop when ||b − A x|| <= tol ||b|| (FP64) or at the iteration
// limit.  Eager: rounds while the check says so.  Captured: `rounds` rounds, each with `budget`
// gated inner iterations, then a last FP64 check; a solve still above the tolerance overflows
// (the cycle is rolled back and stepped eagerly).
int implicit_pcg_refine(tdgl_handle *h) {
  const tdgl_desc &d = h->d;
  const bool extrap = d.pcg_extrap && d.pcg_fast;
  const int *nz = h->pd.fl + FL_NZ;
  PcgArgs a = pcg_args(h);
  a.warm = h->warm;
  a.extrap = extrap;
  a.have2 = h->have2;
  a.nparts_in = field_blocks(h->g);
  PL(K_VEC, launch_pcg_start(a, h->s));
  if (!h->warm) {  // cold start: x = M⁻¹ b in Float64
    SepOut o{&h->x, false, 0};
    CHK(sep_solve(h, h->B, o, nz, false, Conv{CONV_NONE, nullptr, nullptr, 0}));
  }
  h->have2 = h->warm && extrap;
  h->warm = true;
  if (h->capturing) {
    for (int r = 0; r < h->rounds; ++r) CHK(refine_round(h, r));
  } else {
    for (int r = 0;; ++r) {
      if (r > 0) {
        HIPCHK(hipMemcpyAsync(h->flh, h->pd.fl, FL_COUNT * sizeof(int), hipMemcpyDeviceToHost, h->s));
        HIPCHK(hipStreamSynchronize(h->s));
        if (!h->flh[FL_OACT] || h->flh[FL_IT] >= d.pcg_maxit) break;
      }
      CHK(refine_round(h, r));
    }
  }
  // the last check: ||b − A x|| of the final x (OACT set when still above the tolerance)
  a = pcg_args(h);
  a.gate_fl = h->capturing ? FL_OACT : FL_NZ;
  a.norm_only = 1;
  PL(K_APPLYA, launch_pcg_resid32(a, h->s));
  RefFold rf{};
  rf.f = fold_pass(h, h->f32v.r);
  rf.gate = h->pd.fl + a.gate_fl;
  rf.nparts = pcg_box_blocks(h->g, h->box_kc);
  rf.fold = 0;
  rf.d = h->pd;
  PL(K_FOLD, launch_ref_fold(rf, h->s));
  if (!h->capturing) {
    HIPCHK(hipMemcpyAsync(h->flh, h->pd.fl, FL_COUNT * sizeof(int), hipMemcpyDeviceToHost, h->s));
    HIPCHK(hipStreamSynchronize(h->s));
    h->last_iters = h->flh[FL_IT];
  }
  a.ovf_fl = h->capturing ? FL_OACT : 0;
  a.act_slot = 0;
  if (!h->capturing) a.ovf_fl = FL_ACT1;  // eager: the solve ran to its end (ACT1 is 0 after a round's check)
  PL(K_SCATTER, launch_pcg_finish(a, h->s));
  return 0;
}

// implicit_curl!(s, ::PCGSolver) / _implicit_curl_fast! (implicit.jl:279-446), after the field
// kernel has written b = φₑ − θ and the partials of b·b, and scaled φ by D.  Eager: the host
// reads the flags between batches of iterations.  Captured: `budget` iterations, each gated.
int implicit_pcg(tdgl_handle *h) {
  if (h->refine) return implicit_pcg_refine(h);
  const tdgl_desc &d = h->d;
  const bool fast = d.pcg_fast || d.pcg_f32;
  const bool f32 = d.pcg_f32 != 0;
  const bool extrap = d.pcg_extrap && d.pcg_fast;
  const bool flexible = fast && f32;
  const int *nz = h->pd.fl + FL_NZ;
  PcgArgs a = pcg_args(h);
  a.warm = h->warm;
  a.extrap = fast && extrap;
  a.have2 = h->have2;
  a.nparts_in = field_blocks(h->g);
  PL(K_VEC, launch_pcg_start(a, h->s));
  if (!h->warm) {  // cold start: x = M⁻¹ b in Float64
    SepOut o{&h->x, false, 0};
    CHK(sep_solve(h, h->B, o, nz, false, Conv{CONV_NONE, nullptr, nullptr, 0}));
  }
  h->have2 = fast && h->warm && extrap;
  h->warm = true;
  PL(K_APPLYA, launch_pcg_resid(a, h->s));
  {
    Conv cv{CONV_RR0, nz, h->pd.fl + FL_ACT0, pcg_box_blocks(h->g, h->box_kc)};
    SepOut o{&h->z, false, 1};
    CHK(sep_solve(h, h->r, o, nullptr, flexible, cv));
  }
  int q = 0;
  if (h->capturing) {
    const int budget = h->stepwise ? h->cbudget[h->cur_cls] : h->budget;
    for (; q < budget; ++q) CHK(pcg_iteration(h, q, flexible));
  } else {
    int batch = h->last_iters > 0 ? h->last_iters : 4;
    for (;;) {
      HIPCHK(hipMemcpyAsync(h->flh, h->pd.fl, FL_COUNT * sizeof(int), hipMemcpyDeviceToHost, h->s));
      HIPCHK(hipStreamSynchronize(h->s));
      if (!h->flh[FL_ACT0 + (q & 1)] || q >= d.pcg_maxit) break;
      const int nbt = std::min(batch, d.pcg_maxit - q);
      for (int t = 0; t < nbt; ++t, ++q) CHK(pcg_iteration(h, q, flexible));
      batch = 2;
    }
    h->last_iters = h->flh[FL_IT];
  }
  a.act_slot = q & 1;
  PL(K_SCATTER, launch_pcg_finish(a, h->s));
  return 0;
}

// ---------------------------------------------------------------------------- open boundary
// the smallest 2^a 3^b 5^c >= n
int smooth_size(int n) {
  for (int m = std::max(1, n);; ++m) {
    int r = m;
    for (int f : {2, 3, 5})
      while (r % f == 0) r /= f;
    if (r == 1) return m;
  }
}
// the radices of M in the order the line FFT applies them (4s, then 2, 3s, 5s)
void fft_radices(int M, FftAxis &ax) {
  ax.M = M;
  ax.nrad = 0;
  int m = M;
  for (int r : {4, 2, 3, 5})
    while (m % r == 0) {
      ax.rad[ax.nrad++] = r;
      m /= r;
    }
}

// the five passes of one lattice's convolution: x forward from δJ, y forward, z forward × Ĝ and
// back, y back, x back into θ_δ (the padding is never stored)
int open_convolve(tdgl_handle *h, int c, const int *gate) {
  const int n0 = h->on[c][0], n1 = h->on[c][1], n2 = h->on[c][2];
  const int M0 = h->oM[c][0], M1 = h->oM[c][1], M2 = h->oM[c][2];
  const long long P = (long long)M0 * M1;
  FftPass p{};
  // x forward: lines (j < n1, k < n2) of δJ
  p.in = h->odj[c];
  p.in_real = 1;
  p.in_s = 1;
  p.in_l0 = n0;
  p.in_l1 = (long long)n0 * n1;
  p.out = h->obuf[c];
  p.scale = 1.0;
  p.out_s = 1;
  p.out_l0 = M0;
  p.out_l1 = P;
  p.nl0 = n1;
  p.nl1 = n2;
  p.nin = n0;
  p.nout = M0;
  p.dir = -1;
  p.ax = h->oax[c][0];
  PL(K_CURL, launch_fft(p, gate, h->s));
  // y forward: lines (i < M0, k < n2)
  FftPass q{};
  q.in = q.out = h->obuf[c];
  q.scale = 1.0;
  q.in_s = q.out_s = M0;
  q.in_l0 = q.out_l0 = 1;
  q.in_l1 = q.out_l1 = P;
  q.nl0 = M0;
  q.nl1 = n2;
  q.nin = n1;
  q.nout = M1;
  q.dir = -1;
  q.lines_fast = 1;
  q.ax = h->oax[c][1];
  PL(K_CURL, launch_fft(q, gate, h->s));
  // z: forward, × Ĝ, back; lines (i < M0, j < M1), only the box's planes kept
  FftPass z = q;
  z.in_s = z.out_s = P;
  z.in_l0 = z.out_l0 = 1;
  z.in_l1 = z.out_l1 = M0;
  z.nl0 = M0;
  z.nl1 = M1;
  z.nin = n2;
  z.nout = n2;
  z.ghat = h->oghat[c];
  z.g_s = P;
  z.g_l0 = 1;
  z.g_l1 = M0;
  z.ax = h->oax[c][2];
  PL(K_CURL, launch_fft(z, gate, h->s));
  // y back
  FftPass yb = q;
  yb.nin = M1;
  yb.nout = n1;
  yb.dir = 1;
  PL(K_CURL, launch_fft(yb, gate, h->s));
  // x back into θ_δ, normalized
  FftPass xb{};
  xb.in = h->obuf[c];
  xb.in_s = 1;
  xb.in_l0 = M0;
  xb.in_l1 = P;
  xb.out = h->oth[c];
  xb.out_real = 1;
  xb.scale = 1.0 / ((double)M0 * M1 * M2);
  xb.out_s = 1;
  xb.out_l0 = n0;
  xb.out_l1 = (long long)n0 * n1;
  xb.nl0 = n1;
  xb.nl1 = n2;
  xb.nin = M0;
  xb.nout = n0;
  xb.dir = 1;
  xb.ax = h->oax[c][0];
  PL(K_CURL, launch_fft(xb, gate, h->s));
  return 0;
}

// open_update! (openbc.jl:196-215) on the device: the clock, and on its steps h_δ anew
int open_update(tdgl_handle *h) {
  const tdgl_desc &d = h->d;
  const int *re = h->ofl + OF_RECOMP, *first = h->ofl + OF_FIRST;
  PL(K_CURL, launch_open_clock(h->ofl, d.open_every, h->s));
  // a step the host knows not to recompute (the class of a step after the recomputation, its
  // copy of the clock read at this call) leaves out the launches the clock would gate off
  if (h->stepwise && h->clock_known && h->cur_cls != 0) return 0;
  CopyList keep{};
  for (int c = 0; c < 3; ++c) {
    keep.src[c] = h->ohnew[c];
    keep.dst[c] = h->ohold[c];
    keep.n[c] = h->g.n;
  }
  keep.cnt = 3;
  keep.gate = re;
  PL(K_CURL, launch_copies(keep, h->s));
  OpenDJ dj{};
  for (int c = 0; c < 3; ++c) {
    dj.h[c] = h->h[c];
    dj.dj[c] = h->odj[c];
    for (int a = 0; a < 3; ++a) dj.n[c][a] = h->on[c][a];
  }
  dj.Jbar = h->oJbar;
  dj.HJp = h->hp + 1;
  dj.ofl = h->ofl;
  dj.k2 = d.open_kappa2;
  dj.dx = d.dx;
  dj.dy = d.dy;
  dj.dz = d.dz;
  PL(K_CURL, launch_open_dj(h->g, dj, h->s));
  for (int c = 0; c < 3; ++c) CHK(open_convolve(h, c, re));
  OpenFace f{};
  for (int c = 0; c < 3; ++c) {
    f.theta[c] = h->oth[c];
    f.hnew[c] = h->ohnew[c];
    for (int a = 0; a < 3; ++a) f.n[c][a] = h->on[c][a];
  }
  f.ofl = h->ofl;
  f.dx = d.dx;
  f.dy = d.dy;
  f.dz = d.dz;
  PL(K_CURL, launch_open_face(h->g, f, h->s));
  keep.gate = first;  // the first time there is no ramp
  PL(K_CURL, launch_copies(keep, h->s));
  return 0;
}

// Ĝ of each lattice, once: fac·G on the padded box, its three forward transforms, the real part
int open_green(tdgl_handle *h) {
  const tdgl_desc &d = h->d;
  for (int c = 0; c < 3; ++c) {
    const int M0 = h->oM[c][0], M1 = h->oM[c][1], M2 = h->oM[c][2];
    const size_t n = (size_t)M0 * M1 * M2;
    double *g = nullptr;
    cplx *G = nullptr;
    Block bg{}, bG{};
    HIPCHK(dblock(h, n * sizeof(double), &bg));
    if (dblock(h, n * sizeof(cplx), &bG) != hipSuccess) {
      cache_give(bg);
      return fail(TDGL_E_NOMEM, "hipMalloc of the padded Green function failed");
    }
    g = (double *)bg.p;
    G = (cplx *)bG.p;
    OpenGreen og{};
    og.out = g;
    og.M[0] = M0;
    og.M[1] = M1;
    og.M[2] = M2;
    og.lgf = h->olgf;
    og.R = d.open_R;
    og.fac = d.open_fac;
    int rc = 0;
    auto run = [&]() -> int {
      LCHK(launch_open_green(og, h->s));
      const long long P = (long long)M0 * M1;
      FftPass p{};
      p.in = g;
      p.in_real = 1;
      p.in_s = 1;
      p.in_l0 = M0;
      p.in_l1 = P;
      p.out = G;
      p.scale = 1.0;
      p.out_s = 1;
      p.out_l0 = M0;
      p.out_l1 = P;
      p.nl0 = M1;
      p.nl1 = M2;
      p.nin = p.nout = M0;
      p.dir = -1;
      p.ax = h->oax[c][0];
      LCHK(launch_fft(p, nullptr, h->s));
      FftPass q{};
      q.in = q.out = G;
      q.scale = 1.0;
      q.in_s = q.out_s = M0;
      q.in_l0 = q.out_l0 = 1;
      q.in_l1 = q.out_l1 = P;
      q.nl0 = M0;
      q.nl1 = M2;
      q.nin = q.nout = M1;
      q.dir = -1;
      q.lines_fast = 1;
      q.ax = h->oax[c][1];
      LCHK(launch_fft(q, nullptr, h->s));
      FftPass z = q;
      z.out = h->oghat[c];
      z.out_real = 1;
      z.in_s = z.out_s = P;
      z.in_l0 = z.out_l0 = 1;
      z.in_l1 = z.out_l1 = M0;
      z.nl0 = M0;
      z.nl1 = M1;
      z.nin = z.nout = M2;
      z.ax = h->oax[c][2];
      LCHK(launch_fft(z, nullptr, h->s));
      HIPCHK(hipStreamSynchronize(h->s));
      return 0;
    };
    rc = run();
    // run() waits for its stream on success; on failure, wait before the blocks go back
    if (rc) (void)hipStreamSynchronize(h->s);
    cache_give(bg);
    cache_give(bG);
    if (rc) return rc;
  }
  return 0;
}

// ---------------------------------------------------------------------------- normal fluids
// qp_potential! (exchange.jl:344-389, 477-516) from the state at the start of the step: the
// sources, then for every mode the x-y plane launch (W_x, W_y), the z launch with the division
// and the inverse z, and the inverse x-y plane launch (V_x, V_y) into the modes the stages read
int qp_potential(tdgl_handle *h) {
  const tdgl_desc &d = h->d;
  const int K = d.qp_nmodes, n1 = d.Nx - 1, nj = d.qp_nj, nk = d.qp_nk;
  QpConv c{};
  for (int mu = 0; mu < d.nbands; ++mu) c.G[mu] = h->G[mu];
  for (int a = 0; a < 3; ++a) {
    c.V[a] = h->V[a];
    c.tau[a] = h->tau[a];
    for (int mu = 0; mu < TDGL_MAX_BANDS; ++mu) c.wa[a][mu] = d.wa[a][mu];
  }
  for (int k = 0; k < K; ++k) c.S[k] = h->qS[k];
  for (int mu = 0; mu < TDGL_MAX_BANDS; ++mu) {
    c.s[mu] = d.qp_s[mu];
    for (int nu = 0; nu < TDGL_MAX_BANDS; ++nu) {
      c.gam[mu][nu] = d.qp_gamma[mu][nu];
      c.Pinv[mu][nu] = d.qp_Pinv[mu][nu];
    }
  }
  c.rx2 = 1 / (d.dx * d.dx);
  c.ry2 = 1 / (d.dy * d.dy);
  c.rz2 = 1 / (d.dz * d.dz);
  c.nb = d.nbands;
  c.mode = d.qp_mode;
  c.K = K;
  c.n1 = n1;
  c.nj = nj;
  c.nk = nk;
  c.j0 = d.qp_j0;
  c.k0 = d.qp_k0;
  PL(K_QP, launch_qp_conv(h->g, c, h->s));
  const size_t n = (size_t)h->qn;
  PlanePass f{};
  f.ncomp = K;
  for (int k = 0; k < K; ++k) {
    f.e[k][0] = n1;
    f.e[k][1] = nj;
    f.e[k][2] = nk;
    f.in[k] = View{(void *)h->qS[k], n1, n1 * nj};
    f.out[k] = View{(void *)(h->qT1 + k * n), n1, n1 * nj};
    f.X[k] = h->qW[0];
    f.Y[k] = h->qW[1];
  }
  PL(K_QP, (launch_plane<double, double, double>(f, nullptr, h->s)));
  QpZ z{};
  z.in = h->qT1;
  z.out = h->qT2;
  z.Wz = h->qW[2];
  z.Vz = h->qV[2];
  z.ex = h->qe[0];
  z.ey = h->qe[1];
  z.ez = h->qe[2];
  z.sigma = d.qp_sigma;
  for (int k = 0; k < K; ++k) z.lam[k] = d.qp_lambda[k];
  z.K = K;
  z.n1 = n1;
  z.nj = nj;
  z.nk = nk;
  PL(K_QP, launch_qp_zcol(z, h->s));
  PlanePass b = f;
  for (int k = 0; k < K; ++k) {
    b.in[k] = View{(void *)(h->qT2 + k * n), n1, n1 * nj};
    b.out[k] = View{(void *)h->qS[k], n1, n1 * nj};
    b.X[k] = h->qV[0];
    b.Y[k] = h->qV[1];
  }
  PL(K_QP, (launch_plane<double, double, double>(b, nullptr, h->s)));
  return 0;
}

// ---------------------------------------------------------------------------- one step
int step_once(tdgl_handle *h) {
  const tdgl_desc &d = h->d;
  if (d.selffield_bc == TDGL_BC_OPEN) CHK(open_update(h));
  if (d.qp_mode != TDGL_QP_NONE) CHK(qp_potential(h));
  CHK(update_bands(h));
  FieldArgs fa{};
  fa.xcd = h->stxcd;
  for (int mu = 0; mu < d.nbands; ++mu) fa.G[mu] = h->G[mu];
  for (int c = 0; c < 3; ++c) {
    fa.V[c] = h->V[c];
    fa.th[c] = h->th[c];
    fa.ph[c] = h->ph[c];
    fa.J[c] = h->J[c];
    fa.h[c] = h->h[c];
    fa.tau[c] = h->tau[c];
    fa.D[c] = h->D[c];
    fa.B[c] = h->B.c[c];
  }
  fa.rg = h->rg;
  for (int ax = 0; ax < 3; ++ax)
    for (int mu = 0; mu < TDGL_MAX_BANDS; ++mu) fa.wa[ax][mu] = d.wa[ax][mu];
  fa.d1 = d.d1;
  fa.sxy = d.sxy;
  fa.sxz = d.sxz;
  fa.syx = d.syx;
  fa.syz = d.syz;
  fa.szx = d.szx;
  fa.szy = d.szy;
  fa.nb = d.nbands;
  fa.scale = d.field_mode == TDGL_FIELD_PCG || d.field_mode == TDGL_FIELD_EXPLICIT_SCALED;
  fa.gather = d.field_mode == TDGL_FIELD_PCG || d.field_mode == TDGL_FIELD_SEPARABLE;
  fa.part = d.field_mode == TDGL_FIELD_PCG ? h->pd.part + (size_t)PART_NB * 2 * MAX_PART : nullptr;
  if (h->noise && d.noise_field) {
    // field_kick! (noise.jl:184-186): c = √(T̃ dt/(σn dV)), the kick of a link d_l c √D_l n
    const double c = std::sqrt(d.noise_T * d.dt / (d.noise_sigma_n * d.dx * d.dy * d.dz));
    fa.nkick = 1;
    fa.kax = c * d.dx;
    fa.kay = c * d.dy;
    fa.kaz = c * d.dz;
    fa.nz = noise_args(h);
  }
  PL(K_FIELD, launch_field(h->g, fa, h->s));
  if (d.field_mode == TDGL_FIELD_SEPARABLE) {
    SepOut o{nullptr, true, 0};
    CHK(sep_solve(h, h->B, o, nullptr, false, Conv{CONV_NONE, nullptr, nullptr, 0}));
  } else if (d.field_mode == TDGL_FIELD_PCG) {
    CHK(implicit_pcg(h));
  }
  CurlArgs ca{};
  ca.xcd = h->stxcd;
  for (int c = 0; c < 3; ++c) {
    ca.ph[c] = h->ph[c];
    ca.U[c] = h->U[c];
    ca.h[c] = h->h[c];
  }
  ca.hp = h->hp;
  ca.bar = d.selffield_bc == TDGL_BC_BAR || d.selffield_bc == TDGL_BC_OPEN;
  if (d.selffield_bc == TDGL_BC_OPEN) {
    ca.open = 1;
    for (int c = 0; c < 3; ++c) {
      ca.hold[c] = h->ohold[c];
      ca.hnew[c] = h->ohnew[c];
    }
    ca.ofl = h->ofl;
    ca.every = d.open_every;
  }
  ca.bar_hy_z = h->bar[0];
  ca.bar_hy_y = h->bar[1];
  ca.bar_hz_z = h->bar[2];
  ca.bar_hz_y = h->bar[3];
  ca.bar_hz_x = h->bar[4];
  ca.bar_hy_x = h->bar[5];
  ca.dydz = d.dy * d.dz;
  ca.dzdx = d.dz * d.dx;
  ca.dxdy = d.dx * d.dy;
  PL(K_CURL, launch_curl(h->g, ca, h->s));
  GhostArgs ga{};
  for (int mu = 0; mu < d.nbands; ++mu) ga.F[mu] = h->F[mu];
  for (int c = 0; c < 3; ++c) ga.U[c] = h->U[c];
  ga.nb = d.nbands;
  PL(K_GHOSTS, launch_ghosts(h->g, ga, h->s));
  if (h->noise) PL(K_VEC, launch_noise_tick(h->nctr, h->s));
  // advance! (field.jl:248-263): F equals G outside the block, so G <- F is a swap
  for (int mu = 0; mu < d.nbands; ++mu) std::swap(h->G[mu], h->F[mu]);
  for (int c = 0; c < 3; ++c) {
    std::swap(h->th[c], h->ph[c]);
    std::swap(h->V[c], h->U[c]);
  }
  h->parity ^= 1;
  return 0;
}

// ---------------------------------------------------------------------------- a ladder cycle
// The arrays a step reads at its start (the rest is rewritten before it is read)
CopyList backup_list(tdgl_handle *h, bool save) {
  CopyList c{};
  const int n = h->g.n;
  auto add = [&](double *live, double *bk, int len) {
    c.src[c.cnt] = save ? live : bk;
    c.dst[c.cnt] = save ? bk : live;
    c.n[c.cnt] = len;
    ++c.cnt;
  };
  int q = 0;
  for (int mu = 0; mu < h->d.nbands; ++mu) add((double *)h->G[mu], h->bk[q++], 2 * n);
  for (int k = 0; k < 3; ++k) {
    add(h->th[k], h->bk[q++], n);
    add(h->ph[k], h->bk[q++], n);
    add((double *)h->U[k], h->bk[q++], 2 * n);
    add((double *)h->V[k], h->bk[q++], 2 * n);
    add(h->J[k], h->bk[q++], n);
    add(h->h[k], h->bk[q++], n);
  }
  if (h->d.field_mode == TDGL_FIELD_PCG)
    for (int k = 0; k < 3; ++k) {
      add(h->x.c[k], h->bk[q++], n);
      add(h->xprev.c[k], h->bk[q++], n);
    }
  if (h->d.field_mode == TDGL_FIELD_PCG && h->d.selffield_bc == TDGL_BC_OPEN) {
    for (int k = 0; k < 3; ++k) {
      add(h->ohold[k], h->bk[q++], n);
      add(h->ohnew[k], h->bk[q++], n);
    }
    add((double *)h->ofl, h->bk[q++], 2);  // the clock: four ints
  }
  if (h->d.field_mode == TDGL_FIELD_PCG && h->noise) add((double *)h->nctr, h->bk[q++], 2);  // the stream's clock
  return c;
}

HostSnap snap(tdgl_handle *h) {
  HostSnap s{};
  for (int mu = 0; mu < TDGL_MAX_BANDS; ++mu) {
    s.G[mu] = h->G[mu];
    s.F[mu] = h->F[mu];
  }
  for (int c = 0; c < 3; ++c) {
    s.th[c] = h->th[c];
    s.ph[c] = h->ph[c];
    s.U[c] = h->U[c];
    s.V[c] = h->V[c];
  }
  s.parity = h->parity;
  s.warm = h->warm;
  s.have2 = h->have2;
  return s;
}
void restore(tdgl_handle *h, const HostSnap &s) {
  for (int mu = 0; mu < TDGL_MAX_BANDS; ++mu) {
    h->G[mu] = s.G[mu];
    h->F[mu] = s.F[mu];
  }
  for (int c = 0; c < 3; ++c) {
    h->th[c] = s.th[c];
    h->ph[c] = s.ph[c];
    h->U[c] = s.U[c];
    h->V[c] = s.V[c];
  }
  h->parity = s.parity;
  h->warm = s.warm;
  h->have2 = s.have2;
}

// the class of the next step, the host's copy of the clock advanced by it as k_open_clock
// advances the device's (a recomputation when the count reaches `every`, or on the first step)
int step_class(tdgl_handle *h) {
  if (!h->stepwise) return 0;
  const int c = h->oclock + 1;
  const bool re = c >= h->d.open_every || h->oclock < 0;
  h->oclock = re ? 0 : c;
  return std::min(h->oclock, PCG_NCLS - 1);
}
int steps_eager(tdgl_handle *h, int n) {
  for (int q = 0; q < n; ++q) {
    if (h->stepwise && h->clock_known && !h->capturing) h->cur_cls = step_class(h);
    CHK(step_once(h));
  }
  return 0;
}

void drop_graph(CachedGraph &g) {
  if (g.exec) (void)hipGraphExecDestroy(g.exec);
  if (g.graph) (void)hipGraphDestroy(g.graph);
  g = CachedGraph{};
}

// every captured graph (they hold the addresses of the handle's buffers)
int destroy_graph(tdgl_handle *h) {
  for (CachedGraph &g : h->gc) drop_graph(g);
  return 0;
}

// n steps as one captured graph: captured once per key (n, pointer parity, warm, have2, budget)
// and replayed after; the last GRAPH_CACHE graphs are kept, so cycles that alternate between two
// budgets or parities replay instead of capturing again.  With the PCG, the graph holds `budget`
// gated iterations per solve; a cycle in which a solve was still active after them is rolled back
// from a device copy of its start and stepped eagerly, and the budget grows.
// The budget's way down: BUDGET_RUN cycles in a row whose longest solve stayed BUDGET_SLACK or more
// iterations under it, and none in the BUDGET_HOLD cycles after a roll back.
constexpr int BUDGET_SLACK = 6, BUDGET_RUN = 2, BUDGET_HOLD = 4;
int cycle_graph(tdgl_handle *h, int n) {
  const bool pcg = h->d.field_mode == TDGL_FIELD_PCG;
  const GraphKey key{n, h->parity, h->warm ? 1 : 0, h->have2 ? 1 : 0, pcg ? h->budget : 0, 0};
  const HostSnap start = snap(h);
  CachedGraph *g = nullptr;
  CachedGraph *const gc0 = h->gc, *const gc1 = h->gc + GRAPH_CACHE_CALL;
  for (CachedGraph *c = gc0; c < gc1; ++c)
    if (c->live && c->key == key) g = c;
  if (!g) {
    g = gc0;
    for (CachedGraph *c = gc0; c < gc1; ++c)
      if (!c->live || (g->live && c->used < g->used)) g = c;
    drop_graph(*g);
    h->capturing = true;
    // relaxed: other handles, in other host threads, allocate, free and copy on their own streams
    // meanwhile; this capture records only the kernels launched on h->s
    HIPCHK(hipStreamBeginCapture(h->s, hipStreamCaptureModeRelaxed));
    int rc = steps_eager(h, n);
    hipGraph_t gph = nullptr;
    hipError_t e = hipStreamEndCapture(h->s, &gph);
    h->capturing = false;
    if (rc || e != hipSuccess) {
      // nothing ran on the device: the host-side pointers and flags go back to the start
      if (gph) (void)hipGraphDestroy(gph);
      restore(h, start);
      return rc ? rc : fail(TDGL_E_DEVICE, "hipStreamEndCapture: %s", hipGetErrorString(e));
    }
    g->graph = gph;
    e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
      drop_graph(*g);
      restore(h, start);
      return fail(TDGL_E_DEVICE, "hipGraphInstantiate: %s", hipGetErrorString(e));
    }
    g->key = key;
    g->end = snap(h);
    g->live = true;
    h->captures += 1;
    restore(h, start);
  }
  g->used = ++h->gtick;
  if (pcg) {
    PL(K_VEC, launch_copies(backup_list(h, true), h->s));
    PL(K_VEC, launch_pcg_cycle_reset(h->pd, h->s));
  }
  HIPCHK(hipGraphLaunch(g->exec, h->s));
  restore(h, g->end);
  if (pcg) {
    HIPCHK(hipMemcpyAsync(h->flh, h->pd.fl, FL_COUNT * sizeof(int), hipMemcpyDeviceToHost, h->s));
    HIPCHK(hipStreamSynchronize(h->s));
    if (h->flh[FL_OVF]) {
      // the roll back: the start state into the start pointers, then the eager steps
      restore(h, start);
      PL(K_VEC, launch_copies(backup_list(h, false), h->s));
      h->budget = std::min(h->d.pcg_maxit, std::max(2 * h->budget, h->flh[FL_ITMAX] + 4));
      h->rollbacks += 1;
      PL(K_VEC, launch_pcg_cycle_reset(h->pd, h->s));
      CHK(steps_eager(h, n));
      HIPCHK(hipMemcpyAsync(h->flh, h->pd.fl, FL_COUNT * sizeof(int), hipMemcpyDeviceToHost, h->s));
      h->shrink_run = 0;
      h->shrink_max = 0;
      h->shrink_hold = BUDGET_HOLD;
    } else if (h->shrink_hold > 0) {
      h->shrink_hold -= 1;
    } else if (!h->budget_fixed) {
      // the budget follows the solves down: after BUDGET_RUN cycles (in a row) whose longest solve
      // stayed BUDGET_SLACK or more iterations under it, it becomes that longest + 3 (the graph is
      // captured again at the next call).  The first budget comes from the first cycle, whose
      // solves start cold and run longest; every gated iteration past a converged solve still
      // launches its five kernels.
      const int itmax = h->flh[FL_ITMAX];
      if (itmax + BUDGET_SLACK <= h->budget) {
        h->shrink_max = std::max(h->shrink_max, itmax);
        if (++h->shrink_run >= BUDGET_RUN) {
          h->budget = std::max(4, h->shrink_max + 3);
          h->shrink_run = 0;
          h->shrink_max = 0;
        }
      } else {
        h->shrink_run = 0;
        h->shrink_max = 0;
      }
    }
  }
  return 0;
}

int d_maxit(tdgl_handle *h) { return h->d.pcg_maxit; }

// A call of n steps of a stepwise handle: step s replays the one-step graph of its class c_s
// (captured once per key (1, parity, warm, have2, budget of c_s, c_s)); the whole call rolls back
// as cycle_graph's when a solve outran its budget, and the budgets of the classes whose longest
// solve reached theirs grow.  Each class's budget comes down as the single budget does.
int stepwise_graphs(tdgl_handle *h, int n) {
  const HostSnap start = snap(h);
  const int clock0 = h->oclock;
  PL(K_VEC, launch_copies(backup_list(h, true), h->s));
  PL(K_VEC, launch_pcg_cycle_reset(h->pd, h->s));
  for (int q = 0; q < n; ++q) {
    const int c = step_class(h);
    h->cur_cls = c;
    const GraphKey key{1, h->parity, h->warm ? 1 : 0, h->have2 ? 1 : 0, h->cbudget[c], c};
    const HostSnap before = snap(h);
    CachedGraph *g = nullptr;
    for (CachedGraph &cg : h->gc)
      if (cg.live && cg.key == key) g = &cg;
    if (!g) {
      g = &h->gc[0];
      for (CachedGraph &cg : h->gc)
        if (!cg.live || (g->live && cg.used < g->used)) g = &cg;
      drop_graph(*g);
      h->capturing = true;
      HIPCHK(hipStreamBeginCapture(h->s, hipStreamCaptureModeRelaxed));
      int rc = step_once(h);
      hipGraph_t gph = nullptr;
      hipError_t e = hipStreamEndCapture(h->s, &gph);
      h->capturing = false;
      if (rc || e != hipSuccess) {
        if (gph) (void)hipGraphDestroy(gph);
        restore(h, before);
        return rc ? rc : fail(TDGL_E_DEVICE, "hipStreamEndCapture: %s", hipGetErrorString(e));
      }
      g->graph = gph;
      e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
      if (e != hipSuccess) {
        drop_graph(*g);
        restore(h, before);
        return fail(TDGL_E_DEVICE, "hipGraphInstantiate: %s", hipGetErrorString(e));
      }
      g->key = key;
      g->end = snap(h);
      g->live = true;
      h->captures += 1;
      restore(h, before);
    }
    g->used = ++h->gtick;
    HIPCHK(hipGraphLaunch(g->exec, h->s));
    restore(h, g->end);
  }
  HIPCHK(hipMemcpyAsync(h->flh, h->pd.fl, FL_COUNT * sizeof(int), hipMemcpyDeviceToHost, h->s));
  HIPCHK(hipStreamSynchronize(h->s));
  const int *itc = h->flh + FL_ITMAXC;
  if (h->flh[FL_OVF]) {
    restore(h, start);
    h->oclock = clock0;
    PL(K_VEC, launch_copies(backup_list(h, false), h->s));
    for (int c = 0; c < PCG_NCLS; ++c) {
      if (itc[c] >= h->cbudget[c]) h->cbudget[c] = std::min(h->d.pcg_maxit, std::max(2 * h->cbudget[c], itc[c] + 4));
      h->cshrink_run[c] = 0;
      h->cshrink_max[c] = 0;
      h->cshrink_hold[c] = BUDGET_HOLD;
    }
    h->rollbacks += 1;
    PL(K_VEC, launch_pcg_cycle_reset(h->pd, h->s));
    CHK(steps_eager(h, n));
    HIPCHK(hipMemcpyAsync(h->flh, h->pd.fl, FL_COUNT * sizeof(int), hipMemcpyDeviceToHost, h->s));
    return 0;
  }
  for (int c = 0; c < PCG_NCLS; ++c) h->cseen[c] = h->cseen[c] || itc[c] > 0;
  if (h->budget_fixed) return 0;
  for (int c = 0; c < PCG_NCLS; ++c) {
    if (h->cshrink_hold[c] > 0) {
      h->cshrink_hold[c] -= 1;
      continue;
    }
    if (itc[c] == 0) continue;  // no step of this class in the call
    if (itc[c] + BUDGET_SLACK <= h->cbudget[c]) {
      h->cshrink_max[c] = std::max(h->cshrink_max[c], itc[c]);
      if (++h->cshrink_run[c] >= BUDGET_RUN) {
        h->cbudget[c] = std::max(4, h->cshrink_max[c] + 3);
        h->cshrink_run[c] = 0;
        h->cshrink_max[c] = 0;
      }
    } else {
      h->cshrink_run[c] = 0;
      h->cshrink_max[c] = 0;
    }
  }
  return 0;
}

// the most steps one captured graph holds (even, so that the replays of a long call keep the
// parity of advance!'s swaps and reuse one graph)
constexpr int GRAPH_MAX_STEPS = 128;

int check_desc(const tdgl_desc *d, char *why, size_t n) {
  auto no = [&](const char *m) {
    if (why && n) {
      std::strncpy(why, m, n - 1);
      why[n - 1] = 0;
    }
    return 0;
  };
  if (d->abi_version != TDGL_ABI_VERSION) return no("ABI version");
  if (d->nbands < 1 || d->nbands > TDGL_MAX_BANDS) return no("number of bands");
  if (d->Nx < 4 || d->Ny < 4 || d->Nz < 4) return no("grid smaller than 4 cells with padding per axis");
  if (!(d->x1 >= 2 && d->x2 < d->Nx && d->y1 >= 2 && d->y2 < d->Ny && d->z1 >= 2 && d->z2 < d->Nz && d->x1 <= d->x2 &&
        d->y1 <= d->y2 && d->z1 <= d->z2))
    return no("block outside the padded grid");
  for (int mu = 0; mu < d->nbands; ++mu) {
    if (d->stages[mu] < 0 || d->stages[mu] > TDGL_MAX_STAGES) return no("stage count");
    if (d->stages[mu] == 0 && d->nsub[mu] < 1) return no("substep count");
    if (d->stages[mu] > 0 && d->rkc[mu] && !(d->divisor[mu] > 0)) return no("RKC divisor");
  }
  if (d->rotframe && !d->extrap) return no("rotframe without extrap");
  if (d->field_mode < 0 || d->field_mode > 3) return no("field mode");
  if (d->field_mode == TDGL_FIELD_PCG && (d->pcg_maxit < 1 || d->pcg_maxit > (1 << 20) || !(d->pcg_tol >= 0)))
    return no("PCG iteration limit (1..2^20) or tolerance");
  if (d->pcg_refine && (d->field_mode != TDGL_FIELD_PCG || !d->pcg_f32 || d->pcg_rounds < 1 || d->pcg_rounds > 16 ||
                        !(d->pcg_inner_tol > 0 && d->pcg_inner_tol < 1)))
    return no("pcg_refine needs the PCG with pcg_f32, pcg_rounds in 1..16 and pcg_inner_tol in (0, 1)");
  if ((d->field_mode == TDGL_FIELD_PCG || d->field_mode == TDGL_FIELD_EXPLICIT_SCALED) && !(d->Dx && d->Dy && d->Dz))
    return no("link conductivities D");
  if ((d->field_mode == TDGL_FIELD_SEPARABLE || d->field_mode == TDGL_FIELD_PCG) && !(d->dx == d->dy && d->dy == d->dz))
    return no("implicit curl-curl with dx != dy or dz");
  if ((d->selffield_bc == TDGL_BC_BAR || d->selffield_bc == TDGL_BC_OPEN) &&
      !(d->bar_hy_z && d->bar_hy_y && d->bar_hz_z && d->bar_hz_y && d->bar_hz_x && d->bar_hy_x))
    return no("bar profiles");
  if (d->selffield_bc < 0 || d->selffield_bc > 2) return no("self-field boundary");
  if (d->selffield_bc == TDGL_BC_OPEN) {
    if (!(d->dx == d->dy && d->dy == d->dz)) return no("open boundary with dx != dy or dz");
    if (d->open_every < 1 || d->open_R < 0 || !d->open_lgf || !d->open_Jbar) return no("open boundary tables");
    const int nl[3] = {d->Nx + 1, d->Ny + 1, d->Nz + 1};
    for (int a = 0; a < 3; ++a)
      if (smooth_size(2 * nl[a] - 1) > 2048) return no("open boundary larger than the line FFT (2048 per padded axis)");
  }
  if (!(d->f && d->taux && d->tauy && d->tauz)) return no("masks");
  if (d->qp_mode != TDGL_QP_NONE) {
    if (d->qp_mode != TDGL_QP_TWO && d->qp_mode != TDGL_QP_N) return no("normal-fluid model");
    if (d->qp_mode == TDGL_QP_TWO && (d->nbands != 2 || d->qp_nmodes != 1)) return no("two normal fluids need two bands");
    if (d->qp_mode == TDGL_QP_N && (d->qp_nmodes != d->nbands || d->nbands < 2)) return no("normal-fluid modes");
    if (d->qp_j0 < 0 || d->qp_nj < 1 || d->qp_j0 + d->qp_nj > d->Ny + 1 || d->qp_j0 > d->y1 - 1 ||
        d->qp_j0 + d->qp_nj < d->y2 + 1 || d->qp_k0 < 0 || d->qp_nk < 1 || d->qp_k0 + d->qp_nk > d->Nz + 1 ||
        d->qp_k0 > d->z1 - 1 || d->qp_k0 + d->qp_nk < d->z2 + 1)
      return no("normal-fluid region");
    if (d->qp_nk > 256 || d->qp_nj > PLANE_LMAX || d->Nx - 1 < 1) return no("normal-fluid region larger than the transform tiles");
    if (!(d->qp_Vx && d->qp_Wx && d->qp_ex && d->qp_Vy && d->qp_Wy && d->qp_ey && d->qp_Vz && d->qp_Wz && d->qp_ez))
      return no("normal-fluid bases");
    if (!(d->qp_sigma > 0)) return no("normal-fluid σn");
  }
  const double cells = (double)(((d->Nx + 1 + 7) / 8) * 8) * (d->Ny + 1) * (d->Nz + 1);
  if (cells > 2.0e9) return no("grid too large for 32-bit indexing");
  return 1;
}

}  // namespace

extern "C" {

const char *tdgl_last_error(void) { return g_err.c_str(); }

int tdgl_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

int tdgl_set_device(int32_t dev) {
  HIPCHK(hipSetDevice(dev));
  return 0;
}

int tdgl_supported(const tdgl_desc *d, char *why, size_t n) {
  if (!d) return 0;
  if (why && n) why[0] = 0;
  return check_desc(d, why, n);
}

int tdgl_create(const tdgl_desc *desc, tdgl_handle **out) {
  if (!desc || !out) return fail(TDGL_E_ARG, "null argument");
  char why[256] = {0};
  if (!check_desc(desc, why, sizeof why)) return fail(TDGL_E_UNSUPPORTED, "unsupported stepper: %s", why);
  tdgl_handle *h = new tdgl_handle();
  *out = nullptr;
  auto bail = [&](int rc) {
    tdgl_destroy(h);
    return rc;
  };
  h->d = *desc;
  const tdgl_desc &d = h->d;
  (void)hipGetDevice(&h->dev);
  if (hipStreamCreateWithFlags(&h->s, hipStreamNonBlocking) != hipSuccess) return bail(fail(TDGL_E_DEVICE, "stream"));
  if (hipEventCreate(&h->e0) != hipSuccess || hipEventCreate(&h->e1) != hipSuccess) return bail(fail(TDGL_E_DEVICE, "events"));
  Geo &g = h->g;
  g.Nx = d.Nx;
  g.Ny = d.Ny;
  g.Nz = d.Nz;
  g.x1 = d.x1;
  g.x2 = d.x2;
  g.y1 = d.y1;
  g.y2 = d.y2;
  g.z1 = d.z1;
  g.z2 = d.z2;
  g.px = ((d.Nx + 1 + 7) / 8) * 8;
  g.py = d.Ny + 1;
  g.n = g.px * g.py * (d.Nz + 1);
  Ranges &rg = h->rg;
  const int ex[3][3] = {{d.Nx - 2, d.Ny - 1, d.Nz - 1}, {d.Nx - 1, d.Ny - 2, d.Nz - 1}, {d.Nx - 1, d.Ny - 1, d.Nz - 2}};
  int off = 0;
  for (int c = 0; c < 3; ++c) {
    for (int q = 0; q < 3; ++q) rg.e[c][q] = ex[c][q];
    rg.off[c] = off;
    off += ex[c][0] * ex[c][1] * ex[c][2];
  }
  rg.total = off;
  auto run = [&]() -> int {
    const int n = g.n;
    for (int mu = 0; mu < d.nbands; ++mu) {
      ALLOC(h->G[mu], n);
      ALLOC(h->F[mu], n);
    }
    for (int c = 0; c < 3; ++c) {
      ALLOC(h->th[c], n);
      ALLOC(h->ph[c], n);
      ALLOC(h->U[c], n);
      ALLOC(h->V[c], n);
      ALLOC(h->J[c], n);
      ALLOC(h->h[c], n);
      ALLOC(h->tau[c], n);
    }
    ALLOC(h->f, n);
    bool stages = false, subs = false;
    for (int mu = 0; mu < d.nbands; ++mu) {
      stages |= d.stages[mu] > 0;
      subs |= d.stages[mu] == 0 && d.nsub[mu] > 1;
    }
    if (stages || subs) ALLOC(h->W, n);
    if (stages) ALLOC(h->W2, n);
    if (d.rotframe) ALLOC(h->Om, n);
    h->noise = d.noise != 0;
    if (h->noise) {
      ALLOC(h->nctr, 2);
      bool sk = false;
      for (int mu = 0; mu < d.nbands; ++mu) sk = sk || d.noise_sk[mu];
      if (sk) {
        ALLOC(h->Sp, n);
        ALLOC(h->Qth, n);
      }
      if (d.rotframe) ALLOC(h->Om2, n);
    }
    CHK(put(h, h->f, d.f, d.Nx + 1, d.Ny + 1, d.Nz + 1, 8));
    const double *taus[3] = {d.taux, d.tauy, d.tauz};
    for (int c = 0; c < 3; ++c) CHK(put(h, h->tau[c], taus[c], d.Nx + 1, d.Ny + 1, d.Nz + 1, 8));
    if (d.field_mode == TDGL_FIELD_PCG || d.field_mode == TDGL_FIELD_EXPLICIT_SCALED) {
      const double *Ds[3] = {d.Dx, d.Dy, d.Dz};
      const int sh[3][3] = {{d.Nx, d.Ny + 1, d.Nz + 1}, {d.Nx + 1, d.Ny, d.Nz + 1}, {d.Nx + 1, d.Ny + 1, d.Nz}};
      for (int c = 0; c < 3; ++c) {
        ALLOC(h->D[c], n);
        CHK(put(h, h->D[c], Ds[c], sh[c][0], sh[c][1], sh[c][2], 8));
      }
    }
    // the open boundary adds h_δ to the bar self-field of the transport current (openbc.jl:5-8, field.jl:221)
    if (d.selffield_bc == TDGL_BC_BAR || d.selffield_bc == TDGL_BC_OPEN) {
      const size_t nn[6] = {(size_t)(d.Ny + 1) * 2, (size_t)2 * d.Nz, (size_t)d.Ny * 2, (size_t)2 * (d.Nz + 1),
                            (size_t)d.Ny * (d.Nz + 1), (size_t)(d.Ny + 1) * d.Nz};
      const double *src[6] = {d.bar_hy_z, d.bar_hy_y, d.bar_hz_z, d.bar_hz_y, d.bar_hz_x, d.bar_hy_x};
      for (int q = 0; q < 6; ++q) {
        ALLOC(h->bar[q], nn[q]);
        CHK(upload_vec(h, h->bar[q], std::vector<double>(src[q], src[q] + nn[q])));
      }
    }
    if (d.field_mode == TDGL_FIELD_SEPARABLE || d.field_mode == TDGL_FIELD_PCG) {
      const int nn[3] = {d.Nx - 1, d.Ny - 1, d.Nz - 1};
      for (int ax = 0; ax < 3; ++ax) {
        Bases b = make_bases(nn[ax]);
        for (int which = 0; which < 2; ++which) {
          const std::vector<double> &M = which == 0 ? b.V : b.U;
          const int Lm = which == 0 ? nn[ax] : nn[ax] - 1;
          std::vector<double> fwd((size_t)Lm * Lm), inv((size_t)Lm * Lm);
          for (int q = 0; q < Lm; ++q)
            for (int m = 0; m < Lm; ++m) {
              fwd[(size_t)q * Lm + m] = M[q + (size_t)Lm * m];  // M(q, m)
              inv[(size_t)q * Lm + m] = M[m + (size_t)Lm * q];  // M(m, q)
            }
          for (int t = 0; t < 2; ++t) {
            const std::vector<double> &src = t == 0 ? fwd : inv;
            ALLOC(h->Ad[ax][which][t], src.size());
            CHK(upload_vec(h, h->Ad[ax][which][t], src));
            std::vector<float> s32(src.begin(), src.end());
            ALLOC(h->Af[ax][which][t], s32.size());
            CHK(upload_vec(h, h->Af[ax][which][t], s32));
          }
        }
        ALLOC(h->sd[ax], nn[ax]);
        CHK(upload_vec(h, h->sd[ax], b.s));
        std::vector<float> s32(b.s.begin(), b.s.end());
        ALLOC(h->sf[ax], nn[ax]);
        CHK(upload_vec(h, h->sf[ax], s32));
      }
      for (int c = 0; c < 3; ++c) ALLOC(h->B.c[c], n);
      // T1 and T2 also hold the FP64 quarters (a little more than the region: ceil halves)
      int qe[12][3];
      h->t12cap = std::max((size_t)rg.total, (size_t)quarters_of(rg, qe).total);
      ALLOC(h->T1, h->t12cap);
      ALLOC(h->T2, h->t12cap);
    }
    if (d.field_mode == TDGL_FIELD_PCG) {
      Vec3 *vs[8] = {&h->x, &h->r, &h->z, &h->P, &h->P2, &h->Qv, &h->xprev, &h->zold};
      for (Vec3 *v : vs)
        for (int c = 0; c < 3; ++c) ALLOC(v->c[c], n);
      ALLOC(h->T32a, rg.total);
      ALLOC(h->T32b, rg.total);
      ALLOC(h->pd.sc, SC_COUNT);
      ALLOC(h->pd.fl, FL_COUNT);
      ALLOC(h->pd.part, (size_t)PART_N * 2 * MAX_PART);
      // TDGL_HIP_BOX_KC=k: k levels per block of the box kernels instead of box_kc's choice
      h->box_kc = box_kc(g);
      const char *bk = std::getenv("TDGL_HIP_BOX_KC");
      if (bk && std::atoi(bk) >= 1) h->box_kc = std::min(std::atoi(bk), d.Nz - 1);
      if (pcg_box_blocks(g, h->box_kc) > MAX_PART || field_blocks(g) > MAX_PART) return fail(TDGL_E_UNSUPPORTED, "grid too large for the PCG reductions");
      std::vector<double> sc(SC_COUNT, 0.0);
      sc[SC_TOL] = d.pcg_tol;
      sc[SC_TOLI] = d.pcg_inner_tol;
      CHK(upload_vec(h, h->pd.sc, sc));
      std::vector<int> fl(FL_COUNT, 0);
      fl[FL_MAXIT] = d.pcg_maxit;
      CHK(upload_vec(h, h->pd.fl, fl));
    }
    if (d.qp_mode != TDGL_QP_NONE) {
      const int n1 = d.Nx - 1, nj = d.qp_nj, nk = d.qp_nk, K = d.qp_nmodes;
      h->qn = n1 * nj * nk;
      for (int k = 0; k < K; ++k) ALLOC(h->qS[k], (size_t)h->qn);
      ALLOC(h->qT1, (size_t)h->qn * K);
      ALLOC(h->qT2, (size_t)h->qn * K);
      const int nn[3] = {n1, nj, nk};
      const double *V[3] = {d.qp_Vx, d.qp_Vy, d.qp_Vz}, *W[3] = {d.qp_Wx, d.qp_Wy, d.qp_Wz},
                   *E[3] = {d.qp_ex, d.qp_ey, d.qp_ez};
      for (int a = 0; a < 3; ++a) {
        const size_t m = (size_t)nn[a] * nn[a];
        ALLOC(h->qV[a], m);
        ALLOC(h->qW[a], m);
        ALLOC(h->qe[a], (size_t)nn[a]);
        CHK(upload_vec(h, h->qV[a], std::vector<double>(V[a], V[a] + m)));
        CHK(upload_vec(h, h->qW[a], std::vector<double>(W[a], W[a] + m)));
        CHK(upload_vec(h, h->qe[a], std::vector<double>(E[a], E[a] + nn[a])));
      }
      PlanePass t{};
      t.ncomp = K;
      for (int k = 0; k < K; ++k) {
        t.e[k][0] = n1;
        t.e[k][1] = nj;
        t.e[k][2] = nk;
      }
      if (!plane_fits<double>(t.e, K) || plane_parts<double>(t) > MAX_PART)
        return fail(TDGL_E_UNSUPPORTED, "normal-fluid region larger than the transform tiles");
    }
    if (d.selffield_bc == TDGL_BC_OPEN) {
      const int nl[3][3] = {{d.Nx, d.Ny + 1, d.Nz + 1}, {d.Nx + 1, d.Ny, d.Nz + 1}, {d.Nx + 1, d.Ny + 1, d.Nz}};
      const double pi = 3.14159265358979323846;
      for (int c = 0; c < 3; ++c) {
        for (int a = 0; a < 3; ++a) {
          h->on[c][a] = nl[c][a];
          const int M = smooth_size(2 * nl[c][a] - 1);
          h->oM[c][a] = M;
          fft_radices(M, h->oax[c][a]);
          std::vector<double> tw(2 * (size_t)M);
          for (int k = 0; k < M; ++k) {
            tw[2 * k] = std::cos(2 * pi * k / M);
            tw[2 * k + 1] = -std::sin(2 * pi * k / M);
          }
          ALLOC(h->otw[c][a], (size_t)M);
          CHK(upload_vec(h, (double *)h->otw[c][a], tw));
          h->oax[c][a].tw = h->otw[c][a];
        }
        const size_t nb = (size_t)nl[c][0] * nl[c][1] * nl[c][2];
        ALLOC(h->odj[c], nb);
        ALLOC(h->oth[c], nb);
        ALLOC(h->obuf[c], (size_t)h->oM[c][0] * h->oM[c][1] * nl[c][2]);
        ALLOC(h->oghat[c], (size_t)h->oM[c][0] * h->oM[c][1] * h->oM[c][2]);
        ALLOC(h->ohold[c], n);
        ALLOC(h->ohnew[c], n);
      }
      ALLOC(h->ofl, 4);
      CHK(upload_vec(h, h->ofl, std::vector<int>{-1, 0, 0, 0}));
      const size_t nt = (size_t)(d.open_R + 1) * (d.open_R + 1) * (d.open_R + 1);
      ALLOC(h->olgf, nt);
      CHK(upload_vec(h, h->olgf, std::vector<double>(d.open_lgf, d.open_lgf + nt)));
      const size_t nj = (size_t)(d.Ny + 1) * (d.Nz + 1);
      ALLOC(h->oJbar, nj);
      CHK(upload_vec(h, h->oJbar, std::vector<double>(d.open_Jbar, d.open_Jbar + nj)));
      CHK(open_green(h));
    }
    ALLOC(h->hp, 2);
    Block b1{}, b2{};
    HIPCHK(hblock(2 * sizeof(double), &b1));
    h->pins.push_back(b1);
    HIPCHK(hblock(FL_COUNT * sizeof(int), &b2));
    h->pins.push_back(b2);
    h->hph = (double *)b1.p;
    h->flh = (int *)b2.p;
    // the copy of a cycle's start for the roll back of a captured PCG cycle
    if (d.field_mode == TDGL_FIELD_PCG) {
      int q = 0;
      for (int mu = 0; mu < d.nbands; ++mu) ALLOC(h->bk[q++], 2 * (size_t)n);
      for (int k = 0; k < 3; ++k) {
        ALLOC(h->bk[q++], n);
        ALLOC(h->bk[q++], n);
        ALLOC(h->bk[q++], 2 * (size_t)n);
        ALLOC(h->bk[q++], 2 * (size_t)n);
        ALLOC(h->bk[q++], n);
        ALLOC(h->bk[q++], n);
      }
      for (int k = 0; k < 6; ++k) ALLOC(h->bk[q++], n);
      if (d.selffield_bc == TDGL_BC_OPEN) {
        for (int k = 0; k < 6; ++k) ALLOC(h->bk[q++], n);
        ALLOC(h->bk[q++], 2);
      }
      if (h->noise) ALLOC(h->bk[q++], 2);
    }
    // TDGL_HIP_GRAPH=0: every step launched from the host; TDGL_HIP_BUDGET=k: the captured PCG
    // starts with k iterations per solve instead of learning it from an eager first cycle
    const char *ge = std::getenv("TDGL_HIP_GRAPH");
    h->graphs = !(ge && ge[0] == '0');
    const char *gb = std::getenv("TDGL_HIP_BUDGET");
    if (gb && d.field_mode == TDGL_FIELD_PCG) {
      h->budget = std::max(1, std::min(d.pcg_maxit, std::atoi(gb)));
      h->budget_fixed = true;
    }
    // TDGL_HIP_PLANE=0: the separable solve as six single-axis passes even where the fused
    // x-y and z-mode launches fit
    // TDGL_HIP_PLANE1=0: the chunked Float32 x-y kernel also where the one-round kernel fits;
    // =2 or =3: the one-round kernel with that many m tiles per block wherever it fits
    const char *pt = std::getenv("TDGL_HIP_PAPTILE");
    h->paptile = pt && pt[0] == '0' ? 0 : 1;
    const char *bx = std::getenv("TDGL_HIP_BOXXCD");
    h->boxxcd = bx && bx[0] == '0' ? 0 : 1;
    const char *zs = std::getenv("TDGL_HIP_ZSWAP");
    h->zswap = zs && zs[0] == '0' ? 0 : 1;
    const char *zc = std::getenv("TDGL_HIP_ZCOL");
    h->zcols = zc && zc[0] == '1' ? 1 : 0;
    const char *qp = std::getenv("TDGL_HIP_QPLANE");
    h->qplane = qp && qp[0] == '0' ? 0 : 1;
    const char *qs = std::getenv("TDGL_HIP_QPARTS");
    if (qs && qs[0] >= '1' && qs[0] <= '3' && !qs[1]) h->qparts = qs[0] - '0';
    // TDGL_HIP_ZQ_WAVES=4 or 8: the quarter z pass on the vector ALUs with that many waves a block
    // (8 by default), 0: on the matrix cores
    const char *sx = std::getenv("TDGL_HIP_STENCILXCD");
    h->stxcd = sx && sx[0] == '0' ? 0 : 1;
    const char *zw = std::getenv("TDGL_HIP_ZQ_WAVES");
    if (zw && (std::atoi(zw) == 0 || std::atoi(zw) == 4 || std::atoi(zw) == 8)) h->zq_waves = std::atoi(zw);
    const char *g1 = std::getenv("TDGL_HIP_PLANE1");
    if (g1 && (g1[0] == '0' || g1[0] == '2' || g1[0] == '3')) h->plane1 = g1[0] - '0';
    const char *gp = std::getenv("TDGL_HIP_PLANE");
    const bool zfits = d.Nz - 1 >= 2 && d.Nz - 1 <= 256;
    if (!(gp && gp[0] == '0') && zfits) {
      h->plane32 = plane_fits<float>(rg.e, 3);
      h->plane64 = plane_fits<double>(rg.e, 3);
    }
    // the parity-folded Float32 preconditioner (TDGL_HIP_FOLD=0: the full-length transforms)
    const char *gf = std::getenv("TDGL_HIP_FOLD");
    if (d.field_mode == TDGL_FIELD_PCG && d.pcg_f32 && !(gf && gf[0] == '0') && !(gp && gp[0] == '0') &&
        d.Nz - 1 >= 2 && d.Nz - 1 <= ZQ_NZ_MAX) {
      int qe[12][3];
      const Quarters q = quarters_of(rg, qe);
      if (plane_fits<float>(qe, 12) && fold_blocks(q) <= MAX_PART) {
        h->q = q;
        ALLOC(h->qa, q.total);
        ALLOC(h->qb, q.total);
        const int nn[2] = {d.Nx - 1, d.Ny - 1};
        for (int ax = 0; ax < 2; ++ax) {
          const Bases b = make_bases(nn[ax]);
          for (int which = 0; which < 2; ++which) {
            const std::vector<double> &M = which == 0 ? b.V : b.U;
            const int L = which == 0 ? nn[ax] : nn[ax] - 1, Lh = (L + 1) / 2;
            for (int par = 0; par < 2; ++par) {
              // the columns of parity `par` on the first ceil(L/2) rows: Mh(i, cc) = M(i, 2 cc + par)
              std::vector<float> hf((size_t)Lh * Lh, 0.0f), hi((size_t)Lh * Lh, 0.0f);
              for (int i = 0; i < Lh; ++i)
                for (int cc = 0; cc < Lh; ++cc) {
                  const int col = 2 * cc + par;
                  const float v = col < L ? (float)M[i + (size_t)L * col] : 0.0f;
                  hf[(size_t)i * Lh + cc] = v;  // Mh(i, cc)
                  hi[(size_t)cc * Lh + i] = v;  // Hi[q][m] = Mh(m, q)
                }
              ALLOC(h->Hq[ax][which][par][0], hf.size());
              CHK(upload_vec(h, h->Hq[ax][which][par][0], hf));
              ALLOC(h->Hq[ax][which][par][1], hi.size());
              CHK(upload_vec(h, h->Hq[ax][which][par][1], hi));
              // the same, zero-padded to whole 16 x 16 tiles (k_qplane)
              const int LP = 16 * cdiv(Lh, 16);
              for (int t = 0; t < 2; ++t) {
                const std::vector<float> &src = t == 0 ? hf : hi;
                std::vector<float> pd((size_t)LP * LP, 0.0f);
                for (int r = 0; r < Lh; ++r)
                  for (int c = 0; c < Lh; ++c) pd[(size_t)r * LP + c] = src[(size_t)r * Lh + c];
                ALLOC(h->Hp[ax][which][par][t], pd.size());
                CHK(upload_vec(h, h->Hp[ax][which][par][t], pd));
              }
              h->HpL[ax][which] = Lh;
            }
          }
        }
        {
          // the z pass's padded bases: V(n, mz), and U(n, mz - 1) with its zero column and row
          const int nz = d.Nz - 1, Lu = nz - 1;
          const Bases b = make_bases(nz);
          std::vector<float> zb[4];
          for (int k = 0; k < 4; ++k) zb[k].assign((size_t)nz * nz + ZQ_SLACK, 0.0f);
          for (int n = 0; n < nz; ++n)
            for (int mz = 0; mz < nz; ++mz) {
              const float v = (float)b.V[n + (size_t)nz * mz];
              const float u = (mz >= 1 && n < Lu) ? (float)b.U[n + (size_t)Lu * (mz - 1)] : 0.0f;
              zb[0][(size_t)n * nz + mz] = v;
              zb[1][(size_t)n * nz + mz] = u;
              zb[2][(size_t)mz * nz + n] = v;
              zb[3][(size_t)mz * nz + n] = u;
            }
          for (int k = 0; k < 4; ++k) {
            ALLOC(h->zqb[k], zb[k].size());
            CHK(upload_vec(h, h->zqb[k], zb[k]));
          }
          // k_zmodes_mf's image: Vt, Ut (by columns), V, U (by rows), NZP rows of NZP + 4, zero beyond nz
          const int nzp = 16 * ((nz + 15) / 16), sp = nzp + 4;
          std::vector<float> img((size_t)zmodes_mf_image(nz), 0.0f);
          for (int b = 0; b < 4; ++b) {
            const std::vector<float> &src = zb[b == 0 ? 2 : b == 1 ? 3 : b == 2 ? 0 : 1];
            for (int r = 0; r < nz; ++r)
              for (int k = 0; k < nz; ++k) img[((size_t)b * nzp + r) * sp + k] = src[(size_t)r * nz + k];
          }
          ALLOC(h->zqmf, img.size());
          CHK(upload_vec(h, h->zqmf, img));
        }
        h->fold = true;
      }
    }
    // the parity-folded FP64 separable solve (TDGL_HIP_FOLD64=0: the full-length transforms)
    // the folds by default where the full-length FP64 planes are too large for the one-round
    // kernel (wider than 160 x 96: they would fall to the vector ALUs); on smaller planes the two
    // launches the folds add cost more than their halved products save (TDGL_HIP_FOLD64=1: always,
    // =0: never).  The separable handles' z pass takes whole z-columns (TDGL_HIP_ZCOL=0: x-runs).
    const char *gf64 = std::getenv("TDGL_HIP_FOLD64");
    int l0max = 0, l1max = 0;
    for (int c = 0; c < 3; ++c) {
      l0max = std::max(l0max, rg.e[c][0]);
      l1max = std::max(l1max, rg.e[c][1]);
    }
    const bool fold64_pays = l0max > 160 || l1max > 96;
    const bool fold64_on = gf64 ? gf64[0] == '1' : fold64_pays;
    if (d.field_mode == TDGL_FIELD_SEPARABLE && !(zc && zc[0] == '0')) h->zcols = 1;
    if (d.field_mode == TDGL_FIELD_SEPARABLE && h->plane64 && fold64_on && d.Nz - 1 >= 2 && d.Nz - 1 <= ZQ64_NZ_MAX) {
      int qe[12][3];
      const Quarters q = quarters_of(rg, qe);
      if (plane_fits<double>(qe, 12) && fold_blocks(q) <= MAX_PART && (size_t)q.total <= h->t12cap) {
        h->q = q;
        const int nn[2] = {d.Nx - 1, d.Ny - 1};
        for (int ax = 0; ax < 2; ++ax) {
          const Bases b = make_bases(nn[ax]);
          for (int which = 0; which < 2; ++which) {
            const std::vector<double> &M = which == 0 ? b.V : b.U;
            const int L = which == 0 ? nn[ax] : nn[ax] - 1, Lh = (L + 1) / 2;
            for (int par = 0; par < 2; ++par) {
              // Mh(i, cc) = M(i, 2 cc + par) on the first ceil(L/2) rows, and its transpose
              std::vector<double> hf((size_t)Lh * Lh, 0.0), hi((size_t)Lh * Lh, 0.0);
              for (int i = 0; i < Lh; ++i)
                for (int cc = 0; cc < Lh; ++cc) {
                  const int col = 2 * cc + par;
                  const double v = col < L ? M[i + (size_t)L * col] : 0.0;
                  hf[(size_t)i * Lh + cc] = v;
                  hi[(size_t)cc * Lh + i] = v;
                }
              ALLOC(h->Hq64[ax][which][par][0], hf.size());
              CHK(upload_vec(h, h->Hq64[ax][which][par][0], hf));
              ALLOC(h->Hq64[ax][which][par][1], hi.size());
              CHK(upload_vec(h, h->Hq64[ax][which][par][1], hi));
              const int LP = 16 * cdiv(Lh, 16);
              for (int t = 0; t < 2; ++t) {
                const std::vector<double> &src = t == 0 ? hf : hi;
                std::vector<double> pd((size_t)LP * LP, 0.0);
                for (int r = 0; r < Lh; ++r)
                  for (int c = 0; c < Lh; ++c) pd[(size_t)r * LP + c] = src[(size_t)r * Lh + c];
                ALLOC(h->Hp64[ax][which][par][t], pd.size());
                CHK(upload_vec(h, h->Hp64[ax][which][par][t], pd));
              }
              h->HpL[ax][which] = Lh;
            }
          }
        }
        const int nz = d.Nz - 1, Lu = nz - 1;
        const Bases b = make_bases(nz);
        std::vector<double> zb[4];
        for (int k = 0; k < 4; ++k) zb[k].assign((size_t)nz * nz + ZQ_SLACK, 0.0);
        for (int n = 0; n < nz; ++n)
          for (int mz = 0; mz < nz; ++mz) {
            const double v = b.V[n + (size_t)nz * mz];
            const double u = (mz >= 1 && n < Lu) ? b.U[n + (size_t)Lu * (mz - 1)] : 0.0;
            zb[0][(size_t)n * nz + mz] = v;
            zb[1][(size_t)n * nz + mz] = u;
            zb[2][(size_t)mz * nz + n] = v;
            zb[3][(size_t)mz * nz + n] = u;
          }
        for (int k = 0; k < 4; ++k) {
          ALLOC(h->zqb64[k], zb[k].size());
          CHK(upload_vec(h, h->zqb64[k], zb[k]));
        }
        h->fold64 = true;
      }
    }
    if (d.field_mode == TDGL_FIELD_PCG && !(std::getenv("TDGL_HIP_DPAL") && std::getenv("TDGL_HIP_DPAL")[0] == '0')) {
      // D as a palette when each component takes at most DPAL_MAX values: one 32-bit code per
      // node instead of three doubles in the PCG operator (the same values, so the same bits)
      std::vector<double> pal((size_t)3 * DPAL_MAX, 0.0);
      std::vector<uint32_t> code((size_t)n, 0u);
      bool fits = true;
      for (int c = 0; c < 3 && fits; ++c) {
        std::vector<double> dd((size_t)n);
        CHK(hcopy(h, dd.data(), h->D[c], (size_t)n * sizeof(double), hipMemcpyDeviceToHost));
        std::vector<double> vals;
        for (size_t e = 0; e < (size_t)n && fits; ++e) {
          auto it = std::find(vals.begin(), vals.end(), dd[e]);
          size_t ix = (size_t)(it - vals.begin());
          if (it == vals.end()) {
            if (vals.size() == (size_t)DPAL_MAX) {
              fits = false;
              break;
            }
            vals.push_back(dd[e]);
          }
          code[e] |= (uint32_t)ix << (8 * c);
        }
        for (size_t k = 0; k < vals.size(); ++k) pal[(size_t)DPAL_MAX * c + k] = vals[k];
      }
      if (fits) {
        ALLOC(h->Dcode, n);
        CHK(upload_vec(h, h->Dcode, code));
        ALLOC(h->Dpal, pal.size());
        CHK(upload_vec(h, h->Dpal, pal));
        ALLOC(h->Dpal32, pal.size());
        CHK(upload_vec(h, h->Dpal32, std::vector<float>(pal.begin(), pal.end())));
      }
    }
    if (d.field_mode == TDGL_FIELD_PCG && d.pcg_refine) {
      // FP64 iterative refinement around Float32 inner solves: their vectors, D in Float32
      if (!h->fold || !d.pcg_f32 || !(d.pcg_fast || d.pcg_f32))
        return fail(TDGL_E_UNSUPPORTED, "the refinement needs the folded Float32 preconditioner (pcg_f32)");
      if (d.pcg_rounds < 1 || d.pcg_rounds > 16) return fail(TDGL_E_ARG, "pcg_rounds must be 1..16");
      Vec3f *fv[7] = {&h->f32v.d, &h->f32v.r, &h->f32v.z, &h->f32v.zold, &h->f32v.Pnew, &h->f32v.Pold, &h->f32v.Q};
      for (Vec3f *v : fv)
        for (int c = 0; c < 3; ++c) {
          ALLOC(v->c[c], n);
          CHK(upload_vec(h, v->c[c], std::vector<float>((size_t)n, 0.0f)));
        }
      for (int c = 0; c < 3; ++c) {
        std::vector<double> dd((size_t)n);
        CHK(hcopy(h, dd.data(), h->D[c], (size_t)n * sizeof(double), hipMemcpyDeviceToHost));
        ALLOC(h->D32[c], n);
        CHK(upload_vec(h, h->D32[c], std::vector<float>(dd.begin(), dd.end())));
        h->f32v.D[c] = h->D32[c];
      }
      h->refine = true;
      h->rounds = d.pcg_rounds;
    }
    // budgets per step class with the open boundary (its recomputations of h_δ lengthen the next
    // few solves; TDGL_HIP_STEPWISE=0: one budget and one graph a call)
    const char *sw = std::getenv("TDGL_HIP_STEPWISE");
    if (d.field_mode == TDGL_FIELD_PCG && d.selffield_bc == TDGL_BC_OPEN && h->ofl && !h->refine && !(sw && sw[0] == '0')) {
      Block b{};
      HIPCHK(hblock(sizeof(int), &b));
      h->pins.push_back(b);
      h->oclk_h = (int *)b.p;
      h->stepwise = true;
      for (int c = 0; c < PCG_NCLS; ++c) h->cbudget[c] = h->budget;  // TDGL_HIP_BUDGET's, else learnt
    }
    return 0;
  };
  int rc = run();
  if (rc) return bail(rc);
  if (hipStreamSynchronize(h->s) != hipSuccess) return bail(fail(TDGL_E_DEVICE, "the fills and uploads of the allocations"));
  stages_back(h);
  *out = h;
  return 0;
}

int tdgl_destroy(tdgl_handle *h) {
  if (!h) return 0;
  if (h->s) (void)hipStreamSynchronize(h->s);
  destroy_graph(h);
  // the stream has drained (above): the blocks are idle
  for (const Block &b : h->allocs) cache_give(b);
  for (const Block &b : h->mallocs) cache_give(b);
  if (h->ndraw) cache_give(h->ndraw_b);
  for (const Block &b : h->pins) cache_give(b);
  stages_back(h);
  for (hipEvent_t e : h->evs) (void)hipEventDestroy(e);
  if (h->e0) (void)hipEventDestroy(h->e0);
  if (h->e1) (void)hipEventDestroy(h->e1);
  if (h->s) (void)hipStreamDestroy(h->s);
  delete h;
  return 0;
}

int tdgl_upload_state(tdgl_handle *h, const tdgl_state *s) {
  if (!h || !s) return fail(TDGL_E_ARG, "null argument");
  const tdgl_desc &d = h->d;
  const int Nx = d.Nx, Ny = d.Ny, Nz = d.Nz;
  const Geo &g = h->g;
  HIPCHK(hipStreamSynchronize(h->s));
  for (int mu = 0; mu < d.nbands; ++mu) {
    if (!s->G[mu] || !s->F[mu]) {
      (void)flush(h);
      return fail(TDGL_E_ARG, "null band array");
    }
    CHK(put_async(h, h->G[mu], s->G[mu], Nx + 1, Ny + 1, Nz + 1, 16));
    // F takes G's values outside the block (they are never stepped and advance! copies only
    // the block), which makes the advance a swap
    std::vector<double> Fh((size_t)(Nx + 1) * (Ny + 1) * (Nz + 1) * 2);
    std::memcpy(Fh.data(), s->F[mu], Fh.size() * 8);
    for (int k = 1; k <= Nz + 1; ++k)
      for (int j = 1; j <= Ny + 1; ++j)
        for (int i = 1; i <= Nx + 1; ++i) {
          const bool inb = i >= d.x1 && i <= d.x2 + 1 && j >= d.y1 && j <= d.y2 + 1 && k >= d.z1 && k <= d.z2 + 1;
          if (inb) continue;
          const size_t q = (size_t)(i - 1) + (size_t)(Nx + 1) * ((j - 1) + (size_t)(Ny + 1) * (k - 1));
          Fh[2 * q] = s->G[mu][2 * q];
          Fh[2 * q + 1] = s->G[mu][2 * q + 1];
        }
    CHK(put_async(h, h->F[mu], Fh.data(), Nx + 1, Ny + 1, Nz + 1, 16));
    h->Fup[mu].assign(s->F[mu], s->F[mu] + Fh.size());
  }
  h->stepped = false;
  (void)g;
  const int lx[3] = {Nx, Ny + 1, Nz + 1}, ly[3] = {Nx + 1, Ny, Nz + 1}, lz[3] = {Nx + 1, Ny + 1, Nz};
  const int *ls[3] = {lx, ly, lz};
  const int hx[3] = {Nx + 1, Ny, Nz}, hy[3] = {Nx, Ny + 1, Nz}, hz[3] = {Nx, Ny, Nz + 1};
  const int *hs[3] = {hx, hy, hz};
  const double *th[3] = {s->thx, s->thy, s->thz}, *ph[3] = {s->phx, s->phy, s->phz};
  const double *U[3] = {s->Ux, s->Uy, s->Uz}, *V[3] = {s->Vx, s->Vy, s->Vz};
  const double *J[3] = {s->Jx, s->Jy, s->Jz}, *hh[3] = {s->hx, s->hy, s->hz};
  for (int c = 0; c < 3; ++c) {
    if (!th[c] || !ph[c] || !U[c] || !V[c] || !J[c] || !hh[c]) {
      (void)flush(h);
      return fail(TDGL_E_ARG, "null state array");
    }
    CHK(put_async(h, h->th[c], th[c], ls[c][0], ls[c][1], ls[c][2], 8));
    CHK(put_async(h, h->ph[c], ph[c], ls[c][0], ls[c][1], ls[c][2], 8));
    CHK(put_async(h, h->U[c], U[c], ls[c][0], ls[c][1], ls[c][2], 16));
    CHK(put_async(h, h->V[c], V[c], ls[c][0], ls[c][1], ls[c][2], 16));
    CHK(put_async(h, h->J[c], J[c], ls[c][0], ls[c][1], ls[c][2], 8));
    CHK(put_async(h, h->h[c], hh[c], hs[c][0], hs[c][1], hs[c][2], 8));
  }
  return flush(h);
}

int tdgl_download_state(tdgl_handle *h, tdgl_state *s) {
  if (!h || !s) return fail(TDGL_E_ARG, "null argument");
  const tdgl_desc &d = h->d;
  const int Nx = d.Nx, Ny = d.Ny, Nz = d.Nz;
  for (int mu = 0; mu < d.nbands; ++mu)
    if (!s->G[mu] || !s->F[mu]) return fail(TDGL_E_ARG, "null band array");
  const int lx[3] = {Nx, Ny + 1, Nz + 1}, ly[3] = {Nx + 1, Ny, Nz + 1}, lz[3] = {Nx + 1, Ny + 1, Nz};
  const int *ls[3] = {lx, ly, lz};
  const int hx[3] = {Nx + 1, Ny, Nz}, hy[3] = {Nx, Ny + 1, Nz}, hz[3] = {Nx, Ny, Nz + 1};
  const int *hs[3] = {hx, hy, hz};
  double *th[3] = {s->thx, s->thy, s->thz}, *ph[3] = {s->phx, s->phy, s->phz};
  double *U[3] = {s->Ux, s->Uy, s->Uz}, *V[3] = {s->Vx, s->Vy, s->Vz};
  double *J[3] = {s->Jx, s->Jy, s->Jz}, *hh[3] = {s->hx, s->hy, s->hz};
  for (int c = 0; c < 3; ++c)
    if (!th[c] || !ph[c] || !U[c] || !V[c] || !J[c] || !hh[c]) return fail(TDGL_E_ARG, "null state array");
  // every copy queued, one wait
  for (int mu = 0; mu < d.nbands; ++mu) {
    CHK(get_async(h, s->G[mu], h->G[mu], Nx + 1, Ny + 1, Nz + 1, 16));
    if (!h->stepped) CHK(get_async(h, s->F[mu], h->F[mu], Nx + 1, Ny + 1, Nz + 1, 16));
  }
  for (int c = 0; c < 3; ++c) {
    CHK(get_async(h, th[c], h->th[c], ls[c][0], ls[c][1], ls[c][2], 8));
    CHK(get_async(h, ph[c], h->ph[c], ls[c][0], ls[c][1], ls[c][2], 8));
    CHK(get_async(h, U[c], h->U[c], ls[c][0], ls[c][1], ls[c][2], 16));
    CHK(get_async(h, V[c], h->V[c], ls[c][0], ls[c][1], ls[c][2], 16));
    CHK(get_async(h, J[c], h->J[c], ls[c][0], ls[c][1], ls[c][2], 8));
    CHK(get_async(h, hh[c], h->h[c], hs[c][0], hs[c][1], hs[c][2], 8));
  }
  CHK(get_flush(h));
  for (int mu = 0; mu < d.nbands; ++mu) {
    // F as the reference leaves it (advance!, field.jl:248-255): after a step it equals G on the
    // block (the swap left the step before in the device's F), and outside the block it is the
    // F that was uploaded
    const size_t ne = (size_t)(Nx + 1) * (Ny + 1) * (Nz + 1) * 2;
    if (h->stepped) std::memcpy(s->F[mu], s->G[mu], ne * 8);
    if (h->Fup[mu].size() == ne)
      for (int k = 1; k <= Nz + 1; ++k)
        for (int j = 1; j <= Ny + 1; ++j)
          for (int i = 1; i <= Nx + 1; ++i) {
            if (i >= d.x1 && i <= d.x2 + 1 && j >= d.y1 && j <= d.y2 + 1 && k >= d.z1 && k <= d.z2 + 1) continue;
            const size_t q = (size_t)(i - 1) + (size_t)(Nx + 1) * ((j - 1) + (size_t)(Ny + 1) * (k - 1));
            s->F[mu][2 * q] = h->Fup[mu][2 * q];
            s->F[mu][2 * q + 1] = h->Fup[mu][2 * q + 1];
          }
  }
  return 0;
}

// the PCG counters of the cycle whose flags are in flh
static void pcg_account(tdgl_handle *h) {
  h->pcg_iters += h->flh[FL_ITSUM];
  h->last_iters = h->flh[FL_IT];
  // the graph's iterations per solve: the most one solve took, and a margin (per class: the
  // classes no step of the first call had take the overall one)
  if (h->budget == 0) {
    h->budget = std::min(d_maxit(h), std::max(4, h->flh[FL_ITMAX] + 3));
    for (int c = 0; c < PCG_NCLS; ++c) {
      const int m = h->flh[FL_ITMAXC + c];
      h->cbudget[c] = (m > 0 && !h->budget_fixed) ? std::min(d_maxit(h), std::max(4, m + 3)) : h->budget;
      h->cseen[c] = h->cseen[c] || m > 0;
    }
  }
}

int tdgl_step(tdgl_handle *h, int32_t n, double H, double HJ) {
  if (!h) return fail(TDGL_E_ARG, "null handle");
  if (n < 0) return fail(TDGL_E_ARG, "negative step count");
  if (!std::isfinite(H) || !std::isfinite(HJ)) return fail(TDGL_E_ARG, "non-finite H or HJ");
  const bool pcg = h->d.field_mode == TDGL_FIELD_PCG;
  if (h->noise && h->d.noise_host && n > 0 && h->nctr_h + (uint64_t)n > h->nbase_h + (uint64_t)h->nslots)
    return fail(TDGL_E_ARG, "noise: the caller's draws cover %lld more steps, %d asked",
                (long long)(h->nbase_h + (uint64_t)h->nslots - std::min<uint64_t>(h->nctr_h, h->nbase_h + (uint64_t)h->nslots)),
                n);
  if (n > 0) h->stepped = true;
  h->hph[0] = H;
  h->hph[1] = HJ;
  HIPCHK(hipMemcpyAsync(h->hp, h->hph, 2 * sizeof(double), hipMemcpyHostToDevice, h->s));
  HIPCHK(hipEventRecord(h->e0, h->s));
  if (n > 0 && h->stepwise) {
    // the device clock's count: the classes of this call's steps
    HIPCHK(hipMemcpyAsync(h->oclk_h, h->ofl + OF_COUNT, sizeof(int), hipMemcpyDeviceToHost, h->s));
    HIPCHK(hipStreamSynchronize(h->s));
    h->oclock = *h->oclk_h;
    h->clock_known = true;
  }
  // the classes hold for this call's steps only (the tuner and the snapshots step otherwise)
  struct ClockScope {
    tdgl_handle *h;
    ~ClockScope() {
      h->clock_known = false;
      h->cur_cls = 0;
    }
  } clock_scope{h};
  if (n > 0) {
    const bool eager = h->prof || !h->graphs || (pcg && h->budget == 0);
    if (!eager && h->stepwise) {
      CHK(stepwise_graphs(h, n));
      h->graph_cycles += 1;
    } else if (eager) {
      if (pcg) PL(K_VEC, launch_pcg_cycle_reset(h->pd, h->s));
      CHK(steps_eager(h, n));
      if (pcg) HIPCHK(hipMemcpyAsync(h->flh, h->pd.fl, FL_COUNT * sizeof(int), hipMemcpyDeviceToHost, h->s));
    } else {
      // a long call replays a graph of at most GRAPH_MAX_STEPS steps: a captured stream is one
      // chain of dependent nodes (about 15 a step), which the runtime walks when it instantiates
      // and launches it, so its length stays bounded
      for (int left = n; left > 0;) {
        const int m = std::min(left, GRAPH_MAX_STEPS);
        CHK(cycle_graph(h, m));
        h->graph_cycles += 1;
        left -= m;
        if (pcg && left > 0) {
          HIPCHK(hipStreamSynchronize(h->s));
          pcg_account(h);
        }
      }
    }
  }
  HIPCHK(hipEventRecord(h->e1, h->s));
  HIPCHK(hipStreamSynchronize(h->s));
  if (h->prof) CHK(prof_flush(h));
  if (pcg && n > 0) pcg_account(h);
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, h->e0, h->e1));
  h->device_ms += ms;
  if (h->noise) h->nctr_h += (uint64_t)n;
  h->steps += n;
  h->calls += 1;
  return 0;
}

// ---------------------------------------------------------------------------- no allocation
// The PCG's reduction partials, scalars and flags for tdgl_step_noalloc: a file-scope device
// array in the layout of a handle's own (PART_N regions of 2 MAX_PART doubles, then SC_COUNT
// doubles, then FL_COUNT ints), one copy per GPU for the life of the process.  Host code uses the
// address hipGetSymbolAddress gives on each device, never the array's name.
}  // extern "C"
constexpr size_t SCRATCH_PART = (size_t)PART_N * 2 * MAX_PART;  // doubles
constexpr size_t SCRATCH_BYTES = (SCRATCH_PART + SC_COUNT) * sizeof(double) + FL_COUNT * sizeof(int);
static __device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];
namespace {
constexpr int SCRATCH_DEVICES = 64;
void *g_scratch[SCRATCH_DEVICES] = {};
std::mutex g_scratch_mu;

// the scratch's address on device `dev`, looked up once per device with that device current
int scratch_on(int dev, void **out) {
  if (dev < 0 || dev >= SCRATCH_DEVICES) return fail(TDGL_E_ARG, "device %d beyond the scratch table", dev);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch[dev]) {
    int cur = 0;
    HIPCHK(hipGetDevice(&cur));
    if (cur != dev) HIPCHK(hipSetDevice(dev));
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != dev) (void)hipSetDevice(cur);
    if (e != hipSuccess || !p) return fail(TDGL_E_DEVICE, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
    g_scratch[dev] = p;
  }
  *out = g_scratch[dev];
  return 0;
}
}  // namespace
extern "C" {

int tdgl_step_noalloc(tdgl_handle *h, int32_t n, double H, double HJ) {
  if (!h) return fail(TDGL_E_ARG, "null handle");
  if (h->prof) return fail(TDGL_E_ARG, "tdgl_step_noalloc: the per-kernel profile is on (it records events per launch)");
  const bool pcg = h->d.field_mode == TDGL_FIELD_PCG;
  if (pcg) {
    const bool f32 = h->d.pcg_f32 && (h->d.pcg_fast || h->d.pcg_f32);
    const int need = std::max({field_blocks(h->g), pcg_box_blocks(h->g, h->box_kc), sep_last_parts(h, f32),
                               sep_last_parts(h, false)});
    if (need > MAX_PART)
      return fail(TDGL_E_ARG, "tdgl_step_noalloc: the grid's reductions need %d partials, the fixed scratch holds %d",
                  need, MAX_PART);
  }
  void *scr = nullptr;
  CHK(scratch_on(h->dev, &scr));
  const PcgDev own = h->pd;
  if (pcg) {
    // the solver's scalars and flags in, every call: the array keeps the last call's contents
    PcgDev sp{};
    sp.part = (double *)scr;
    sp.sc = sp.part + SCRATCH_PART;
    sp.fl = (int *)(sp.sc + SC_COUNT);
    HIPCHK(hipMemcpyAsync(sp.sc, own.sc, SC_COUNT * sizeof(double), hipMemcpyDeviceToDevice, h->s));
    HIPCHK(hipMemcpyAsync(sp.fl, own.fl, FL_COUNT * sizeof(int), hipMemcpyDeviceToDevice, h->s));
    h->pd = sp;
  }
  // launched from the host: a captured graph would be instantiated (an allocation)
  const bool graphs = h->graphs;
  h->graphs = false;
  const int rc = tdgl_step(h, n, H, HJ);
  h->graphs = graphs;
  if (pcg) {
    const PcgDev sp = h->pd;
    h->pd = own;
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(own.sc, sp.sc, SC_COUNT * sizeof(double), hipMemcpyDeviceToDevice, h->s));
    HIPCHK(hipMemcpyAsync(own.fl, sp.fl, FL_COUNT * sizeof(int), hipMemcpyDeviceToDevice, h->s));
    HIPCHK(hipStreamSynchronize(h->s));
  }
  return rc;
}

// ---------------------------------------------------------------------------- the tuner
// Launch shapes chosen by timing them on the handle's own state: per candidate one time step,
// launched from the host with HIP events around every launch, from a device copy of the state (the
// roll-back buffers), which is put back after the last.  The PCG's box levels per block (box_kc)
// decide how the block partials of its reductions fall, the others only how the work is cut.
int tdgl_autotune(tdgl_handle *h, double H, double HJ, tdgl_tuning *out) {
  if (!h || !out) return fail(TDGL_E_ARG, "null argument");
  if (h->d.field_mode != TDGL_FIELD_PCG) return fail(TDGL_E_ARG, "the tuner times the PCG's kernels (no PCG here)");
  if (h->prof) return fail(TDGL_E_ARG, "the per-kernel profile is on");
  if (h->noise && h->d.noise_host) return fail(TDGL_E_ARG, "the tuner does not step on the caller's draws");
  if (!std::isfinite(H) || !std::isfinite(HJ)) return fail(TDGL_E_ARG, "non-finite H or HJ");
  *out = tdgl_tuning{};
  HIPCHK(hipStreamSynchronize(h->s));
  h->hph[0] = H;
  h->hph[1] = HJ;
  HIPCHK(hipMemcpyAsync(h->hp, h->hph, 2 * sizeof(double), hipMemcpyHostToDevice, h->s));
  const HostSnap start = snap(h);
  const int last_iters = h->last_iters;
  PL(K_VEC, launch_copies(backup_list(h, true), h->s));
  // what a step rewrites before reading it but the getters return (F of every band, the normal
  // fluids' potentials) goes to the host and back
  std::vector<std::pair<void *, std::vector<char>>> kept;
  for (int mu = 0; mu < h->d.nbands; ++mu) kept.push_back({(void *)h->F[mu], std::vector<char>(2 * sizeof(double) * h->g.n)});
  for (int k = 0; k < h->d.qp_nmodes && h->d.qp_mode != TDGL_QP_NONE; ++k)
    kept.push_back({(void *)h->qS[k], std::vector<char>(sizeof(double) * (size_t)h->qn)});
  for (auto &kv : kept) CHK(hcopy(h, kv.second.data(), kv.first, kv.second.size(), hipMemcpyDeviceToHost));
  // one step of a candidate; the class times of it (ms per launch of `cls`, and of the second)
  auto trial = [&](int kind, int value, int cls, int cls2, double *us) -> int {
    restore(h, start);
    PL(K_VEC, launch_copies(backup_list(h, false), h->s));
    h->last_iters = last_iters;
    PL(K_VEC, launch_pcg_cycle_reset(h->pd, h->s));
    CHK(tdgl_profile_enable(h, 1));
    const int rc = steps_eager(h, 1);
    tdgl_profile p{};
    const int rc2 = rc ? rc : tdgl_profile_get(h, &p);
    (void)tdgl_profile_enable(h, 0);
    if (rc2) return rc2;
    if (cls < 0) {
      // a whole PCG iteration: the kernels of the iterations (and the preconditioner's passes)
      // over the step, per iteration (one operator launch each)
      const double n = (double)std::max<int64_t>(1, p.launches[K_APPLYA]);
      *us = 1e3 * (p.ms[K_APPLYA] + p.ms[K_VEC] + p.ms[K_TX32] + p.ms[K_MODES] + p.ms[K_FOLD]) / n;
    } else {
      const double n = (double)std::max<int64_t>(1, p.launches[cls]);
      *us = 1e3 * (p.ms[cls] + (cls2 >= 0 ? p.ms[cls2] : 0.0)) / n;
    }
    if (out->ncand < TDGL_TUNE_MAX) out->cand[out->ncand++] = tdgl_tune_cand{kind, value, *us};
    return 0;
  };
  int rc = 0;
  // the PCG's levels per block: the operator and the vector update per iteration
  const int kc0 = h->box_kc;
  int best_kc = kc0;
  double best = 1e300;
  for (int kc = 1; kc <= 32 && rc == 0; kc *= 2) {
    if (kc > h->g.Nz - 1 || pcg_box_blocks(h->g, kc) > MAX_PART) continue;
    h->box_kc = kc;
    double us = 0;
    rc = trial(TDGL_TUNE_BOX_KC, kc, K_APPLYA, K_VEC, &us);
    if (rc == 0 && us < best) {
      best = us;
      best_kc = kc;
    }
  }
  h->box_kc = best_kc;
  // m tiles per block of the one-round x-y kernel, timed on the full-length path that launches it
  // (a folded handle's x-y launches are the quarter kernel's), unless a refinement keeps the folds
  const int mt0 = h->plane1;
  int best_mt = mt0;
  best = 1e300;
  const bool fold_in = h->fold;
  if (!h->refine) h->fold = false;
  for (int mt = 2; mt <= 3 && rc == 0; ++mt) {
    h->plane1 = mt;
    double us = 0;
    rc = trial(TDGL_TUNE_PLANE_MT, mt, K_TX32, -1, &us);
    if (rc == 0 && us < best) {
      best = us;
      best_mt = mt;
    }
  }
  h->fold = fold_in;
  h->plane1 = best_mt;
  // waves (groups of z modes) per block of the quarter z pass
  const int zw0 = h->zq_waves;
  int best_zw = zw0;
  if (h->fold) {
    best = 1e300;
    for (int zw : {8, 4, 0}) {
      if (rc) break;
      h->zq_waves = zw;
      double us = 0;
      rc = trial(TDGL_TUNE_ZMODES_WAVES, zw, K_MODES, -1, &us);
      if (rc == 0 && us < best) {
        best = us;
        best_zw = zw;
      }
    }
    h->zq_waves = best_zw;
  }
  // blocks a folded quarter plane's x-y transform is split over (its m tiles)
  const int qp0 = h->qparts;
  if (h->fold && h->qplane) {
    int best_qp = qp0;
    best = 1e300;
    for (int qp = 1; qp <= 2 && rc == 0; ++qp) {
      h->qparts = qp;
      double us = 0;
      rc = trial(TDGL_TUNE_QPARTS, qp, K_TX32, -1, &us);
      if (rc == 0 && us < best) {
        best = us;
        best_qp = qp;
      }
    }
    h->qparts = best_qp;
  }
  // the preconditioner's path (a whole PCG iteration each): the folds, or the full-length
  // transforms with the z pass over x-runs or whole columns; the folds stay where a refinement
  // needs them
  const bool fold0 = h->fold;
  const int zcol0 = h->zcols;
  bool best_fold = fold0;
  int best_zc = zcol0;
  if (h->d.pcg_f32 && rc == 0) {
    best = 1e300;
    struct Path {
      bool fold;
      int zcol;
    };
    std::vector<Path> paths;
    if (fold0) paths.push_back({true, zcol0});
    if (!h->refine) {
      paths.push_back({false, 0});
      paths.push_back({false, 1});
    }
    for (const Path &pa : paths) {
      if (rc) break;
      h->fold = pa.fold;
      h->zcols = pa.zcol;
      double us = 0;
      rc = pa.fold ? trial(TDGL_TUNE_FOLD, 1, -1, -1, &us) : trial(TDGL_TUNE_ZCOL, pa.zcol, -1, -1, &us);
      if (rc == 0 && us < best) {
        best = us;
        best_fold = pa.fold;
        best_zc = pa.zcol;
      }
    }
    h->fold = best_fold;
    h->zcols = best_zc;
  }
  // the state back where it was, whatever happened
  restore(h, start);
  h->last_iters = last_iters;
  const hipError_t e = launch_copies(backup_list(h, false), h->s);
  hipError_t e2 = hipStreamSynchronize(h->s);
  for (auto &kv : kept)
    if (e2 == hipSuccess) e2 = hipMemcpy(kv.first, kv.second.data(), kv.second.size(), hipMemcpyHostToDevice);
  CHK(destroy_graph(h));  // captured cycles hold the old shapes
  if (rc) {
    h->box_kc = kc0;
    h->plane1 = mt0;
    h->zq_waves = zw0;
    h->qparts = qp0;
    h->fold = fold0;
    h->zcols = zcol0;
    return rc;
  }
  if (e != hipSuccess || e2 != hipSuccess) return fail(TDGL_E_DEVICE, "the tuner's restore of the state");
  out->box_kc = h->box_kc;
  out->plane_mt = h->plane1;
  out->zmodes_waves = h->zq_waves;
  out->fold = h->fold ? 1 : 0;
  out->zcol = h->zcols;
  out->qparts = h->qparts;
  return 0;
}

int tdgl_get_pcg(tdgl_handle *h, tdgl_pcg *p) {
  if (!h || !p) return fail(TDGL_E_ARG, "null argument");
  if (h->d.field_mode != TDGL_FIELD_PCG) return fail(TDGL_E_ARG, "the stepper has no PCG solver");
  HIPCHK(hipStreamSynchronize(h->s));
  std::vector<double> node((size_t)h->g.n);
  for (int c = 0; c < 3; ++c) {
    double *dst[2] = {p->x[c], p->xprev[c]};
    const double *src[2] = {h->x.c[c], h->xprev.c[c]};
    for (int w = 0; w < 2; ++w) {
      if (!dst[w]) continue;
      CHK(hcopy(h, node.data(), src[w], node.size() * 8, hipMemcpyDeviceToHost));
      const int e0 = h->rg.e[c][0], e1 = h->rg.e[c][1], e2 = h->rg.e[c][2];
      for (int k = 0; k < e2; ++k)
        for (int j = 0; j < e1; ++j)
          for (int i = 0; i < e0; ++i) dst[w][i + (size_t)e0 * (j + (size_t)e1 * k)] = node[h->g.at(i + 2, j + 2, k + 2)];
    }
  }
  p->warm = h->warm;
  p->have2 = h->have2;
  p->iters = h->last_iters;
  return 0;
}

int tdgl_set_pcg(tdgl_handle *h, const tdgl_pcg *p) {
  if (!h || !p) return fail(TDGL_E_ARG, "null argument");
  if (h->d.field_mode != TDGL_FIELD_PCG) return fail(TDGL_E_ARG, "the stepper has no PCG solver");
  HIPCHK(hipStreamSynchronize(h->s));
  std::vector<double> node((size_t)h->g.n);
  for (int c = 0; c < 3; ++c) {
    const double *src[2] = {p->x[c], p->xprev[c]};
    double *dst[2] = {h->x.c[c], h->xprev.c[c]};
    for (int w = 0; w < 2; ++w) {
      if (!src[w]) continue;
      std::fill(node.begin(), node.end(), 0.0);
      const int e0 = h->rg.e[c][0], e1 = h->rg.e[c][1], e2 = h->rg.e[c][2];
      for (int k = 0; k < e2; ++k)
        for (int j = 0; j < e1; ++j)
          for (int i = 0; i < e0; ++i) node[h->g.at(i + 2, j + 2, k + 2)] = src[w][i + (size_t)e0 * (j + (size_t)e1 * k)];
      CHK(hcopy(h, dst[w], node.data(), node.size() * 8, hipMemcpyHostToDevice));
    }
  }
  h->warm = p->warm != 0;
  h->have2 = p->have2 != 0;
  h->last_iters = p->iters;
  return 0;
}

int tdgl_noise_get(tdgl_handle *h, uint64_t *counter) {
  if (!h || !counter) return fail(TDGL_E_ARG, "null argument");
  if (!h->noise) return fail(TDGL_E_ARG, "the stepper has no noise");
  HIPCHK(hipStreamSynchronize(h->s));
  uint64_t c[2];
  CHK(hcopy(h, c, h->nctr, sizeof c, hipMemcpyDeviceToHost));
  *counter = c[0];
  return 0;
}

int tdgl_noise_set(tdgl_handle *h, uint64_t counter) {
  if (!h) return fail(TDGL_E_ARG, "null handle");
  if (!h->noise) return fail(TDGL_E_ARG, "the stepper has no noise");
  HIPCHK(hipStreamSynchronize(h->s));
  CHK(hcopy(h, h->nctr, &counter, sizeof counter, hipMemcpyHostToDevice));
  h->nctr_h = counter;
  return 0;
}

int tdgl_noise_draws(tdgl_handle *h, int32_t nsteps, const double *draws) {
  if (!h) return fail(TDGL_E_ARG, "null handle");
  const tdgl_desc &d = h->d;
  if (!h->noise || !d.noise_host) return fail(TDGL_E_ARG, "the stepper takes no draws of the caller (noise_host)");
  if (nsteps < 0 || (nsteps > 0 && !draws)) return fail(TDGL_E_ARG, "bad draws");
  HIPCHK(hipStreamSynchronize(h->s));
  // the captured graph holds the old buffer's address
  CHK(destroy_graph(h));
  if (h->ndraw) {
    cache_give(h->ndraw_b);
    h->ndraw = nullptr;
  }
  const int n = h->g.n;
  const int Nx = d.Nx, Ny = d.Ny, Nz = d.Nz;
  const long long stride = (long long)d.nbands * 2 * n + (d.noise_field ? 3LL * n : 0);
  h->nstride = stride;
  h->nslots = nsteps;
  if (nsteps > 0) {
    if (dblock(h, (size_t)stride * nsteps * sizeof(double), &h->ndraw_b) != hipSuccess)
      return fail(TDGL_E_NOMEM, "hipMalloc of the noise draws failed");
    h->ndraw = (double *)h->ndraw_b.p;
    CHK(hfill0(h, h->ndraw, (size_t)stride * nsteps * sizeof(double)));
    const size_t nn = (size_t)(Nx + 1) * (Ny + 1) * (Nz + 1);
    const int sh[3][3] = {{Nx, Ny + 1, Nz + 1}, {Nx + 1, Ny, Nz + 1}, {Nx + 1, Ny + 1, Nz}};
    const double *p = draws;
    for (int t = 0; t < nsteps; ++t) {
      double *slot = h->ndraw + (size_t)t * stride;
      for (int mu = 0; mu < d.nbands; ++mu, p += 2 * nn) CHK(put(h, slot + (size_t)mu * 2 * n, p, Nx + 1, Ny + 1, Nz + 1, 16));
      if (d.noise_field)
        for (int c = 0; c < 3; ++c) {
          CHK(put(h, slot + (size_t)d.nbands * 2 * n + (size_t)c * n, p, sh[c][0], sh[c][1], sh[c][2], 8));
          p += (size_t)sh[c][0] * sh[c][1] * sh[c][2];
        }
    }
  }
  // the first of these draws is for the next step
  uint64_t c[2];
  CHK(hcopy(h, c, h->nctr, sizeof c, hipMemcpyDeviceToHost));
  c[1] = c[0];
  CHK(hcopy(h, h->nctr, c, sizeof c, hipMemcpyHostToDevice));
  h->nctr_h = c[0];
  h->nbase_h = c[0];
  return 0;
}

int tdgl_get_open(tdgl_handle *h, tdgl_open *o) {
  if (!h || !o) return fail(TDGL_E_ARG, "null handle or open state");
  if (h->d.selffield_bc != TDGL_BC_OPEN) return fail(TDGL_E_ARG, "the stepper has no open boundary");
  const tdgl_desc &d = h->d;
  const int sh[3][3] = {{d.Nx + 1, d.Ny, d.Nz}, {d.Nx, d.Ny + 1, d.Nz}, {d.Nx, d.Ny, d.Nz + 1}};
  for (int c = 0; c < 3; ++c) {
    if (!o->hold[c] || !o->hnew[c]) return fail(TDGL_E_ARG, "null open-boundary array");
    CHK(get(h, o->hold[c], h->ohold[c], sh[c][0], sh[c][1], sh[c][2], 8));
    CHK(get(h, o->hnew[c], h->ohnew[c], sh[c][0], sh[c][1], sh[c][2], 8));
  }
  int fl[4];
  CHK(hcopy(h, fl, h->ofl, sizeof fl, hipMemcpyDeviceToHost));
  o->count = fl[OF_COUNT];
  return 0;
}

int tdgl_set_open(tdgl_handle *h, const tdgl_open *o) {
  if (!h || !o) return fail(TDGL_E_ARG, "null handle or open state");
  if (h->d.selffield_bc != TDGL_BC_OPEN) return fail(TDGL_E_ARG, "the stepper has no open boundary");
  const tdgl_desc &d = h->d;
  const int sh[3][3] = {{d.Nx + 1, d.Ny, d.Nz}, {d.Nx, d.Ny + 1, d.Nz}, {d.Nx, d.Ny, d.Nz + 1}};
  for (int c = 0; c < 3; ++c) {
    if (!o->hold[c] || !o->hnew[c]) return fail(TDGL_E_ARG, "null open-boundary array");
    CHK(put(h, h->ohold[c], o->hold[c], sh[c][0], sh[c][1], sh[c][2], 8));
    CHK(put(h, h->ohnew[c], o->hnew[c], sh[c][0], sh[c][1], sh[c][2], 8));
  }
  const int fl[4] = {o->count, 0, 0, 0};
  CHK(hcopy(h, h->ofl, fl, sizeof fl, hipMemcpyHostToDevice));
  return 0;
}

int tdgl_get_qp(tdgl_handle *h, tdgl_qp *q) {
  if (!h || !q) return fail(TDGL_E_ARG, "null handle or normal-fluid potentials");
  const tdgl_desc &d = h->d;
  if (d.qp_mode == TDGL_QP_NONE) return fail(TDGL_E_ARG, "the stepper has one normal fluid (no potentials)");
  const int npot = d.qp_mode == TDGL_QP_TWO ? 1 : d.nbands;
  for (int p = 0; p < npot; ++p)
    if (!q->pot[p]) return fail(TDGL_E_ARG, "null potential array %d", p);
  const int K = d.qp_nmodes, n1 = d.Nx - 1, nj = d.qp_nj, nk = d.qp_nk;
  const size_t nr = (size_t)h->qn;
  std::vector<double> M((size_t)K * nr);
  for (int k = 0; k < K; ++k) CHK(hcopy(h, M.data() + k * nr, h->qS[k], nr * sizeof(double), hipMemcpyDeviceToHost));
  const size_t NX = (size_t)d.Nx + 1, NY = (size_t)d.Ny + 1, NZ = (size_t)d.Nz + 1;
  for (int p = 0; p < npot; ++p) {
    double *pot = q->pot[p];
    std::fill(pot, pot + NX * NY * NZ, 0.0);
    for (int kz = 0; kz < nk; ++kz)
      for (int jy = 0; jy < nj; ++jy)
        for (int ix = 0; ix < n1; ++ix) {
          const size_t r = (size_t)ix + (size_t)n1 * (jy + (size_t)nj * kz);
          double v;
          if (d.qp_mode == TDGL_QP_TWO) {
            v = M[r];
          } else {
            v = 0.0;  // μ_µ = Σ_k P[µ][k] m_k, in the order of exchange.jl:506-510
            for (int kk = 0; kk < K; ++kk) v += d.qp_P[p][kk] * M[kk * nr + r];
          }
          pot[(size_t)(ix + 1) + NX * ((size_t)(jy + d.qp_j0) + NY * (size_t)(kz + d.qp_k0))] = v;
        }
  }
  return 0;
}

// ---------------------------------------------------------------------------- the ladder's meter
extern "C++" {
namespace {
// the meter's arguments with the state's current arrays (the advance swaps them)
MeterArgs meter_args(tdgl_handle *h, int first) {
  MeterArgs a = h->ma;
  for (int mu = 0; mu < h->d.nbands; ++mu) a.G[mu] = h->G[mu];
  a.Vx = h->V[0];
  a.phx = h->ph[0];
  a.hz = h->h[2];
  a.hy = h->h[1];
  a.first = first ? 1 : 0;
  a.init = 0;
  return a;
}
// the record of a cycle from the device's MR_N values
void to_record(const double *r, tdgl_cycle_record *out) {
  *out = tdgl_cycle_record{};
  out->V = r[MR_V];
  out->TA = r[MR_TA];
  out->Tf1 = r[MR_TF1];
  for (int mu = 0; mu < TDGL_MAX_BANDS; ++mu) {
    out->T[mu] = r[MR_T + mu];
    out->Tabs[mu] = r[MR_TABS + mu];
    out->dif[mu] = r[MR_DIF + mu];
  }
  out->dmax = r[MR_DMAX];
  out->nbig = (int32_t)r[MR_NBIG];
  out->It = r[MR_IT];
  out->Is = r[MR_IS];
  out->amax = r[MR_AMAX];
  out->diffh = r[MR_DIFFH];
}
template <class T>
int malloc_meter(tdgl_handle *h, T **p, size_t n) {
  Block b{};
  if (dblock(h, (n ? n : 1) * sizeof(T), &b) != hipSuccess) return fail(TDGL_E_NOMEM, "hipMalloc of the meter failed");
  h->mallocs.push_back(b);
  *p = (T *)b.p;
  return 0;
}
}  // namespace
}  // extern "C++"

int tdgl_meter_begin(tdgl_handle *h, const tdgl_meter *m) {
  if (!h || !m || !m->Min || !m->Mou || !m->line_ok || !m->line_f1) return fail(TDGL_E_ARG, "null handle or meter");
  const tdgl_desc &d = h->d;
  const Geo &g = h->g;
  if (m->vi < d.x1 || m->vo > d.x2 || m->vi >= m->vo) return fail(TDGL_E_ARG, "voltmeter planes %d, %d outside x1..x2", m->vi, m->vo);
  HIPCHK(hipStreamSynchronize(h->s));
  for (const Block &b : h->mallocs) cache_give(b);
  h->mallocs.clear();
  h->meter = false;
  MeterArgs a{};
  a.g = g;
  a.N = d.nbands;
  a.vi = m->vi;
  a.vo = m->vo;
  a.nl1 = d.y2 - d.y1 + 2;
  a.nl2 = d.z2 - d.z1 + 2;
  const int nl = a.nl1 * a.nl2;
  for (int q = 0; q < nl; ++q) {
    a.nin += m->Min[q] != 0;
    a.nou += m->Mou[q] != 0;
    a.nlin += m->line_ok[q] != 0;
    a.nlin_f1 += m->line_f1[q] != 0;
  }
  if (a.nlin == 0) return fail(TDGL_E_ARG, "no complete line between the voltmeter planes");
  a.mid = (d.Nx + 1) / 2;  // oround(Nx / 2)
  a.ks2 = m->kappa2;
  a.dy = d.dy;
  a.dz = d.dz;
  const int nb = (d.x2 - d.x1 + 2) * nl;
  a.nlb = cdiv(nl, 4);
  a.nnb = cdiv(nb, 256);
  a.nhb = cdiv(d.Nx * d.Ny * (d.Nz + 1), 256);
  unsigned char *mk[4];
  const uint8_t *src[4] = {m->Min, m->Mou, m->line_ok, m->line_f1};
  // the masks staged and queued, the fills queued: one wait at the end
  for (int q = 0; q < 4; ++q) {
    CHK(malloc_meter(h, &mk[q], (size_t)nl));
    CHK(upload_vec(h, mk[q], std::vector<unsigned char>(src[q], src[q] + nl)));
  }
  a.Min = mk[0];
  a.Mou = mk[1];
  a.lok = mk[2];
  a.lf1 = mk[3];
  for (int mu = 0; mu < d.nbands; ++mu) {
    CHK(malloc_meter(h, &a.Dprev[mu], (size_t)nl));
    CHK(malloc_meter(h, &a.T[mu], (size_t)nl));
    CHK(malloc_meter(h, &a.cooper[mu], (size_t)nb));
    HIPCHK(hipMemsetAsync(a.Dprev[mu], 0, nl * sizeof(double), h->s));
    HIPCHK(hipMemsetAsync(a.T[mu], 0, nl * sizeof(double), h->s));
  }
  CHK(malloc_meter(h, &a.hold, (size_t)g.n));
  CHK(malloc_meter(h, &a.part, (size_t)meter_blocks(a) * METER_Q));
  CHK(malloc_meter(h, &h->mrec, (size_t)MR_N));
  CHK(malloc_meter(h, &h->mrung, 2));
  HIPCHK(hipMemsetAsync(h->mrung, 0, 2 * sizeof(double), h->s));
  h->ma = a;
  h->meter = true;
  // Cooper0 and fieldOld of the current state
  MeterArgs b = meter_args(h, 0);
  b.init = 1;
  LCHK(launch_meter(b, h->s));
  return flush(h);
}

int tdgl_meter_cycle(tdgl_handle *h, int32_t first, tdgl_cycle_record *out) {
  if (!h || !out) return fail(TDGL_E_ARG, "null handle or record");
  if (!h->meter) return fail(TDGL_E_ARG, "tdgl_meter_begin has not been called");
  const MeterArgs a = meter_args(h, first);
  PL(K_METER, launch_meter(a, h->s));
  PL(K_METER, launch_meter_finish(a, h->mrec, h->mrung, h->s));
  double r[MR_N];
  HIPCHK(hipMemcpyAsync(r, h->mrec, sizeof r, hipMemcpyDeviceToHost, h->s));
  HIPCHK(hipStreamSynchronize(h->s));
  if (h->prof) CHK(prof_flush(h));
  to_record(r, out);
  return 0;
}

int tdgl_get_stats(tdgl_handle *h, tdgl_stats *st) {
  if (!h || !st) return fail(TDGL_E_ARG, "null argument");
  st->graph_cycles = h->graph_cycles;
  st->rollbacks = h->rollbacks;
  st->pcg_budget = h->budget;
  if (h->stepwise) {  // the largest budget of the classes that have run (the single one is not used)
    int m = 0;
    for (int c = 0; c < PCG_NCLS; ++c)
      if (h->cseen[c]) m = std::max(m, h->cbudget[c]);
    if (m > 0) st->pcg_budget = m;
  }
  st->captures = h->captures;
  st->folded = (h->fold || h->fold64) ? 1 : 0;
  st->steps = h->steps;
  st->calls = h->calls;
  st->pcg_iters = h->pcg_iters;
  st->last_iters = h->last_iters;
  st->device_ms = h->device_ms;
  st->device_bytes = h->bytes;
  return 0;
}

int tdgl_profile_enable(tdgl_handle *h, int32_t on) {
  if (!h) return fail(TDGL_E_ARG, "null handle");
  CHK(prof_flush(h));
  if (on && !h->prof)
    for (int k = 0; k < K_COUNT; ++k) {
      h->launches[k] = 0;
      h->kms[k] = 0;
    }
  h->prof = on != 0;
  return 0;
}

int tdgl_profile_get(tdgl_handle *h, tdgl_profile *p) {
  if (!h || !p) return fail(TDGL_E_ARG, "null argument");
  CHK(prof_flush(h));
  for (int k = 0; k < K_COUNT; ++k) {
    p->launches[k] = h->launches[k];
    p->ms[k] = h->kms[k];
  }
  return 0;
}

const char *tdgl_kernel_name(int32_t k) { return (k >= 0 && k < K_COUNT) ? K_NAMES[k] : ""; }

}  // extern "C"

// ---------------------------------------------------------------------------- batched runs
// one batched launch: its kind, blocks, runs and where its tables sit in the batch's block
struct LaunchPlan {
  int kind, variant, blocks, n;
  size_t lds;
  size_t off, dims, g, args;
  int pos;  // the position in the step sequence: the instances of one position are independent
};
// The steps of a call are recorded once per run (the launches its own step would make, with the
// single-axis passes of the separable solve), checked to be the same sequence of kernels in
// every run, laid out as one table per launch (block offsets, each run's grid, Geo and argument
// struct) and captured as a graph of one batched launch per position; the graph is replayed
// while the steps per call and every run's parity stay the same.
struct tdgl_batch {
  std::vector<tdgl_handle *> hs;
  int dev = 0;
  hipStream_t s = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  int gn = -1;                 // steps of the captured graph
  std::vector<int> gpar;       // the parity of each run it was captured from
  std::vector<HostSnap> gend;  // each run's pointers after it
  Block tab{};                 // device tables of the graph's launches
  std::vector<LaunchPlan> plans;  // the launches of gn steps (replayed eagerly under TDGL_HIP_GRAPH=0)
  bool graphs = true;
  bool tx_only = false;           // the runs differ in their transform passes: single-axis ones in all
  Block mtab{}, mstage{};      // device tables of the meter's launches, their pinned staging
  Block dH{}, hH{};            // (H, HJ) of every run: device, pinned host
  Block drec{}, hrec{};        // the meter's records: device, pinned host
  // the graph's parallel branches: the kernel instances of one position run side by side, each
  // on a side stream that forks from s at an event and joins it again (default priority)
  std::vector<hipStream_t> side;
  std::vector<hipEvent_t> join;
  hipEvent_t fork = nullptr;
};

namespace {

struct TabBuild {
  std::vector<unsigned char> bytes;
  size_t put(const void *p, size_t n) {
    const size_t o = (bytes.size() + 255) / 256 * 256;
    bytes.resize(o + n);
    std::memcpy(&bytes[o], p, n);
    return o;
  }
};

// the table of one batched launch of `kind` (instance `variant`, dynamic LDS `lds`): run r with
// Geo gs[r], argument struct args[r] and the grid grids[r] of its own launch
int plan_launch(int kind, int variant, size_t lds, const std::vector<Geo> &gs, const std::vector<const void *> &args,
                const std::vector<dim3> &grids, TabBuild &tb, LaunchPlan &lp) {
  const int n = (int)gs.size();
  const size_t es = batch_args_size(kind);
  if (es == 0) return fail(TDGL_E_ARG, "batch: unknown launch kind %d", kind);
  std::vector<int> off(n + 1, 0), dims(3 * n, 1);
  std::vector<unsigned char> ab((size_t)n * es);
  for (int r = 0; r < n; ++r) {
    const dim3 gd = grids[r];
    const long long nb = (long long)gd.x * gd.y * gd.z;
    dims[3 * r] = gd.x > 0 ? (int)gd.x : 1;
    dims[3 * r + 1] = gd.y > 0 ? (int)gd.y : 1;
    dims[3 * r + 2] = gd.z > 0 ? (int)gd.z : 1;
    if ((long long)off[r] + nb > 0x7fffffffLL) return fail(TDGL_E_ARG, "batch: too many blocks in one launch");
    off[r + 1] = off[r] + (int)nb;
    std::memcpy(&ab[(size_t)r * es], args[r], es);
  }
  lp.kind = kind;
  lp.variant = variant;
  lp.lds = lds;
  lp.n = n;
  lp.blocks = off[n];
  lp.off = tb.put(off.data(), off.size() * sizeof(int));
  lp.dims = tb.put(dims.data(), dims.size() * sizeof(int));
  lp.g = tb.put(gs.data(), gs.size() * sizeof(Geo));
  lp.args = tb.put(ab.data(), ab.size());
  return 0;
}
BatchTab tab_of(const Block &dev, const LaunchPlan &lp) {
  const unsigned char *base = (const unsigned char *)dev.p;
  return BatchTab{(const int *)(base + lp.off), (const int *)(base + lp.dims), (const Geo *)(base + lp.g),
                  (const void *)(base + lp.args), lp.n};
}
// the tables into a device block of the batch (grown as needed) through pinned staging
int upload_tabs(tdgl_batch *b, const TabBuild &tb, Block &dst, Block &stage) {
  const size_t n = tb.bytes.size() ? tb.bytes.size() : 1;
  if (!dst.p || dst.cap < n) {
    if (dst.p) cache_give(dst);
    dst = Block{};
    HIPCHK(dblock_on(b->dev, n, &dst));
  }
  if (!stage.p || stage.cap < n) {
    if (stage.p) cache_give(stage);
    stage = Block{};
    HIPCHK(hblock(n, &stage));
  }
  std::memcpy(stage.p, tb.bytes.data(), tb.bytes.size());
  HIPCHK(hipMemcpyAsync(dst.p, stage.p, tb.bytes.size(), hipMemcpyHostToDevice, b->s));
  HIPCHK(hipStreamSynchronize(b->s));
  return 0;
}

void batch_drop_graph(tdgl_batch *b) {
  if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->gexec = nullptr;
  b->graph = nullptr;
  b->gn = -1;
}

// the graph of m steps of every run, recorded and captured unless the one there fits
int batch_graph(tdgl_batch *b, int m) {
  const int R = (int)b->hs.size();
  std::vector<int> par(R);
  for (int r = 0; r < R; ++r) par[r] = b->hs[r]->parity;
  if (b->gn == m && b->gpar == par && (b->gexec || !b->graphs)) return 0;
  batch_drop_graph(b);
  std::vector<Recorder> rec(R);
  b->gend.assign(R, HostSnap{});
  for (int r = 0; r < R; ++r) {
    tdgl_handle *h = b->hs[r];
    const HostSnap start = snap(h);
    h->tx_only = b->tx_only;
    g_rec = &rec[r];
    const int rc = steps_eager(h, m);
    g_rec = nullptr;
    h->tx_only = false;
    b->gend[r] = snap(h);
    restore(h, start);
    if (rc) return fail(rc, "batch: run %d cannot be batched: %s", r, g_err.c_str());
  }
  const size_t P = rec[0].recs.size();
  for (int r = 1; r < R; ++r) {
    if (rec[r].recs.size() != P) return fail(TDGL_E_ARG, "batch: run %d makes %zu launches a call, run 0 %zu", r, rec[r].recs.size(), P);
    for (size_t q = 0; q < P; ++q)
      if (rec[r].recs[q].kind != rec[0].recs[q].kind) return fail(TDGL_E_ARG, "batch: run %d differs from run 0 at launch %zu", r, q);
  }
  // one launch per position and kernel instance (the plane and z-mode passes have several)
  TabBuild tb;
  std::vector<LaunchPlan> plans;
  std::vector<CurlArgs> curl(R);
  for (size_t q = 0; q < P; ++q) {
    const int kind = rec[0].recs[q].kind;
    std::vector<int> variants;
    for (int r = 0; r < R; ++r) {
      const BatchRec &br = rec[r].recs[q];
      if (br.args.size() != batch_args_size(kind)) return fail(TDGL_E_ARG, "batch: launch %zu has the wrong argument size", q);
      if (std::find(variants.begin(), variants.end(), br.variant) == variants.end()) variants.push_back(br.variant);
    }
    for (int v : variants) {
      std::vector<Geo> gs;
      std::vector<const void *> args;
      std::vector<dim3> grids;
      size_t lds = 0;
      for (int r = 0; r < R; ++r) {
        const BatchRec &br = rec[r].recs[q];
        if (br.variant != v) continue;
        const void *a = br.args.data();
        if (kind == BK_CURL) {
          // the field boundary reads the run's (H, HJ) from the batch's array
          std::memcpy(&curl[r], br.args.data(), sizeof(CurlArgs));
          curl[r].hp = (double *)b->dH.p + 2 * r;
          a = &curl[r];
        }
        gs.push_back(br.g);
        args.push_back(a);
        grids.push_back(br.has_grid ? br.grid : batch_grid(kind, br.g, a));
        lds = std::max(lds, br.lds);
      }
      LaunchPlan lp{};
      CHK(plan_launch(kind, v, lds, gs, args, grids, tb, lp));
      lp.pos = (int)q;
      plans.push_back(lp);
    }
  }
  Block stage{};
  const int rc = upload_tabs(b, tb, b->tab, stage);
  if (stage.p) cache_give(stage);
  if (rc) return rc;
  b->plans = plans;
  if (!b->graphs) {
    b->gn = m;
    b->gpar = par;
    return 0;
  }
  // side streams and events for the widest position (TDGL_HIP_BATCH_FORK=0: one chain)
  const char *bf = std::getenv("TDGL_HIP_BATCH_FORK");
  const bool forks = !(bf && bf[0] == '0');
  size_t width = 1;
  for (size_t q = 0; q < plans.size();) {
    size_t e = q;
    while (e < plans.size() && plans[e].pos == plans[q].pos) ++e;
    width = std::max(width, e - q);
    q = e;
  }
  if (forks) {
    if (!b->fork) HIPCHK(hipEventCreateWithFlags(&b->fork, hipEventDisableTiming));
    while (b->side.size() + 1 < width) {
      hipStream_t st = nullptr;
      hipEvent_t ev = nullptr;
      HIPCHK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
      b->side.push_back(st);
      HIPCHK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
      b->join.push_back(ev);
    }
  }
  HIPCHK(hipStreamBeginCapture(b->s, hipStreamCaptureModeRelaxed));
  hipError_t le = hipSuccess;
  for (size_t q = 0; q < plans.size() && le == hipSuccess;) {
    size_t e = q;
    while (e < plans.size() && plans[e].pos == plans[q].pos) ++e;
    if (!forks || e - q == 1) {
      for (; q < e && le == hipSuccess; ++q)
        le = launch_batch(plans[q].kind, plans[q].variant, plans[q].lds, tab_of(b->tab, plans[q]), plans[q].blocks, b->s);
      continue;
    }
    // fork: instance k >= 1 on side stream k - 1 after the fork event, the first on s; join
    le = hipEventRecord(b->fork, b->s);
    for (size_t k = 1; k < e - q && le == hipSuccess; ++k) {
      const LaunchPlan &lp = plans[q + k];
      hipStream_t st = b->side[k - 1];
      le = hipStreamWaitEvent(st, b->fork, 0);
      if (le == hipSuccess) le = launch_batch(lp.kind, lp.variant, lp.lds, tab_of(b->tab, lp), lp.blocks, st);
      if (le == hipSuccess) le = hipEventRecord(b->join[k - 1], st);
    }
    if (le == hipSuccess)
      le = launch_batch(plans[q].kind, plans[q].variant, plans[q].lds, tab_of(b->tab, plans[q]), plans[q].blocks, b->s);
    for (size_t k = 1; k < e - q && le == hipSuccess; ++k) le = hipStreamWaitEvent(b->s, b->join[k - 1], 0);
    q = e;
  }
  hipGraph_t gph = nullptr;
  const hipError_t ee = hipStreamEndCapture(b->s, &gph);
  if (le != hipSuccess || ee != hipSuccess) {
    if (gph) (void)hipGraphDestroy(gph);
    return fail(TDGL_E_DEVICE, "batch: capture: %s", hipGetErrorString(le != hipSuccess ? le : ee));
  }
  b->graph = gph;
  HIPCHK(hipGraphInstantiate(&b->gexec, b->graph, nullptr, nullptr, 0));
  b->gn = m;
  b->gpar = par;
  return 0;
}

}  // namespace

extern "C" {

int tdgl_batch_create(tdgl_handle *const *hs, int32_t n, tdgl_batch **out) {
  if (!hs || !out || n < 1 || n > TDGL_BATCH_MAX) return fail(TDGL_E_ARG, "batch: 1..%d handles", TDGL_BATCH_MAX);
  *out = nullptr;
  for (int r = 0; r < n; ++r) {
    const tdgl_handle *h = hs[r];
    if (!h) return fail(TDGL_E_ARG, "batch: null handle %d", r);
    for (int q = 0; q < r; ++q)
      if (hs[q] == h) return fail(TDGL_E_ARG, "batch: handle %d given twice", r);
    if (h->dev != hs[0]->dev) return fail(TDGL_E_ARG, "batch: the handles are on different devices");
    const tdgl_desc &d = h->d;
    if (d.field_mode != TDGL_FIELD_SEPARABLE) return fail(TDGL_E_UNSUPPORTED, "batch: run %d: only the separable field is batched", r);
    if (d.selffield_bc == TDGL_BC_OPEN) return fail(TDGL_E_UNSUPPORTED, "batch: run %d: the open boundary is not batched", r);
    if (d.qp_mode != TDGL_QP_NONE) return fail(TDGL_E_UNSUPPORTED, "batch: run %d: normal-fluid exchange is not batched", r);
    if (h->noise) return fail(TDGL_E_UNSUPPORTED, "batch: run %d: thermal noise is not batched", r);
  }
  tdgl_batch *b = new tdgl_batch();
  b->hs.assign(hs, hs + n);
  b->dev = hs[0]->dev;
  const char *ge = std::getenv("TDGL_HIP_GRAPH");  // =0: the batched launches from the host (profiling)
  b->graphs = !(ge && ge[0] == '0');
  for (int r = 1; r < n; ++r) b->tx_only |= hs[r]->plane64 != hs[0]->plane64 || hs[r]->fold64 != hs[0]->fold64;
  auto bail = [&](int rc) {
    tdgl_batch_destroy(b);
    return rc;
  };
  if (hipStreamCreateWithFlags(&b->s, hipStreamNonBlocking) != hipSuccess) return bail(fail(TDGL_E_DEVICE, "batch: stream"));
  if (dblock_on(b->dev, 2 * sizeof(double) * n, &b->dH) != hipSuccess || hblock(2 * sizeof(double) * n, &b->hH) != hipSuccess ||
      dblock_on(b->dev, MR_N * sizeof(double) * n, &b->drec) != hipSuccess || hblock(MR_N * sizeof(double) * n, &b->hrec) != hipSuccess)
    return bail(fail(TDGL_E_NOMEM, "batch: allocation"));
  *out = b;
  return 0;
}

int tdgl_batch_step(tdgl_batch *b, int32_t n, const double *H, const double *HJ) {
  if (!b) return fail(TDGL_E_ARG, "null batch");
  if (n < 0) return fail(TDGL_E_ARG, "negative step count");
  if (n == 0) return 0;
  if (!H || !HJ) return fail(TDGL_E_ARG, "batch: null H or HJ");
  const int R = (int)b->hs.size();
  double *hh = (double *)b->hH.p;
  for (int r = 0; r < R; ++r) {
    if (!std::isfinite(H[r]) || !std::isfinite(HJ[r])) return fail(TDGL_E_ARG, "batch: non-finite H or HJ of run %d", r);
    if (b->hs[r]->prof) return fail(TDGL_E_ARG, "batch: run %d has its profile on", r);
    hh[2 * r] = H[r];
    hh[2 * r + 1] = HJ[r];
  }
  HIPCHK(hipMemcpyAsync(b->dH.p, hh, 2 * sizeof(double) * R, hipMemcpyHostToDevice, b->s));
  for (int left = n; left > 0;) {
    const int m = std::min(left, GRAPH_MAX_STEPS);
    CHK(batch_graph(b, m));
    if (b->graphs) {
      HIPCHK(hipGraphLaunch(b->gexec, b->s));
    } else {
      for (const LaunchPlan &lp : b->plans) HIPCHK(launch_batch(lp.kind, lp.variant, lp.lds, tab_of(b->tab, lp), lp.blocks, b->s));
    }
    for (int r = 0; r < R; ++r) restore(b->hs[r], b->gend[r]);
    left -= m;
  }
  HIPCHK(hipStreamSynchronize(b->s));
  for (tdgl_handle *h : b->hs) {
    h->stepped = true;
    h->steps += n;
    h->calls += 1;
  }
  return 0;
}

int tdgl_batch_meter_cycle(tdgl_batch *b, int32_t first, tdgl_cycle_record *out) {
  if (!b || !out) return fail(TDGL_E_ARG, "null batch or records");
  const int R = (int)b->hs.size();
  std::vector<Geo> gs(R);
  std::vector<MeterArgs> ma(R);
  std::vector<MeterFinArgs> mf(R);
  std::vector<const void *> pa(R), pf(R);
  for (int r = 0; r < R; ++r) {
    tdgl_handle *h = b->hs[r];
    if (!h->meter) return fail(TDGL_E_ARG, "batch: run %d: tdgl_meter_begin has not been called", r);
    ma[r] = meter_args(h, first);
    mf[r] = MeterFinArgs{ma[r], (double *)b->drec.p + (size_t)MR_N * r, h->mrung};
    gs[r] = h->g;
    pa[r] = &ma[r];
    pf[r] = &mf[r];
  }
  std::vector<dim3> gm(R), gf(R, dim3(1));
  for (int r = 0; r < R; ++r) gm[r] = meter_batch_grid(BK_METER, &ma[r]);
  TabBuild tb;
  LaunchPlan lm{}, lf{};
  CHK(plan_launch(BK_METER, 0, 0, gs, pa, gm, tb, lm));
  CHK(plan_launch(BK_METERFIN, 0, 0, gs, pf, gf, tb, lf));
  CHK(upload_tabs(b, tb, b->mtab, b->mstage));
  HIPCHK(launch_batch(BK_METER, 0, 0, tab_of(b->mtab, lm), lm.blocks, b->s));
  HIPCHK(launch_batch(BK_METERFIN, 0, 0, tab_of(b->mtab, lf), lf.blocks, b->s));
  HIPCHK(hipMemcpyAsync(b->hrec.p, b->drec.p, MR_N * sizeof(double) * R, hipMemcpyDeviceToHost, b->s));
  HIPCHK(hipStreamSynchronize(b->s));
  for (int r = 0; r < R; ++r) to_record((const double *)b->hrec.p + (size_t)MR_N * r, &out[r]);
  return 0;
}

int tdgl_batch_destroy(tdgl_batch *b) {
  if (!b) return 0;
  if (b->s) (void)hipStreamSynchronize(b->s);
  batch_drop_graph(b);
  for (Block *x : {&b->tab, &b->mtab, &b->mstage, &b->dH, &b->hH, &b->drec, &b->hrec})
    if (x->p) cache_give(*x);
  for (hipStream_t st : b->side) (void)hipStreamDestroy(st);
  for (hipEvent_t ev : b->join) (void)hipEventDestroy(ev);
  if (b->fork) (void)hipEventDestroy(b->fork);
  if (b->s) (void)hipStreamDestroy(b->s);
  delete b;
  return 0;
}

}  // extern "C"
