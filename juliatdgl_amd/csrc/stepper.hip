// stepper.hip — the C ABI of include/tdglhip.h: device state, the order of one step
// (gtdgl/src/step.jl:13-35) and of the band update (bands.jl:330-521), the separable solve
// and the device-resident PCG (implicit.jl:62-104, 279-446).
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
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

}  // namespace

struct tdgl_handle {
  tdgl_desc d;
  Geo g;
  Ranges rg;
  int dev = 0;
  hipStream_t s = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  std::vector<void *> allocs;
  int64_t bytes = 0;
  // state (state.jl:51-60); advance! swaps G/F, θ/φ, V/U
  cplx *G[TDGL_MAX_BANDS] = {}, *F[TDGL_MAX_BANDS] = {};
  double *th[3] = {}, *ph[3] = {};
  cplx *U[3] = {}, *V[3] = {};
  double *J[3] = {}, *h[3] = {};
  // workspace
  cplx *W = nullptr, *W2 = nullptr, *L[3] = {}, *Q[3] = {};
  double *Om = nullptr, *R[3] = {};
  // masks and tables
  double *f = nullptr, *tau[3] = {}, *D[3] = {};
  double *bar[6] = {};
  // separable solve: per axis, forward/inverse A for V and U, double and float; s
  double *Ad[3][2][2] = {};  // [axis][0=V,1=U][0=fwd,1=inv]
  float *Af[3][2][2] = {};
  double *sd[3] = {};
  float *sf[3] = {};
  double *B = nullptr, *T1 = nullptr, *T2 = nullptr;
  float *T32a = nullptr, *T32b = nullptr;
  // PCG
  double *x = nullptr, *r = nullptr, *z = nullptr, *P = nullptr, *Qv = nullptr, *xprev = nullptr, *zold = nullptr;
  PcgDev pd{};
  bool warm = false, have2 = false;
  int last_iters = 0;
  // counters
  int64_t steps = 0, calls = 0, pcg_iters = 0;
  double device_ms = 0;
  // host staging
  std::vector<unsigned char> stage;
};

namespace {

template <class T>
int dalloc(tdgl_handle *h, T **p, size_t n) {
  void *q = nullptr;
  if (n == 0) n = 1;
  if (hipMalloc(&q, n * sizeof(T)) != hipSuccess) return fail(TDGL_E_NOMEM, "hipMalloc of %zu bytes failed", n * sizeof(T));
  if (hipMemset(q, 0, n * sizeof(T)) != hipSuccess) return fail(TDGL_E_DEVICE, "hipMemset failed");
  h->allocs.push_back(q);
  h->bytes += (int64_t)(n * sizeof(T));
  *p = (T *)q;
  return 0;
}

#define ALLOC(ptr, n)                        \
  do {                                       \
    int rc_ = dalloc(h, &(ptr), (size_t)(n)); \
    if (rc_) return rc_;                     \
  } while (0)

template <class T>
int upload_vec(T *dst, const std::vector<T> &v) {
  HIPCHK(hipMemcpy(dst, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
  return 0;
}

// Julia-shaped (n1, n2, n3) host array of `es`-byte elements <-> node-shaped device array
int put(tdgl_handle *h, void *dev, const void *host, int n1, int n2, int n3, size_t es) {
  const Geo &g = h->g;
  h->stage.assign((size_t)g.n * es, 0);
  const unsigned char *src = (const unsigned char *)host;
  for (int k = 0; k < n3; ++k)
    for (int j = 0; j < n2; ++j)
      std::memcpy(&h->stage[((size_t)g.px * (j + (size_t)g.py * k)) * es], src + ((size_t)n1 * (j + (size_t)n2 * k)) * es,
                  (size_t)n1 * es);
  HIPCHK(hipMemcpy(dev, h->stage.data(), (size_t)g.n * es, hipMemcpyHostToDevice));
  return 0;
}
int get(tdgl_handle *h, void *host, const void *dev, int n1, int n2, int n3, size_t es) {
  const Geo &g = h->g;
  h->stage.resize((size_t)g.n * es);
  HIPCHK(hipMemcpy(h->stage.data(), dev, (size_t)g.n * es, hipMemcpyDeviceToHost));
  unsigned char *dst = (unsigned char *)host;
  for (int k = 0; k < n3; ++k)
    for (int j = 0; j < n2; ++j)
      std::memcpy(dst + ((size_t)n1 * (j + (size_t)n2 * k)) * es, &h->stage[((size_t)g.px * (j + (size_t)g.py * k)) * es],
                  (size_t)n1 * es);
  return 0;
}

#define CHK(expr)          \
  do {                     \
    int rc_ = (expr);      \
    if (rc_) return rc_;   \
  } while (0)
#define LCHK(expr)                                                                                   \
  do {                                                                                               \
    hipError_t e_ = (expr);                                                                          \
    if (e_ != hipSuccess) return fail(TDGL_E_DEVICE, "launch %s: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

// ---------------------------------------------------------------------------- band update
StageArgs stage_base(tdgl_handle *h, int mu, const tdgl_bandcoef &cf) {
  StageArgs a{};
  for (int nu = 0; nu < h->d.nbands; ++nu) a.G[nu] = h->G[nu];
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
  return a;
}

LinkArgs link_base(tdgl_handle *h, int mode) {
  LinkArgs a{};
  a.mode = mode;
  for (int c = 0; c < 3; ++c) {
    a.L[c] = h->L[c];
    a.Q[c] = h->Q[c];
    a.R[c] = h->R[c];
    a.V[c] = h->V[c];
    a.U[c] = h->U[c];
  }
  a.Om = h->Om;
  a.dt = h->d.dt;
  return a;
}

// _band_rkl! (bands.jl:464-521): stage j writes buf[(j - ns) mod 3] of (F, W, W2) so that the
// last stage lands in F and no stage writes a buffer it reads.
int band_rkl(tdgl_handle *h, int mu, int ns) {
  const tdgl_desc &d = h->d;
  const bool kc = d.rkc[mu] != 0;
  const double q = kc ? d.divisor[mu] : (double)(ns * (ns + 1) / 2);
  const double hh = d.dt / q;
  const bool rot = d.rotframe, ex = d.extrap;
  cplx *buf[3] = {h->F[mu], h->W, h->W2};
  const cplx *A = h->G[mu], *B = h->G[mu];
  for (int j = 1; j <= ns; ++j) {
    cplx *C = buf[(((j - ns) % 3) + 3) % 3];
    StageArgs a = stage_base(h, mu, d.bcs[mu]);
    a.D = C;
    a.S = B;
    a.A = A;
    const bool own = ex && j >= 2;
    a.Lx = own ? h->L[0] : h->V[0];
    a.Ly = own ? h->L[1] : h->V[1];
    a.Lz = own ? h->L[2] : h->V[2];
    a.ghost_fly = j >= 2;
    a.frame = rot ? (j == 1 ? FRAME_RATE : FRAME_TERM) : FRAME_NONE;
    a.h = hh;
    if (j >= 2) {
      a.combine = 1;
      a.cmu = kc ? d.rkc_mu[mu][j - 1] : (2.0 * j - 1) / j;
      a.cnu = kc ? d.rkc_nu[mu][j - 1] : (1.0 - j) / j;
    }
    a.turn = rot && j == ns;
    LCHK(launch_stage(h->g, a, h->s));
    if (rot && j == 1 && ex && ns > 1) LCHK(launch_links(h->g, link_base(h, LINK_RATE), h->s));
    if (j < ns && ex) {
      LinkArgs la = link_base(h, rot ? LINK_ROTLINKS : LINK_ROTFRAC);
      la.c = kc ? d.rkc_c[mu][j - 1] : (double)(j * (j + 1)) / (double)(ns * (ns + 1));
      LCHK(launch_links(h->g, la, h->s));
    }
    A = B;
    B = C;
  }
  return 0;
}

// _band_substeps! (bands.jl:378-406)
int band_substeps(tdgl_handle *h, int mu, int n) {
  const tdgl_desc &d = h->d;
  const bool rot = d.rotframe, ex = d.extrap;
  const double hh = d.dt / n;
  if (ex) {
    LinkArgs la = link_base(h, LINK_ROOT);
    la.n = n;
    LCHK(launch_links(h->g, la, h->s));
  }
  const cplx *src = h->G[mu];
  for (int m = 1; m <= n; ++m) {
    cplx *dst = ((n - m) % 2 == 0) ? h->F[mu] : h->W;
    StageArgs a = stage_base(h, mu, d.bcs[mu]);
    a.D = dst;
    a.S = src;
    a.Lx = ex ? h->L[0] : h->V[0];
    a.Ly = ex ? h->L[1] : h->V[1];
    a.Lz = ex ? h->L[2] : h->V[2];
    a.ghost_fly = m >= 2;
    a.frame = rot ? (m == 1 ? FRAME_RATE : FRAME_TERM) : FRAME_NONE;
    a.h = hh;
    a.turn = rot && m == n;
    LCHK(launch_stage(h->g, a, h->s));
    if (m < n && ex) {
      LinkArgs la = link_base(h, LINK_ADVANCE);
      la.frame = rot;
      la.c = hh;
      LCHK(launch_links(h->g, la, h->s));
    }
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
      a.Lx = h->V[0];
      a.Ly = h->V[1];
      a.Lz = h->V[2];
      a.frame = d.rotframe ? FRAME_RATE : FRAME_NONE;
      a.h = d.dt;
      a.turn = d.rotframe;
      LCHK(launch_stage(h->g, a, h->s));
    } else {
      CHK(band_substeps(h, mu, d.nsub[mu]));
    }
  }
  return 0;
}

// ---------------------------------------------------------------------------- separable solve
TxPass tx_pass(tdgl_handle *h, const void *in, void *out, int axis, bool inv, bool f32) {
  TxPass p{};
  p.axis = axis;
  for (int c = 0; c < 3; ++c) {
    const size_t off = (size_t)h->rg.off[c];
    const size_t es_in = 8, es_out = 8;
    (void)es_in;
    (void)es_out;
    p.in[c] = in;
    p.out[c] = out;
    for (int q = 0; q < 3; ++q) p.e[c][q] = h->rg.e[c][q];
    const int which = (axis == c) ? 1 : 0;  // U on the component's own axis, V elsewhere
    p.A[c] = f32 ? (const void *)h->Af[axis][which][inv] : (const void *)h->Ad[axis][which][inv];
    (void)off;
  }
  return p;
}

// offsets applied per component with the element size of each side
void tx_offsets(TxPass &p, const Ranges &rg, size_t es_in, size_t es_out) {
  for (int c = 0; c < 3; ++c) {
    p.in[c] = (const unsigned char *)p.in[c] + (size_t)rg.off[c] * es_in;
    p.out[c] = (unsigned char *)p.out[c] + (size_t)rg.off[c] * es_out;
  }
}

// out = (I + aK)^{-1} in on range vectors (_sep_solve!, implicit.jl:62-83; out may be in)
int sep_solve(tdgl_handle *h, const double *in, double *out, const int *gate) {
  const double *src[6] = {in, h->T1, h->T2, h->T1, h->T2, h->T1};
  double *dst[6] = {h->T1, h->T2, h->T1, h->T2, h->T1, out};
  for (int q = 0; q < 6; ++q) {
    const int axis = q % 3;
    const bool inv = q >= 3;
    TxPass p = tx_pass(h, src[q], dst[q], axis, inv, false);
    tx_offsets(p, h->rg, 8, 8);
    LCHK(launch_tx<double>(p, gate, h->s));
    if (q == 2) {
      ModeArgs m{h->T1 + h->rg.off[0], h->T1 + h->rg.off[1], h->T1 + h->rg.off[2], h->sd[0], h->sd[1], h->sd[2],
                 h->g.Nx - 1, h->g.Ny - 1, h->g.Nz - 1, h->d.a};
      LCHK(launch_modes<double>(m, gate, h->s));
    }
  }
  return 0;
}

// The Float32 separable inverse (_precond32!, implicit.jl:353-373)
int sep_solve32(tdgl_handle *h, const double *in, double *out, const int *gate) {
  const void *src[6] = {in, h->T32a, h->T32b, h->T32a, h->T32b, h->T32a};
  void *dst[6] = {h->T32a, h->T32b, h->T32a, h->T32b, h->T32a, out};
  for (int q = 0; q < 6; ++q) {
    const int axis = q % 3;
    const bool inv = q >= 3;
    TxPass p = tx_pass(h, src[q], dst[q], axis, inv, true);
    if (q == 0) {
      tx_offsets(p, h->rg, 8, 4);
      LCHK(launch_tx_d2f(p, gate, h->s));
    } else if (q == 5) {
      tx_offsets(p, h->rg, 4, 8);
      LCHK(launch_tx_f2d(p, gate, h->s));
    } else {
      tx_offsets(p, h->rg, 4, 4);
      LCHK(launch_tx<float>(p, gate, h->s));
    }
    if (q == 2) {
      ModeArgs m{h->T32a + h->rg.off[0], h->T32a + h->rg.off[1], h->T32a + h->rg.off[2], h->sf[0], h->sf[1], h->sf[2],
                 h->g.Nx - 1, h->g.Ny - 1, h->g.Nz - 1, h->d.a};
      LCHK(launch_modes<float>(m, gate, h->s));
    }
  }
  return 0;
}

// ---------------------------------------------------------------------------- PCG
int pcg_iteration(tdgl_handle *h, bool flexible, bool f32) {
  const tdgl_desc &d = h->d;
  const int n = h->rg.total;
  const int *act = h->pd.fl + FL_ACTIVE;
  LCHK(launch_apply_A(h->g, h->rg, h->P, h->Qv, h->D, d.sxy, d.sxz, d.syx, d.syz, d.szx, d.szy, d.dx, d.dy, d.dz, act, h->s));
  LCHK(launch_dot(h->P, h->Qv, nullptr, n, h->pd.part, act, h->s));
  LCHK(launch_finish(FIN_ALPHA, h->pd, 0, act, h->s));
  VecArgs v{};
  v.n = n;
  v.x = h->x;
  v.r = h->r;
  v.P = h->P;
  v.Q = h->Qv;
  v.z = h->z;
  v.zold = h->zold;
  v.sc = h->pd.sc;
  v.flexible = flexible;
  v.op = V_XR;
  LCHK(launch_vec(v, act, h->s));
  LCHK(launch_dot(h->r, h->r, nullptr, n, h->pd.part, act, h->s));
  LCHK(launch_finish(FIN_RR, h->pd, 0, act, h->s));
  CHK(f32 ? sep_solve32(h, h->r, h->z, act) : sep_solve(h, h->r, h->z, act));
  LCHK(launch_dot(h->r, h->z, flexible ? h->zold : nullptr, n, h->pd.part, act, h->s));
  LCHK(launch_finish(FIN_BETA, h->pd, flexible, act, h->s));
  v.op = V_P;
  LCHK(launch_vec(v, act, h->s));
  return 0;
}

// implicit_curl!(s, ::PCGSolver) / _implicit_curl_fast! (implicit.jl:279-446), after the field
// kernel has written b = φₑ − θ (h->B) and scaled φ by D.
int implicit_pcg(tdgl_handle *h) {
  const tdgl_desc &d = h->d;
  const bool fast = d.pcg_fast || d.pcg_f32;
  const bool f32 = d.pcg_f32 != 0;
  const bool extrap = d.pcg_extrap && d.pcg_fast;
  const bool flexible = fast && f32;
  const int n = h->rg.total;
  const int *nz = h->pd.fl + FL_COLD;
  const int *act = h->pd.fl + FL_ACTIVE;
  LCHK(launch_dot(h->B, h->B, nullptr, n, h->pd.part, nullptr, h->s));
  LCHK(launch_finish(FIN_NB, h->pd, 0, nullptr, h->s));
  VecArgs v{};
  v.n = n;
  v.x = h->x;
  v.r = h->r;
  v.P = h->P;
  v.Q = h->Qv;
  v.z = h->z;
  v.b = h->B;
  v.zold = h->zold;
  v.xprev = h->xprev;
  v.sc = h->pd.sc;
  v.flexible = flexible;
  v.warm = h->warm;
  v.extrap = fast && extrap;
  v.have2 = h->have2;
  v.op = V_START;
  LCHK(launch_vec(v, nullptr, h->s));
  if (!h->warm) CHK(sep_solve(h, h->B, h->x, nz));  // cold start: x = M⁻¹ b in Float64
  const bool warm_before = h->warm;
  h->have2 = fast && warm_before && extrap;
  h->warm = true;
  LCHK(launch_apply_A(h->g, h->rg, h->x, h->Qv, h->D, d.sxy, d.sxz, d.syx, d.syz, d.szx, d.szy, d.dx, d.dy, d.dz, nz, h->s));
  v.op = V_RESID;
  LCHK(launch_vec(v, nz, h->s));
  LCHK(launch_dot(h->r, h->r, nullptr, n, h->pd.part, nz, h->s));
  LCHK(launch_finish(FIN_RR0, h->pd, 0, nz, h->s));
  CHK(flexible ? sep_solve32(h, h->r, h->z, act) : sep_solve(h, h->r, h->z, act));
  v.op = V_COPY;
  LCHK(launch_vec(v, act, h->s));
  LCHK(launch_dot(h->r, h->z, nullptr, n, h->pd.part, act, h->s));
  LCHK(launch_finish(FIN_RZ0, h->pd, 0, act, h->s));
  // iterations in batches; the host reads the flags between batches only
  int launched = 0;
  int batch = h->last_iters > 0 ? h->last_iters : 4;
  int fl[FL_COUNT];
  for (;;) {
    HIPCHK(hipMemcpyAsync(fl, h->pd.fl, sizeof fl, hipMemcpyDeviceToHost, h->s));
    HIPCHK(hipStreamSynchronize(h->s));
    if (!fl[FL_ACTIVE] || launched >= d.pcg_maxit) break;
    const int nb = std::min(batch, d.pcg_maxit - launched);
    for (int q = 0; q < nb; ++q) CHK(pcg_iteration(h, flexible, flexible));
    launched += nb;
    batch = 2;
  }
  h->last_iters = fl[FL_IT];
  h->pcg_iters += fl[FL_IT];
  const double *th[3] = {h->th[0], h->th[1], h->th[2]};
  double *ph[3] = {h->ph[0], h->ph[1], h->ph[2]};
  LCHK(launch_scatter(h->g, h->rg, h->x, th, ph, h->s));
  return 0;
}

// ---------------------------------------------------------------------------- one step
int step_once(tdgl_handle *h, double H, double HJ) {
  const tdgl_desc &d = h->d;
  CHK(update_bands(h));
  FieldArgs fa{};
  for (int mu = 0; mu < d.nbands; ++mu) fa.G[mu] = h->G[mu];
  for (int c = 0; c < 3; ++c) {
    fa.V[c] = h->V[c];
    fa.th[c] = h->th[c];
    fa.ph[c] = h->ph[c];
    fa.J[c] = h->J[c];
    fa.h[c] = h->h[c];
    fa.tau[c] = h->tau[c];
    fa.D[c] = h->D[c];
  }
  fa.B = h->B;
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
  LCHK(launch_field(h->g, fa, h->s));
  if (d.field_mode == TDGL_FIELD_SEPARABLE) {
    CHK(sep_solve(h, h->B, h->B, nullptr));
    const double *th[3] = {h->th[0], h->th[1], h->th[2]};
    double *ph[3] = {h->ph[0], h->ph[1], h->ph[2]};
    LCHK(launch_scatter(h->g, h->rg, h->B, th, ph, h->s));
  } else if (d.field_mode == TDGL_FIELD_PCG) {
    CHK(implicit_pcg(h));
  }
  CurlArgs ca{};
  for (int c = 0; c < 3; ++c) {
    ca.ph[c] = h->ph[c];
    ca.U[c] = h->U[c];
    ca.h[c] = h->h[c];
  }
  ca.H = H;
  ca.HJ = HJ;
  ca.bar = d.selffield_bc == TDGL_BC_BAR;
  ca.bar_hy_z = h->bar[0];
  ca.bar_hy_y = h->bar[1];
  ca.bar_hz_z = h->bar[2];
  ca.bar_hz_y = h->bar[3];
  ca.bar_hz_x = h->bar[4];
  ca.bar_hy_x = h->bar[5];
  ca.dydz = d.dy * d.dz;
  ca.dzdx = d.dz * d.dx;
  ca.dxdy = d.dx * d.dy;
  LCHK(launch_curl(h->g, ca, h->s));
  GhostArgs ga{};
  for (int mu = 0; mu < d.nbands; ++mu) ga.F[mu] = h->F[mu];
  for (int c = 0; c < 3; ++c) ga.U[c] = h->U[c];
  ga.nb = d.nbands;
  LCHK(launch_ghosts(h->g, ga, h->s));
  // advance! (field.jl:248-263): F equals G outside the block, so G <- F is a swap
  for (int mu = 0; mu < d.nbands; ++mu) std::swap(h->G[mu], h->F[mu]);
  for (int c = 0; c < 3; ++c) {
    std::swap(h->th[c], h->ph[c]);
    std::swap(h->V[c], h->U[c]);
  }
  return 0;
}

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
  if ((d->field_mode == TDGL_FIELD_PCG || d->field_mode == TDGL_FIELD_EXPLICIT_SCALED) && !(d->Dx && d->Dy && d->Dz))
    return no("link conductivities D");
  if ((d->field_mode == TDGL_FIELD_SEPARABLE || d->field_mode == TDGL_FIELD_PCG) && !(d->dx == d->dy && d->dy == d->dz))
    return no("implicit curl-curl with dx != dy or dz");
  if (d->selffield_bc == TDGL_BC_BAR && !(d->bar_hy_z && d->bar_hy_y && d->bar_hz_z && d->bar_hz_y && d->bar_hz_x && d->bar_hy_x))
    return no("bar profiles");
  if (!(d->f && d->taux && d->tauy && d->tauz)) return no("masks");
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
    if (d.extrap && (stages || subs))
      for (int c = 0; c < 3; ++c) ALLOC(h->L[c], n);
    if (d.extrap && subs)
      for (int c = 0; c < 3; ++c) ALLOC(h->Q[c], n);
    if (d.rotframe) {
      ALLOC(h->Om, n);
      for (int c = 0; c < 3; ++c) ALLOC(h->R[c], n);
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
    if (d.selffield_bc == TDGL_BC_BAR) {
      const size_t nn[6] = {(size_t)(d.Ny + 1) * 2, (size_t)2 * d.Nz, (size_t)d.Ny * 2, (size_t)2 * (d.Nz + 1),
                            (size_t)d.Ny * (d.Nz + 1), (size_t)(d.Ny + 1) * d.Nz};
      const double *src[6] = {d.bar_hy_z, d.bar_hy_y, d.bar_hz_z, d.bar_hz_y, d.bar_hz_x, d.bar_hy_x};
      for (int q = 0; q < 6; ++q) {
        ALLOC(h->bar[q], nn[q]);
        CHK(upload_vec(h->bar[q], std::vector<double>(src[q], src[q] + nn[q])));
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
            CHK(upload_vec(h->Ad[ax][which][t], src));
            std::vector<float> s32(src.begin(), src.end());
            ALLOC(h->Af[ax][which][t], s32.size());
            CHK(upload_vec(h->Af[ax][which][t], s32));
          }
        }
        ALLOC(h->sd[ax], nn[ax]);
        CHK(upload_vec(h->sd[ax], b.s));
        std::vector<float> s32(b.s.begin(), b.s.end());
        ALLOC(h->sf[ax], nn[ax]);
        CHK(upload_vec(h->sf[ax], s32));
      }
      ALLOC(h->B, rg.total);
      ALLOC(h->T1, rg.total);
      ALLOC(h->T2, rg.total);
    }
    if (d.field_mode == TDGL_FIELD_PCG) {
      ALLOC(h->x, rg.total);
      ALLOC(h->r, rg.total);
      ALLOC(h->z, rg.total);
      ALLOC(h->P, rg.total);
      ALLOC(h->Qv, rg.total);
      ALLOC(h->xprev, rg.total);
      ALLOC(h->zold, rg.total);
      ALLOC(h->T32a, rg.total);
      ALLOC(h->T32b, rg.total);
      ALLOC(h->pd.sc, SC_COUNT);
      ALLOC(h->pd.fl, FL_COUNT);
      ALLOC(h->pd.part, 2 * dot_blocks());
      std::vector<double> sc(SC_COUNT, 0.0);
      sc[SC_TOLNB] = d.pcg_tol;
      CHK(upload_vec(h->pd.sc, sc));
      std::vector<int> fl(FL_COUNT, 0);
      fl[FL_MAXIT] = d.pcg_maxit;
      CHK(upload_vec(h->pd.fl, fl));
    }
    return 0;
  };
  int rc = run();
  if (rc) return bail(rc);
  *out = h;
  return 0;
}

int tdgl_destroy(tdgl_handle *h) {
  if (!h) return 0;
  if (h->s) (void)hipStreamSynchronize(h->s);
  for (void *p : h->allocs) (void)hipFree(p);
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
    if (!s->G[mu] || !s->F[mu]) return fail(TDGL_E_ARG, "null band array");
    CHK(put(h, h->G[mu], s->G[mu], Nx + 1, Ny + 1, Nz + 1, 16));
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
    CHK(put(h, h->F[mu], Fh.data(), Nx + 1, Ny + 1, Nz + 1, 16));
  }
  (void)g;
  const int lx[3] = {Nx, Ny + 1, Nz + 1}, ly[3] = {Nx + 1, Ny, Nz + 1}, lz[3] = {Nx + 1, Ny + 1, Nz};
  const int *ls[3] = {lx, ly, lz};
  const int hx[3] = {Nx + 1, Ny, Nz}, hy[3] = {Nx, Ny + 1, Nz}, hz[3] = {Nx, Ny, Nz + 1};
  const int *hs[3] = {hx, hy, hz};
  const double *th[3] = {s->thx, s->thy, s->thz}, *ph[3] = {s->phx, s->phy, s->phz};
  const double *U[3] = {s->Ux, s->Uy, s->Uz}, *V[3] = {s->Vx, s->Vy, s->Vz};
  const double *J[3] = {s->Jx, s->Jy, s->Jz}, *hh[3] = {s->hx, s->hy, s->hz};
  for (int c = 0; c < 3; ++c) {
    if (!th[c] || !ph[c] || !U[c] || !V[c] || !J[c] || !hh[c]) return fail(TDGL_E_ARG, "null state array");
    CHK(put(h, h->th[c], th[c], ls[c][0], ls[c][1], ls[c][2], 8));
    CHK(put(h, h->ph[c], ph[c], ls[c][0], ls[c][1], ls[c][2], 8));
    CHK(put(h, h->U[c], U[c], ls[c][0], ls[c][1], ls[c][2], 16));
    CHK(put(h, h->V[c], V[c], ls[c][0], ls[c][1], ls[c][2], 16));
    CHK(put(h, h->J[c], J[c], ls[c][0], ls[c][1], ls[c][2], 8));
    CHK(put(h, h->h[c], hh[c], hs[c][0], hs[c][1], hs[c][2], 8));
  }
  return 0;
}

int tdgl_download_state(tdgl_handle *h, tdgl_state *s) {
  if (!h || !s) return fail(TDGL_E_ARG, "null argument");
  const tdgl_desc &d = h->d;
  const int Nx = d.Nx, Ny = d.Ny, Nz = d.Nz;
  HIPCHK(hipStreamSynchronize(h->s));
  for (int mu = 0; mu < d.nbands; ++mu) {
    if (!s->G[mu] || !s->F[mu]) return fail(TDGL_E_ARG, "null band array");
    CHK(get(h, s->G[mu], h->G[mu], Nx + 1, Ny + 1, Nz + 1, 16));
    CHK(get(h, s->F[mu], h->F[mu], Nx + 1, Ny + 1, Nz + 1, 16));
  }
  const int lx[3] = {Nx, Ny + 1, Nz + 1}, ly[3] = {Nx + 1, Ny, Nz + 1}, lz[3] = {Nx + 1, Ny + 1, Nz};
  const int *ls[3] = {lx, ly, lz};
  const int hx[3] = {Nx + 1, Ny, Nz}, hy[3] = {Nx, Ny + 1, Nz}, hz[3] = {Nx, Ny, Nz + 1};
  const int *hs[3] = {hx, hy, hz};
  double *th[3] = {s->thx, s->thy, s->thz}, *ph[3] = {s->phx, s->phy, s->phz};
  double *U[3] = {s->Ux, s->Uy, s->Uz}, *V[3] = {s->Vx, s->Vy, s->Vz};
  double *J[3] = {s->Jx, s->Jy, s->Jz}, *hh[3] = {s->hx, s->hy, s->hz};
  for (int c = 0; c < 3; ++c) {
    if (!th[c] || !ph[c] || !U[c] || !V[c] || !J[c] || !hh[c]) return fail(TDGL_E_ARG, "null state array");
    CHK(get(h, th[c], h->th[c], ls[c][0], ls[c][1], ls[c][2], 8));
    CHK(get(h, ph[c], h->ph[c], ls[c][0], ls[c][1], ls[c][2], 8));
    CHK(get(h, U[c], h->U[c], ls[c][0], ls[c][1], ls[c][2], 16));
    CHK(get(h, V[c], h->V[c], ls[c][0], ls[c][1], ls[c][2], 16));
    CHK(get(h, J[c], h->J[c], ls[c][0], ls[c][1], ls[c][2], 8));
    CHK(get(h, hh[c], h->h[c], hs[c][0], hs[c][1], hs[c][2], 8));
  }
  return 0;
}

int tdgl_step(tdgl_handle *h, int32_t n, double H, double HJ) {
  if (!h) return fail(TDGL_E_ARG, "null handle");
  if (n < 0) return fail(TDGL_E_ARG, "negative step count");
  if (!std::isfinite(H) || !std::isfinite(HJ)) return fail(TDGL_E_ARG, "non-finite H or HJ");
  HIPCHK(hipEventRecord(h->e0, h->s));
  for (int q = 0; q < n; ++q) CHK(step_once(h, H, HJ));
  HIPCHK(hipEventRecord(h->e1, h->s));
  HIPCHK(hipStreamSynchronize(h->s));
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, h->e0, h->e1));
  h->device_ms += ms;
  h->steps += n;
  h->calls += 1;
  return 0;
}

int tdgl_get_pcg(tdgl_handle *h, tdgl_pcg *p) {
  if (!h || !p) return fail(TDGL_E_ARG, "null argument");
  if (h->d.field_mode != TDGL_FIELD_PCG) return fail(TDGL_E_ARG, "the stepper has no PCG solver");
  HIPCHK(hipStreamSynchronize(h->s));
  for (int c = 0; c < 3; ++c) {
    const size_t n = (size_t)h->rg.e[c][0] * h->rg.e[c][1] * h->rg.e[c][2];
    if (p->x[c]) HIPCHK(hipMemcpy(p->x[c], h->x + h->rg.off[c], n * 8, hipMemcpyDeviceToHost));
    if (p->xprev[c]) HIPCHK(hipMemcpy(p->xprev[c], h->xprev + h->rg.off[c], n * 8, hipMemcpyDeviceToHost));
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
  for (int c = 0; c < 3; ++c) {
    const size_t n = (size_t)h->rg.e[c][0] * h->rg.e[c][1] * h->rg.e[c][2];
    if (p->x[c]) HIPCHK(hipMemcpy(h->x + h->rg.off[c], p->x[c], n * 8, hipMemcpyHostToDevice));
    if (p->xprev[c]) HIPCHK(hipMemcpy(h->xprev + h->rg.off[c], p->xprev[c], n * 8, hipMemcpyHostToDevice));
  }
  h->warm = p->warm != 0;
  h->have2 = p->have2 != 0;
  h->last_iters = p->iters;
  return 0;
}

int tdgl_get_stats(tdgl_handle *h, tdgl_stats *st) {
  if (!h || !st) return fail(TDGL_E_ARG, "null argument");
  st->steps = h->steps;
  st->calls = h->calls;
  st->pcg_iters = h->pcg_iters;
  st->last_iters = h->last_iters;
  st->device_ms = h->device_ms;
  st->device_bytes = h->bytes;
  return 0;
}

}  // extern "C"
