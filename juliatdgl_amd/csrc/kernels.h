// kernels.h — launch interface of the HIP kernels of one step (kernels.hip).  Every launcher
// enqueues on `stream` and returns the launch status; none synchronises.
#pragma once

#include "common.h"

namespace tdgl {

// ---------------------------------------------------------------- band update (bands.jl)
enum FrameMode { FRAME_NONE = 0, FRAME_RATE = 1, FRAME_TERM = 2 };

struct StageArgs {
  cplx *D;                     // output (interior nodes)
  const cplx *S;               // source Y_{j-1}
  const cplx *A;               // Y_{j-2} (combine) or null
  const cplx *G[TDGL_MAX_BANDS];  // the bands at the start of the step (coupling)
  const cplx *Lx, *Ly, *Lz;    // links of this stage
  const double *f;             // condensation mask
  double *Om;                  // rotating-frame rate (written by FRAME_RATE, read otherwise)
  double kx, ky, kz, c, eps, beta, gamma2;
  double cpl[TDGL_MAX_BANDS];
  int mu, nb;
  int ghost_fly;               // S's face ghosts are L·(centre), not read from S
  int frame;                   // FrameMode
  double h, wmax;              // Euler step of the stage, cap of the rate
  int combine;                 // C = cmu·C + cnu·A
  double cmu, cnu;
  int turn;                    // F = C·cis(Ω dt) (end of the band's step)
  double dt;
  int hard;
};
hipError_t launch_stage(const Geo &g, const StageArgs &a, hipStream_t s);

enum LinkMode {
  LINK_ROTFRAC = 0,   // L = V (V Ū)^c                     (_rotfrac!, bands.jl:409-423)
  LINK_ROTLINKS = 1,  // L = V _cis(R c)                    (_rotlinks!, :311-316)
  LINK_RATE = 2,      // R = arg(V Ū) + (Ω[+e] − Ω) dt       (_link_rate!, :303-310)
  LINK_ROOT = 3,      // L = V, Q = (V Ū)^(1/n)              (_rot!, :355-369)
  LINK_ADVANCE = 4    // L = L·Q [· _cis((Ω[+e] − Ω) τ)]      (_advance_links! + _frame_links!)
};
struct LinkArgs {
  int mode;
  cplx *L[3];
  cplx *Q[3];            // substep roots (complex)
  double *R[3];          // rotframe link rates (real)
  const cplx *V[3];
  const cplx *U[3];
  const double *Om;
  double c;              // fraction of the step / τ
  double dt;
  int n;                 // substeps
  int frame;             // ADVANCE: also the frame links
};
hipError_t launch_links(const Geo &g, const LinkArgs &a, hipStream_t s);

// ---------------------------------------------------------------- field (field.jl)
struct FieldArgs {
  const cplx *G[TDGL_MAX_BANDS];
  const cplx *V[3];
  const double *th[3];
  double *ph[3];
  double *J[3];
  const double *h[3];
  const double *tau[3];
  const double *D[3];    // scaled modes: φ = θ + (φₑ − θ)/D on every link
  double *B;             // implicit modes: φₑ − θ on the ranges (compact)
  Ranges rg;
  double wa[3][TDGL_MAX_BANDS];
  double d1, sxy, sxz, syx, syz, szx, szy;
  int nb;
  int scale;             // apply the D scaling
  int gather;            // write B
};
hipError_t launch_field(const Geo &g, const FieldArgs &a, hipStream_t s);

struct CurlArgs {
  const double *ph[3];
  cplx *U[3];
  double *h[3];
  double H, HJ;
  int bar;
  const double *bar_hy_z, *bar_hy_y, *bar_hz_z, *bar_hz_y, *bar_hz_x, *bar_hy_x;
  double idydz, idzdx, idxdy;  // 1/(dy dz) etc. are NOT used: the reference divides by (dy*dz)
  double dydz, dzdx, dxdy;
};
hipError_t launch_curl(const Geo &g, const CurlArgs &a, hipStream_t s);

struct GhostArgs {
  cplx *F[TDGL_MAX_BANDS];
  const cplx *U[3];
  int nb;
};
hipError_t launch_ghosts(const Geo &g, const GhostArgs &a, hipStream_t s);

// φ = θ + B on the ranges (implicit_curl!, implicit.jl:98-102 and :328-332)
hipError_t launch_scatter(const Geo &g, const Ranges &rg, const double *B, const double *const th[3],
                          double *const ph[3], hipStream_t s);

// ---------------------------------------------------------------- separable solve (implicit.jl)
// One pass of the 1D transforms along `axis` for the three components: out = M' in (the
// forward product with the basis of each component on that axis) with A[n*L + m] = M(n, m).
struct TxPass {
  const void *in[3];
  void *out[3];
  const void *A[3];
  int e[3][3];
  int axis;
};
template <class T>
hipError_t launch_tx(const TxPass &p, const int *gate, hipStream_t s);
// The same pass from double input to T (first pass of the Float32 solve) and from T to double.
hipError_t launch_tx_d2f(const TxPass &p, const int *gate, hipStream_t s);
hipError_t launch_tx_f2d(const TxPass &p, const int *gate, hipStream_t s);

struct ModeArgs {
  void *X, *Y, *Z;
  const void *sx, *sy, *sz;
  int nx, ny, nz;  // basis sizes (nodes) per axis
  double a;
};
template <class T>
hipError_t launch_modes(const ModeArgs &m, const int *gate, hipStream_t s);

// ---------------------------------------------------------------- PCG (implicit.jl:279-446)
// Device scalars of one solve
enum {
  SC_NB = 0, SC_RZ, SC_PAP, SC_ALPHA, SC_RR, SC_RZ2, SC_RZO, SC_BETA, SC_TOLNB,
  SC_COUNT
};
enum { FL_ACTIVE = 0, FL_COLD, FL_IT, FL_MAXIT, FL_COUNT };
struct PcgDev {
  double *sc;       // SC_*
  int *fl;          // FL_*: ACTIVE (iterations run while 1), COLD gate, iteration counter
  double *part;     // partial sums (NPART x 2)
};
hipError_t launch_apply_A(const Geo &g, const Ranges &rg, const double *P, double *Q, const double *const D[3],
                          double sxy, double sxz, double syx, double syz, double szx, double szy, double dx, double dy,
                          double dz, const int *gate, hipStream_t s);
// partial dot products over the whole range vector: part[b] = a·b (and a·c if c)
hipError_t launch_dot(const double *a, const double *b, const double *c, int n, double *part, const int *gate,
                      hipStream_t s);
int dot_blocks();
// what to do with the reduced sums
enum {
  FIN_NB = 0,     // nb = sqrt(Σ); ACTIVE = nb > 0 (COLD gate too)
  FIN_RR0,        // first residual: ACTIVE = nb > 0 && sqrt(Σ) > tol nb
  FIN_RZ0,        // rz = Σ
  FIN_ALPHA,      // pAp = Σ; α = rz/pAp; it++
  FIN_RR,         // rr = Σ; converged or maxit -> ACTIVE = 0
  FIN_BETA,       // rz2 = Σ0 (rzo = Σ1): β; rz = rz2
};
hipError_t launch_finish(int what, PcgDev d, int flexible, const int *gate, hipStream_t s);
// elementwise vector ops over the whole range vector
enum {
  V_COPY = 0,       // y = x
  V_XR,             // x += α P, r −= α Q
  V_P,              // P = z + β P  (and zold = z if flexible)
  V_RESID,          // r = b − Q
  V_START           // warm start: extrapolation or zero source
};
struct VecArgs {
  int op, n;
  double *x, *r, *P, *zold, *xprev;
  const double *Q, *z, *b;
  const double *sc;
  int flexible;
  int warm, extrap, have2;
};
hipError_t launch_vec(const VecArgs &v, const int *gate, hipStream_t s);

}  // namespace tdgl
