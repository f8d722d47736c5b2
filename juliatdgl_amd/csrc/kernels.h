// kernels.h — launch interface of the HIP kernels of one step (kernels.hip).  Every launcher
// enqueues on `stream` and returns the launch status; none synchronises.
#pragma once

#include "common.h"

#include <hip/hip_ext.h>

#include <vector>

namespace tdgl {

// Every launch of the library goes through TDGL_LAUNCH.  While a profile pair is set (the per-kernel
// profile, eager launches only) the runtime records it at the kernel's own start and end
// (hipExtLaunchKernelGGL), so the pair times the kernel and not the gap before it.
struct ProfPair {
  hipEvent_t a, b;
  bool used;
};
extern thread_local ProfPair *prof_pair;
#define TDGL_LAUNCH(kern, grid, block, shmem, stream, ...)                                                 \
  do {                                                                                                     \
    if (::tdgl::prof_pair) {                                                                               \
      hipExtLaunchKernelGGL(kern, grid, block, shmem, stream, ::tdgl::prof_pair->a, ::tdgl::prof_pair->b, 0, \
                            __VA_ARGS__);                                                                  \
      ::tdgl::prof_pair->used = true;                                                                      \
    } else {                                                                                               \
      hipLaunchKernelGGL(kern, grid, block, shmem, stream, __VA_ARGS__);                                   \
    }                                                                                                      \
  } while (0)


// ---------------------------------------------------------------- band update (bands.jl)
// With noise (bands.jl:210-211, 237-281) the first stage's rate is low-passed before its term
// (FRAME_RAW: the rate only; k_frame_smooth does the rest) and the later terms are rotations
// (FRAME_TURNTERM: C e^{−iω̂h}).
enum FrameMode { FRAME_NONE = 0, FRAME_RATE = 1, FRAME_TERM = 2, FRAME_RAW = 3, FRAME_TURNTERM = 4 };

// Thermal noise (noise.jl): where a kernel takes the normals of the current step.  ctr[0] counts
// the steps taken with noise (advanced on the device at the end of each step, so a replayed graph
// draws new numbers), ctr[1] is the step of the first draw in `hd`.  Generator mode: Philox4x32-10
// keyed by `seed` at counter (index, step, tag); host mode: hd + (ctr[0] − ctr[1])·hstride, each
// slot N complex node arrays then 3 real link arrays in the device layout.
struct NoiseArgs {
  const uint64_t *ctr;
  const double *hd;
  long long hstride, nslots;
  uint64_t seed;
  int host;
  int nb;
  double Tdt, dV;        // T̃ dt and dx dy dz
};

// How a stage gets its links (computed per use from V, U and Ω, never stored):
enum LinkSource {
  LS_DIRECT = 0,    // V                                         (first stage; extrap off)
  LS_ROTFRAC = 1,   // V (V Ū)^c                                 (_rotfrac!, bands.jl:409-423)
  LS_ROTLINKS = 2,  // V _cis(R c), R = arg(V Ū) + (Ω[+e] − Ω) dt (_link_rate! :303-310, _rotlinks! :311-316)
  LS_SUBSTEP = 3    // V Q^(m−1) with Q = (V Ū)^(1/n), each power times the frame's _cis((Ω[+e] − Ω) h)
                    //  (_rot! :355-369, _advance_links! :371-376, _frame_links! :292-297)
};

struct StageArgs {
  cplx *D;                     // output (interior nodes)
  const cplx *S;               // source Y_{j-1}
  const cplx *A;               // Y_{j-2} (combine) or null
  const cplx *G[TDGL_MAX_BANDS];  // the bands at the start of the step (coupling)
  const cplx *V[3], *U[3];     // links of this step and of the previous one
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
  int lsrc;                    // LinkSource
  double lc;                   // fraction c of the step (ROTFRAC, ROTLINKS)
  int msub, nsub;              // substep m of n (SUBSTEP)
  int lframe;                  // SUBSTEP: the frame's factor on each power
  // normal fluids: the band's last stage turns it by its potential (qp_turn!, exchange.jl:400-412,
  // 533-543), e^{-i m_µ dt δ} (TWO) or e^{-i dt Σ_k P[µ][k] m_k} (N), m read from the solved
  // modes on the region of the unknowns (nodes i = 2..Nx, j0+1.., k0+1..)
  int qp;                      // TDGL_QP_*
  const double *qpM[TDGL_MAX_BANDS];
  int qpK, qn1, qnj, qj0, qk0;
  double qpc[TDGL_MAX_BANDS];  // TWO: [0] = m_µ dt; N: P[µ][k]
  // noise: the KWT increment as a rotation (kwt_update_rot, bands.jl:150-158), the stages
  // combined in log-polar form (_combine_log!, :439-445), SK-ROCK's turn of the first stage by
  // e^{i(κ₁−ν₁)δθ} (:487), and after the band's last stage the kick of the step (noise_kick! /
  // noise_apply!, noise.jl:44-65, 127-153) with the band's Γ and γ̃² (full step)
  int kwt_rot, logcomb;
  const double *Qth;           // SK-ROCK's δθ of this band (sk_on, and the kick with nsk)
  int sk_on;
  double sk_post;              // κ₁ − ν₁
  int noise, nsk;
  double nGm, ng2;
  NoiseArgs nz;
  double *Om2;                 // k_frame_smooth: the rate between the two sweeps
  int xcd;  // the blocks to the XCDs in contiguous runs of (x, y, k) order (each block's nodes unchanged)
};
hipError_t launch_stage(const Geo &g, const StageArgs &a, hipStream_t s);
// With noise and the rotating frame, after the first stage (FRAME_RAW): the two low-pass sweeps
// of the rate, the frame term as a rotation, and what k_stage would have done after it
// (_frame_rate_smooth!, bands.jl:243-270)
hipError_t launch_frame_smooth(const Geo &g, const StageArgs &a, hipStream_t s);
// SK-ROCK's draw (noise_draw!, noise.jl:99-124, and _sk_turn!, bands.jl:447-453): δθ of the band
// from Ψⁿ and its normals into Qth, and the first stage's source Ψⁿ e^{iν₁δθ} into Sp
struct SkArgs {
  const cplx *G;
  const double *f;
  double *Qth;
  cplx *Sp;
  int mu, hard;
  double c, g2, Gm, nu1;
  NoiseArgs nz;
};
hipError_t launch_sk_src(const Geo &g, const SkArgs &a, hipStream_t s);
// the stream's step counter, after every step with noise
hipError_t launch_noise_tick(uint64_t *ctr, hipStream_t s);

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
  double *B[3];          // implicit modes: φₑ − θ on the ranges (node-shaped, zero elsewhere)
  double *part;          // PCG: block partials of b·b (null otherwise)
  Ranges rg;
  double wa[3][TDGL_MAX_BANDS];
  double d1, sxy, sxz, syx, syz, szx, szy;
  int nb;
  int scale;             // apply the D scaling
  int gather;            // write B
  // field_kick! (noise.jl:182-198): φ += (a_c √D) n on the links the step evolves, before B
  int nkick;
  double kax, kay, kaz;
  NoiseArgs nz;
  int xcd;  // the blocks to the XCDs in contiguous runs of (x, y, k) order (each block's nodes unchanged)
};
hipError_t launch_field(const Geo &g, const FieldArgs &a, hipStream_t s);
int field_blocks(const Geo &g);

struct CurlArgs {
  const double *ph[3];
  cplx *U[3];
  double *h[3];
  const double *hp;  // (H, HJ) in device memory, set before each call of the step
  int bar;
  const double *bar_hy_z, *bar_hy_y, *bar_hz_z, *bar_hz_y, *bar_hz_x, *bar_hy_x;
  double dydz, dzdx, dxdy;
  // open boundary: h_δ of the last two recomputations (h-shaped), ramped on the six faces after
  // the bar field (_open_add!, openbc.jl:218-231) by f = min((count + 1)/every, 1)
  int open;
  const double *hold[3], *hnew[3];
  const int *ofl;
  int every;
  int xcd;  // the blocks to the XCDs in contiguous runs of (x, y, k) order (each block's nodes unchanged)
};
hipError_t launch_curl(const Geo &g, const CurlArgs &a, hipStream_t s);

struct GhostArgs {
  cplx *F[TDGL_MAX_BANDS];
  const cplx *U[3];
  int nb;
};
hipError_t launch_ghosts(const Geo &g, const GhostArgs &a, hipStream_t s);

// device-to-device copies of up to 48 arrays in one launch
struct CopyList {
  const double *src[48];
  double *dst[48];
  int n[48];
  int cnt;
  const int *gate;   // copy only while *gate (null: always)
};
hipError_t launch_copies(const CopyList &c, hipStream_t s);

// ---------------------------------------------------------------- implicit solve (implicit.jl)
// Scalars, flags and reduction partials of one PCG solve
// SC_TOLI, SC_NBI: the relative tolerance and the starting residual norm of an inner Float32
// solve of the refinement (d.pcg_refine)
enum { SC_NB = 0, SC_RZ0, SC_RZ1, SC_TOL, SC_TOLI, SC_NBI, SC_COUNT };
// classes of a step for the captured PCG's budgets (PcgArgs::cls): with the open boundary, the
// steps since its last recomputation of h_δ (0 .. PCG_NCLS - 2, then the rest), whose solves start
// further from their extrapolated guesses the nearer they are to it; without it, one class
constexpr int PCG_NCLS = 6;
enum {
  FL_NZ = 0,      // the source is not zero
  FL_ACT0,        // the iterations run while set: iteration q reads ACT[q % 2], its convergence
  FL_ACT1,        //   check writes ACT[(q + 1) % 2]
  FL_IT,          // iterations of this solve
  FL_MAXIT,
  FL_OVF,         // a solve of the cycle was still active after its launched iterations
  FL_ITSUM,       // iterations summed over the cycle
  FL_ITMAX,       // the most iterations of one solve in the cycle
  FL_OACT,        // refinement: the FP64 residual is still above the tolerance (the rounds run while set)
  FL_ITMAXC,      // the most iterations of one solve of each step class in the cycle (PCG_NCLS slots)
  FL_COUNT = FL_ITMAXC + PCG_NCLS
};
// partial sums, two per block, one region per reduction (every consumer re-sums them in block
// order: the same value in every block, no fence and no extra launch)
enum { PART_NB = 0, PART_RR, PART_PQ, PART_RZ, PART_N };
constexpr int MAX_PART = 16384;
struct PcgDev {
  double *sc;
  int *fl;
  double *part;   // PART_N regions of 2 * MAX_PART
};

// One component of a range vector: element (i0, i1, i2) at p + i0 + s1 i1 + s2 i2
struct View {
  void *p;
  int s1, s2;
};

// The convergence check of a first preconditioner pass (FIN of implicit.jl:305, 319, 426)
enum { CONV_NONE = 0, CONV_RR0 = 1, CONV_RR = 2 };
struct Conv {
  int mode;
  const int *in;   // the incoming flag (NZ for RR0, ACT[q % 2] for RR)
  int *out;        // ACT of the iterations that follow
  int nparts;      // partials of r·r
  int inner;       // against SC_TOLI x SC_NBI (an inner solve of the refinement), else SC_TOL x SC_NB
};

// One pass of the 1D transforms along `axis` for the three components, out = M' in with
// A[n*L + m] = M(n, m).  `addth`: out is φ and receives θ + value (implicit.jl:98-102).
// `dot`: the pass writes z and the partials of r·z (and r·zold) at its elements.
struct TxPass {
  View in[3], out[3];
  const void *A[3];
  int e[3][3];
  int axis;
  int addth;
  View th[3];
  int dot;          // 1: r·z, 2: r·z and r·zold
  View r[3], zold[3];
  Conv conv;
  PcgDev d;
};
template <class TI, class TA, class TO>
hipError_t launch_tx(const TxPass &p, const int *gate, hipStream_t s);
int tx_parts(const TxPass &p);

// The last forward pass (axis 2) fused with the per-wavevector projector: one thread per mode.
struct Tx2Modes {
  const void *in;        // compact range vector after the axis-0 and axis-1 passes (T)
  void *out;             // compact range vector of projected modes (T)
  const void *Az[2];     // axis-2 forward matrices: [0] V_z (X and Y), [1] U_z (Z)
  const void *sx, *sy, *sz;
  int off[3];
  int nx, ny, nz;
  double a;
};
template <class T>
hipError_t launch_tx2_modes(const Tx2Modes &m, const int *gate, hipStream_t s);

// Two passes of the transforms in one launch, for each z-plane of each component: along x with
// X, then along y with Y (out(m, q) = Σ_j Y(j, q) Σ_i X(i, m) in(i, j), each sum in increasing
// order, the x result rounded to T as a pass of its own would store it).  One block per
// (component, plane, slice of PLANE_MS(T) rows of m).  The y extent is at most PLANE_LMAX.
constexpr int PLANE_LMAX = 256;
constexpr int PLANE_MAXC = 12;  // components (the three of a link vector, the modes of a potential, or the
                                // four parity quarters of each of the three components)
struct PlanePass {
  int ncomp;                   // 0 means 3
  int one;                     // Float32: the one-round kernel where it pays (1), forced with 2 or 3 m tiles, off (0)
  View in[PLANE_MAXC], out[PLANE_MAXC];
  const void *X[PLANE_MAXC], *Y[PLANE_MAXC];
  const void *XT[PLANE_MAXC], *YT[PLANE_MAXC];  // their transposes (the one-round kernel)
  int e[PLANE_MAXC][3];
  int addth;
  View th[3];
  int dot;
  View r[3], zold[3];
  Conv conv;
  PcgDev d;
};
template <class TI, class T, class TO>
hipError_t launch_plane(const PlanePass &p, const int *gate, hipStream_t s);
template <class T>
int plane_parts(const PlanePass &p);
template <class T>
bool plane_fits(const int e[][3], int ncomp);

// ---------------------------------------------------------------- the parity-folded transforms
// Every column of the separable bases is even or odd under the mirror i -> L-1-i of its axis (the
// DCT-II and DST-I columns of parity index p carry the sign (-1)^p), so a transform of length L
// splits into two of length ceil(L/2): the even columns act on the folded sums a[i] + a[L-1-i],
// the odd ones on the differences (one radix-2 stage of the symmetric transform, exact for every
// length, also the prime ones).  Folding x and y turns each component's x-y transform into four
// quarter transforms with half-size bases.  A component's quarter (px, py) holds the element
// (i, j, k) of its fold (or, after the forward transforms, the mode (2i+px, 2j+py, k)) at
// off[c][2 px + py] + i + h[c][0] (j + h[c][1] k), i < h[c][0] = ceil(Lx/2), j < h[c][1], k < e2[c].
struct Quarters {
  int L[3][2];     // the components' x and y extents
  int h[3][2];     // ceil halves
  int e2[3];       // planes
  int off[3][4];
  int total;
};
// fold: the three components of a node-shaped FP64 vector into the quarters (FP32, or FP64 into
// out64 where it is set: the FP64 separable solve); the convergence check of a preconditioner's
// first pass is the forward x-y launch's
struct FoldPass {
  View in[3];
  float *out;
  Quarters q;
  int f32;         // the input is Float32 (an inner solve's residual), else FP64
  double *out64;   // FP64 quarters (then out is unused)
};
hipError_t launch_fold(const FoldPass &f, const int *gate, hipStream_t s);
// The check of a refinement round and the fold of its Float32 residual: from the partials of
// ||b − A x||² (PART_RR, nparts of them) the round goes on while the residual is above SC_TOL x
// SC_NB and the iterations below FL_MAXIT (FL_OACT; SC_NBI = the norm; ACT0 of the inner solve),
// and then, with `fold`, r32 is folded into the quarters.  Off unless *gate.
struct RefFold {
  FoldPass f;
  const int *gate;
  int nparts, fold;
  PcgDev d;
};
hipError_t launch_ref_fold(const RefFold &r, hipStream_t s);
// unfold: the quarters of the inverse x-y transforms back into a node-shaped FP64 vector, with the
// partials of r·z (dot 1) and r·zold (dot 2) in PART_RZ
struct UnfoldPass {
  const float *in;
  Quarters q;
  View out[3], r[3], zold[3];
  int dot;
  int f32;         // out, r and zold are Float32 (an inner solve), else FP64
  PcgDev d;
  const double *in64;  // FP64 quarters (the FP64 separable solve; then in is unused)
  int addth;           // out = th + the value (φ = θ + the solve)
  View th[3];
};
hipError_t launch_unfold(const UnfoldPass &u, const int *gate, hipStream_t s);
int fold_blocks(const Quarters &q);
// the z pass, projector and inverse z pass on modes held in quarters, in place
struct ZModesQ {
  float *modes;
  Quarters q;
  const float *zb[4];    // the padded z bases (ZQ_BASES): V and U by rows (position, mode), then by columns
  const float *sx, *sy, *sz;
  int nx, ny, nz;
  double a;
  int waves;             // k_zmodes_q: waves per block, each a group of z modes (4 or 8); 0: k_zmodes_mf (matrix cores)
  const float *zmf;      // k_zmodes_mf's padded image of the bases (zmodes_mf_image(nz) floats)
};
hipError_t launch_zmodes_q(const ZModesQ &m, const int *gate, hipStream_t s);
// the floats of k_zmodes_mf's image: Vt (mz, n), Ut (mz, n), V (n, mz), U (n, mz), each NZP = 16 ceil(nz / 16)
// rows of NZP + 4 (zero beyond nz)
int zmodes_mf_image(int nz);
// the same in FP64 (the FP64 separable solve): the bases by rows and columns as ZQ_BASES, double
struct ZModesQ64 {
  double *modes;
  Quarters q;
  const double *zb[4];
  const double *sx, *sy, *sz;
  int nx, ny, nz;
  double a;
  int waves;             // 4 or 8
};
hipError_t launch_zmodes_q64(const ZModesQ64 &m, hipStream_t s);

// The x-y transforms of the twelve quarters (three components by four parities) in one launch:
// out(m, q) = Σ_i Σ_j X(i, m) in(i, j) Y(j, q) per z plane, with the half-size bases given
// transposed and zero-padded to whole 16 x 16 tiles: XT[m QS_x + i] = X(i, m), YT[q QS_y + j] =
// Y(j, q), QS = 16 ceil(L/16) for a basis of L = e rows.  A block per (quarter, plane), a
// wavefront per 16 rows of m.  x extents up to 16 QP_MT, y extents up to 16 QP_JT.
constexpr int QP_MT = 10, QP_JT = 6;
// T = float: the Float32 preconditioner's quarters; T = double: the FP64 separable solve's
template <class T>
struct QPassT {
  View in[12], out[12];
  const T *XT[12], *YT[12];
  int e[12][3];
  Conv conv;
  PcgDev d;
  int e2max, mtx, jtx;  // the launch's planes and largest tiles (set by the launcher)
  int parts;            // blocks a quarter plane's m tiles are split over (1: one block a plane; Float32 only)
};
typedef QPassT<float> QPass;
typedef QPassT<double> QPass64;
bool qplane_fits(const QPass &p);
hipError_t launch_qplane(const QPass &p, const int *gate, hipStream_t s);
bool qplane64_fits(const QPass64 &p);
hipError_t launch_qplane64(const QPass64 &p, hipStream_t s);
constexpr int ZQ_NZ_MAX = 64;  // z extents the quarter z pass takes (16 modes a wave at 4 waves a block)
constexpr int ZQ64_NZ_MAX = 48;  // in FP64, with its bases in LDS beside the columns (3 nz 64 + 4 (nz² + nz + 32) doubles)
// The quarter z pass's bases, float, each nz x nz followed by ZQ_SLACK zeros: [0] V(n, mz) at
// n nz + mz, [1] U(n, mz - 1) at n nz + mz (zero for mz = 0 or n = nz - 1), [2] and [3] the same at
// mz nz + n.  The slack covers the reads of the last group past nz and of the prefetched row nz.
constexpr int ZQ_SLACK = 128;

// The z pass, the per-wavevector projector and the inverse z pass in one launch: a block holds
// whole z-columns of modes (floor(256 / nz) columns of nz threads), so the inverse pass reads
// the projected modes from LDS.  nz <= 256.
struct ZModes {
  const void *in;        // compact range vector after the forward x and y passes (T)
  void *out;             // compact range vector before the inverse x and y passes (T)
  const void *Az[2];     // forward z: [0] V_z (X and Y), [1] U_z (Z)
  const void *Ai[2];     // inverse z
  const void *sx, *sy, *sz;
  int off[3];
  int nx, ny, nz;
  double a;
  int columns;           // 1: one block per few whole z-columns (k_zmodes), 0: x-runs (k_zmodes_run) where nz fits
};
template <class T>
hipError_t launch_zmodes(const ZModes &m, const int *gate, hipStream_t s);

// ---------------------------------------------------------------- open boundary (openbc.jl)
// the clock: COUNT (−1 before the first step), RECOMP and FIRST of this step
enum { OF_COUNT = 0, OF_RECOMP, OF_FIRST, OF_N };
hipError_t launch_open_clock(int *ofl, int every, hipStream_t s);

struct OpenDJ {
  const double *h[3];      // hx, hy, hz (node-shaped)
  double *dj[3];           // δJ per link lattice, compact (n1, n2, n3)
  int n[3][3];
  const double *Jbar;      // (Ny+1, Nz+1)
  const double *HJp;       // HJ of the step (device)
  const int *ofl;
  double k2, dx, dy, dz;
};
hipError_t launch_open_dj(const Geo &g, const OpenDJ &a, hipStream_t s);

// One axis of the FFTs: length M = Π rad, twiddles ω^k = e^{−2πik/M}
struct FftAxis {
  int M, nrad;
  int rad[24];
  const cplx *tw;
};
// Lines of length M along one axis: line (l0, l1) starts at l0·l0stride + l1·l1stride, element e at
// e·s.  The first nin elements are read (zeros beyond), the first nout written (times `scale`, the
// real part when out_real).  With ghat, the forward transform is multiplied by Ĝ (real, same line
// geometry) and transformed back.
struct FftPass {
  const void *in;
  int in_real;
  long long in_s, in_l0, in_l1;
  void *out;
  int out_real;
  double scale;
  long long out_s, out_l0, out_l1;
  int nl0, nl1, nin, nout, dir;
  int lines_fast;          // consecutive threads over lines (a line's elements are far apart)
  const double *ghat;
  long long g_s, g_l0, g_l1;
  int lpb;                 // set by the launcher
  FftAxis ax;
};
hipError_t launch_fft(FftPass p, const int *gate, hipStream_t s);

struct OpenGreen {         // fac·G on the padded box M0 x M1 x M2 (real)
  double *out;
  int M[3];
  const double *lgf;
  int R;
  double fac;
};
hipError_t launch_open_green(const OpenGreen &a, hipStream_t s);

struct OpenFace {
  const double *theta[3];  // θ_δ per lattice, compact
  int n[3][3];
  double *hnew[3];
  const int *ofl;
  double dx, dy, dz;
};
hipError_t launch_open_face(const Geo &g, const OpenFace &a, hipStream_t s);

// ---------------------------------------------------------------- normal fluids (exchange.jl)
// The sources of the potentials on the region of the unknowns (compact (Nx-1) x nj x nk, zero
// off the block): TWO, S = C_2/s_2 − C_1/s_1; N, m_k = Σ_µ Pinv[k][µ] (C_µ/s_µ − Σ_ν C_ν)
// (qp_potential!, exchange.jl:344-360, 477-498)
struct QpConv {
  const cplx *G[TDGL_MAX_BANDS];
  const cplx *V[3];
  const double *tau[3];
  double *S[TDGL_MAX_BANDS];
  double wa[3][TDGL_MAX_BANDS];
  double gam[TDGL_MAX_BANDS][TDGL_MAX_BANDS];
  double s[TDGL_MAX_BANDS];
  double Pinv[TDGL_MAX_BANDS][TDGL_MAX_BANDS];
  double rx2, ry2, rz2;
  int nb, mode, K;
  int n1, nj, nk, j0, k0;
};
hipError_t launch_qp_conv(const Geo &g, const QpConv &a, hipStream_t s);

// The z pass, the division by σn(e_x + e_y + e_z) − λ_k and the inverse z pass of each mode
// (_qp_sep!, exchange.jl:295-302; the x and y passes are plane launches)
struct QpZ {
  const double *in;      // K compact region vectors, mode k at k * n
  double *out;
  const double *Wz, *Vz, *ex, *ey, *ez;
  double sigma;
  double lam[TDGL_MAX_BANDS];
  int K, n1, nj, nk;
};
hipError_t launch_qp_zcol(const QpZ &a, hipStream_t s);

// PCG kernels over the box i, j, k in 2..N (the union of the ranges), node-shaped vectors
struct Vec3 {
  double *c[3];
};
struct Vec3f {
  float *c[3];
};
// the Float32 vectors of an inner solve of the refinement (d.pcg_refine): its solution increment
// d, residual r, preconditioned z and z_old, the directions P (new and old) and Q = A P, and D
struct Pcg32 {
  Vec3f d, r, z, zold, Pnew, Pold, Q;
  const float *D[3];
};
struct PcgArgs {
  Geo g;
  Ranges rg;
  PcgDev d;
  int cls;  // the step's class (k_pcg_finish's slot of FL_ITMAXC)
  const double *D[3];
  double sxy, sxz, syx, syz, szx, szy, dydz, dzdx, dxdy;
  Vec3 x, r, b, z, Pold, Pnew, Q, zold, xprev;
  int q;             // iteration of the solve (ap, xr)
  int flexible;
  int warm, extrap, have2;   // start
  int nparts_in;     // partials the kernel consumes
  int kc;            // levels of k per block of the box kernels (box_kc)
  const double *th[3];       // finish: φ = θ + x
  double *ph[3];
  int act_slot;      // finish: the flag that tells an unfinished solve
  int tile;          // pap: the LDS-tiled kernel (k_pcg_pap_tile), else one gather per stencil value
  int zcopy;         // pap, flexible: store zold = z (0: the caller alternates z and zold instead)
  int xcd;           // box kernels: blocks to the XCDs in contiguous runs (box_idx)
  Pcg32 f;           // the Float32 vectors (the refinement's inner solves)
  int gate_fl;       // resid32, xupdate: the flag that gates them
  int norm_only;     // resid32: only the partials of ||b − A x||²
  int ovf_fl;        // finish: the flag that tells an unfinished solve (ACT[act_slot] when 0)
  // D as a palette (pap): byte k of Dcode[node] indexes Dpal[k] (FP64) / Dpal32[k] (Float32); null
  // without (each component of D takes at most DPAL_MAX values: the bar's σn and the padding's σvac)
  const uint32_t *Dcode;
  const double *Dpal;
  const float *Dpal32;
};
constexpr int DPAL_MAX = 16;  // values per component of D in the palette (held in LDS by the operator)
int pcg_box_blocks(const Geo &g, int kc);
int box_kc(const Geo &g);
hipError_t launch_pcg_start(const PcgArgs &a, hipStream_t s);    // nb, warm start or zero source
hipError_t launch_pcg_resid(const PcgArgs &a, hipStream_t s);    // Q = A x, r = b − Q, r·r
hipError_t launch_pcg_p(const PcgArgs &a, hipStream_t s);        // β, P = z + βP, zold = z
hipError_t launch_pcg_ap(const PcgArgs &a, hipStream_t s);       // Q = A P, p·Ap
// Raw buffer loads: the base in scalar registers and a 32-bit byte offset per lane (one address
// register per load instead of two); the hardware returns 0 beyond `bytes`.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const void *base, int bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), (short)0, bytes, 0x00020000);
}
template <class T> __device__ __forceinline__ T buf_load(__amdgpu_buffer_rsrc_t r, int off);
template <> __device__ __forceinline__ double buf_load<double>(__amdgpu_buffer_rsrc_t r, int off) {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
}
template <> __device__ __forceinline__ float buf_load<float>(__amdgpu_buffer_rsrc_t r, int off) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
}
__device__ __forceinline__ double2 buf_load_d2(__amdgpu_buffer_rsrc_t r, int off) {
  return __builtin_bit_cast(double2, __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0));
}
// with a scalar (wave-uniform) byte offset added
template <class T> __device__ __forceinline__ T buf_load_s(__amdgpu_buffer_rsrc_t r, int off, int soff);
template <> __device__ __forceinline__ double buf_load_s<double>(__amdgpu_buffer_rsrc_t r, int off, int soff) {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, off, soff, 0));
}
template <> __device__ __forceinline__ float buf_load_s<float>(__amdgpu_buffer_rsrc_t r, int off, int soff) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, off, soff, 0));
}

// ---------------------------------------------------------------- the ladder's per-cycle measurements
// partial slots of a line block (and, from 0, of the other blocks' partials)
enum { METER_PIN = 0, METER_POUT, METER_TA, METER_TF1, METER_DMAX, METER_NBIG, METER_T,
       METER_TABS = METER_T + TDGL_MAX_BANDS, METER_Q = METER_TABS + TDGL_MAX_BANDS };
// the record's slots (tdgl_cycle_record in this order, nbig as a double)
enum { MR_V = 0, MR_TA, MR_TF1, MR_T, MR_TABS = MR_T + TDGL_MAX_BANDS, MR_DMAX = MR_TABS + TDGL_MAX_BANDS, MR_NBIG,
       MR_IT, MR_IS, MR_DIF, MR_AMAX = MR_DIF + TDGL_MAX_BANDS, MR_DIFFH, MR_N };
struct MeterArgs {
  Geo g;
  int N;
  const cplx *G[TDGL_MAX_BANDS];  // Ψ of the state (the reference's F on the block)
  const cplx *Vx;                 // V on x links
  const double *phx, *hz, *hy;    // the reference's φx, hz, hy
  double *cooper[TDGL_MAX_BANDS]; // |Ψ_µ| of the last cycle, block-compact (x fastest)
  double *hold;                   // hz of the last cycle (node layout)
  double *Dprev[TDGL_MAX_BANDS], *T[TDGL_MAX_BANDS];  // per line: last path phase, unwrapped phase
  const unsigned char *Min, *Mou, *lok, *lf1;
  int vi, vo, nl1, nl2;
  int nin, nou, nlin, nlin_f1;
  int mid;                        // oround(Nx / 2)
  double ks2, dy, dz;
  int first, init;                // first cycle of a rung; init: only take |Ψ| and hz
  double *part;
  int nlb, nnb, nhb;              // line, block-node and hz blocks (then one mid-plane block)
};
int meter_blocks(const MeterArgs &a);
hipError_t launch_meter(const MeterArgs &a, hipStream_t s);
hipError_t launch_meter_finish(const MeterArgs &a, double *rec, double *rung, hipStream_t s);

// ---------------------------------------------------------------------------- batched launches
// One launch runs a kernel over the blocks of many runs: run r owns the blocks off[r] ..
// off[r+1]-1 of a 1D grid, which it takes as its own grid of dims[3r] x dims[3r+1] x dims[3r+2]
// blocks (the grid its single-run launch would have), with its own Geo and argument struct.
enum { BK_STAGE = 0, BK_FIELD, BK_TX, BK_TX2, BK_CURL, BK_GHOSTS, BK_METER, BK_METERFIN, BK_PLANE, BK_ZMODES, BK_FOLD, BK_UNFOLD,
       BK_ZMODESQ, BK_QPLANE, BK_N };
struct BatchTab {
  const int *off;    // n + 1 block offsets
  const int *dims;   // 3 per run
  const Geo *g;      // per run
  const void *args;  // per run: the kind's argument struct
  int n;
};
struct MeterFinArgs {
  MeterArgs a;
  double *rec, *rung;
};
// While a Recorder is set (host thread-locally), the launchers of the kinds above append their
// launch to it instead of launching; any other launcher refuses (hipErrorNotSupported).
struct BatchRec {
  int kind;
  int variant;    // the kernel instance of kinds that have several (the plane and z-mode passes)
  bool has_grid;  // grid and lds as the launcher chose them (else batch_grid, no dynamic LDS)
  dim3 grid;
  size_t lds;
  Geo g;
  std::vector<unsigned char> args;
};
struct Recorder {
  std::vector<BatchRec> recs;
};
extern thread_local Recorder *g_rec;
size_t batch_args_size(int kind);
dim3 batch_grid(int kind, const Geo &g, const void *args);  // the single-run grid of a record
hipError_t launch_batch(int kind, int variant, size_t lds, const BatchTab &t, int blocks, hipStream_t s);
dim3 meter_batch_grid(int kind, const void *args);
hipError_t launch_batch_meter(int kind, const BatchTab &t, int blocks, hipStream_t s);

hipError_t launch_pcg_pap(const PcgArgs &a, hipStream_t s);      // β, P = z + βP, zold = z, Q = A P, p·Ap
hipError_t launch_pcg_xr(const PcgArgs &a, hipStream_t s);       // α, x += αP, r −= αQ, r·r
hipError_t launch_pcg_xr_fold(const PcgArgs &a, const FoldPass &f, hipStream_t s);  // the same, and the fold of r
// the refinement's Float32 inner iteration: pap and x/r update with fold on a.f (x is a.f.d)
hipError_t launch_pcg32_pap(const PcgArgs &a, hipStream_t s);
hipError_t launch_pcg32_xr_fold(const PcgArgs &a, const FoldPass &f, hipStream_t s);
// r32 = b − A x (FP64 operator, rounded) and the partials of ||b − A x||²; gated by fl[gate_fl]
hipError_t launch_pcg_resid32(const PcgArgs &a, hipStream_t s);
// x += d32, gated by fl[gate_fl]
hipError_t launch_pcg_xupdate(const PcgArgs &a, hipStream_t s);
hipError_t launch_pcg_finish(const PcgArgs &a, hipStream_t s);   // φ = θ + x, the cycle's counts
hipError_t launch_pcg_cycle_reset(PcgDev d, hipStream_t s);

}  // namespace tdgl
