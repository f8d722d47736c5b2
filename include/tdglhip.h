/*
 * tdglhip.h — C ABI of the MI355X ladder-cycle stepper for the N-band TDGL solver (release 2.0).
 *
 * The library replaces the device half of gpu/TDGLGPU/src/TDGLGPU.jl: a caller builds a stepper
 * description from its own Stepper (the coefficients that gtdgl/src/scheme.jl:250-434 derives),
 * uploads a State, steps it n times on the GPU and downloads it.  Every pointer is a plain host
 * pointer; arrays are column-major (Julia order, i fastest) in exactly the shapes of
 * gtdgl/src/state.jl:51-60, complex values interleaved (re, im) as ComplexF64.
 *
 * Grid conventions (gtdgl/src/grid.jl:109-127): Nx, Ny, Nz cells with padding, nodes 1..N+1 per
 * axis, the block x1..x2+1 (nodes), 1-based indices as in the reference.
 *
 * Errors: every entry point returns 0 on success and a negative TDGL_E_* code otherwise; the
 * message of the last failure on the calling thread is tdgl_last_error().  Calls are synchronous:
 * when tdgl_step returns, the device state is final.  One handle must not be used from two
 * threads at once.
 */
#ifndef TDGLHIP_H
#define TDGLHIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TDGL_ABI_VERSION 5
#define TDGL_MAX_BANDS 4
#define TDGL_MAX_STAGES 64

enum {
  TDGL_OK = 0,
  TDGL_E_ARG = -1,         /* a bad argument (sizes, pointers, versions) */
  TDGL_E_UNSUPPORTED = -2, /* an option this version does not run (tdgl_supported says which) */
  TDGL_E_DEVICE = -3,      /* a HIP runtime error */
  TDGL_E_NOMEM = -4        /* device allocation failed */
};

/* Field solve of the step (gtdgl/src/step.jl:25-29). */
enum {
  TDGL_FIELD_EXPLICIT = 0,      /* φ from field_increment! only */
  TDGL_FIELD_SEPARABLE = 1,     /* implicit curl-curl, SepSolver (implicit.jl:62-104) */
  TDGL_FIELD_PCG = 2,           /* implicit with conductivity per link, PCGSolver (implicit.jl:279-446) */
  TDGL_FIELD_EXPLICIT_SCALED = 3 /* explicit, scale_increment! by D (implicit.jl:337-342) */
};

/* Self-field boundary (scheme.jl:414-416, field.jl:189-228): OPEN is the bar field plus h_δ, the
 * free-space field of the box's own currents beyond it (gtdgl/src/openbc.jl). */
enum { TDGL_BC_SLAB = 0, TDGL_BC_BAR = 1, TDGL_BC_OPEN = 2 };

/* Coefficients of one band for one (sub)step: BandCoef (gtdgl/src/scheme.jl:21-28). */
typedef struct {
  double kx, ky, kz; /* dt/(m_i Γ̂ dx_i²) */
  double c;          /* dt/Γ */
  double eps;        /* ε = a − T */
  double beta;       /* β */
  double gamma2;     /* γ̃² of KWT (0 = plain TDGL) */
  double cpl[TDGL_MAX_BANDS]; /* (dt/Γ̂) γ_µν, cpl[µ] = 0 */
} tdgl_bandcoef;

/* The stepper: what Stepper{N} (scheme.jl:167-198) holds and the step reads. */
typedef struct {
  int32_t abi_version;           /* TDGL_ABI_VERSION */
  int32_t nbands;                /* N, 1..TDGL_MAX_BANDS */
  int32_t Nx, Ny, Nz;            /* cells, padding included */
  int32_t x1, x2, y1, y2, z1, z2;
  double dx, dy, dz;
  double dt;

  tdgl_bandcoef bc[TDGL_MAX_BANDS];  /* full-step coefficients (st.bc) */
  tdgl_bandcoef bcs[TDGL_MAX_BANDS]; /* per substep / stage (st.bcs) */
  int32_t nsub[TDGL_MAX_BANDS];      /* Euler substeps (st.nsub) */
  int32_t stages[TDGL_MAX_BANDS];    /* RKL1/RKC1 stages (st.rkl; 0 = Euler) */
  int32_t rkc[TDGL_MAX_BANDS];       /* 1: the stage weights below are RKC1 (RKCCoef), 0: RKL1 */
  double divisor[TDGL_MAX_BANDS];    /* q of the stages (s(s+1)/2 or RKCCoef.q), or nsub */
  double rkc_mu[TDGL_MAX_BANDS][TDGL_MAX_STAGES]; /* RKCCoef.μ (index j-1) */
  double rkc_nu[TDGL_MAX_BANDS][TDGL_MAX_STAGES]; /* RKCCoef.ν */
  double rkc_c[TDGL_MAX_BANDS][TDGL_MAX_STAGES];  /* RKCCoef.c */
  double wa[3][TDGL_MAX_BANDS];      /* supercurrent weight g_µ along x, y, z (st.wa) */

  /* FieldCoef (scheme.jl:149-153) */
  double d1, sxy, sxz, syx, syz, szx, szy, a;
  int32_t field_mode;                /* TDGL_FIELD_* */
  int32_t selffield_bc;              /* TDGL_BC_* */
  int32_t extrap, rotframe, hard;    /* Stepper flags */

  /* PCGSolver (implicit.jl:142-189) */
  double pcg_tol;
  int32_t pcg_maxit;
  int32_t pcg_fast;                  /* PCGFast present (warm extrapolation + flexible β) */
  int32_t pcg_f32;                   /* Float32 separable preconditioner */
  int32_t pcg_extrap;                /* 2x_n − x_{n−1} first guess (with pcg_fast) */

  /* Host arrays, copied at creation.  Masks (grid.jl:135-143) are node-shaped
   * (Nx+1, Ny+1, Nz+1); D are link-shaped; bar profiles as bar_field (field.jl:151-180). */
  const double *f;
  const double *taux, *tauy, *tauz;
  const double *Dx, *Dy, *Dz;        /* field_mode PCG / EXPLICIT_SCALED, else NULL */
  const double *bar_hy_z;            /* (Ny+1, 2) */
  const double *bar_hy_y;            /* (2, Nz) */
  const double *bar_hz_z;            /* (Ny, 2) */
  const double *bar_hz_y;            /* (2, Nz+1) */
  const double *bar_hz_x;            /* (Ny, Nz+1) */
  const double *bar_hy_x;            /* (Ny+1, Nz) */

  /* Normal fluids with a finite interband exchange (qp_exchange, gtdgl/src/exchange.jl):
   * TDGL_QP_NONE (one normal fluid), TDGL_QP_TWO (QPExchange: one potential δ = μ_π − μ_σ, solved
   * from σn∇²δ − λ0 δ = C_π/s_π − C_σ/s_σ) or TDGL_QP_N (QPExchangeN: N potentials through the
   * qp_nmodes = N modes of S⁻¹L, each σn∇²m − λ_k m = r_k).  The unknowns are the nodes
   * i = 2..Nx, j = qp_j0+1..qp_j0+qp_nj, k = qp_k0+1..qp_k0+qp_nk (QPSolve's region:
   * exchange.jl:40-47); their 1D operators' eigenbases (_lap1d, exchange.jl:148-166) are
   * the caller's own: qp_Vx (Nx-1)², qp_Wx = inv(qp_Vx), qp_ex (Nx-1), and likewise for y (qp_nj)
   * and z (qp_nk), column-major.  Uniform σn and λ only (no materials, no sculpted vacuum, no
   * drag). */
  int32_t qp_mode;
  int32_t qp_nmodes;
  int32_t qp_j0, qp_nj, qp_k0, qp_nk;
  double qp_sigma;                        /* σn */
  double qp_s[TDGL_MAX_BANDS];            /* shares s_µ of σn */
  double qp_turn[TDGL_MAX_BANDS];         /* TWO: m_µ dt with m = (−s_π, s_σ) (qp_turn!, :400-412) */
  double qp_lambda[TDGL_MAX_BANDS];       /* TWO: λ0; N: the eigenvalues λ_k of S⁻¹L */
  double qp_P[TDGL_MAX_BANDS][TDGL_MAX_BANDS];    /* N: μ_µ = Σ_k P[µ][k] m_k */
  double qp_Pinv[TDGL_MAX_BANDS][TDGL_MAX_BANDS]; /* N: m_k = Σ_µ Pinv[k][µ] R_µ */
  double qp_gamma[TDGL_MAX_BANDS][TDGL_MAX_BANDS]; /* the Josephson matrix γ_µν (_conv, :316-337) */
  const double *qp_Vx, *qp_Wx, *qp_ex;
  const double *qp_Vy, *qp_Wy, *qp_ey;
  const double *qp_Vz, *qp_Wz, *qp_ez;

  /* Open boundary (selffield_bc = TDGL_BC_OPEN; OpenBC, openbc.jl:109-140): h_δ recomputed every
   * open_every steps from δJ = κ² ∇_h × h − HJ·J_bar and ramped in between.  The bar profiles
   * above are required too. */
  int32_t open_every;
  int32_t open_R;              /* the table covers offsets 0..R per axis (1/4πr beyond) */
  double open_fac;             /* dx³/κ² */
  double open_kappa2;          /* κ² */
  const double *open_lgf;      /* lattice_green(R): (R+1)³, column-major */
  const double *open_Jbar;     /* (Ny+1, Nz+1): κ² curl of the x-face bar field per unit HJ */

  /* Thermal noise (Langevin; gtdgl/src/noise.jl, Stepper's noise / noise_seed / noise_field,
   * scheme.jl:239-244, 336-360, 424-431).  With noise every band gets the kick
   * of each step after its update (noise_kick!, noise.jl:40-65: Euler–Maruyama, the polar KWT kick
   * with its Itô drift), a KWT band on RKC stages its phase kick inside the first stage (SK-ROCK,
   * noise_draw! / noise_apply!, :99-153), a KWT band steps in log-polar form and the rotating frame
   * low-passes its rate (bands.jl:210-211, 237-270, 432-455); with noise_field the link phases get
   * their kick before the implicit solve (field_kick!, :182-198).
   * The normals: noise_host = 0 draws them on the device from the counter-based Philox4x32-10
   * generator (Salmon et al., SC'11) keyed by noise_seed, with the counter (node or link index,
   * step count, band or link family) and the Box–Muller transform; noise_host = 1 takes the
   * caller's own draws (tdgl_noise_draws), e.g. the reference's stream in its loop order. */
  int32_t noise;                         /* 1: the Stepper has noise (at T̃ = 0 too: the scheme changes) */
  int32_t noise_field;                   /* 1: the field kick too */
  int32_t noise_host;                    /* 1: draws from tdgl_noise_draws */
  int32_t noise_sk[TDGL_MAX_BANDS];      /* _skrock(st, µ) (bands.jl:446) */
  double noise_T;                        /* T̃ */
  uint64_t noise_seed;
  double noise_sk_w[TDGL_MAX_BANDS][2];  /* RKCCoef.sk = (ν₁, κ₁) = (sω₁/2, sω₁/ω₀) (scheme.jl:80, 99) */
  double noise_Gamma[TDGL_MAX_BANDS];    /* Γ_µ */
  double noise_sigma_n;                  /* σn of the field kick */

  /* The precision of the PCG (field_mode PCG; an extension): with pcg_refine the FP64 operator,
   * residual and solution sit around inner PCG solves in Float32 (their vectors, operator and the
   * folded Float32 preconditioner): each round solves A d = b − A x to pcg_inner_tol of its starting
   * residual, x += d, until ||b − A x|| <= pcg_tol ||b|| in FP64.  A captured cycle holds pcg_rounds
   * rounds (1..16).  Needs pcg_f32 and the folded preconditioner; 0 keeps the reference's solver. */
  int32_t pcg_refine;
  int32_t pcg_rounds;
  double pcg_inner_tol;
} tdgl_desc;

/* Normal-fluid models (tdgl_desc.qp_mode) */
enum { TDGL_QP_NONE = 0, TDGL_QP_TWO = 1, TDGL_QP_N = 2 };

/* Host views of a State (gtdgl/src/state.jl:51-60), Julia shapes:
 *   G[µ], F[µ]      ComplexF64 (Nx+1, Ny+1, Nz+1)
 *   thx, phx, Jx    Float64 (Nx, Ny+1, Nz+1);  Ux, Vx ComplexF64, same shape
 *   thy, phy, Jy    Float64 (Nx+1, Ny, Nz+1);  Uy, Vy
 *   thz, phz, Jz    Float64 (Nx+1, Ny+1, Nz);  Uz, Vz
 *   hx (Nx+1, Ny, Nz), hy (Nx, Ny+1, Nz), hz (Nx, Ny, Nz+1)          Float64 */
typedef struct {
  double *G[TDGL_MAX_BANDS];
  double *F[TDGL_MAX_BANDS];
  double *thx, *thy, *thz;
  double *phx, *phy, *phz;
  double *Ux, *Uy, *Uz;
  double *Vx, *Vy, *Vz;
  double *Jx, *Jy, *Jz;
  double *hx, *hy, *hz;
} tdgl_state;

/* Stepper-side PCG state that a checkpoint holds (implicit.jl:142-169, _pcg_snap :196-197):
 * x and xprev on the ranges rx = (2:Nx-1, 2:Ny, 2:Nz), ry = (2:Nx, 2:Ny-1, 2:Nz),
 * rz = (2:Nx, 2:Ny, 2:Nz-1), column-major. */
typedef struct {
  double *x[3];
  double *xprev[3];
  int32_t warm, have2, iters;
} tdgl_pcg;

typedef struct {
  int64_t steps;        /* steps taken by this handle */
  int64_t calls;        /* tdgl_step calls (cycles) */
  int64_t pcg_iters;    /* PCG iterations summed over steps */
  int32_t last_iters;   /* iterations of the last solve (PCGSolver.iters) */
  double device_ms;     /* device time of the steps, HIP events */
  int64_t device_bytes; /* device memory held by the handle */
  int64_t graph_cycles; /* tdgl_step calls replayed from a captured graph */
  int64_t rollbacks;    /* captured cycles redone eagerly (a solve needed more iterations) */
  int32_t pcg_budget;   /* PCG iterations per solve in the captured cycle */
  int64_t captures;     /* graphs captured and instantiated (the last four are kept and replayed) */
  int32_t folded;       /* 1: the Float32 preconditioner runs its parity-folded transforms */
} tdgl_stats;

typedef struct tdgl_handle tdgl_handle;

/* Replaces supported(st) (TDGLGPU.jl:39-44): 1 if the description runs, else 0 and the
 * first missing option in why (NUL-terminated, at most why_len bytes). */
int tdgl_supported(const tdgl_desc *desc, char *why, size_t why_len);

/* Replaces GPUStepper(st, backend) (TDGLGPU.jl:1299-1316): device copies of the coefficients,
 * masks and work arrays, on the current HIP device. */
int tdgl_create(const tdgl_desc *desc, tdgl_handle **out);
int tdgl_destroy(tdgl_handle *h);

/* Replaces device_state / device_state! (TDGLGPU.jl:64-74) and host_state! : the whole State. */
int tdgl_upload_state(tdgl_handle *h, const tdgl_state *s);
int tdgl_download_state(tdgl_handle *h, tdgl_state *s);

/* Replaces gpu_steps!(g, gs, n; H, HJ) (TDGLGPU.jl:2065-2071): n steps of step!
 * (gtdgl/src/step.jl:13-35), synchronous. */
int tdgl_step(tdgl_handle *h, int32_t n, double H, double HJ);

/* No counterpart in the reference: tdgl_step for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops).  The same steps and the same results as
 * tdgl_step, launched from the host (no graph is captured or instantiated), with the PCG's
 * reduction partials, scalars and flags in a fixed device array of the library (about 1 MiB per
 * GPU, held until the process ends) instead of the handle's; grids whose reductions need more than
 * its 16384 partials per region are refused with TDGL_E_ARG.  Calls that overlap on one GPU share
 * that array and race: serialise them.  The per-kernel profile must be off. */
int tdgl_step_noalloc(tdgl_handle *h, int32_t n, double H, double HJ);

/* No counterpart in the reference (its kernels have fixed shapes): the library's tuner.  For a PCG
 * stepper it times, on the handle's own state and at (H, HJ), one time step per candidate launch
 * shape and keeps the fastest: the PCG box kernels' k levels per block (box_kc: 1, 2, 4, ... while
 * the reductions' partials fit), the m tiles per block of the one-round x-y transform (2 or 3), and
 * the folded preconditioner's z pass: on the matrix cores (0) or with 4 or 8 waves per block
 * (groups of z modes) on the vector ALUs, the blocks a folded quarter plane's x-y transform is split
 * over (qparts 1 or 2, by its m tiles); then the Float32 preconditioner's path, timed as whole
 * PCG iterations: the parity folds (fold 1) or the full-length transforms (fold 0, where a
 * refinement does not need the folds), and for the latter the z pass over whole z-columns (zcol 1)
 * or over x-runs of modes (zcol 0).  The state (F and the normal-fluid potentials included), the
 * PCG history, the open boundary and the noise position are put back; captured cycles are dropped.
 * box_kc changes how the PCG's block partials fall, the z pass's kernel how its sums are grouped
 * and the folds how the preconditioner rounds (each agrees with the others to the PCG tolerance);
 * the m tiles, the vector ALUs' waves, qparts and zcol change no result.  cand[] lists every candidate
 * timed (us per launch of the kernels it shapes; per PCG iteration for fold and zcol). */
#define TDGL_TUNE_MAX 16
enum { TDGL_TUNE_BOX_KC = 1, TDGL_TUNE_PLANE_MT = 2, TDGL_TUNE_ZMODES_WAVES = 3, TDGL_TUNE_FOLD = 4,
       TDGL_TUNE_ZCOL = 5, TDGL_TUNE_QPARTS = 6 };
typedef struct {
  int32_t kind, value;
  double us;
} tdgl_tune_cand;
typedef struct {
  int32_t box_kc, plane_mt, zmodes_waves;
  int32_t ncand;
  tdgl_tune_cand cand[TDGL_TUNE_MAX];
  int32_t fold, zcol;  /* the preconditioner's path chosen */
  int32_t qparts;      /* blocks a folded quarter plane's x-y transform is split over */
} tdgl_tuning;
int tdgl_autotune(tdgl_handle *h, double H, double HJ, tdgl_tuning *out);

/* Replaces _pcg_up! / _pcg_down! (TDGLGPU.jl:217-224): PCG history in and out. */
int tdgl_get_pcg(tdgl_handle *h, tdgl_pcg *p);
int tdgl_set_pcg(tdgl_handle *h, const tdgl_pcg *p);

/* Stepper-side state of the open boundary that a checkpoint holds (OpenBC count, old, new;
 * open_extra / restore_open!, openbc.jl:109-118): the clock (−1 before the first step) and h_δ of
 * the last two recomputations, h-shaped: hx (Nx+1, Ny, Nz), hy (Nx, Ny+1, Nz), hz (Nx, Ny, Nz+1).
 * Only the face values are meaningful.  TDGL_E_ARG for a stepper without the open boundary. */
typedef struct {
  int32_t count;
  double *hold[3];
  double *hnew[3];
} tdgl_open;
int tdgl_get_open(tdgl_handle *h, tdgl_open *o);
int tdgl_set_open(tdgl_handle *h, const tdgl_open *o);

/* The normal-fluid potentials of the last step, what the host reads from the Stepper after a
 * cycle (TDGLGPU.jl:2136-2140: st.qpx.δ of QPExchange, st.qpx.μ of QPExchangeN): node-shaped
 * (Nx+1, Ny+1, Nz+1) Float64 arrays, set on the region of the unknowns (x = 2..Nx,
 * y = qp_j0+1..qp_j0+qp_nj, z = qp_k0+1..qp_k0+qp_nk) and zero elsewhere.  TDGL_QP_TWO: pot[0] = δ =
 * μ_π − μ_σ (exchange.jl:383-388); TDGL_QP_N: pot[µ] = μ_µ = Σ_k P[µ][k] m_k for every band
 * (exchange.jl:506-512).  Zero before the first step.  TDGL_E_ARG for one normal fluid. */
typedef struct {
  double *pot[TDGL_MAX_BANDS];
} tdgl_qp;
int tdgl_get_qp(tdgl_handle *h, tdgl_qp *q);

/* The noise stream's position: the steps taken with noise since the stream began (0 at
 * tdgl_create).  With noise_seed it is the whole state of the generator, what a checkpoint holds
 * in place of the reference's Xoshiro state (_noise_snap / _noise_back!, noise.jl:200-201;
 * protocol.jl:382, 659), so that resuming is continuing to the bit. */
int tdgl_noise_get(tdgl_handle *h, uint64_t *counter);
int tdgl_noise_set(tdgl_handle *h, uint64_t counter);

/* noise_host = 1: the normals of the next nsteps steps (replacing any not yet used), step after
 * step, each: the N bands' complex node arrays (Nx+1, Ny+1, Nz+1), ξ = n₁ + i n₂ at every node the
 * kick visits (interior, and not a hard f = 0 node; anything elsewhere), then with noise_field the
 * real link arrays (Nx, Ny+1, Nz+1), (Nx+1, Ny, Nz+1), (Nx+1, Ny+1, Nz) with n on the links
 * field_kick! visits.  Column-major, doubles.  The reference draws these from its seeded Xoshiro
 * stream on the host in its loop order (noise.jl:40-198, TDGLGPU.jl:714-870), which is how it
 * runs noise on a GPU: a stream in that order maps onto these arrays one to one.  tdgl_step fails
 * with TDGL_E_ARG for a step without draws. */
int tdgl_noise_draws(tdgl_handle *h, int32_t nsteps, const double *draws);

/* The voltmeter of a ladder (ladder, gtdgl/src/protocol.jl:341-352): the x-link planes vi and vo
 * (1-based, x1 + delta_v and x2 - delta_v) and four masks over the lines (j, k), j in y1..y2+1
 * fastest, k in z1..z2+1, one byte each: the live links of plane vi (Min) and of plane vo (Mou), the
 * lines whose links vi..vo-1 are all live (line_ok), and those of them that are condensed
 * throughout (line_f1); κ² of the model for the mid-plane current. */
typedef struct {
  int32_t vi, vo;
  double kappa2;
  const uint8_t *Min, *Mou, *line_ok, *line_f1;
} tdgl_meter;

/* One cycle's measurements of the ladder, from the state on the device (protocol.jl:416-464; the
 * host reference reads them from the whole downloaded state every cycle):
 *   V = Σ_Mou φx[vo] / |Mou| − Σ_Min φx[vi] / |Min|                              (:416, _plane_sum :148-155)
 *   TA = the mean over line_ok of Σ_{ii=vi}^{vo-1} φx[ii]                          (:417, _theta_A :157-166)
 *   T[µ] = the mean over line_ok of the unwrapped path phase of band µ, Tabs[µ] the sum of its |value| over
 *          line_ok (_crossings_now :193-199), Tf1 the mean of band 1 over line_f1,
 *          dmax and nbig the largest wrapped change of band 1 and the number above 2 since the rung began
 *                                                                          (:418-440, _path_phase! :168-179, _unwrap! :181-189)
 *   It, Is = the current through the mid plane, total and through the sample   (:442, observables.jl:102-114)
 *   dif[µ] = max |(|Ψ_µ| − its value at the last cycle)| over the block, amax = max |Ψ_µ|   (:445-454)
 *   diffh = max |hz − its value at the last cycle|                               (:459-463)
 * Maxima propagate NaN, as Julia's max does (a diverged state shows in amax).
 * `first` starts a rung (m = 1: the path phases become the reference, T, Tf1, dmax and nbig zero). */
typedef struct {
  double V, TA, Tf1;
  double T[TDGL_MAX_BANDS];
  double Tabs[TDGL_MAX_BANDS];
  double dmax;
  int32_t nbig;
  int32_t pad;
  double It, Is;
  double dif[TDGL_MAX_BANDS];
  double amax, diffh;
} tdgl_cycle_record;
/* Uploads the voltmeter and takes |Ψ_µ| on the block and hz of the current state as the last
 * cycle's (Cooper0 and fieldOld, protocol.jl:371-372). */
int tdgl_meter_begin(tdgl_handle *h, const tdgl_meter *m);
/* The record of the cycle that just ended (after tdgl_step); TDGL_E_ARG before tdgl_meter_begin. */
int tdgl_meter_cycle(tdgl_handle *h, int32_t first, tdgl_cycle_record *out);

/* Counters (execution_stats, TDGLGPU.jl:2086-2087, per handle). */
int tdgl_get_stats(tdgl_handle *h, tdgl_stats *st);

/* Per-kernel device time (no counterpart in the reference: the measurement of this library).
 * While enabled, steps are launched from the host and every launch records a pair of HIP events
 * at the kernel's own start and end; the totals are per kernel class (tdgl_kernel_name).
 * Enabling clears them. */
#define TDGL_NKERNELS 15
typedef struct {
  int64_t launches[TDGL_NKERNELS];
  double ms[TDGL_NKERNELS];
} tdgl_profile;
int tdgl_profile_enable(tdgl_handle *h, int32_t on);
int tdgl_profile_get(tdgl_handle *h, tdgl_profile *p);
const char *tdgl_kernel_name(int32_t k);

/* Batched runs (no counterpart in the reference, whose runs are separate processes; the
 * batched and server variants its GPU notes leave for later).  A batch advances many handles of
 * one kind together: every step of every run goes through one launch per kernel over the runs'
 * own grids (uneven sizes), the steps of a call replayed as one captured graph.  The handles keep
 * their state, PCG history and meter; after a batch call each is where its own tdgl_step would
 * have left it, to the bit: each run's kernels are the instances its own step takes.  (Only when
 * the runs differ in the form of their separable solve do all take its single-axis passes, the
 * same values to rounding, ~1e-15 a step.)
 * Supported: the separable field (TDGL_FIELD_SEPARABLE), bar or slab boundary, one normal fluid,
 * no thermal noise; the runs must have the same bands and stage structure.  A handle belongs to at
 * most one live batch, is not used by itself while the batch runs (calls are synchronous) and is
 * destroyed only after the batch: the batch's graph holds the addresses of its arrays.
 * tdgl_batch_step: n steps of every run, run r at (H[r], HJ[r]).  tdgl_batch_meter_cycle: one
 * tdgl_meter_cycle of every run (each after its own tdgl_meter_begin), out[r] for run r. */
typedef struct tdgl_batch tdgl_batch;
#define TDGL_BATCH_MAX 1024
int tdgl_batch_create(tdgl_handle *const *hs, int32_t n, tdgl_batch **out);
int tdgl_batch_step(tdgl_batch *b, int32_t n, const double *H, const double *HJ);
int tdgl_batch_meter_cycle(tdgl_batch *b, int32_t first, tdgl_cycle_record *out);
int tdgl_batch_destroy(tdgl_batch *b);

/* Message of the last failed call on this thread ("" if none). */
const char *tdgl_last_error(void);

/* Number of HIP devices visible (0 if none); selects device `dev` for later tdgl_create. */
int tdgl_device_count(void);
int tdgl_set_device(int32_t dev);

#ifdef __cplusplus
}
#endif

#endif /* TDGLHIP_H */
