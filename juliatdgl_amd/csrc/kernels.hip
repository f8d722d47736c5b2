// kernels.hip — the HIP kernels of one release-2.0 time step on gfx950.
//
// Every kernel keeps the reference's order of operations per node (gtdgl/src/bands.jl,
// field.jl, implicit.jl); the file is compiled with -ffp-contract=off, so products and sums
// round as the CPU step's do.  Transcendentals (sincos, atan2, sqrt) come from the device math
// library.  Node-shaped layout and ranges: common.h.
#include "kernels.h"

namespace tdgl {

thread_local ProfPair *prof_pair = nullptr;
thread_local Recorder *g_rec = nullptr;

static hipError_t record(int kind, const Geo &g, const void *args, size_t bytes, int variant = 0, bool has_grid = false,
                         dim3 grid = dim3(0, 0, 0), size_t lds = 0) {
  BatchRec r;
  r.kind = kind;
  r.variant = variant;
  r.has_grid = has_grid;
  r.grid = grid;
  r.lds = lds;
  r.g = g;
  r.args.assign((const unsigned char *)args, (const unsigned char *)args + bytes);
  g_rec->recs.push_back(std::move(r));
  return hipSuccess;
}

// A kernel body's block: its place in its run's grid and that grid's size
struct Blk {
  int x, y, z, nx, ny, nz;
};
__device__ __forceinline__ Blk hw_blk() {
  return Blk{(int)blockIdx.x, (int)blockIdx.y, (int)blockIdx.z, (int)gridDim.x, (int)gridDim.y, (int)gridDim.z};
}
// the hardware block index to the block's place in a logical order (the plane kernels' (m tile,
// plane, component), the stencils' (x, y, k)): the hardware deals blocks to the 8 XCDs in turn,
// XCD x receives a contiguous run of that order, so blocks that share halo planes share an L2
__device__ __forceinline__ int xcd_run(int hw, int nb) {
  const int q = nb / 8, r = nb % 8, x = hw % 8, s = hw / 8;
  return (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + s;
}
// hw_blk with the blocks in XCD runs of the (x, y, z) order
__device__ __forceinline__ Blk hw_blk_xcd() {
  const int nx = gridDim.x, ny = gridDim.y, nz = gridDim.z;
  const int l = xcd_run((int)blockIdx.x + nx * ((int)blockIdx.y + ny * (int)blockIdx.z), nx * ny * nz);
  return Blk{l % nx, (l / nx) % ny, l / (nx * ny), nx, ny, nz};
}
// The run of a batched launch's block b (the last r with off[r] <= b: runs of no blocks are
// passed over) and the block's place in that run's grid
struct BatchIdx {
  int r;
  Blk b;
};
__device__ __forceinline__ BatchIdx batch_idx(const BatchTab &t, int b = -1) {
  if (b < 0) b = blockIdx.x;
  int lo = 0, hi = t.n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (t.off[mid] <= b)
      lo = mid;
    else
      hi = mid - 1;
  }
  const int l = b - t.off[lo];
  const int nx = t.dims[3 * lo], ny = t.dims[3 * lo + 1], nz = t.dims[3 * lo + 2];
  return BatchIdx{lo, Blk{l % nx, (l / nx) % ny, l / (nx * ny), nx, ny, nz}};
}

// ============================================================================ band stage
// Rate of the frame at a block node: the faces copy the interior (precedence x, y, z), so the
// value at any block node is the one of the nearest interior node (bands.jl:224-232).
__device__ __forceinline__ double om_at(const Geo &g, const double *Om, int i, int j, int k) {
  i = i < g.x1 + 1 ? g.x1 + 1 : (i > g.x2 ? g.x2 : i);
  j = j < g.y1 + 1 ? g.y1 + 1 : (j > g.y2 ? g.y2 : j);
  k = k < g.z1 + 1 ? g.z1 + 1 : (k > g.z2 ? g.z2 : k);
  return Om[g.at(i, j, k)];
}

// The link of family `ax` at node (i, j, k) (array index c) for this stage, by the reference's
// expressions for the stage links, evaluated where the stencil uses it.
__device__ __forceinline__ cplx stage_link(const Geo &g, const StageArgs &a, int ax, int c, int i, int j, int k) {
  const cplx v = a.V[ax][c];
  if (a.lsrc == LS_DIRECT) return v;
  const int di = ax == 0, dj = ax == 1, dk = ax == 2;
  const cplx q = cmul(v, cconj(a.U[ax][c]));
  const double sf = -q.y;
  const bool small = q.x > 0 && fabs(sf) <= 0.05;
  if (a.lsrc == LS_ROTFRAC) {
    if (small) {
      const double t = (sf + sf * sf * sf / 6) * a.lc;
      const double t2 = t * t;
      return cmul(v, mk(1 - t2 / 2, -t * (1 - t2 / 6)));
    }
    return cmul(v, dcis(dangle(q) * a.lc));
  }
  if (a.lsrc == LS_ROTLINKS) {
    const double t = small ? -(sf + sf * sf * sf / 6) : dangle(q);
    const double R = t + (om_at(g, a.Om, i + di, j + dj, k + dk) - om_at(g, a.Om, i, j, k)) * a.dt;
    return cmul(v, scis(R * a.lc));
  }
  // LS_SUBSTEP
  cplx root;
  if (small) {
    const double t = (sf + sf * sf * sf / 6) / a.nsub;
    const double t2 = t * t;
    root = mk(1 - t2 / 2, -t * (1 - t2 / 6));
  } else {
    root = dcis(dangle(q) / a.nsub);
  }
  cplx fr = mk(1.0, 0.0);
  if (a.lframe) fr = scis((om_at(g, a.Om, i + di, j + dj, k + dk) - om_at(g, a.Om, i, j, k)) * a.h);
  cplx l = v;
  for (int m = 2; m <= a.msub; ++m) {
    l = cmul(l, root);
    if (a.lframe) l = cmul(l, fr);
  }
  return l;
}

// ============================================================================ noise
// Philox4x32-10 (Salmon, Moraes, Dror & Shaw, "Parallel random numbers: as easy as 1, 2, 3",
// SC'11): ten rounds of two 32x32->64 products, the key bumped by the Weyl constants between them.
__device__ __forceinline__ void philox10(uint32_t &c0, uint32_t &c1, uint32_t &c2, uint32_t &c3, uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    const uint32_t lo0 = 0xD2511F53u * c0, hi0 = __umulhi(0xD2511F53u, c0);
    const uint32_t lo1 = 0xCD9E8D57u * c2, hi1 = __umulhi(0xCD9E8D57u, c2);
    const uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
    c0 = n0;
    c1 = lo1;
    c2 = n2;
    c3 = lo0;
  }
}

// Two standard normals (Box & Muller, 1958) from the block at counter (q, step, tag) under the
// seed: u in (0, 1) and v in [0, 1) from 52 bits each, √(−2 log u) (cos 2πv, sin 2πv)
__device__ __forceinline__ cplx normal2(uint64_t seed, uint64_t t, uint32_t tag, uint32_t q) {
  uint32_t c0 = q, c1 = (uint32_t)t, c2 = (uint32_t)(t >> 32), c3 = tag;
  philox10(c0, c1, c2, c3, (uint32_t)seed, (uint32_t)(seed >> 32));
  const double e52 = 2.220446049250313e-16;  // 2^-52
  const double u = ((double)(((uint64_t)(c0 >> 6) << 26) | (uint64_t)(c1 >> 6)) + 0.5) * e52;
  const double v = (double)(((uint64_t)(c2 >> 6) << 26) | (uint64_t)(c3 >> 6)) * e52;
  const double r = sqrt(-2.0 * log(u));
  double sn, cs;
  sincos(6.283185307179586 * v, &sn, &cs);
  return mk(r * cs, r * sn);
}

// The slot of the caller's draws for this step, or -1 past them (the host refuses such a step
// before it is launched; this keeps every read inside the buffer regardless)
__device__ __forceinline__ long long noise_slot(const NoiseArgs &z) {
  const uint64_t t = z.ctr[0], b = z.ctr[1];
  return (t >= b && t - b < (uint64_t)z.nslots) ? (long long)(t - b) : -1;
}

// ξ of band µ at node (i, j, k) (array index c) for this step
__device__ __forceinline__ cplx noise_xi(const NoiseArgs &z, const Geo &g, int mu, int c, int i, int j, int k) {
  if (z.host) {
    const long long sl = noise_slot(z);
    if (sl < 0) return mk(0.0, 0.0);
    const double *p = z.hd + sl * z.hstride + (long long)mu * 2 * g.n + 2 * (long long)c;
    return mk(p[0], p[1]);
  }
  const uint32_t q = (uint32_t)((i - 1) + (g.Nx + 1) * ((j - 1) + (g.Ny + 1) * (k - 1)));
  return normal2(z.seed, z.ctr[0], (uint32_t)mu, q);
}

// n at the link of family ax at (i, j, k) (array index c) for this step
__device__ __forceinline__ double noise_link(const NoiseArgs &z, const Geo &g, int ax, int c, int i, int j, int k) {
  if (z.host) {
    const long long sl = noise_slot(z);
    if (sl < 0) return 0.0;
    return z.hd[sl * z.hstride + (long long)z.nb * 2 * g.n + (long long)ax * g.n + c];
  }
  const int e0 = ax == 0 ? g.Nx : g.Nx + 1, e1 = ax == 1 ? g.Ny : g.Ny + 1;
  const uint32_t q = (uint32_t)((i - 1) + e0 * ((j - 1) + e1 * (k - 1)));
  return normal2(z.seed, z.ctr[0], 16u + (uint32_t)ax, q).x;
}

// _kwt_kick (noise.jl:75-87): the polar kick of a KWT node at ψ = Ψⁿ, the Cartesian one where
// |ψ|² or |F|² is under _ROT_MIN2
__device__ __forceinline__ cplx kwt_kick(cplx F, cplx psi, cplx xi, double a, double c, double dr, double g2, double Gm,
                                         double d0) {
  const double r2 = abs2(psi), sk = sqrt(1 + g2 * r2);
  if (!(r2 > 1e-24)) return cadd(F, rmul(a, xi));
  const double sr = sqrt(r2);
  const cplx u = mk(psi.x / sr, psi.y / sr);
  const cplx w = cmul(cconj(u), xi);
  if (r2 > 1e-3 && abs2(F) > 1e-3) {
    const double rho = sqrt(r2), aF = hypot(F.x, F.y);
    const double drho = c * w.x / sqrt(Gm * sk) + d0 / (rho * (sk * sk * sk));
    const double dth = c * w.y * sqrt(sk / Gm) / rho;
    return cmul(mk(F.x / aF, F.y / aF), rmul(aF + drho, dcis(dth)));
  }
  const cplx t = cmul(rmul(c, u), mk(w.x / sqrt(Gm * sk), w.y * sqrt(sk / Gm)));
  return csub(cadd(F, t), rmul(dr * (1 + sk * sk) / (sk * sk * sk), psi));
}

// The kick of one node after its band's last stage: noise_kick! (noise.jl:44-65), or after
// SK-ROCK stages the amplitude part of noise_apply! (:127-153) where the phase went in with them
__device__ __forceinline__ cplx noise_node(const StageArgs &a, const Geo &g, cplx F, int c, int i, int j, int k) {
  const NoiseArgs &z = a.nz;
  const cplx xi = noise_xi(z, g, a.mu, c, i, j, k);
  const double Tdt = z.Tdt, dV = z.dV, Gm = a.nGm, g2 = a.ng2;
  const double aa = sqrt(Tdt / (Gm * dV));
  if (g2 == 0) return cadd(F, rmul(aa, xi));
  const double cc = sqrt(Tdt / dV);
  if (a.nsk && a.Qth[c] != 0.0) {
    const cplx psi = a.G[a.mu][c];
    const double rho = hypot(psi.x, psi.y), sk = sqrt(1 + g2 * (rho * rho));
    const cplx u = mk(psi.x / rho, psi.y / rho);
    const double drho = cc * cmul(cconj(u), xi).x / sqrt(Gm * sk) + Tdt / (2 * dV * Gm) / (rho * (sk * sk * sk));
    const double aF = hypot(F.x, F.y);
    return aF > sqrt(1e-3) ? rmul((aF + drho) / aF, F) : cadd(F, rmul(drho, u));
  }
  return kwt_kick(F, a.G[a.mu][c], xi, aa, cc, Tdt / (2 * dV) * g2 / Gm, g2, Gm, Tdt / (2 * dV * Gm));
}

// What follows the frame term and the combination on a stage's node, in the reference's order:
// SK-ROCK's turn of the first stage (_sk_turn!, bands.jl:487), the turn back of the frame
// (_frame_turn!, :321-327), the normal-fluid turn (qp_turn!) and the kick of the step
__device__ __forceinline__ cplx stage_tail(const Geo &g, const StageArgs &a, int i, int j, int k, int c, cplx p, double om) {
  if (a.sk_on) {
    const double q = a.Qth[c];
    if (q != 0.0) p = cmul(p, dcis(a.sk_post * q));
  }
  if (a.turn) p = cmul(p, dcis(om * a.dt));
  if (a.qp) {
    const int q = (i - 2) + a.qn1 * ((j - a.qj0 - 1) + a.qnj * (k - a.qk0 - 1));
    double ang;
    if (a.qp == TDGL_QP_TWO) {
      ang = -a.qpc[0] * a.qpM[0][q];
    } else {
      double pot = 0.0;
      for (int kk = 0; kk < a.qpK; ++kk) pot += a.qpc[kk] * a.qpM[kk][q];
      ang = -a.dt * pot;
    }
    p = cmul(p, dcis(ang));
  }
  if (a.noise && !(a.hard && a.f[c] == 0.0)) p = noise_node(a, g, p, c, i, j, k);
  return p;
}

// One (sub)stage of band µ on the interior nodes, fused: the band pass (_band_row!,
// bands.jl:71-97) with KWT (:135-140), the rotating-frame rate or term (_frame_rate!
// :213-234 / _frame_term! :274-286), the RKL/RKC combination (_combine! :425-430) and the
// turn back of the frame (_frame_turn! :321-327), in that order, as the reference applies them.
// Stage ghosts: from the second (sub)stage on, the face values of S are the covariant Neumann
// ghosts of S with this stage's links (ghosts!, :524-539), which for the nodes a stencil reads are
// the link times the stencil's own centre: computed here instead of stored.
__device__ __forceinline__ void stage_body(const Geo &g, const StageArgs &a, const Blk bk) {
  const int i = g.x1 + 1 + bk.x * blockDim.x + threadIdx.x;
  const int j = g.y1 + 1 + bk.y * blockDim.y + threadIdx.y;
  const int k = g.z1 + 1 + bk.z;
  if (i > g.x2 || j > g.y2 || k > g.z2) return;
  const int c = g.at(i, j, k);
  const int sx = 1, sy = g.px, sz = g.px * g.py;
  const cplx gc = a.S[c];
  const double fm = a.f[c];
  cplx p;
  if (a.hard && fm == 0.0) {
    p = mk(0.0, 0.0);
  } else {
    const cplx lxp = stage_link(g, a, 0, c, i, j, k), lyp = stage_link(g, a, 1, c, i, j, k),
               lzp = stage_link(g, a, 2, c, i, j, k);
    const cplx lxm = stage_link(g, a, 0, c - sx, i - 1, j, k), lym = stage_link(g, a, 1, c - sy, i, j - 1, k),
               lzm = stage_link(g, a, 2, c - sz, i, j, k - 1);
    cplx sxp, sxm, syp, sym, szp, szm;
    if (a.ghost_fly) {
      sxp = (i == g.x2) ? cmul(cconj(lxp), gc) : a.S[c + sx];
      sxm = (i == g.x1 + 1) ? cmul(lxm, gc) : a.S[c - sx];
      syp = (j == g.y2) ? cmul(cconj(lyp), gc) : a.S[c + sy];
      sym = (j == g.y1 + 1) ? cmul(lym, gc) : a.S[c - sy];
      szp = (k == g.z2) ? cmul(cconj(lzp), gc) : a.S[c + sz];
      szm = (k == g.z1 + 1) ? cmul(lzm, gc) : a.S[c - sz];
    } else {
      sxp = a.S[c + sx];
      sxm = a.S[c - sx];
      syp = a.S[c + sy];
      sym = a.S[c - sy];
      szp = a.S[c + sz];
      szm = a.S[c - sz];
    }
    const cplx g2 = rmul(2.0, gc);
    const cplx tx = rmul(a.kx, cadd(csub(cmul(lxp, sxp), g2), cmul(cconj(lxm), sxm)));
    const cplx ty = rmul(a.ky, cadd(csub(cmul(lyp, syp), g2), cmul(cconj(lym), sym)));
    const cplx tz = rmul(a.kz, cadd(csub(cmul(lzp, szp), g2), cmul(cconj(lzm), szm)));
    const cplx nl = rmul(fm * a.eps - a.beta * abs2(gc), rmul(a.c, gc));
    cplx acc = cadd(cadd(cadd(cadd(gc, tx), ty), tz), nl);
    for (int nu = 0; nu < a.nb; ++nu) {
      if (nu == a.mu) continue;
      acc = cadd(acc, rmul(a.cpl[nu], a.G[nu][c]));
    }
    p = a.gamma2 > 0 ? (a.kwt_rot ? kwt_update_rot(gc, csub(acc, gc), a.gamma2) : kwt_update(gc, csub(acc, gc), a.gamma2))
                     : acc;
  }
  double om = 0.0;
  if (a.frame == FRAME_RATE || a.frame == FRAME_RAW) {
    const double a2 = abs2(gc);
    om = a2 > 1e-3 ? jclamp(cmul(cconj(gc), csub(p, gc)).y / (a2 * a.h), -a.wmax, a.wmax) : 0.0;
    a.Om[c] = om;
    if (a.frame == FRAME_RAW) {
      // the low-passed rate and everything after it: k_frame_smooth
      a.D[c] = p;
      return;
    }
    p = csub(p, imx_mul(om * a.h, gc));
  } else if (a.frame == FRAME_TERM) {
    om = a.Om[c];
    p = csub(p, imx_mul(om * a.h, gc));
  } else if (a.frame == FRAME_TURNTERM) {
    om = a.Om[c];
    p = cmul(p, dcis(-(om * a.h)));
  } else if (a.turn) {
    om = a.Om[c];
  }
  if (a.combine) {
    if (a.logcomb) {
      const cplx A = a.A[c];
      p = abs2(A) > 1e-3 && abs2(p) > 1e-3 ? cmul(A, cexpz(rmul(a.cmu, clogz(cdivz(p, A)))))
                                            : cadd(rmul(a.cmu, p), rmul(1 - a.cmu, A));
    } else {
      p = cadd(rmul(a.cmu, p), rmul(a.cnu, a.A[c]));
    }
  }
  a.D[c] = stage_tail(g, a, i, j, k, c, p, om);
}

__global__ void __launch_bounds__(256) k_stage(Geo g, StageArgs a) { stage_body(g, a, a.xcd ? hw_blk_xcd() : hw_blk()); }
// batched: the whole launch's blocks in XCD runs, so each XCD takes whole runs' neighbouring planes
__global__ void __launch_bounds__(256) k_stage_b(BatchTab t) {
  const BatchIdx q = batch_idx(t, xcd_run((int)blockIdx.x, (int)gridDim.x));
  stage_body(t.g[q.r], ((const StageArgs *)t.args)[q.r], q.b);
}

hipError_t launch_stage(const Geo &g, const StageArgs &a, hipStream_t s) {
  if (g_rec) return record(BK_STAGE, g, &a, sizeof a);
  const int ni = g.x2 - g.x1, nj = g.y2 - g.y1, nk = g.z2 - g.z1;
  if (ni <= 0 || nj <= 0 || nk <= 0) return hipSuccess;
  dim3 b(64, 4, 1), gr(cdiv(ni, 64), cdiv(nj, 4), nk);
  TDGL_LAUNCH(k_stage, gr, b, 0, s, g, a);
  return hipGetLastError();
}

// One low-pass sweep of the frame's rate, Ω ← Ω/2 + (sum of the 6 neighbours)/12 with the faces
// copying the interior (_frame_rate_smooth!, bands.jl:256-262): Om -> Om2, then Om2 -> Om; the
// second sweep also turns the stage's output by e^{−iω̂h} (:266-268) and finishes it (stage_tail).
__global__ void __launch_bounds__(256) k_frame_smooth(Geo g, StageArgs a, int pass) {
  const int i = g.x1 + 1 + blockIdx.x * blockDim.x + threadIdx.x;
  const int j = g.y1 + 1 + blockIdx.y * blockDim.y + threadIdx.y;
  const int k = g.z1 + 1 + blockIdx.z;
  if (i > g.x2 || j > g.y2 || k > g.z2) return;
  const int c = g.at(i, j, k);
  const double *T = pass == 1 ? a.Om : a.Om2;
  const double nb = (om_at(g, T, i + 1, j, k) + om_at(g, T, i - 1, j, k)) +
                    (om_at(g, T, i, j + 1, k) + om_at(g, T, i, j - 1, k)) +
                    (om_at(g, T, i, j, k + 1) + om_at(g, T, i, j, k - 1));
  const double om = 0.5 * T[c] + nb / 12;
  if (pass == 1) {
    a.Om2[c] = om;
    return;
  }
  a.Om[c] = om;
  const cplx p = cmul(a.D[c], dcis(-(om * a.h)));
  a.D[c] = stage_tail(g, a, i, j, k, c, p, om);
}

hipError_t launch_frame_smooth(const Geo &g, const StageArgs &a, hipStream_t s) {
  if (g_rec) return hipErrorNotSupported;
  const int ni = g.x2 - g.x1, nj = g.y2 - g.y1, nk = g.z2 - g.z1;
  if (ni <= 0 || nj <= 0 || nk <= 0) return hipSuccess;
  if (!a.Om || !a.Om2) return hipErrorInvalidValue;
  dim3 b(64, 4, 1), gr(cdiv(ni, 64), cdiv(nj, 4), nk);
  TDGL_LAUNCH(k_frame_smooth, gr, b, 0, s, g, a, 1);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  TDGL_LAUNCH(k_frame_smooth, gr, b, 0, s, g, a, 2);
  return hipGetLastError();
}

// noise_draw! for one SK-ROCK band (noise.jl:108-121) and its first stage's source
// Ψⁿ e^{iν₁δθ} (_sk_turn!, bands.jl:484), on the interior
__global__ void __launch_bounds__(256) k_sk_src(Geo g, SkArgs a) {
  const int i = g.x1 + 1 + blockIdx.x * blockDim.x + threadIdx.x;
  const int j = g.y1 + 1 + blockIdx.y * blockDim.y + threadIdx.y;
  const int k = g.z1 + 1 + blockIdx.z;
  if (i > g.x2 || j > g.y2 || k > g.z2) return;
  const int c = g.at(i, j, k);
  const cplx psi = a.G[c];
  double q = 0.0;
  if (!(a.hard && a.f[c] == 0.0)) {
    const double r2 = abs2(psi);
    if (a.g2 > 0 && r2 > 1e-3) {
      const cplx xi = noise_xi(a.nz, g, a.mu, c, i, j, k);
      const double sk = sqrt(1 + a.g2 * r2), rho = sqrt(r2);
      q = a.c * cmul(cconj(mk(psi.x / rho, psi.y / rho)), xi).y * sqrt(sk / a.Gm) / rho;
    }
  }
  a.Qth[c] = q;
  a.Sp[c] = q == 0.0 ? psi : cmul(psi, dcis(a.nu1 * q));
}

hipError_t launch_sk_src(const Geo &g, const SkArgs &a, hipStream_t s) {
  if (g_rec) return hipErrorNotSupported;
  const int ni = g.x2 - g.x1, nj = g.y2 - g.y1, nk = g.z2 - g.z1;
  if (ni <= 0 || nj <= 0 || nk <= 0) return hipSuccess;
  dim3 b(64, 4, 1), gr(cdiv(ni, 64), cdiv(nj, 4), nk);
  TDGL_LAUNCH(k_sk_src, gr, b, 0, s, g, a);
  return hipGetLastError();
}

// the step count of the noise stream, after all of the step's draws (one lane writes)
__global__ void __launch_bounds__(64) k_noise_tick(uint64_t *ctr) {
  if (threadIdx.x == 0) ctr[0] = ctr[0] + 1;
}

hipError_t launch_noise_tick(uint64_t *ctr, hipStream_t s) {
  if (g_rec) return hipErrorNotSupported;
  TDGL_LAUNCH(k_noise_tick, dim3(1), dim3(64), 0, s, ctr);
  return hipGetLastError();
}

// ============================================================================ field
// J_link = Σ_µ w_µ Im((Ψ̄_µ V) Ψ_µ[+e]) (_jlink, field.jl:38-46)
__device__ __forceinline__ double jlink(const FieldArgs &a, int ax, cplx v, int c, int c2) {
  double acc = a.wa[ax][0] * cmul(cmul(cconj(a.G[0][c]), v), a.G[0][c2]).y;
  for (int mu = 1; mu < a.nb; ++mu) acc += a.wa[ax][mu] * cmul(cmul(cconj(a.G[mu][c]), v), a.G[mu][c2]).y;
  return acc;
}

// Supercurrent (supercurrent!, field.jl:19-36), field increment (field_increment!, :84-102)
// and, per mode, the D scaling of every link (implicit.jl:291-293, 337-342) and the right-hand
// side φₑ − θ of the implicit solve on its ranges (implicit.jl:93-95, 286-288).
__device__ __forceinline__ double block_sum_t(double v, double *sh, int tid);

__device__ __forceinline__ void field_body(const Geo &g, const FieldArgs &a, const Blk bk) {
  __shared__ double sh[256];
  const int tid = threadIdx.x + blockDim.x * threadIdx.y;
  const int i = 1 + bk.x * blockDim.x + threadIdx.x;
  const int j = 1 + bk.y * blockDim.y + threadIdx.y;
  const int k = 1 + bk.z;
  double bb = 0.0;
  if (i <= g.Nx + 1 && j <= g.Ny + 1 && k <= g.Nz + 1) {
    const int Nx = g.Nx, Ny = g.Ny, Nz = g.Nz;
    const int c = g.at(i, j, k);
    const int sy = g.px, sz = g.px * g.py;
    // x links (Nx, Ny+1, Nz+1)
    if (i <= Nx) {
      double ph;
      const bool inr = j >= 2 && j <= Ny && k >= 2 && k <= Nz;
      if (inr) {
        const double J = jlink(a, 0, a.V[0][c], c, c + 1);
        a.J[0][c] = J;
        ph = ((a.th[0][c] + (a.d1 * a.tau[0][c]) * J) - a.sxy * (a.h[2][c] - a.h[2][c - sy])) +
             a.sxz * (a.h[1][c] - a.h[1][c - sz]);
        if (a.nkick) ph += (a.D[0] ? a.kax * sqrt(a.D[0][c]) : a.kax) * noise_link(a.nz, g, 0, c, i, j, k);
        if (a.gather && a.rg.in(0, i, j, k)) {
          const double b = ph - a.th[0][c];
          a.B[0][c] = b;
          bb += b * b;
        }
      } else {
        ph = a.ph[0][c];
      }
      if (a.scale) ph = a.th[0][c] + (ph - a.th[0][c]) / a.D[0][c];
      if (inr || a.scale) a.ph[0][c] = ph;
    }
    // y links (Nx+1, Ny, Nz+1)
    if (j <= Ny) {
      double ph;
      const bool inr = i >= 2 && i <= Nx && k >= 2 && k <= Nz;
      if (inr) {
        const double J = jlink(a, 1, a.V[1][c], c, c + sy);
        a.J[1][c] = J;
        ph = ((a.th[1][c] + (a.d1 * a.tau[1][c]) * J) - a.syz * (a.h[0][c] - a.h[0][c - sz])) +
             a.syx * (a.h[2][c] - a.h[2][c - 1]);
        if (a.nkick) ph += (a.D[1] ? a.kay * sqrt(a.D[1][c]) : a.kay) * noise_link(a.nz, g, 1, c, i, j, k);
        if (a.gather && a.rg.in(1, i, j, k)) {
          const double b = ph - a.th[1][c];
          a.B[1][c] = b;
          bb += b * b;
        }
      } else {
        ph = a.ph[1][c];
      }
      if (a.scale) ph = a.th[1][c] + (ph - a.th[1][c]) / a.D[1][c];
      if (inr || a.scale) a.ph[1][c] = ph;
    }
    // z links (Nx+1, Ny+1, Nz)
    if (k <= Nz) {
      double ph;
      const bool inr = i >= 2 && i <= Nx && j >= 2 && j <= Ny;
      if (inr) {
        const double J = jlink(a, 2, a.V[2][c], c, c + sz);
        a.J[2][c] = J;
        ph = ((a.th[2][c] + (a.d1 * a.tau[2][c]) * J) - a.szx * (a.h[1][c] - a.h[1][c - 1])) +
             a.szy * (a.h[0][c] - a.h[0][c - sy]);
        if (a.nkick) ph += (a.D[2] ? a.kaz * sqrt(a.D[2][c]) : a.kaz) * noise_link(a.nz, g, 2, c, i, j, k);
        if (a.gather && a.rg.in(2, i, j, k)) {
          const double b = ph - a.th[2][c];
          a.B[2][c] = b;
          bb += b * b;
        }
      } else {
        ph = a.ph[2][c];
      }
      if (a.scale) ph = a.th[2][c] + (ph - a.th[2][c]) / a.D[2][c];
      if (inr || a.scale) a.ph[2][c] = ph;
    }
  }
  if (a.part) {
    const double r = block_sum_t(bb, sh, tid);
    if (tid == 0) {
      const int bid = bk.x + bk.nx * (bk.y + bk.ny * bk.z);
      a.part[2 * bid] = r;
      a.part[2 * bid + 1] = 0.0;
    }
  }
}

__global__ void __launch_bounds__(256) k_field(Geo g, FieldArgs a) { field_body(g, a, a.xcd ? hw_blk_xcd() : hw_blk()); }
__global__ void __launch_bounds__(256) k_field_b(BatchTab t) {
  const BatchIdx q = batch_idx(t, xcd_run((int)blockIdx.x, (int)gridDim.x));
  field_body(t.g[q.r], ((const FieldArgs *)t.args)[q.r], q.b);
}

int field_blocks(const Geo &g) { return cdiv(g.Nx + 1, 64) * cdiv(g.Ny + 1, 4) * (g.Nz + 1); }

hipError_t launch_field(const Geo &g, const FieldArgs &a, hipStream_t s) {
  if (a.part && field_blocks(g) > MAX_PART) return hipErrorInvalidValue;
  if (g_rec) return a.part ? hipErrorNotSupported : record(BK_FIELD, g, &a, sizeof a);
  dim3 b(64, 4, 1), gr(cdiv(g.Nx + 1, 64), cdiv(g.Ny + 1, 4), g.Nz + 1);
  TDGL_LAUNCH(k_field, gr, b, 0, s, g, a);
  return hipGetLastError();
}

// Links U = e^{−iφ} (links!, field.jl:105-121), field h = curl φ (curl_h!, :124-137) and the
// field boundary (field_bc!, :189-228: applied H, slab or bar self-field; the later assignment
// of the reference wins where two cover one element).
__device__ __forceinline__ void curl_body(const Geo &g, const CurlArgs &a, const Blk bk) {
  const int i = 1 + bk.x * blockDim.x + threadIdx.x;
  const int j = 1 + bk.y * blockDim.y + threadIdx.y;
  const int k = 1 + bk.z;
  if (i > g.Nx + 1 || j > g.Ny + 1 || k > g.Nz + 1) return;
  const int Nx = g.Nx, Ny = g.Ny, Nz = g.Nz;
  const int c = g.at(i, j, k);
  const int sy = g.px, sz = g.px * g.py;
  const double *phx = a.ph[0], *phy = a.ph[1], *phz = a.ph[2];
  const double H = a.hp[0], HJ = a.hp[1];
  // the point of the open boundary's ramp (after this step's clock)
  const double openf = a.open ? fmin((double)(a.ofl[OF_COUNT] + 1) / (double)a.every, 1.0) : 0.0;
  if (i <= Nx && j >= 2 && j <= Ny && k >= 2 && k <= Nz) a.U[0][c] = dcis(-phx[c]);
  if (i >= 2 && i <= Nx && j <= Ny && k >= 2 && k <= Nz) a.U[1][c] = dcis(-phy[c]);
  if (i >= 2 && i <= Nx && j >= 2 && j <= Ny && k <= Nz) a.U[2][c] = dcis(-phz[c]);
  // hx (Nx+1, Ny, Nz)
  if (j <= Ny && k <= Nz) {
    bool set = false;
    double v = 0.0;
    if (i >= 2 && i <= Nx && j >= 2 && j <= Ny - 1 && k >= 2 && k <= Nz - 1) {
      v = ((phz[c + sy] - phz[c]) - phy[c + sz] + phy[c]) / a.dydz;
      set = true;
    }
    if (i == 1 || i == Nx + 1 || j == 1 || j == Ny) {
      v = 0.0;
      set = true;
    }
    if (a.open && i >= 2 && i <= Nx && (j == 1 || j == Ny || k == 1 || k == Nz)) {
      const double o = a.hold[0][c];
      v = o + openf * (a.hnew[0][c] - o);
      set = true;
    }
    if (set) a.h[0][c] = v;
  }
  // hy (Nx, Ny+1, Nz)
  if (i <= Nx && k <= Nz) {
    bool set = false;
    double v = 0.0;
    if (i >= 2 && i <= Nx - 1 && j >= 2 && j <= Ny && k >= 2 && k <= Nz - 1) {
      v = ((phx[c + sz] - phx[c]) - phz[c + 1] + phz[c]) / a.dzdx;
      set = true;
    }
    if (j == 1 || j == Ny + 1) {
      v = 0.0;
      set = true;
    }
    if (!a.bar) {
      if (k == 1) {
        v = -HJ;
        set = true;
      }
      if (k == Nz) {
        v = +HJ;
        set = true;
      }
    } else {
      if (k == 1 || k == Nz) {
        v = HJ * a.bar_hy_z[(j - 1) + (Ny + 1) * (k == 1 ? 0 : 1)];
        if (k == 1 && k == Nz) v = HJ * a.bar_hy_z[(j - 1) + (Ny + 1)];
        set = true;
      }
      if (j == 1 || j == Ny + 1) {
        v = HJ * a.bar_hy_y[(j == 1 ? 0 : 1) + 2 * (k - 1)];
        set = true;
      }
      if (i == 1 || i == Nx) {
        v = HJ * a.bar_hy_x[(j - 1) + (Ny + 1) * (k - 1)];
        set = true;
      }
    }
    if (a.open && (((k == 1 || k == Nz) && i >= 2 && i <= Nx - 1) || i == 1 || i == Nx)) {
      const double o = a.hold[1][c];
      v += o + openf * (a.hnew[1][c] - o);
    }
    if (set) a.h[1][c] = v;
  }
  // hz (Nx, Ny, Nz+1)
  if (i <= Nx && j <= Ny) {
    bool set = false;
    double v = 0.0;
    if (i >= 2 && i <= Nx - 1 && j >= 2 && j <= Ny - 1 && k >= 2 && k <= Nz) {
      v = ((phy[c + 1] - phy[c]) - phx[c + sy] + phx[c]) / a.dxdy;
      set = true;
    }
    if (k == 1 || k == Nz + 1 || i == 1 || i == Nx || j == 1 || j == Ny) {
      v = H;
      set = true;
    }
    if (a.bar) {
      if (k == 1 || k == Nz + 1) v = H + HJ * a.bar_hz_z[(j - 1) + Ny * (k == 1 ? 0 : 1)];
      if (j == 1 || j == Ny) v = H + HJ * a.bar_hz_y[(j == 1 ? 0 : 1) + 2 * (k - 1)];
      if (i == 1 || i == Nx) v = H + HJ * a.bar_hz_x[(j - 1) + Ny * (k - 1)];
    }
    if (a.open && (((j == 1 || j == Ny) && i >= 2 && i <= Nx - 1) || i == 1 || i == Nx)) {
      const double o = a.hold[2][c];
      v += o + openf * (a.hnew[2][c] - o);
    }
    if (set) a.h[2][c] = v;
  }
}

__global__ void __launch_bounds__(256) k_curl(Geo g, CurlArgs a) { curl_body(g, a, a.xcd ? hw_blk_xcd() : hw_blk()); }
__global__ void __launch_bounds__(256) k_curl_b(BatchTab t) {
  const BatchIdx q = batch_idx(t, xcd_run((int)blockIdx.x, (int)gridDim.x));
  curl_body(t.g[q.r], ((const CurlArgs *)t.args)[q.r], q.b);
}

hipError_t launch_curl(const Geo &g, const CurlArgs &a, hipStream_t s) {
  // every table the boundary reads must be there: a missing one is an error, not a device fault
  if (a.bar && !(a.bar_hy_z && a.bar_hy_y && a.bar_hz_z && a.bar_hz_y && a.bar_hz_x && a.bar_hy_x))
    return hipErrorInvalidValue;
  if (a.open && !(a.ofl && a.hold[0] && a.hold[1] && a.hold[2] && a.hnew[0] && a.hnew[1] && a.hnew[2]))
    return hipErrorInvalidValue;
  if (g_rec) return a.open ? hipErrorNotSupported : record(BK_CURL, g, &a, sizeof a);
  dim3 b(64, 4, 1), gr(cdiv(g.Nx + 1, 64), cdiv(g.Ny + 1, 4), g.Nz + 1);
  TDGL_LAUNCH(k_curl, gr, b, 0, s, g, a);
  return hipGetLastError();
}

// Covariant Neumann ghosts of every band on the six faces of the block with the new links
// (neumann!, field.jl:231-245; ghosts!, bands.jl:524-539).  The passes run x, then y, then z,
// each reading what the earlier left: a face node takes the value of the last pass that writes
// it, computed here by following that chain back to the interior.
__device__ __forceinline__ cplx ghost_value(const Geo &g, const cplx *X, const GhostArgs &a, int i, int j, int k) {
  // the chain of multiplications, applied from the interior outwards
  cplx m[3];
  int nm = 0;
  if (k == g.z1) {
    m[nm++] = a.U[2][g.at(i, j, g.z1)];
    k = g.z1 + 1;
  } else if (k == g.z2 + 1) {
    m[nm++] = cconj(a.U[2][g.at(i, j, g.z2)]);
    k = g.z2;
  }
  if (j == g.y1) {
    m[nm++] = a.U[1][g.at(i, g.y1, k)];
    j = g.y1 + 1;
  } else if (j == g.y2 + 1) {
    m[nm++] = cconj(a.U[1][g.at(i, g.y2, k)]);
    j = g.y2;
  }
  if (i == g.x1) {
    m[nm++] = a.U[0][g.at(g.x1, j, k)];
    i = g.x1 + 1;
  } else if (i == g.x2 + 1) {
    m[nm++] = cconj(a.U[0][g.at(g.x2, j, k)]);
    i = g.x2;
  }
  cplx v = X[g.at(i, j, k)];
  for (int q = nm - 1; q >= 0; --q) v = cmul(m[q], v);
  return v;
}

// Threads over the six faces: face x (two planes j,k), face y (i,k), face z (i,j); edges are
// computed by each face that holds them, to the same value.
__device__ __forceinline__ void ghosts_body(const Geo &g, const GhostArgs &a, const Blk bk) {
  const int ni = g.x2 - g.x1 + 2, nj = g.y2 - g.y1 + 2, nk = g.z2 - g.z1 + 2;
  int t = bk.x * blockDim.x + threadIdx.x;
  const int fx = 2 * nj * nk, fy = 2 * ni * nk, fz = 2 * ni * nj;
  int i, j, k;
  if (t < fx) {
    const int side = t / (nj * nk);
    t -= side * nj * nk;
    j = g.y1 + t % nj;
    k = g.z1 + t / nj;
    i = side ? g.x2 + 1 : g.x1;
  } else if ((t -= fx) < fy) {
    const int side = t / (ni * nk);
    t -= side * ni * nk;
    i = g.x1 + t % ni;
    k = g.z1 + t / ni;
    j = side ? g.y2 + 1 : g.y1;
  } else if ((t -= fy) < fz) {
    const int side = t / (ni * nj);
    t -= side * ni * nj;
    i = g.x1 + t % ni;
    j = g.y1 + t / ni;
    k = side ? g.z2 + 1 : g.z1;
  } else {
    return;
  }
  const int c = g.at(i, j, k);
  for (int mu = 0; mu < a.nb; ++mu) {
    cplx *X = a.F[mu];
    X[c] = ghost_value(g, X, a, i, j, k);
  }
}

__global__ void __launch_bounds__(256) k_ghosts(Geo g, GhostArgs a) { ghosts_body(g, a, hw_blk()); }
__global__ void __launch_bounds__(256) k_ghosts_b(BatchTab t) {
  const BatchIdx q = batch_idx(t);
  ghosts_body(t.g[q.r], ((const GhostArgs *)t.args)[q.r], q.b);
}

hipError_t launch_ghosts(const Geo &g, const GhostArgs &a, hipStream_t s) {
  if (g_rec) return record(BK_GHOSTS, g, &a, sizeof a);
  const int ni = g.x2 - g.x1 + 2, nj = g.y2 - g.y1 + 2, nk = g.z2 - g.z1 + 2;
  const int n = 2 * (nj * nk + ni * nk + ni * nj);
  TDGL_LAUNCH(k_ghosts, dim3(cdiv(n, 256)), dim3(256), 0, s, g, a);
  return hipGetLastError();
}

__global__ void __launch_bounds__(256) k_copies(CopyList c) {
  const int a = blockIdx.y;
  if (a >= c.cnt) return;
  if (c.gate && !*c.gate) return;
  const double *src = c.src[a];
  double *dst = c.dst[a];
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < c.n[a]; t += gridDim.x * blockDim.x) dst[t] = src[t];
}

hipError_t launch_copies(const CopyList &c, hipStream_t s) {
  if (g_rec) return hipErrorNotSupported;
  if (c.cnt <= 0) return hipSuccess;
  int m = 0;
  for (int a = 0; a < c.cnt; ++a) m = c.n[a] > m ? c.n[a] : m;
  const int bx = cdiv(m, 256) < 1024 ? cdiv(m, 256) : 1024;
  TDGL_LAUNCH(k_copies, dim3(bx, c.cnt), dim3(256), 0, s, c);
  return hipGetLastError();
}

// ============================================================================ reductions
// Deterministic sums: a fixed assignment of elements to threads, a fixed LDS tree per block,
// block partials in block order.  A consumer sums a producer's partials itself (every block the
// same value) after the producer's kernel has ended.
constexpr int RED_T = 256;  // threads of every reducing block: four wavefronts

// Σ over a wavefront by xor butterflies: every lane ends with the same bits (each step adds the
// same two values in both lanes of a pair)
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Σ over a block of RED_T threads: the wavefront sums, then the four in a fixed order.  `sh`
// holds at least 8 doubles.
__device__ __forceinline__ double block_sum_t(double v, double *sh, int tid) {
  v = wave_sum(v);
  if ((tid & 63) == 0) sh[tid >> 6] = v;
  __syncthreads();
  const double r = (sh[0] + sh[1]) + (sh[2] + sh[3]);
  __syncthreads();
  return r;
}
__device__ __forceinline__ void block_sum2_t(double &a, double &b, double *sh, int tid) {
  a = wave_sum(a);
  b = wave_sum(b);
  if ((tid & 63) == 0) {
    sh[tid >> 6] = a;
    sh[4 + (tid >> 6)] = b;
  }
  __syncthreads();
  a = (sh[0] + sh[1]) + (sh[2] + sh[3]);
  b = (sh[4] + sh[5]) + (sh[6] + sh[7]);
  __syncthreads();
}

// Σ over the partials [slot][which] of a region, identical in every block.  The loads of a
// thread are issued together, eight at a time (one memory latency for up to 2048 partials).
constexpr int PART_UNROLL = 8;
__device__ __forceinline__ double sum_parts(const double *part, int n, int which, double *sh, int tid) {
  double s = 0.0;
  for (int b0 = 0; b0 < n; b0 += PART_UNROLL * RED_T) {
    double v[PART_UNROLL];
#pragma unroll
    for (int u = 0; u < PART_UNROLL; ++u) {
      v[u] = part[2 * min(b0 + tid + u * RED_T, n - 1) + which];  // clamped: no load waits under a condition
    }
#pragma unroll
    for (int u = 0; u < PART_UNROLL; ++u) s += (b0 + tid + u * RED_T < n) ? v[u] : 0.0;
  }
  return block_sum_t(s, sh, tid);
}
// both partials of every slot
__device__ __forceinline__ void sum_parts2(const double *part, int n, double &s0, double &s1, double *sh, int tid) {
  s0 = 0.0;
  s1 = 0.0;
  for (int b0 = 0; b0 < n; b0 += PART_UNROLL * RED_T) {
    double2 v[PART_UNROLL];
#pragma unroll
    for (int u = 0; u < PART_UNROLL; ++u) {
      v[u] = reinterpret_cast<const double2 *>(part)[min(b0 + tid + u * RED_T, n - 1)];
    }
#pragma unroll
    for (int u = 0; u < PART_UNROLL; ++u) {
      const bool in = b0 + tid + u * RED_T < n;
      s0 += in ? v[u].x : 0.0;
      s1 += in ? v[u].y : 0.0;
    }
  }
  block_sum2_t(s0, s1, sh, tid);
}

__device__ __forceinline__ const double *region(const PcgDev &d, int r) { return d.part + (size_t)r * 2 * MAX_PART; }
__device__ __forceinline__ double *region_w(const PcgDev &d, int r) { return d.part + (size_t)r * 2 * MAX_PART; }

// ============================================================================ separable solve
// out[m along axis] = Σ_n A[n*L + m] in[n along axis] for the three components (blockIdx.z):
// a small GEMM per component, Out(m, o) = Σ_n A(n, m) In(n, o) over the "other" index o of the
// range.  LDS tiles of 32 (m) x 32 (o) outputs with 32-deep chunks of n, each thread a 2 x 2
// register tile; every output sums over n in increasing order (the order of the CPU step).
constexpr int TXM = 32, TXO = 32;
// contraction depth staged per round: 128 (Float32) or 64 (Float64) values of n, 33.8 KB of LDS
template <class TA> struct TxDepth { static constexpr int K = 64; };
template <> struct TxDepth<float> { static constexpr int K = 128; };

// (i0, i1, i2) of (n along the axis, o)
__device__ __forceinline__ void tx_coords(int axis, int e0, int e1, int n, int o, int &i0, int &i1, int &i2) {
  if (axis == 0) {
    i0 = n;
    i1 = o % e1;
    i2 = o / e1;
  } else if (axis == 1) {
    i0 = o % e0;
    i1 = n;
    i2 = o / e0;
  } else {
    i0 = o % e0;
    i1 = o / e0;
    i2 = n;
  }
}
__device__ __forceinline__ int voff(const View &v, int i0, int i1, int i2) { return i0 + v.s1 * i1 + v.s2 * i2; }

template <class TI, class TA, class TO>
__device__ __forceinline__ void tx_body(const TxPass &p, const int *gate, const Blk bk) {
  constexpr int TXK = TxDepth<TA>::K;
  constexpr int NL = TXK * TXM / 256;  // loads per thread of each operand per chunk
  static_assert(TXM == 32 && TXO == 32 && NL * 256 == TXK * TXM, "tile shape");
  __shared__ TA As[TXK][TXM + 1];
  __shared__ TA Is[TXK][TXO + 1];
  __shared__ double sh[8];
  __shared__ int colb[TXO];
  const int tid = threadIdx.x;
  const int bid = bk.x + bk.nx * (bk.y + bk.ny * bk.z);
  if (p.conv.mode != CONV_NONE) {
    // the convergence check that decides whether this solve goes on (implicit.jl:305, 319, 426)
    if (!*p.conv.in) {
      if (bid == 0 && tid == 0) *p.conv.out = 0;
      return;
    }
    const double rr = sum_parts(region(p.d, PART_RR), p.conv.nparts, 0, sh, tid);
    const double tol = p.d.sc[p.conv.inner ? SC_TOLI : SC_TOL], nb = p.d.sc[p.conv.inner ? SC_NBI : SC_NB];
    int active;
    if (p.conv.mode == CONV_RR0)
      active = (nb > 0 && sqrt(rr) > tol * nb) ? 1 : 0;
    else
      active = (sqrt(rr) <= tol * nb || p.d.fl[FL_IT] >= p.d.fl[FL_MAXIT]) ? 0 : 1;
    if (bid == 0 && tid == 0) *p.conv.out = active;
    if (!active) return;
  } else if (gate && !*gate) {
    return;
  }
  const int comp = bk.z;
  const int e0 = p.e[comp][0], e1 = p.e[comp][1], e2 = p.e[comp][2];
  const int axis = p.axis;
  const int L = axis == 0 ? e0 : (axis == 1 ? e1 : e2);
  const int O = (e0 * e1 * e2) / L;
  const int eo = axis == 0 ? e1 : e0;  // the first "other" extent: o = c1 + eo c2
  const int m0 = bk.y * TXM, o0 = bk.x * TXO;
  double d0 = 0.0, d1 = 0.0;
  if (m0 < L && o0 < O) {
    const TI *in = (const TI *)p.in[comp].p;
    const TA *A = (const TA *)p.A[comp];
    const View vin = p.in[comp];
    // where this thread's input elements start and how they step along n
    int ibase, istep, ikk, ioo;
    if (axis == 0) {
      // column oo of the tile starts at colb[oo]; the thread loads n = kk, oo = tid/TXK + r 256/TXK
      if (tid < TXO) {
        const int o = o0 + tid;
        colb[tid] = o < O ? voff(vin, 0, o % e1, o / e1) : -1;
      }
      __syncthreads();
      ikk = tid % TXK;
      ioo = tid / TXK;
      ibase = 0;
      istep = 1;
    } else {
      // one column per thread (oo = tid % 32), n = kk = tid/32 + 8 r
      ioo = tid % TXO;
      ikk = tid / TXO;
      const int o = o0 + ioo;
      const int c1 = o % e0, c2 = o / e0;
      ibase = o < O ? (axis == 1 ? voff(vin, c1, 0, c2) : voff(vin, c1, c2, 0)) : -1;
      istep = axis == 1 ? vin.s1 : vin.s2;
    }
    const int amm = tid % TXM, akk = tid / TXM;
    TA ra[NL];
    TI ri[NL];  // converted when stored to LDS, so no load waits inside the fetch
    auto fetch = [&](int n0) {
#pragma unroll
      for (int r = 0; r < NL; ++r) {
        const int n = n0 + akk + r * (256 / TXM);
        ra[r] = A[min(n, L - 1) * L + min(m0 + amm, L - 1)];  // clamped; padded at the LDS store
      }
#pragma unroll
      for (int r = 0; r < NL; ++r) {
        int n, base;
        if (axis == 0) {
          n = n0 + ikk;
          base = colb[ioo + r * (256 / TXK)];
        } else {
          n = n0 + ikk + r * (256 / TXO);
          base = ibase;
        }
        ri[r] = in[(base >= 0 ? base : 0) + min(n, L - 1) * istep];
      }
    };
    // thread tile: m fastest for the x pass (its output is contiguous along m), o fastest else
    const int tm = axis == 0 ? (tid % 16) * 2 : (tid / 16) * 2;
    const int to = axis == 0 ? (tid / 16) * 2 : (tid % 16) * 2;
    TA acc[2][2] = {{0, 0}, {0, 0}};
    fetch(0);
    for (int n0 = 0; n0 < L; n0 += TXK) {
#pragma unroll
      for (int r = 0; r < NL; ++r) {
        const int na = n0 + akk + r * (256 / TXM);
        As[akk + r * (256 / TXM)][amm] = (na < L && m0 + amm < L) ? ra[r] : TA(0);
        if (axis == 0) {
          const bool ok = n0 + ikk < L && colb[ioo + r * (256 / TXK)] >= 0;
          Is[ikk][ioo + r * (256 / TXK)] = ok ? (TA)ri[r] : TA(0);
        } else {
          const bool ok = n0 + ikk + r * (256 / TXO) < L && ibase >= 0;
          Is[ikk + r * (256 / TXO)][ioo] = ok ? (TA)ri[r] : TA(0);
        }
      }
      __syncthreads();
      if (n0 + TXK < L) fetch(n0 + TXK);  // the next chunk's loads overlap this chunk's products
      const int kmax = (L - n0) < TXK ? (L - n0) : TXK;
#pragma unroll 4
      for (int kk = 0; kk < kmax; ++kk) {
        const TA a0 = As[kk][tm], a1 = As[kk][tm + 1];
        const TA i0 = Is[kk][to], i1 = Is[kk][to + 1];
        acc[0][0] += a0 * i0;
        acc[0][1] += a0 * i1;
        acc[1][0] += a1 * i0;
        acc[1][1] += a1 * i1;
      }
      __syncthreads();
    }
    // epilogue: the coordinates of the four outputs, every load, then the stores
    int qo[2][2];
    bool ok[2][2];
    double tv[2][2] = {}, rv[2][2] = {}, zv[2][2] = {};
    for (int dout = 0; dout < 2; ++dout) {
      const int o = o0 + to + dout;
      const int c1 = o % eo, c2 = o / eo;
      for (int dm = 0; dm < 2; ++dm) {
        const int m = m0 + tm + dm;
        ok[dm][dout] = m < L && o < O;
        const int i0 = axis == 0 ? m : c1, i1 = axis == 0 ? c1 : (axis == 1 ? m : c2), i2 = axis == 2 ? m : c2;
        qo[dm][dout] = voff(p.out[comp], i0, i1, i2);
        if (!ok[dm][dout]) continue;
        if (p.addth) tv[dm][dout] = ((const double *)p.th[comp].p)[voff(p.th[comp], i0, i1, i2)];
        if (p.dot) rv[dm][dout] = ((const double *)p.r[comp].p)[voff(p.r[comp], i0, i1, i2)];
        if (p.dot == 2) zv[dm][dout] = ((const double *)p.zold[comp].p)[voff(p.zold[comp], i0, i1, i2)];
      }
    }
    for (int dout = 0; dout < 2; ++dout)
      for (int dm = 0; dm < 2; ++dm) {
        if (!ok[dm][dout]) continue;
        if (p.addth) {
          ((double *)p.out[comp].p)[qo[dm][dout]] = tv[dm][dout] + (double)acc[dm][dout];
        } else {
          const TO v = (TO)acc[dm][dout];
          ((TO *)p.out[comp].p)[qo[dm][dout]] = v;
          if (p.dot) {
            d0 += rv[dm][dout] * (double)v;
            if (p.dot == 2) d1 += rv[dm][dout] * zv[dm][dout];
          }
        }
      }
  }
  if (p.dot) {
    block_sum2_t(d0, d1, sh, tid);
    if (tid == 0) {
      double *pr = region_w(p.d, PART_RZ);
      pr[2 * bid] = d0;
      pr[2 * bid + 1] = d1;
    }
  }
}

template <class TI, class TA, class TO>
__global__ void __launch_bounds__(256) k_tx(TxPass p, const int *gate) { tx_body<TI, TA, TO>(p, gate, hw_blk()); }
__global__ void __launch_bounds__(256) k_tx_b(BatchTab t) {
  const BatchIdx q = batch_idx(t);
  tx_body<double, double, double>(((const TxPass *)t.args)[q.r], nullptr, q.b);
}

static dim3 tx_grid(const TxPass &p) {
  int mo = 1, ml = 1;
  for (int c = 0; c < 3; ++c) {
    const int L = p.e[c][p.axis];
    const int O = p.e[c][0] * p.e[c][1] * p.e[c][2] / L;
    mo = cdiv(O, TXO) > mo ? cdiv(O, TXO) : mo;
    ml = cdiv(L, TXM) > ml ? cdiv(L, TXM) : ml;
  }
  return dim3(mo, ml, 3);
}
int tx_parts(const TxPass &p) {
  dim3 g = tx_grid(p);
  return (int)(g.x * g.y * g.z);
}

template <class TI, class TA, class TO>
hipError_t launch_tx(const TxPass &p, const int *gate, hipStream_t s) {
  if (tx_parts(p) > MAX_PART) return hipErrorInvalidValue;
  if (g_rec) {
    constexpr bool f64 = sizeof(TI) == 8 && sizeof(TA) == 8 && sizeof(TO) == 8;
    if (!f64 || gate || p.conv.mode != CONV_NONE || p.dot) return hipErrorNotSupported;
    return record(BK_TX, Geo{}, &p, sizeof p);
  }
  TDGL_LAUNCH((k_tx<TI, TA, TO>), tx_grid(p), dim3(256), 0, s, p, gate);
  return hipGetLastError();
}
template hipError_t launch_tx<double, double, double>(const TxPass &, const int *, hipStream_t);
template hipError_t launch_tx<float, float, float>(const TxPass &, const int *, hipStream_t);
template hipError_t launch_tx<double, float, float>(const TxPass &, const int *, hipStream_t);
template hipError_t launch_tx<float, float, double>(const TxPass &, const int *, hipStream_t);

// The per-wavevector projector (I + aK)^{-1} = P_L + P_T/(1 + a|k|²) (implicit.jl:67-78) on the
// modes of one wavevector; x1, y1, z1 are the X, Y, Z coefficients (0 where absent).
template <class T>
__device__ __forceinline__ void project(T k1, T k2, T k3, T a, T &x1, T &y1, T &z1, bool hx, bool hy, bool hz) {
  const T k2s = k1 * k1 + k2 * k2 + k3 * k3;
  const T d = (k1 * x1 + k2 * y1 + k3 * z1) / k2s;
  const T f = T(1) / (T(1) + a * k2s);
  if (hx) x1 = k1 * d + f * (x1 - k1 * d);
  if (hy) y1 = k2 * d + f * (y1 - k2 * d);
  if (hz) z1 = k3 * d + f * (z1 - k3 * d);
}

// The axis-2 pass of the three components and the projector, one thread per mode (mx, my, mz):
// X[mx-1, my, ·], Y[mx, my-1, ·] and Z[mx, my, ·] are transformed along z to the mode, projected
// and written; every output element belongs to exactly one mode.
template <class T>
__device__ __forceinline__ void tx2_modes_body(const Tx2Modes &m, const int *gate, const Blk bk) {
  if (gate && !*gate) return;
  const int t = bk.x * blockDim.x + threadIdx.x;
  const int nx = m.nx, ny = m.ny, nz = m.nz;
  if (t >= nx * ny * nz) return;
  const int mx = t % nx, my = (t / nx) % ny, mz = t / (nx * ny);
  const T *in = (const T *)m.in;
  T *out = (T *)m.out;
  const T *AV = (const T *)m.Az[0], *AU = (const T *)m.Az[1];
  const T k1 = ((const T *)m.sx)[mx], k2 = ((const T *)m.sy)[my], k3 = ((const T *)m.sz)[mz];
  const bool hx = mx > 0, hy = my > 0, hz = mz > 0;
  // X: extents (nx-1, ny, nz); Y: (nx, ny-1, nz); Z: (nx, ny, nz-1).  The columns of an absent
  // component are read at a clamped index and dropped, so every load is unconditional.
  const T *colx = in + m.off[0] + (hx ? mx - 1 : 0) + (nx - 1) * my;
  const T *coly = in + m.off[1] + mx + nx * (hy ? my - 1 : 0);
  const T *colz = in + m.off[2] + mx + nx * my;
  const int stx = (nx - 1) * ny, sty = nx * (ny - 1), stz = nx * ny;
  const int mzu = hz ? mz - 1 : 0;
  T ax = 0, ay = 0, az = 0;
#pragma unroll 16
  for (int n = 0; n < nz - 1; ++n) {
    const T v = AV[n * nz + mz];
    ax += colx[stx * n] * v;
    ay += coly[sty * n] * v;
    az += colz[stz * n] * AU[n * (nz - 1) + mzu];
  }
  {
    const T v = AV[(nz - 1) * nz + mz];
    ax += colx[stx * (nz - 1)] * v;
    ay += coly[sty * (nz - 1)] * v;
  }
  T x1 = hx ? ax : T(0), y1 = hy ? ay : T(0), z1 = hz ? az : T(0);
  const T k2s = k1 * k1 + k2 * k2 + k3 * k3;
  if (k2s != 0) project<T>(k1, k2, k3, (T)m.a, x1, y1, z1, hx, hy, hz);
  if (hx) out[m.off[0] + (mx - 1) + (nx - 1) * (my + ny * mz)] = x1;
  if (hy) out[m.off[1] + mx + nx * ((my - 1) + (ny - 1) * mz)] = y1;
  if (hz) out[m.off[2] + mx + nx * (my + ny * (mz - 1))] = z1;
}

template <class T>
__global__ void __launch_bounds__(256) k_tx2_modes(Tx2Modes m, const int *gate) { tx2_modes_body<T>(m, gate, hw_blk()); }
__global__ void __launch_bounds__(256) k_tx2_modes_b(BatchTab t) {
  const BatchIdx q = batch_idx(t);
  tx2_modes_body<double>(((const Tx2Modes *)t.args)[q.r], nullptr, q.b);
}

template <class T>
hipError_t launch_tx2_modes(const Tx2Modes &m, const int *gate, hipStream_t s) {
  if (g_rec) return (sizeof(T) != 8 || gate) ? hipErrorNotSupported : record(BK_TX2, Geo{}, &m, sizeof m);
  TDGL_LAUNCH((k_tx2_modes<T>), dim3(cdiv(m.nx * m.ny * m.nz, 256)), dim3(256), 0, s, m, gate);
  return hipGetLastError();
}
template hipError_t launch_tx2_modes<double>(const Tx2Modes &, const int *, hipStream_t);
template hipError_t launch_tx2_modes<float>(const Tx2Modes &, const int *, hipStream_t);

// Whether a transform launch runs: the convergence check of a first preconditioner pass
// (implicit.jl:305, 319, 426; it writes the flag of the iterations that follow) or a gate flag.
// Uniform over the block.
__device__ __forceinline__ bool pass_active(const Conv &cv, const int *gate, const PcgDev &d, double *sh, int tid,
                                            int bid) {
  if (cv.mode != CONV_NONE) {
    if (!*cv.in) {
      if (bid == 0 && tid == 0) *cv.out = 0;
      return false;
    }
    const double rr = sum_parts(region(d, PART_RR), cv.nparts, 0, sh, tid);
    const double tol = d.sc[cv.inner ? SC_TOLI : SC_TOL], nb = d.sc[cv.inner ? SC_NBI : SC_NB];
    int active;
    if (cv.mode == CONV_RR0)
      active = (nb > 0 && sqrt(rr) > tol * nb) ? 1 : 0;
    else
      active = (sqrt(rr) <= tol * nb || d.fl[FL_IT] >= d.fl[FL_MAXIT]) ? 0 : 1;
    if (bid == 0 && tid == 0) *cv.out = active;
    return active != 0;
  }
  return !(gate && !*gate);
}

// ---------------------------------------------------------------------------- x and y in one launch
// Per block: rows m0..m0+MS of one plane.  Phase 1 stages IC rows of the input plane (all of y)
// and the matching IC x MS block of X, each thread accumulating RM rows x (y/32) columns of
// S = X' in; S goes to LDS (transposed); phase 2 stages IC rows of Y and accumulates Y' S the
// same way.  The next chunk's loads are in flight while the current one is multiplied.
template <class T> struct PlaneCfg;
template <> struct PlaneCfg<float> { static constexpr int MS = 32, IC = 32, RM = 4; };
template <> struct PlaneCfg<double> { static constexpr int MS = 16, IC = 16, RM = 2; };

// LDS rows are one element longer than the tile (bank-conflict-free staging and S stores)
template <class T>
static size_t plane_lds(int l1) {
  constexpr int MS = PlaneCfg<T>::MS, IC = PlaneCfg<T>::IC;
  const int l1p = (l1 + 31) / 32 * 32;
  return (size_t)(IC * MS + IC * (l1p + 1) + l1p * (MS + 1)) * sizeof(T);
}

// TJ = ceil(y extent / 32) column slots per thread, a compile-time count so that the LDS reads
// of one step of the contraction are issued together
template <class TI, class T, class TO, int TJ>
__device__ __forceinline__ void plane_body(const PlanePass &p, const int *gate, const Blk bk) {
  constexpr int MS = PlaneCfg<T>::MS, IC = PlaneCfg<T>::IC, RM = PlaneCfg<T>::RM;
  static_assert(MS == 8 * RM && 256 % IC == 0 && (IC * MS) % 256 == 0, "plane tile");
  constexpr int NX = IC * MS / 256;     // staged X values per thread
  constexpr int L1P = 32 * TJ;          // staged columns (zero beyond the y extent)
  constexpr int L1S = L1P + 1;          // LDS row of the staged input and of Y
  constexpr int MSS = MS + 1;           // LDS row of S
  constexpr int NB = IC * L1P / 256;    // staged input values per thread (ceil)
  extern __shared__ __align__(16) unsigned char plane_smem[];
  __shared__ double sh[8];
  const int tid = threadIdx.x;
  const int bid = bk.x + bk.nx * (bk.y + bk.ny * bk.z);
  if (!pass_active(p.conv, gate, p.d, sh, tid, bid)) return;
  const int comp = bk.z, plane = bk.y, m0 = bk.x * MS;
  const int e0 = p.e[comp][0], e1 = p.e[comp][1], e2 = p.e[comp][2];
  double d0 = 0.0, d1 = 0.0;
  if (plane < e2 && m0 < e0) {
    T *Xs = (T *)plane_smem;  // [IC][MS]
    T *Bs = Xs + IC * MS;     // [IC][L1S]: input rows (phase 1), rows of Y (phase 2)
    T *St = Bs + IC * L1S;    // [L1P][MSS]: S transposed
    const TI *in = (const TI *)p.in[comp].p + (size_t)p.in[comp].s2 * plane;
    const int s1 = p.in[comp].s1;
    const T *X = (const T *)p.X[comp], *Y = (const T *)p.Y[comp];
    const int rg = tid / 32, lane = tid % 32;
    T rx[NX];
    TI rb[NB];
    // input element (ii, j): ii = tid % IC, j = tid / IC + r (256 / IC), so a row of x is read by
    // consecutive threads
    auto fetch1 = [&](int i0) {  // clamped addresses; the padding is applied at the LDS store
#pragma unroll
      for (int r = 0; r < NX; ++r) {
        const int t = tid + 256 * r, ii = t / MS, mm = t % MS;
        rx[r] = X[min(i0 + ii, e0 - 1) * e0 + min(m0 + mm, e0 - 1)];
      }
#pragma unroll
      for (int r = 0; r < NB; ++r) {
        const int ii = tid % IC, j = tid / IC + r * (256 / IC);
        rb[r] = in[min(i0 + ii, e0 - 1) + s1 * min(j, e1 - 1)];
      }
    };
    T acc[RM][TJ];
#pragma unroll
    for (int u = 0; u < RM; ++u)
#pragma unroll
      for (int t = 0; t < TJ; ++t) acc[u][t] = T(0);
    fetch1(0);
    for (int i0 = 0; i0 < e0; i0 += IC) {
#pragma unroll
      for (int r = 0; r < NX; ++r) {
        const int t = tid + 256 * r, ii = t / MS, mm = t % MS;
        Xs[t] = (i0 + ii < e0 && m0 + mm < e0) ? rx[r] : T(0);
      }
#pragma unroll
      for (int r = 0; r < NB; ++r) {
        const int ii = tid % IC, j = tid / IC + r * (256 / IC);
        if (j < L1P) Bs[ii * L1S + j] = (i0 + ii < e0 && j < e1) ? (T)rb[r] : T(0);
      }
      __syncthreads();
      if (i0 + IC < e0) fetch1(i0 + IC);
      const int kmax = (e0 - i0) < IC ? (e0 - i0) : IC;
#pragma unroll 2
      for (int ii = 0; ii < kmax; ++ii) {
        T xv[RM], bv[TJ];
#pragma unroll
        for (int u = 0; u < RM; ++u) xv[u] = Xs[ii * MS + rg * RM + u];
#pragma unroll
        for (int t = 0; t < TJ; ++t) bv[t] = Bs[ii * L1S + lane + 32 * t];
#pragma unroll
        for (int t = 0; t < TJ; ++t)
#pragma unroll
          for (int u = 0; u < RM; ++u) acc[u][t] += xv[u] * bv[t];
      }
      __syncthreads();
    }
#pragma unroll
    for (int t = 0; t < TJ; ++t)
#pragma unroll
      for (int u = 0; u < RM; ++u) St[(lane + 32 * t) * MSS + rg * RM + u] = acc[u][t];
    // phase 2: out(m, q) = Σ_j Y(j, q) S(m, j); row jj of Y is staged by thread q = tid
    T ry[IC];
    auto fetch2 = [&](int j0) {
#pragma unroll
      for (int r = 0; r < IC; ++r) ry[r] = Y[min(j0 + r, e1 - 1) * e1 + min(tid, e1 - 1)];
    };
#pragma unroll
    for (int u = 0; u < RM; ++u)
#pragma unroll
      for (int t = 0; t < TJ; ++t) acc[u][t] = T(0);
    fetch2(0);
    for (int j0 = 0; j0 < e1; j0 += IC) {
      if (tid < L1P)
#pragma unroll
        for (int r = 0; r < IC; ++r) Bs[r * L1S + tid] = (j0 + r < e1 && tid < e1) ? ry[r] : T(0);
      __syncthreads();
      if (j0 + IC < e1) fetch2(j0 + IC);
      const int kmax = (e1 - j0) < IC ? (e1 - j0) : IC;
#pragma unroll 2
      for (int jj = 0; jj < kmax; ++jj) {
        T sv[RM], bv[TJ];
#pragma unroll
        for (int u = 0; u < RM; ++u) sv[u] = St[(j0 + jj) * MSS + rg * RM + u];
#pragma unroll
        for (int t = 0; t < TJ; ++t) bv[t] = Bs[jj * L1S + lane + 32 * t];
#pragma unroll
        for (int t = 0; t < TJ; ++t)
#pragma unroll
          for (int u = 0; u < RM; ++u) acc[u][t] += sv[u] * bv[t];
      }
      __syncthreads();
    }
    // epilogue: (m, q, plane) for the thread's outputs; every load, then the stores
    const View vo = p.out[comp];
#pragma unroll
    for (int t = 0; t < TJ; ++t) {
      const int q = lane + 32 * t;
      double tv[RM], rv[RM], zv[RM];
      bool ok[RM];
#pragma unroll
      for (int u = 0; u < RM; ++u) {
        const int m = m0 + rg * RM + u;
        ok[u] = m < e0 && q < e1;
        tv[u] = rv[u] = zv[u] = 0.0;
        if (!ok[u]) continue;
        if (p.addth) tv[u] = ((const double *)p.th[comp].p)[voff(p.th[comp], m, q, plane)];
        if (p.dot) rv[u] = ((const double *)p.r[comp].p)[voff(p.r[comp], m, q, plane)];
        if (p.dot == 2) zv[u] = ((const double *)p.zold[comp].p)[voff(p.zold[comp], m, q, plane)];
      }
#pragma unroll
      for (int u = 0; u < RM; ++u) {
        if (!ok[u]) continue;
        const int m = m0 + rg * RM + u;
        const int qo = voff(vo, m, q, plane);
        if (p.addth) {
          ((double *)vo.p)[qo] = tv[u] + (double)acc[u][t];
        } else {
          const TO v = (TO)acc[u][t];
          ((TO *)vo.p)[qo] = v;
          if (p.dot) {
            d0 += rv[u] * (double)v;
            if (p.dot == 2) d1 += rv[u] * zv[u];
          }
        }
      }
    }
  }
  if (p.dot) {
    block_sum2_t(d0, d1, sh, tid);
    if (tid == 0) {
      double *pr = region_w(p.d, PART_RZ);
      pr[2 * bid] = d0;
      pr[2 * bid + 1] = d1;
    }
  }
}

// The Float32 plane pass on the matrix cores: the same two products as k_plane, each as 16 x 16
// tiles of v_mfma_f32_16x16x4f32 (A: lane l holds (m = l%16, k = l/16); B: (k = l/16, n = l%16);
// D: (m = 4(l/16) + r, n = l%16)).  A block's 32 rows x 32 TJ columns are 4 TJ tiles, TJ per
// wavefront.  Sums are grouped by the instruction (four products at a time, rounded once):
// the Float32 preconditioner, whose result the FP64 outer PCG only needs as an approximation.
typedef float v4f __attribute__((ext_vector_type(4)));

// Rows of the staged chunks: 64 while the y extent is at most 128 (fewer dependent load rounds),
// 32 above (the LDS and register budget of the wider tiles)
template <int TJ> struct PlaneMfCfg { static constexpr int IC = TJ <= 4 ? 64 : 32; };

template <class TI, class T, class TO, int TJ>
__global__ void __launch_bounds__(256) k_plane(PlanePass p, const int *gate) { plane_body<TI, T, TO, TJ>(p, gate, hw_blk()); }
template <int TJ>
__global__ void __launch_bounds__(256) k_plane_b(BatchTab t) {
  const BatchIdx q = batch_idx(t);
  plane_body<double, double, double, TJ>(((const PlanePass *)t.args)[q.r], nullptr, q.b);
}

template <class TI, class TO, int TJ>
__global__ void __launch_bounds__(256) k_plane_mf(PlanePass p, const int *gate) {
  constexpr int MS = 32, IC = PlaneMfCfg<TJ>::IC;
  constexpr int NX = IC * MS / 256;     // staged X values per thread
  constexpr int L1P = 32 * TJ;          // staged columns (zero beyond the y extent)
  constexpr int L1S = L1P + 1;          // LDS row of the staged input and of Y
  constexpr int XSS = MS + 1;           // LDS row of X
  constexpr int MSS = MS + 1;           // LDS row of S (transposed)
  constexpr int NB = IC * L1P / 256;    // staged input values per thread
  extern __shared__ __align__(16) unsigned char plane_smem[];
  __shared__ double sh[8];
  const int tid = threadIdx.x;
  const int bid = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
  if (!pass_active(p.conv, gate, p.d, sh, tid, bid)) return;
  const int comp = blockIdx.z, plane = blockIdx.y, m0 = blockIdx.x * MS;
  const int e0 = p.e[comp][0], e1 = p.e[comp][1], e2 = p.e[comp][2];
  double d0 = 0.0, d1 = 0.0;
  if (plane < e2 && m0 < e0) {
    float *Xs = (float *)plane_smem;  // [IC][XSS]
    float *Bs = Xs + IC * XSS;        // [IC][L1S]
    float *St = Bs + IC * L1S;        // [L1P][MSS]
    const TI *in = (const TI *)p.in[comp].p + (size_t)p.in[comp].s2 * plane;
    const int s1 = p.in[comp].s1;
    const float *X = (const float *)p.X[comp], *Y = (const float *)p.Y[comp];
    const int w = tid / 64, l = tid % 64, lr = l % 16, lk = l / 16;
    float rx[NX];
    TI rb[NB];
    float ry[IC];
    // Every load reads a clamped, in-bounds address unconditionally; the padding is applied where
    // the value is stored to LDS (a load under a divergent condition would wait for memory).
    auto fetch1 = [&](int i0) {
#pragma unroll
      for (int r = 0; r < NX; ++r) {
        const int t = tid + 256 * r, ii = t / MS, mm = t % MS;
        rx[r] = X[min(i0 + ii, e0 - 1) * e0 + min(m0 + mm, e0 - 1)];
      }
#pragma unroll
      for (int r = 0; r < NB; ++r) {
        const int ii = tid % IC, j = tid / IC + r * (256 / IC);
        rb[r] = in[min(i0 + ii, e0 - 1) + s1 * min(j, e1 - 1)];
      }
    };
    // row j0 + r of Y, column q = tid (Y does not depend on the data: its first chunk is
    // requested with the first input chunk)
    auto fetch2 = [&](int j0) {
#pragma unroll
      for (int r = 0; r < IC; ++r) ry[r] = Y[min(j0 + r, e1 - 1) * e1 + min(tid, e1 - 1)];
    };
    v4f acc[TJ];
#pragma unroll
    for (int t = 0; t < TJ; ++t) acc[t] = v4f{0.0f, 0.0f, 0.0f, 0.0f};
    fetch1(0);
    fetch2(0);
    for (int i0 = 0; i0 < e0; i0 += IC) {
#pragma unroll
      for (int r = 0; r < NX; ++r) {
        const int t = tid + 256 * r, ii = t / MS, mm = t % MS;
        Xs[ii * XSS + mm] = (i0 + ii < e0 && m0 + mm < e0) ? rx[r] : 0.0f;
      }
#pragma unroll
      for (int r = 0; r < NB; ++r) {
        const int ii = tid % IC, j = tid / IC + r * (256 / IC);
        Bs[ii * L1S + j] = (i0 + ii < e0 && j < e1) ? (float)rb[r] : 0.0f;
      }
      __syncthreads();
      if (i0 + IC < e0) fetch1(i0 + IC);
#pragma unroll 2
      for (int k0 = 0; k0 < IC; k0 += 4) {
        const float a0 = Xs[(k0 + lk) * XSS + lr], a1 = Xs[(k0 + lk) * XSS + 16 + lr];
#pragma unroll
        for (int s = 0; s < TJ; ++s) {
          const int t = w * TJ + s, mt = t & 1, jt = t >> 1;
          const float b = Bs[(k0 + lk) * L1S + jt * 16 + lr];
          acc[s] = __builtin_amdgcn_mfma_f32_16x16x4f32(mt ? a1 : a0, b, acc[s], 0, 0, 0);
        }
      }
      __syncthreads();
    }
    // S(m, j) to LDS, transposed: lane l holds m = mt 16 + 4 lk + r, j = jt 16 + lr
#pragma unroll
    for (int s = 0; s < TJ; ++s) {
      const int t = w * TJ + s, mt = t & 1, jt = t >> 1;
#pragma unroll
      for (int r = 0; r < 4; ++r) St[(jt * 16 + lr) * MSS + mt * 16 + 4 * lk + r] = acc[s][r];
      acc[s] = v4f{0.0f, 0.0f, 0.0f, 0.0f};
    }
    // phase 2: out(m, q) = Σ_j S(m, j) Y(j, q); row jj of Y is staged by thread q = tid
    for (int j0 = 0; j0 < e1; j0 += IC) {
      if (tid < L1P)
#pragma unroll
        for (int r = 0; r < IC; ++r) Bs[r * L1S + tid] = (j0 + r < e1 && tid < e1) ? ry[r] : 0.0f;
      __syncthreads();
      if (j0 + IC < e1) fetch2(j0 + IC);
      // rows of S at or beyond the staged columns are zero
#pragma unroll 2
      for (int k0 = 0; k0 < IC; k0 += 4) {
        const int jr = j0 + k0 + lk;
        const float a0 = jr < L1P ? St[jr * MSS + lr] : 0.0f, a1 = jr < L1P ? St[jr * MSS + 16 + lr] : 0.0f;
#pragma unroll
        for (int s = 0; s < TJ; ++s) {
          const int t = w * TJ + s, mt = t & 1, qt = t >> 1;
          const float b = Bs[(k0 + lk) * L1S + qt * 16 + lr];
          acc[s] = __builtin_amdgcn_mfma_f32_16x16x4f32(mt ? a1 : a0, b, acc[s], 0, 0, 0);
        }
      }
      __syncthreads();
    }
    // epilogue: every load, then the stores
    const View vo = p.out[comp];
#pragma unroll
    for (int s = 0; s < TJ; ++s) {
      const int t = w * TJ + s, mt = t & 1, qt = t >> 1;
      const int q = qt * 16 + lr;
      double tv[4], rv[4], zv[4];
      bool ok[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = m0 + mt * 16 + 4 * lk + r;
        ok[r] = m < e0 && q < e1;
        tv[r] = rv[r] = zv[r] = 0.0;
        if (!ok[r]) continue;
        if (p.addth) tv[r] = ((const double *)p.th[comp].p)[voff(p.th[comp], m, q, plane)];
        if (p.dot) rv[r] = ((const double *)p.r[comp].p)[voff(p.r[comp], m, q, plane)];
        if (p.dot == 2) zv[r] = ((const double *)p.zold[comp].p)[voff(p.zold[comp], m, q, plane)];
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        if (!ok[r]) continue;
        const int m = m0 + mt * 16 + 4 * lk + r;
        const int qo = voff(vo, m, q, plane);
        if (p.addth) {
          ((double *)vo.p)[qo] = tv[r] + (double)acc[s][r];
        } else {
          const TO v = (TO)acc[s][r];
          ((TO *)vo.p)[qo] = v;
          if (p.dot) {
            d0 += rv[r] * (double)v;
            if (p.dot == 2) d1 += rv[r] * zv[r];
          }
        }
      }
    }
  }
  if (p.dot) {
    block_sum2_t(d0, d1, sh, tid);
    if (tid == 0) {
      double *pr = region_w(p.d, PART_RZ);
      pr[2 * bid] = d0;
      pr[2 * bid + 1] = d1;
    }
  }
}

// Planes of at most 32 XC x 32 TJ elements in one memory round.  A block of eight wavefronts
// holds MT 16-row tiles of m of one plane: the whole input plane and its MS columns of X are
// requested before anything waits (after the gate flag; the convergence sum of a first pass is
// taken while they are in flight), at most 63 loads per thread (what one wavefront can have in
// flight); Y is requested next and arrives during the first product.  Both are staged as Float32 in LDS k-contiguous, so that one 16-byte read gives a lane
// its operands of four v_mfma_f32_16x16x4f32 steps (group g, step u, lane k index lk: row
// 16 g + 4 lk + u; the four products of one instruction are rounded once, as in k_plane_mf, but
// grouped differently).  LDS rows are 4 (mod 8) words long: the 16-byte reads of 16 lanes are
// bank-conflict free.  The next group's operands are read while the current group multiplies.
// Blocks go to the XCDs in contiguous runs of the (m tile, plane, component) order, so one
// plane's blocks share an L2.
constexpr int PLANE1_T = 512;
// LDS rows of 4 (mod 8) 32-bit words: E0 + 4 floats or E0 + 2 doubles
template <int XC, int TJ, int MT, class TC> struct PlaneMf1 {
  static constexpr int PADE = 16 / (int)sizeof(TC);
  static constexpr int E0 = 32 * XC, L1P = 32 * TJ, MS = 16 * MT;
  static constexpr int E0S = E0 + PADE, JS = L1P + PADE;
  static constexpr int NJ = 2 * TJ, NT = MT * NJ, S = (NT + 7) / 8;  // j tiles, tiles, tiles per wave
  static constexpr int NB = E0 * L1P / PLANE1_T, NX = E0 * MS / PLANE1_T, NY = L1P * L1P / PLANE1_T;
  static constexpr int BROW = E0S > JS ? E0S : JS;
  // Xt [MS][E0S] | Bt [L1P][E0S], later Yt [L1P][JS] | St [MS][JS]
  static constexpr int BOFF = MS * E0S, SOFF = BOFF + L1P * BROW;
  static constexpr size_t lds = (size_t)(SOFF + MS * JS) * sizeof(TC);
  static_assert(NB + NX <= 63 && NY <= 63, "staging loads per thread");
};
static size_t plane1_lds(int xc, int tj, int mt, int es) {
  const int pade = 16 / es, e0s = 32 * xc + pade, js = 32 * tj + pade, ms = 16 * mt;
  const int brow = e0s > js ? e0s : js;
  return (size_t)(ms * e0s + 32 * tj * brow + ms * js) * es;
}
constexpr size_t PLANE1_LDS_MAX = 160 * 1024;
constexpr int PLANE1_XC = 5, PLANE1_TJ = 3;  // planes up to 160 x 96



// sum_parts in a block of eight wavefronts: the first four load and sum as a block of RED_T
// threads does (the same value, bit for bit, as every other reader of the partials)
__device__ __forceinline__ double sum_parts_8w(const double *part, int n, int which, double *sh, int tid) {
  double s = 0.0;
  if (tid < RED_T) {
    const __amdgpu_buffer_rsrc_t rp = buf_rsrc(part, 16 * n);
    for (int b0 = 0; b0 < n; b0 += PART_UNROLL * RED_T) {
      double v[PART_UNROLL];
#pragma unroll
      for (int u = 0; u < PART_UNROLL; ++u) v[u] = buf_load<double>(rp, 8 * (2 * min(b0 + tid + u * RED_T, n - 1) + which));
#pragma unroll
      for (int u = 0; u < PART_UNROLL; ++u) s += (b0 + tid + u * RED_T < n) ? v[u] : 0.0;
    }
  }
  s = wave_sum(s);
  if ((tid & 63) == 0) sh[tid >> 6] = s;
  __syncthreads();
  const double r = (sh[0] + sh[1]) + (sh[2] + sh[3]);
  __syncthreads();
  return r;
}
// both sums of eight wavefronts in a fixed order
__device__ __forceinline__ void block_sum2_8w(double &a, double &b, double *sh, int tid) {
  a = wave_sum(a);
  b = wave_sum(b);
  if ((tid & 63) == 0) {
    sh[tid >> 6] = a;
    sh[8 + (tid >> 6)] = b;
  }
  __syncthreads();
  a = ((sh[0] + sh[1]) + (sh[2] + sh[3])) + ((sh[4] + sh[5]) + (sh[6] + sh[7]));
  b = ((sh[8] + sh[9]) + (sh[10] + sh[11])) + ((sh[12] + sh[13]) + (sh[14] + sh[15]));
  __syncthreads();
}

typedef double v4d __attribute__((ext_vector_type(4)));
typedef double v2d __attribute__((ext_vector_type(2)));
// the 16x16x4 matrix instruction of each type: A lane l holds (m = l % 16, k = l / 16), B lane l
// holds (k = l / 16, n = l % 16); the accumulator's element r of lane l is row mrow(l / 16, r),
// column l % 16 (FP32: 4 (l / 16) + r; FP64: 4 r + l / 16)
template <class TC> struct Mf;
template <> struct Mf<float> {
  typedef v4f v4;
  static __device__ __forceinline__ v4 mma(float a, float b, v4 c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
  static __device__ __forceinline__ v4 ld4(const float *p) { return *(const v4f *)p; }
  static __device__ __forceinline__ int mrow(int lk, int r) { return 4 * lk + r; }
};
template <> struct Mf<double> {
  typedef v4d v4;
  static __device__ __forceinline__ v4 mma(double a, double b, v4 c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
  static __device__ __forceinline__ v4 ld4(const double *p) {
    const v2d lo = *(const v2d *)p, hi = *(const v2d *)(p + 2);
    return v4d{lo.x, lo.y, hi.x, hi.y};
  }
  static __device__ __forceinline__ int mrow(int lk, int r) { return 4 * r + lk; }
};

// acc[s] += A_s' B_s over G groups of 16 rows for the first NS tiles of a wave (operands in LDS
// at Ab + ao[s], Bb + bo[s], 16 rows further per group); the next group's operands are read
// while the current group multiplies
template <int NS, int S, class TC>
__device__ __forceinline__ void mf_groups(typename Mf<TC>::v4 (&acc)[S], const TC *Ab, const int (&ao)[S], const TC *Bb,
                                          const int (&bo)[S], int G) {
  typedef typename Mf<TC>::v4 V4;
  V4 A[NS], B[NS];
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    A[s] = Mf<TC>::ld4(Ab + ao[s]);
    B[s] = Mf<TC>::ld4(Bb + bo[s]);
  }
  for (int g = 0; g < G; ++g) {
    const int gn = min(g + 1, G - 1) * 16;
    V4 An[NS], Bn[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      An[s] = Mf<TC>::ld4(Ab + ao[s] + gn);
      Bn[s] = Mf<TC>::ld4(Bb + bo[s] + gn);
    }
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int s = 0; s < NS; ++s) acc[s] = Mf<TC>::mma(A[s][u], B[s][u], acc[s]);
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      A[s] = An[s];
      B[s] = Bn[s];
    }
  }
}

// mf_groups for the ns tiles of a wave (0 <= ns <= S, at least S - 2; wave-uniform)
template <int S, class TC>
__device__ __forceinline__ void mf_slots(int ns, typename Mf<TC>::v4 (&acc)[S], const TC *Ab, const int (&ao)[S],
                                         const TC *Bb, const int (&bo)[S], int G) {
  if (ns >= S) {
    mf_groups<S, S, TC>(acc, Ab, ao, Bb, bo, G);
  } else if constexpr (S >= 2) {
    if (ns == S - 1) {
      mf_groups<S - 1, S, TC>(acc, Ab, ao, Bb, bo, G);
    } else if constexpr (S >= 3) {
      if (ns == S - 2) mf_groups<S - 2, S, TC>(acc, Ab, ao, Bb, bo, G);
    }
  }
}

template <class TI, class TC, class TO, int XC, int TJ, int MT>
__device__ __forceinline__ void plane_mf1_body(const PlanePass &p, const int *gate, int gx, int gy, int hwb, int nb) {
  using C = PlaneMf1<XC, TJ, MT, TC>;
  typedef typename Mf<TC>::v4 V4;
  constexpr int NTH = PLANE1_T, MS = C::MS, E0 = C::E0, L1P = C::L1P, E0S = C::E0S, JS = C::JS, S = C::S;
  extern __shared__ __align__(16) unsigned char plane_smem[];
  __shared__ double sh[16];
  const int tid = threadIdx.x;
  const int bid = xcd_run(hwb, nb);
  const int mb = bid % gx, plane = (bid / gx) % gy, comp = bid / (gx * gy), m0 = mb * MS;
  // the gate flag first: a pass the gate turns off loads nothing
  if (p.conv.mode != CONV_NONE) {
    if (!*p.conv.in) {
      if (bid == 0 && tid == 0) *p.conv.out = 0;
      return;
    }
  } else if (gate && !*gate) {
    return;
  }
  const int e0 = p.e[comp][0], e1 = p.e[comp][1], e2 = p.e[comp][2];
  const int pl = min(plane, e2 - 1);  // blocks beyond this component's planes load from its last
  const int s1 = p.in[comp].s1;
  const __amdgpu_buffer_rsrc_t rin = buf_rsrc((const TI *)p.in[comp].p + (size_t)p.in[comp].s2 * pl,
                                              (s1 * (e1 - 1) + e0) * (int)sizeof(TI));
  constexpr int ES = (int)sizeof(TC);
  const __amdgpu_buffer_rsrc_t rX = buf_rsrc(p.XT[comp], e0 * e0 * ES), rY = buf_rsrc(p.YT[comp], e1 * e1 * ES);
  TI rb[C::NB];
  TC rx[C::NX], ry[C::NY];
#pragma unroll
  for (int r = 0; r < C::NB; ++r) {
    const int t = tid + NTH * r, ii = t % E0, j = t / E0;
    rb[r] = buf_load<TI>(rin, (min(ii, e0 - 1) + s1 * min(j, e1 - 1)) * (int)sizeof(TI));
  }
#pragma unroll
  for (int r = 0; r < C::NX; ++r) {
    const int t = tid + NTH * r, ii = t % E0, mm = t / E0;
    rx[r] = buf_load<TC>(rX, (min(m0 + mm, e0 - 1) * e0 + min(ii, e0 - 1)) * ES);
  }
  if (p.conv.mode != CONV_NONE) {
    const double rr = sum_parts_8w(region(p.d, PART_RR), p.conv.nparts, 0, sh, tid);
    const double tol = p.d.sc[p.conv.inner ? SC_TOLI : SC_TOL], nrm = p.d.sc[p.conv.inner ? SC_NBI : SC_NB];
    const int active = p.conv.mode == CONV_RR0 ? ((nrm > 0 && sqrt(rr) > tol * nrm) ? 1 : 0)
                                               : ((sqrt(rr) <= tol * nrm || p.d.fl[FL_IT] >= p.d.fl[FL_MAXIT]) ? 0 : 1);
    if (bid == 0 && tid == 0) *p.conv.out = active;
    if (!active) return;
  }
  double d0 = 0.0, d1 = 0.0;
  if (plane < e2 && m0 < e0) {
    TC *Xt = (TC *)plane_smem, *Bt = Xt + C::BOFF, *St = Xt + C::SOFF;
#pragma unroll
    for (int r = 0; r < C::NX; ++r) {
      const int t = tid + NTH * r, ii = t % E0, mm = t / E0;
      Xt[mm * E0S + ii] = (ii < e0 && m0 + mm < e0) ? rx[r] : TC(0);
    }
#pragma unroll
    for (int r = 0; r < C::NB; ++r) {
      const int t = tid + NTH * r, ii = t % E0, j = t / E0;
      Bt[j * E0S + ii] = (ii < e0 && j < e1) ? (TC)rb[r] : TC(0);
    }
    // Y, for phase 2: in flight during phase 1
#pragma unroll
    for (int r = 0; r < C::NY; ++r) {
      const int t = tid + NTH * r, jj = t % L1P, q = t / L1P;
      ry[r] = buf_load<TC>(rY, (min(q, e1 - 1) * e1 + min(jj, e1 - 1)) * ES);
    }
    __syncthreads();
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6), l = tid % 64, lr = l % 16, lk = l / 16;
    // the tiles: MT m tiles by the 16-column j tiles the y extent needs; wave w takes t = w + 8 s,
    // (mt, jt) = (t mod MT, t / MT), ns of them (wave-uniform)
    const int NTr = MT * ((e1 + 15) / 16);
    const int ns = w < NTr ? min(S, (NTr - w + 7) / 8) : 0;
    int ao[S], bo[S];
#pragma unroll
    for (int s = 0; s < S; ++s) {
      const int t = min(w + 8 * s, NTr - 1);
      ao[s] = ((t % MT) * 16 + lr) * E0S + 4 * lk;
      bo[s] = ((t / MT) * 16 + lr) * E0S + 4 * lk;
    }
    V4 acc[S];
#pragma unroll
    for (int s = 0; s < S; ++s) acc[s] = V4{0, 0, 0, 0};
    // phase 1: S(m, j) = Σ_i X(i, m) in(i, j), groups of 16 rows of i
    const int G1 = (e0 + 15) / 16;
    mf_slots<S, TC>(ns, acc, Xt, ao, Bt, bo, G1);
    __syncthreads();  // every wave is done with the input plane
    // S to LDS (St [m][j]), Y to LDS (Yt [q][j], over the input plane)
#pragma unroll
    for (int s = 0; s < S; ++s) {
      const int t = min(w + 8 * s, NTr - 1), mt = t % MT, jt = t / MT;
      if (s < ns)
#pragma unroll
        for (int r = 0; r < 4; ++r) St[(mt * 16 + Mf<TC>::mrow(lk, r)) * JS + jt * 16 + lr] = acc[s][r];
      acc[s] = V4{0, 0, 0, 0};
    }
    TC *Yt = Bt;
#pragma unroll
    for (int r = 0; r < C::NY; ++r) {
      const int t = tid + NTH * r, jj = t % L1P, q = t / L1P;
      Yt[q * JS + jj] = (jj < e1 && q < e1) ? ry[r] : TC(0);
    }
    // the epilogue's operands, in flight during phase 2 (clamped addresses)
    double ev[S][4], zv[S][4];
    const int ec = p.addth ? 1 : (p.dot ? 2 : 0);
    const View ve = ec == 1 ? p.th[comp < 3 ? comp : 2] : p.r[comp < 3 ? comp : 2];
    const View vz = p.zold[comp < 3 ? comp : 2];
    const int span = (ve.s1 * (e1 - 1) + ve.s2 * plane + e0) * 8, spanz = (vz.s1 * (e1 - 1) + vz.s2 * plane + e0) * 8;
    const __amdgpu_buffer_rsrc_t re = buf_rsrc(ve.p, span), rz = buf_rsrc(vz.p, spanz);
#pragma unroll
    for (int s = 0; s < S; ++s)
#pragma unroll
      for (int r = 0; r < 4; ++r) ev[s][r] = zv[s][r] = 0.0;
    if (ec) {
#pragma unroll
      for (int s = 0; s < S; ++s) {
        const int t = min(w + 8 * s, NTr - 1), mt = t % MT, qt = t / MT;
        const int qc = min(qt * 16 + lr, e1 - 1);
#pragma unroll
        for (int r = 0; r < 4; ++r)
          ev[s][r] = buf_load<double>(re, voff(ve, min(m0 + mt * 16 + Mf<TC>::mrow(lk, r), e0 - 1), qc, plane) * 8);
      }
      if (p.dot == 2)
#pragma unroll
        for (int s = 0; s < S; ++s) {
          const int t = min(w + 8 * s, NTr - 1), mt = t % MT, qt = t / MT;
          const int qc = min(qt * 16 + lr, e1 - 1);
#pragma unroll
          for (int r = 0; r < 4; ++r)
            zv[s][r] = buf_load<double>(rz, voff(vz, min(m0 + mt * 16 + Mf<TC>::mrow(lk, r), e0 - 1), qc, plane) * 8);
        }
    }
    __syncthreads();
    // phase 2: out(m, q) = Σ_j S(m, j) Y(j, q), groups of 16 j
#pragma unroll
    for (int s = 0; s < S; ++s) {
      const int t = min(w + 8 * s, NTr - 1);
      ao[s] = ((t % MT) * 16 + lr) * JS + 4 * lk;
      bo[s] = ((t / MT) * 16 + lr) * JS + 4 * lk;
    }
    const int G2 = (e1 + 15) / 16;
    mf_slots<S, TC>(ns, acc, St, ao, Yt, bo, G2);
    const View vo = p.out[comp];
#pragma unroll
    for (int s = 0; s < S; ++s) {
      if (s >= ns) continue;
      const int t = w + 8 * s, mt = t % MT, qt = t / MT;
      const int q = qt * 16 + lr;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = m0 + mt * 16 + Mf<TC>::mrow(lk, r);
        if (m >= e0 || q >= e1) continue;
        const int qo = voff(vo, m, q, plane);
        if (ec == 1) {
          ((double *)vo.p)[qo] = ev[s][r] + (double)acc[s][r];
        } else {
          const TO v = (TO)acc[s][r];
          ((TO *)vo.p)[qo] = v;
          d0 += ev[s][r] * (double)v;  // zero unless dot
          d1 += ev[s][r] * zv[s][r];
        }
      }
    }
  }
  if (p.dot) {
    block_sum2_8w(d0, d1, sh, tid);
    if (tid == 0) {
      double *pr = region_w(p.d, PART_RZ);
      pr[2 * bid] = d0;
      pr[2 * bid + 1] = d1;
    }
  }
}

template <class TI, class TC, class TO, int XC, int TJ, int MT>
__global__ void __launch_bounds__(PLANE1_T) k_plane_mf1(PlanePass p, const int *gate, int gx, int gy) {
  plane_mf1_body<TI, TC, TO, XC, TJ, MT>(p, gate, gx, gy, blockIdx.x, gridDim.x);
}
// batched: run r's blocks take the run's own (gx, gy) = (dims[3r], dims[3r+1]) and block count
template <int XC, int TJ, int MT>
__global__ void __launch_bounds__(PLANE1_T) k_plane_mf1_b(BatchTab t) {
  const BatchIdx q = batch_idx(t);
  const int l = q.b.x + q.b.nx * (q.b.y + q.b.ny * q.b.z);
  plane_mf1_body<double, double, double, XC, TJ, MT>(((const PlanePass *)t.args)[q.r], nullptr, q.b.nx, q.b.ny, l,
                                                      q.b.nx * q.b.ny * q.b.nz);
}

static size_t plane_mf_lds(int l1) {
  const int l1p = (l1 + 31) / 32 * 32;
  const int ic = l1p <= 128 ? 64 : 32;  // PlaneMfCfg
  return (size_t)(ic * 33 + ic * (l1p + 1) + l1p * 33) * sizeof(float);
}

static int plane_ncomp(const PlanePass &p) { return p.ncomp > 0 ? p.ncomp : 3; }
template <class T>
static dim3 plane_grid(const PlanePass &p) {
  int gx = 1, gy = 1;
  const int nc = plane_ncomp(p);
  for (int c = 0; c < nc; ++c) {
    gx = cdiv(p.e[c][0], PlaneCfg<T>::MS) > gx ? cdiv(p.e[c][0], PlaneCfg<T>::MS) : gx;
    gy = p.e[c][2] > gy ? p.e[c][2] : gy;
  }
  return dim3(gx, gy, nc);
}
// The one-round kernel (p.one != 0), FP32 (the preconditioner) or FP64 (the separable solve):
// planes up to 32 PLANE1_XC x 32 PLANE1_TJ with x extents above 64 and y extents above 32 whose
// staging fits the LDS (smaller planes, and FP64 planes too large for the LDS, keep k_plane_mf /
// k_plane).  Blocks of 2 m tiles, or of 3 where 2 would need more blocks than the chip has CUs
// and 3 needs fewer (p.one = 1), or p.one m tiles (2 or 3).
struct Plane1Pick {
  int use, xc, tj, mt, gx, gy, nc;
};
// the instances of the one-round kernel: (2..5) x (2, 3) x (2, 3) and (5, 3, 1) in both types, and
// in FP64 also those of one 32-row chunk along x or y (TDGL_PLANE1_SMALL)
static bool plane1_has(int xc, int tj, int mt) {
  if (xc >= 2 && xc <= 5 && tj >= 2 && tj <= 3 && mt >= 2 && mt <= 3) return true;
  if (xc == 5 && tj == 3 && mt == 1) return true;
  if (xc == 1 && tj >= 1 && tj <= 3 && mt == 2) return true;
  return xc >= 2 && xc <= 5 && tj == 1 && mt >= 2 && mt <= 3;
}
static Plane1Pick plane1_pick(const PlanePass &p, int es) {
  Plane1Pick k{};
  k.nc = plane_ncomp(p);
  int l0 = 0, l1 = 0;
  k.gy = 1;
  for (int c = 0; c < k.nc; ++c) {
    l0 = p.e[c][0] > l0 ? p.e[c][0] : l0;
    l1 = p.e[c][1] > l1 ? p.e[c][1] : l1;
    k.gy = p.e[c][2] > k.gy ? p.e[c][2] : k.gy;
  }
  bool tr = true;
  for (int c = 0; c < k.nc; ++c) tr = tr && p.XT[c] && p.YT[c];
  // FP64 planes of any extent up to 32 PLANE1_XC x 32 PLANE1_TJ (the parity quarters of the FP64
  // separable solve are mostly below 33); FP32 ones from 65 x 33 (below, the chunked kernel was faster)
  k.use = p.one != 0 && tr && l0 > (es == 8 ? 0 : 64) && l0 <= 32 * PLANE1_XC && l1 > (es == 8 ? 0 : 32) &&
          l1 <= 32 * PLANE1_TJ;
  if (!k.use) return k;
  k.xc = cdiv(l0, 32);
  k.tj = cdiv(l1, 32);
  const int n2 = cdiv(l0, 32) * k.gy * k.nc, n3 = cdiv(l0, 48) * k.gy * k.nc;
  k.mt = p.one >= 2 ? p.one : ((n2 > 256 && n3 < n2) ? 3 : 2);
  if (plane1_lds(k.xc, k.tj, k.mt, es) > PLANE1_LDS_MAX && p.one < 2) k.mt = 2;
  // FP64 planes of up to 160 x 96: one m tile a block where two do not fit the LDS
  if (plane1_lds(k.xc, k.tj, k.mt, es) > PLANE1_LDS_MAX && p.one < 2 && es == 8) k.mt = 1;
  if (plane1_lds(k.xc, k.tj, k.mt, es) > PLANE1_LDS_MAX) k.use = 0;
  if (!plane1_has(k.xc, k.tj, k.mt)) k.use = 0;
  k.gx = cdiv(l0, 16 * k.mt);
  return k;
}
template <class T>
int plane_parts(const PlanePass &p) {
  const Plane1Pick k = plane1_pick(p, (int)sizeof(T));
  if (k.use) return k.gx * k.gy * k.nc;
  const dim3 g = plane_grid<T>(p);
  return (int)(g.x * g.y * g.z);
}
template <class T>
bool plane_fits(const int e[][3], int ncomp) {
  if (ncomp < 1 || ncomp > PLANE_MAXC) return false;
  for (int c = 0; c < ncomp; ++c)
    if (e[c][1] > PLANE_LMAX || e[c][0] < 1 || e[c][1] < 1 || e[c][2] < 1) return false;
  int l1 = 0;
  for (int c = 0; c < ncomp; ++c) l1 = e[c][1] > l1 ? e[c][1] : l1;
  return (sizeof(T) == 4 ? plane_mf_lds(l1) : plane_lds<T>(l1)) <= 64 * 1024;
}

// the FP64 instances of one 32-row chunk along x or y
#define TDGL_PLANE1_SMALL(C) \
  C(1, 1, 2) C(1, 2, 2) C(1, 3, 2) C(2, 1, 2) C(2, 1, 3) C(3, 1, 2) C(3, 1, 3) C(4, 1, 2) C(4, 1, 3) C(5, 1, 2) C(5, 1, 3)

template <class TI, class T, class TO>
hipError_t launch_plane(const PlanePass &p, const int *gate, hipStream_t s) {
  const int nc = plane_ncomp(p);
  if (!plane_fits<T>(p.e, nc) || plane_parts<T>(p) > MAX_PART) return hipErrorInvalidValue;
  int l1 = 0;
  for (int c = 0; c < nc; ++c) l1 = p.e[c][1] > l1 ? p.e[c][1] : l1;
  const dim3 gr = plane_grid<T>(p);
  const bool mf = sizeof(T) == 4;  // Float32 on the matrix cores
  if (g_rec) {
    // the FP64 passes of the separable solve: the instance this launch would take
    constexpr bool f64 = sizeof(TI) == 8 && sizeof(T) == 8 && sizeof(TO) == 8;
    if (!f64 || gate || p.conv.mode != CONV_NONE || p.dot || (nc != 3 && nc != 12) || (nc == 12 && p.addth))
      return hipErrorNotSupported;
    const Plane1Pick k = plane1_pick(p, 8);
    if (k.use)
      return record(BK_PLANE, Geo{}, &p, sizeof p, 1000 + (k.xc * 4 + k.tj) * 4 + k.mt, true, dim3(k.gx, k.gy, k.nc),
                    plane1_lds(k.xc, k.tj, k.mt, 8));
    const int tj = (l1 + 31) / 32;
    if (tj < 1 || tj > 8) return hipErrorInvalidValue;
    return record(BK_PLANE, Geo{}, &p, sizeof p, tj, true, gr, plane_lds<T>(l1));
  }
  {
    const Plane1Pick k = plane1_pick(p, (int)sizeof(T));
    if (k.use) {
      const dim3 g1(k.gx * k.gy * k.nc);
      switch ((k.xc * 4 + k.tj) * 4 + k.mt) {
#define TDGL_PLANE1_CASE(x, t, m)                                                                               \
  case (x * 4 + t) * 4 + m:                                                                                     \
    TDGL_LAUNCH((k_plane_mf1<TI, T, TO, x, t, m>), g1, dim3(PLANE1_T), (PlaneMf1<x, t, m, T>::lds), s, p, gate, \
                k.gx, k.gy);                                                                                    \
    break;
        TDGL_PLANE1_CASE(2, 2, 2) TDGL_PLANE1_CASE(2, 3, 2) TDGL_PLANE1_CASE(2, 2, 3) TDGL_PLANE1_CASE(2, 3, 3)
        TDGL_PLANE1_CASE(3, 2, 2) TDGL_PLANE1_CASE(3, 3, 2) TDGL_PLANE1_CASE(3, 2, 3) TDGL_PLANE1_CASE(3, 3, 3)
        TDGL_PLANE1_CASE(4, 2, 2) TDGL_PLANE1_CASE(4, 3, 2) TDGL_PLANE1_CASE(4, 2, 3) TDGL_PLANE1_CASE(4, 3, 3)
        TDGL_PLANE1_CASE(5, 2, 2) TDGL_PLANE1_CASE(5, 3, 2) TDGL_PLANE1_CASE(5, 2, 3) TDGL_PLANE1_CASE(5, 3, 3)
        TDGL_PLANE1_CASE(5, 3, 1)
        default:
          if constexpr (sizeof(T) == 8) {
            switch ((k.xc * 4 + k.tj) * 4 + k.mt) {
              TDGL_PLANE1_SMALL(TDGL_PLANE1_CASE)
              default: return hipErrorInvalidValue;
            }
          } else {
            return hipErrorInvalidValue;
          }
#undef TDGL_PLANE1_CASE
      }
      return hipGetLastError();
    }
  }
  const size_t lds = mf ? plane_mf_lds(l1) : plane_lds<T>(l1);
  switch ((l1 + 31) / 32) {
#define TDGL_PLANE_CASE(tj)                                                                     \
  case tj:                                                                                      \
    if constexpr (sizeof(T) == 4)                                                               \
      TDGL_LAUNCH((k_plane_mf<TI, TO, tj>), gr, dim3(256), lds, s, p, gate);             \
    else                                                                                        \
      TDGL_LAUNCH((k_plane<TI, T, TO, tj>), gr, dim3(256), lds, s, p, gate);             \
    break;
    TDGL_PLANE_CASE(1)
    TDGL_PLANE_CASE(2)
    TDGL_PLANE_CASE(3)
    TDGL_PLANE_CASE(4)
    TDGL_PLANE_CASE(5)
    TDGL_PLANE_CASE(6)
    TDGL_PLANE_CASE(7)
    TDGL_PLANE_CASE(8)
#undef TDGL_PLANE_CASE
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}
template hipError_t launch_plane<double, float, float>(const PlanePass &, const int *, hipStream_t);
template hipError_t launch_plane<float, float, double>(const PlanePass &, const int *, hipStream_t);
template hipError_t launch_plane<double, double, double>(const PlanePass &, const int *, hipStream_t);
template hipError_t launch_plane<float, float, float>(const PlanePass &, const int *, hipStream_t);

// ---------------------------------------------------------------- the quarter-plane transforms
// k_qplane: one block per (quarter, z plane) holds the whole plane in LDS once (Bt[j][i], rows of
// 4 (mod 8) words), and wave w computes the 16 rows m = 16 w .. 16 w + 15 of both products on
// v_mfma_f32_16x16x4f32: phase 1 S(m, j) = Σ_i X(i, m) in(i, j) with A = X read from the padded
// transposed basis in global memory (16 bytes a lane, every block reads the same few KB: cache
// hits) and B = the plane from LDS; S goes to LDS over the plane (St[m][j]); phase 2 the
// transposed product out(q, m) = Σ_j Y(j, q) S(m, j) with A = Y from global memory and B = S from
// LDS, so that a lane's results lie along m, the fastest index of the output.  The operands of
// one instruction and the order of the instructions are those of k_plane_mf1 (group g, step u,
// lane k index lk: k = 16 g + 4 lk + u), and the padding is zero in both: the same bits.
// The next group's global operands are requested while the current group multiplies.
// the two products of k_qplane for one quarter plane held in LDS (Bt); FULL: the quarter has the
// launch's JT tiles along j, so no tile is guarded.  Mf<T> gives the instruction of each type and
// where its results lie (the FP64 accumulator's element r of lane l is row 4 r + l / 16).
template <class T, int JT, bool FULL>
__device__ __forceinline__ void qplane_products(const QPassT<T> &p, int comp, int plane, int tid, int mt, int jt, int e0,
                                                int e1, int E0S, int JS, int xs, int ys, T *Bt, T *St, int m0t, int mh) {
  typedef typename Mf<T>::v4 V4;
  // wave w takes the m tile m0t + w of the block's mh (its rows of S in LDS are the block's own)
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6), l = tid & 63, lr = l & 15, lk = l >> 4;
  const int wm = m0t + w;
  const bool on = w < mh && wm < mt;
  V4 acc[JT];
#pragma unroll
  for (int c = 0; c < JT; ++c) acc[c] = V4{0, 0, 0, 0};
  if (on) {
    // phase 1: A = X(16 g + 4 lk + u, 16 w + lr) from XT, B = in(16 g + 4 lk + u, 16 c + lr) from Bt
    const T *xa = p.XT[comp] + (size_t)(16 * wm + lr) * xs + 4 * lk;
    V4 A = Mf<T>::ld4(xa);
    for (int g = 0; g < mt; ++g) {
      const V4 An = Mf<T>::ld4(xa + 16 * min(g + 1, mt - 1));
      V4 B[JT];
#pragma unroll
      for (int c = 0; c < JT; ++c)
        B[c] = (FULL || c < jt) ? Mf<T>::ld4(Bt + (16 * c + lr) * E0S + 16 * g + 4 * lk) : V4{0, 0, 0, 0};
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int c = 0; c < JT; ++c)
          if ((FULL || c < jt)) acc[c] = Mf<T>::mma(A[u], B[c][u], acc[c]);
      A = An;
    }
  }
  __syncthreads();  // every wave is done with the plane: S goes over it
  if (on)
#pragma unroll
    for (int c = 0; c < JT; ++c)
      if ((FULL || c < jt))
#pragma unroll
        for (int r = 0; r < 4; ++r) St[(16 * w + Mf<T>::mrow(lk, r)) * JS + 16 * c + lr] = acc[c][r];
  __syncthreads();
  if (!on) return;
  // phase 2: A = Y(16 g + 4 lk + u, 16 c + lr) from YT, B = S(16 w + lr, 16 g + 4 lk + u) from St
#pragma unroll
  for (int c = 0; c < JT; ++c) acc[c] = V4{0, 0, 0, 0};
  const T *ya = p.YT[comp] + (size_t)lr * ys + 4 * lk;
  V4 Y[JT];
#pragma unroll
  for (int c = 0; c < JT; ++c) Y[c] = (FULL || c < jt) ? Mf<T>::ld4(ya + (size_t)16 * c * ys) : V4{0, 0, 0, 0};
  for (int g = 0; g < jt; ++g) {
    const int gn = 16 * min(g + 1, jt - 1);
    V4 Yn[JT];
#pragma unroll
    for (int c = 0; c < JT; ++c) Yn[c] = (FULL || c < jt) ? Mf<T>::ld4(ya + (size_t)16 * c * ys + gn) : V4{0, 0, 0, 0};
    const V4 B = Mf<T>::ld4(St + (16 * w + lr) * JS + 16 * g + 4 * lk);
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int c = 0; c < JT; ++c)
        if ((FULL || c < jt)) acc[c] = Mf<T>::mma(Y[c][u], B[u], acc[c]);
#pragma unroll
    for (int c = 0; c < JT; ++c) Y[c] = Yn[c];
  }
  // out(q, m): element r of lane l is q = 16 c + mrow(l / 16, r), m = 16 w + lr
  const View vo = p.out[comp];
  T *dst = (T *)vo.p + (size_t)vo.s2 * plane;
  const int m = 16 * wm + lr;
#pragma unroll
  for (int c = 0; c < JT; ++c)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int q = 16 * c + Mf<T>::mrow(lk, r);
      if ((FULL || c < jt) && m < e0 && q < e1) dst[m + (size_t)vo.s1 * q] = acc[c][r];
    }
}

// LDS rows of 4 (mod 8) words: 4 floats or 2 doubles of padding
template <class T> constexpr int qplane_pad() { return 16 / (int)sizeof(T); }

// block `bid` of a k_qplane launch: unit u = (quarter, plane) = (u / e2max, u % e2max) and its part h,
// which takes the m tiles h mh .. h mh + mh - 1, mh = ceil(mt / parts), of the plane.  The parts of a
// unit are 8 block indices apart (u = 8 (bid / 8 parts) + bid % 8, h = (bid / 8) % parts): the
// hardware deals blocks to the eight XCDs in turn, so they share one XCD and the later part finds
// the plane in its L2
template <class T, int JT>
__device__ __forceinline__ void qplane_body(const QPassT<T> &p, const int *gate, int bid) {
  extern __shared__ __align__(16) unsigned char qplane_smem[];
  __shared__ double sh[16];
  const int tid = threadIdx.x, nth = blockDim.x;
  const int e2max = p.e2max, mtx = p.mtx, jtx = p.jtx, parts = p.parts > 1 ? p.parts : 1;
  const int grp = bid / (8 * parts), r8 = bid - grp * 8 * parts;
  const int unit = 8 * grp + (r8 & 7), part = r8 >> 3;
  if (unit >= 12 * e2max) return;
  const int comp = unit / e2max, plane = unit % e2max;
  if (p.conv.mode != CONV_NONE) {
    if (!*p.conv.in) {
      if (bid == 0 && tid == 0) *p.conv.out = 0;
      return;
    }
  } else if (gate && !*gate) {
    return;
  }
  const int e0 = p.e[comp][0], e1 = p.e[comp][1], e2 = p.e[comp][2];
  if (p.conv.mode != CONV_NONE) {
    const double rr = sum_parts_8w(region(p.d, PART_RR), p.conv.nparts, 0, sh, tid);
    const double tol = p.d.sc[p.conv.inner ? SC_TOLI : SC_TOL], nrm = p.d.sc[p.conv.inner ? SC_NBI : SC_NB];
    const int active = p.conv.mode == CONV_RR0 ? ((nrm > 0 && sqrt(rr) > tol * nrm) ? 1 : 0)
                                               : ((sqrt(rr) <= tol * nrm || p.d.fl[FL_IT] >= p.d.fl[FL_MAXIT]) ? 0 : 1);
    if (bid == 0 && tid == 0) *p.conv.out = active;
    if (!active) return;
  }
  if (plane >= e2) return;
  const int mt = (e0 + 15) / 16, jt = (e1 + 15) / 16;  // this quarter's tiles along m (and i), j (and q)
  const int E0S = 16 * mtx + qplane_pad<T>(), JS = 16 * jtx + qplane_pad<T>();  // LDS row strides
  const int xs = 16 * mt, ys = 16 * jt;                // the padded bases' row strides
  T *Bt = (T *)qplane_smem, *St = Bt;
  {
    // the plane, zero beyond e0 x e1 up to whole tiles: wave w takes the rows j = w, w + waves, ...
    // four at a time, its lanes along the row (ii = l, l + 64, l + 128: W <= 16 QP_MT = 160), so
    // that no index is divided; twelve loads in flight per lane
    const T *src = (const T *)p.in[comp].p + (size_t)p.in[comp].s2 * plane;
    const int s1 = p.in[comp].s1, W = 16 * mt, H = 16 * jt;
    const int nw = nth >> 6, w = __builtin_amdgcn_readfirstlane(tid >> 6), l = tid & 63;
    constexpr int R = 4, LR = (16 * QP_MT + 63) / 64;
    for (int j0 = w; j0 < H; j0 += R * nw) {
      T v[R][LR];
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int j = j0 + r * nw;
#pragma unroll
        for (int c = 0; c < LR; ++c) {
          const int ii = l + 64 * c;
          v[r][c] = (j < e1 && ii < e0) ? src[ii + (size_t)s1 * j] : T(0);
        }
      }
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int j = j0 + r * nw;
#pragma unroll
        for (int c = 0; c < LR; ++c) {
          const int ii = l + 64 * c;
          if (j < H && ii < W) Bt[j * E0S + ii] = v[r][c];
        }
      }
    }
  }
  __syncthreads();
  const int mh = (mt + parts - 1) / parts;
  if (jt == JT)
    qplane_products<T, JT, true>(p, comp, plane, tid, mt, jt, e0, e1, E0S, JS, xs, ys, Bt, St, part * mh, mh);
  else
    qplane_products<T, JT, false>(p, comp, plane, tid, mt, jt, e0, e1, E0S, JS, xs, ys, Bt, St, part * mh, mh);
}

template <int JT>
__global__ void __launch_bounds__(64 * QP_MT) k_qplane(QPass p, const int *gate) {
  qplane_body<float, JT>(p, gate, (int)blockIdx.x);
}
// FP64: at most 8 waves (quarters up to 128 rows), so that the doubled operands of 4 j tiles fit
// the registers without spilling
constexpr int QP64_MT = 8;
template <int JT>
__global__ void __launch_bounds__(64 * QP64_MT) k_qplane64(QPass64 p) {
  qplane_body<double, JT>(p, nullptr, (int)blockIdx.x);
}
// batched: run r's blocks are its (quarter, plane) blocks; the launch's block has the waves of
// the largest m tile count among its runs (a run's surplus waves stage the plane and wait)
template <int JT>
__global__ void __launch_bounds__(64 * QP64_MT) k_qplane64_b(BatchTab t) {
  const BatchIdx q = batch_idx(t);
  qplane_body<double, JT>(((const QPass64 *)t.args)[q.r], nullptr, q.b.x);
}

template <class T>
static void qplane_tiles(const QPassT<T> &p, int &mtx, int &jtx, int &e2max) {
  mtx = jtx = e2max = 1;
  for (int c = 0; c < 12; ++c) {
    mtx = std::max(mtx, cdiv(p.e[c][0], 16));
    jtx = std::max(jtx, cdiv(p.e[c][1], 16));
    e2max = std::max(e2max, p.e[c][2]);
  }
}
template <class T>
static size_t qplane_lds(int mtx, int jtx) {
  const int pad = qplane_pad<T>();
  return (size_t)std::max(16 * jtx * (16 * mtx + pad), 16 * mtx * (16 * jtx + pad)) * sizeof(T);
}
template <class T>
static bool qplane_fits_t(const QPassT<T> &p) {
  int mtx, jtx, e2max;
  qplane_tiles(p, mtx, jtx, e2max);
  for (int c = 0; c < 12; ++c)
    if (p.e[c][0] < 1 || p.e[c][1] < 1 || p.e[c][2] < 1) return false;
  return mtx <= (sizeof(T) == 8 ? QP64_MT : QP_MT) && jtx <= QP_JT && qplane_lds<T>(mtx, jtx) <= 160 * 1024;
}
bool qplane_fits(const QPass &p) { return qplane_fits_t(p); }
bool qplane64_fits(const QPass64 &p) { return qplane_fits_t(p); }
hipError_t launch_qplane(const QPass &p0, const int *gate, hipStream_t s) {
  if (g_rec) return hipErrorNotSupported;
  if (!qplane_fits(p0)) return hipErrorInvalidValue;
  QPass p = p0;
  qplane_tiles(p, p.mtx, p.jtx, p.e2max);
  p.parts = std::max(1, std::min(p.parts, p.mtx));
  const dim3 grid(p.parts > 1 ? 8 * p.parts * cdiv(12 * p.e2max, 8) : 12 * p.e2max), blk(64 * std::max(4, cdiv(p.mtx, p.parts)));
  const size_t lds = qplane_lds<float>(p.mtx, p.jtx);
  switch (p.jtx) {
    case 1: TDGL_LAUNCH((k_qplane<1>), grid, blk, lds, s, p, gate); break;
    case 2: TDGL_LAUNCH((k_qplane<2>), grid, blk, lds, s, p, gate); break;
    case 3: TDGL_LAUNCH((k_qplane<3>), grid, blk, lds, s, p, gate); break;
    case 4: TDGL_LAUNCH((k_qplane<4>), grid, blk, lds, s, p, gate); break;
    case 5: TDGL_LAUNCH((k_qplane<5>), grid, blk, lds, s, p, gate); break;
    default: TDGL_LAUNCH((k_qplane<6>), grid, blk, lds, s, p, gate); break;
  }
  return hipGetLastError();
}
// FP64: recorded with the instance 100 JT + waves (the batch launches each pair on its own)
hipError_t launch_qplane64(const QPass64 &p0, hipStream_t s) {
  if (!qplane64_fits(p0)) return hipErrorInvalidValue;
  QPass64 p = p0;
  qplane_tiles(p, p.mtx, p.jtx, p.e2max);
  p.parts = 1;
  const int waves = std::max(4, p.mtx);
  const dim3 grid(12 * p.e2max), blk(64 * waves);
  const size_t lds = qplane_lds<double>(p.mtx, p.jtx);
  if (g_rec) return record(BK_QPLANE, Geo{}, &p, sizeof p, 100 * p.jtx + waves, true, grid, lds);
  switch (p.jtx) {
    case 1: TDGL_LAUNCH((k_qplane64<1>), grid, blk, lds, s, p); break;
    case 2: TDGL_LAUNCH((k_qplane64<2>), grid, blk, lds, s, p); break;
    case 3: TDGL_LAUNCH((k_qplane64<3>), grid, blk, lds, s, p); break;
    case 4: TDGL_LAUNCH((k_qplane64<4>), grid, blk, lds, s, p); break;
    case 5: TDGL_LAUNCH((k_qplane64<5>), grid, blk, lds, s, p); break;
    default: TDGL_LAUNCH((k_qplane64<6>), grid, blk, lds, s, p); break;
  }
  return hipGetLastError();
}
template int plane_parts<float>(const PlanePass &);
template int plane_parts<double>(const PlanePass &);
template bool plane_fits<float>(const int e[][3], int);
template bool plane_fits<double>(const int e[][3], int);

// ---------------------------------------------------------------------------- z, modes, inverse z
// Thread (column tc, mode mz) transforms the three z-columns that meet at mode (mx, my, mz)
// (X at (mx-1, my), Y at (mx, my-1), Z at (mx, my)), projects them and leaves the projected modes
// in LDS; then thread (tc, n) takes the inverse z transform of those columns at z index n.  Every
// projected mode of a column comes from the same block, so no other launch is needed between.
// The columns are staged in LDS first (thread (tc, n) loads z index n of its column), and with
// LB the four z bases too, so every global load is issued before the first wait; the sums run in
// the same order as from global memory.
template <class T, bool LB>
__device__ __forceinline__ void zmodes_body(const ZModes &m, const int *gate, const Blk bk) {
  if (gate && !*gate) return;
  extern __shared__ __align__(16) unsigned char zm_smem[];
  const int nx = m.nx, ny = m.ny, nz = m.nz;
  T *Cx = (T *)zm_smem, *Cy = Cx + 256, *Cz = Cy + 256, *Mx = Cz + 256, *My = Mx + 256, *Mz = My + 256;
  T *Lav = Mz + 256, *Lbv = Lav + nz * nz, *Lau = Lbv + nz * nz, *Lbu = Lau + (nz - 1) * (nz - 1);
  const int C = 256 / nz;
  const int t = threadIdx.x;
  const int tc = t / nz, mz = t % nz;
  const int col = bk.x * C + tc;
  const bool live = tc < C && col < nx * ny;
  const int mx = live ? col % nx : 0, my = live ? col / nx : 0;
  const bool hx = mx > 0, hy = my > 0, hz = mz > 0;
  const T *in = (const T *)m.in;
  const int stx = (nx - 1) * ny, sty = nx * (ny - 1), stz = nx * ny;
  if (live) {
    // z index mz of the three columns (clamped addresses, the padding selected after)
    const T *colx = in + m.off[0] + (hx ? mx - 1 : 0) + (nx - 1) * my;
    const T *coly = in + m.off[1] + mx + nx * (hy ? my - 1 : 0);
    const T *colz = in + m.off[2] + mx + nx * my;
    const T vx = colx[stx * mz], vy = coly[sty * mz], vz = colz[stz * min(mz, nz - 2)];
    Cx[t] = vx;
    Cy[t] = vy;
    Cz[t] = mz < nz - 1 ? vz : T(0);
  }
  if (LB) {
    const T *AV = (const T *)m.Az[0], *AU = (const T *)m.Az[1], *BV = (const T *)m.Ai[0], *BU = (const T *)m.Ai[1];
    for (int e = t; e < nz * nz; e += 256) {
      Lav[e] = AV[e];
      Lbv[e] = BV[e];
    }
    for (int e = t; e < (nz - 1) * (nz - 1); e += 256) {
      Lau[e] = AU[e];
      Lbu[e] = BU[e];
    }
  }
  __syncthreads();
  const T *AV = LB ? Lav : (const T *)m.Az[0], *AU = LB ? Lau : (const T *)m.Az[1];
  const T *BV = LB ? Lbv : (const T *)m.Ai[0], *BU = LB ? Lbu : (const T *)m.Ai[1];
  const int b0 = tc * nz;
  if (live) {
    const T k1 = ((const T *)m.sx)[mx], k2 = ((const T *)m.sy)[my], k3 = ((const T *)m.sz)[mz];
    const int mzu = hz ? mz - 1 : 0;
    T ax = 0, ay = 0, az = 0;
#pragma unroll 8
    for (int n = 0; n < nz - 1; ++n) {
      const T v = AV[n * nz + mz];
      ax += Cx[b0 + n] * v;
      ay += Cy[b0 + n] * v;
      az += Cz[b0 + n] * AU[n * (nz - 1) + mzu];
    }
    {
      const T v = AV[(nz - 1) * nz + mz];
      ax += Cx[b0 + nz - 1] * v;
      ay += Cy[b0 + nz - 1] * v;
    }
    T x1 = hx ? ax : T(0), y1 = hy ? ay : T(0), z1 = hz ? az : T(0);
    const T k2s = k1 * k1 + k2 * k2 + k3 * k3;
    if (k2s != 0) project<T>(k1, k2, k3, (T)m.a, x1, y1, z1, hx, hy, hz);
    Mx[t] = x1;
    My[t] = y1;
    Mz[t] = z1;
  }
  __syncthreads();
  if (!live) return;
  const int n = mz;  // the z index this thread writes
  T ox = 0, oy = 0, oz = 0;
#pragma unroll 8
  for (int q = 0; q < nz; ++q) {
    const T v = BV[q * nz + n];
    ox += Mx[b0 + q] * v;
    oy += My[b0 + q] * v;
  }
  const int nu = n < nz - 1 ? n : 0;
#pragma unroll 8
  for (int q = 1; q < nz; ++q) oz += Mz[b0 + q] * BU[(q - 1) * (nz - 1) + nu];
  T *out = (T *)m.out;
  if (hx) out[m.off[0] + (mx - 1) + (nx - 1) * (my + ny * n)] = ox;
  if (hy) out[m.off[1] + mx + nx * ((my - 1) + (ny - 1) * n)] = oy;
  if (n < nz - 1) out[m.off[2] + mx + nx * (my + ny * n)] = oz;
}

// The same z pass, transforms and projector over x-runs of TX modes: a block owns the modes
// (mx0 .. mx0+TX-1, my, all mz) and stages their columns z-row by z-row, each row a contiguous run
// of x (coalesced loads and stores), with the forward z bases once per block (the inverse ones are
// their transposes: the same values, read with the indices swapped).  Thread (p, g) takes column p
// and the modes (and then the output z indices) g, g + G, g + 2G, ...; every sum runs over n (or q)
// in the order of zmodes_body, so the results are the same bits.
constexpr int ZT_NZ_MAX = 64;  // the run kernel holds the z bases in LDS up to this nz
template <class T, int TX>
__global__ void __launch_bounds__(256) k_zmodes_run(ZModes m, const int *gate) {
  if (gate && !*gate) return;
  constexpr int G = 256 / TX;                          // mode groups
  constexpr int MZ = (ZT_NZ_MAX + G - 1) / G;          // modes per thread at most
  extern __shared__ __align__(16) unsigned char zt_smem[];
  const int nx = m.nx, ny = m.ny, nz = m.nz;
  T *Cx = (T *)zt_smem, *Cy = Cx + nz * TX, *Cz = Cy + nz * TX;   // [z][TX]
  T *AV = Cz + nz * TX, *AU = AV + nz * nz;                        // AV[n nz + m] = V(n, m)
  const int t = threadIdx.x, p = t % TX, g = t / TX;
  const int mx = blockIdx.x * TX + p, my = blockIdx.y;
  const bool live = mx < nx;
  const bool hx = live && mx > 0, hy = live && my > 0;
  {
    const T *gV = (const T *)m.Az[0], *gU = (const T *)m.Az[1];
    for (int e = t; e < nz * nz; e += 256) AV[e] = gV[e];
    for (int e = t; e < (nz - 1) * (nz - 1); e += 256) AU[e] = gU[e];
  }
  {
    // clamped, in-bounds addresses, loaded unconditionally; the absent values are selected after
    const T *in = (const T *)m.in;
    const int mc = min(mx, nx - 1);
    const size_t stx = (size_t)(nx - 1) * ny, sty = (size_t)nx * (ny - 1), stz = (size_t)nx * ny;
    const T *colx = in + m.off[0] + max(mc - 1, 0) + (nx - 1) * my;
    const T *coly = in + m.off[1] + mc + nx * max(my - 1, 0);
    const T *colz = in + m.off[2] + mc + nx * my;
    for (int z = g; z < nz; z += G) {
      const T vx = colx[stx * z], vy = coly[sty * z], vz = colz[stz * min(z, nz - 2)];
      Cx[z * TX + p] = hx ? vx : T(0);
      Cy[z * TX + p] = hy ? vy : T(0);
      Cz[z * TX + p] = (live && z < nz - 1) ? vz : T(0);
    }
  }
  __syncthreads();
  T ax[MZ], ay[MZ], az[MZ];
#pragma unroll
  for (int r = 0; r < MZ; ++r) ax[r] = ay[r] = az[r] = T(0);
  for (int n = 0; n < nz - 1; ++n) {
    const T cx = Cx[n * TX + p], cy = Cy[n * TX + p], cz = Cz[n * TX + p];
#pragma unroll
    for (int r = 0; r < MZ; ++r) {
      const int mz = g + G * r;
      if (mz < nz) {
        const T v = AV[n * nz + mz];
        ax[r] += cx * v;
        ay[r] += cy * v;
        az[r] += cz * AU[n * (nz - 1) + (mz > 0 ? mz - 1 : 0)];
      }
    }
  }
  {
    const T cx = Cx[(nz - 1) * TX + p], cy = Cy[(nz - 1) * TX + p];
#pragma unroll
    for (int r = 0; r < MZ; ++r) {
      const int mz = g + G * r;
      if (mz < nz) {
        const T v = AV[(nz - 1) * nz + mz];
        ax[r] += cx * v;
        ay[r] += cy * v;
      }
    }
  }
  __syncthreads();  // the columns are read: their LDS takes the projected modes
  if (live) {
    const T k1 = ((const T *)m.sx)[mx], k2 = ((const T *)m.sy)[my];
#pragma unroll
    for (int r = 0; r < MZ; ++r) {
      const int mz = g + G * r;
      if (mz < nz) {
        const T k3 = ((const T *)m.sz)[mz];
        const bool hz = mz > 0;
        T x1 = hx ? ax[r] : T(0), y1 = hy ? ay[r] : T(0), z1 = hz ? az[r] : T(0);
        const T k2s = k1 * k1 + k2 * k2 + k3 * k3;
        if (k2s != 0) project<T>(k1, k2, k3, (T)m.a, x1, y1, z1, hx, hy, hz);
        Cx[mz * TX + p] = x1;
        Cy[mz * TX + p] = y1;
        Cz[mz * TX + p] = z1;
      }
    }
  }
  __syncthreads();
  if (!live) return;
  T *out = (T *)m.out;
#pragma unroll
  for (int r = 0; r < MZ; ++r) {
    const int n = g + G * r;  // the z index written
    if (n >= nz) continue;
    T ox = 0, oy = 0, oz = 0;
    for (int q = 0; q < nz; ++q) {
      const T v = AV[n * nz + q];
      ox += Cx[q * TX + p] * v;
      oy += Cy[q * TX + p] * v;
    }
    const int nu = n < nz - 1 ? n : 0;
    for (int q = 1; q < nz; ++q) oz += Cz[q * TX + p] * AU[nu * (nz - 1) + (q - 1)];
    if (hx) out[m.off[0] + (mx - 1) + (size_t)(nx - 1) * (my + (size_t)ny * n)] = ox;
    if (hy) out[m.off[1] + mx + (size_t)nx * ((my - 1) + (size_t)(ny - 1) * n)] = oy;
    if (n < nz - 1) out[m.off[2] + mx + (size_t)nx * (my + (size_t)ny * n)] = oz;
  }
}

// ---------------------------------------------------------------------------- parity folds
// Blocks of 64 x 4 threads: 64 consecutive i of one j row of one component, four groups of k.
// Thread (i, j) folds the four mirror images (i, j), (L0-1-i, j), (i, L1-1-j), (L0-1-i, L1-1-j) of
// every k (a centre row or column, on an odd length, is its own image: it goes to the even
// quarter alone).
__device__ __forceinline__ void fold_index(const Blk &b, int &c, int &i, int &j, int &kg) {
  c = b.z;
  j = b.y;
  i = b.x * 64 + (int)(threadIdx.x & 63);
  kg = (int)(threadIdx.x >> 6);
}
__device__ __forceinline__ void fold_index(const Quarters &, int &c, int &i, int &j, int &kg) {
  fold_index(hw_blk(), c, i, j, kg);
}
int fold_blocks(const Quarters &q) {
  int hx = 1, hy = 1;
  for (int c = 0; c < 3; ++c) {
    hx = q.h[c][0] > hx ? q.h[c][0] : hx;
    hy = q.h[c][1] > hy ? q.h[c][1] : hy;
  }
  return cdiv(hx, 64) * hy * 3;
}
static dim3 fold_grid(const Quarters &q) {
  int hx = 1, hy = 1;
  for (int c = 0; c < 3; ++c) {
    hx = q.h[c][0] > hx ? q.h[c][0] : hx;
    hy = q.h[c][1] > hy ? q.h[c][1] : hy;
  }
  return dim3(cdiv(hx, 64), hy, 3);
}

// the fold of block b into quarters of TQ (FP32: the preconditioner's; FP64: the separable solve's)
template <class TQ>
__device__ __forceinline__ void fold_body(const FoldPass &f, TQ *out, const Blk &b) {
  int c, i, j, kg;
  fold_index(b, c, i, j, kg);
  const int L0 = f.q.L[c][0], L1 = f.q.L[c][1], h0 = f.q.h[c][0], h1 = f.q.h[c][1], e2 = f.q.e2[c];
  if (i >= h0 || j >= h1) return;
  const int is = L0 - 1 - i, js = L1 - 1 - j;
  const bool ci = is == i, cj = js == j;
  const View v = f.in[c];
  const size_t o_ij = (size_t)i + (size_t)v.s1 * j, o_sj = (size_t)is + (size_t)v.s1 * j;
  const size_t o_is = (size_t)i + (size_t)v.s1 * js, o_ss = (size_t)is + (size_t)v.s1 * js;
  const size_t qp = (size_t)i + (size_t)h0 * j, qs = (size_t)h0 * h1;
  TQ *q00 = out + f.q.off[c][0] + qp, *q01 = out + f.q.off[c][1] + qp;
  TQ *q10 = out + f.q.off[c][2] + qp, *q11 = out + f.q.off[c][3] + qp;
  for (int k = kg; k < e2; k += 4) {
    const size_t zk = (size_t)v.s2 * k;
    double a, b, cc, d;
    if (f.f32) {
      const float *pk = (const float *)v.p + zk;
      a = pk[o_ij], b = pk[o_sj], cc = pk[o_is], d = pk[o_ss];
    } else {
      const double *pk = (const double *)v.p + zk;
      a = pk[o_ij], b = pk[o_sj], cc = pk[o_is], d = pk[o_ss];
    }
    const double sj = ci ? a : a + b, dj = ci ? 0.0 : a - b;
    const double ss = ci ? cc : cc + d, ds = ci ? 0.0 : cc - d;
    const size_t qk = qs * k;
    q00[qk] = (TQ)(cj ? sj : sj + ss);
    q01[qk] = (TQ)(cj ? 0.0 : sj - ss);
    q10[qk] = (TQ)(cj ? dj : dj + ds);
    q11[qk] = (TQ)(cj ? 0.0 : dj - ds);
  }
}
__global__ void __launch_bounds__(256) k_fold(FoldPass f, const int *gate) {
  if (gate && !*gate) return;
  fold_body<float>(f, f.out, hw_blk());
}
__global__ void __launch_bounds__(256) k_fold64(FoldPass f) { fold_body<double>(f, f.out64, hw_blk()); }
// batched: the FP64 fold of each run's FoldPass
__global__ void __launch_bounds__(256) k_fold_b(BatchTab t) {
  const BatchIdx q = batch_idx(t);
  const FoldPass &f = ((const FoldPass *)t.args)[q.r];
  fold_body<double>(f, f.out64, q.b);
}

// Iteration q of the PCG with the folded preconditioner: α = rz/pAp, x += αP and r −= αQ at the
// four mirror images of each quarter point (each range element is one image of exactly one
// point), the partials of r·r, and the fold of the new r into the quarters, in one pass.  IT += 1.
// The convergence check that k_fold's first pass would take moves to the forward x-y launch, which
// reads these partials (fold_blocks of them).
// T = float: the refinement's inner iteration on a.f, its x the increment d (set at q = 0)
template <class T>
__global__ void __launch_bounds__(256) k_pcg_xr_fold(PcgArgs a, FoldPass f) {
  constexpr bool F32 = sizeof(T) == 4;
  if (!a.d.fl[FL_ACT0 + (a.q & 1)]) return;
  __shared__ double sh[8];
  const int tid = threadIdx.x;
  const int bid = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
  const double pAp = sum_parts(region(a.d, PART_PQ), a.nparts_in, 0, sh, tid);
  const double al = a.d.sc[SC_RZ0 + (a.q & 1)] / pAp;
  if (bid == 0 && tid == 0) a.d.fl[FL_IT] += 1;
  int c, i, j, kg;
  fold_index(f.q, c, i, j, kg);
  const int L0 = f.q.L[c][0], L1 = f.q.L[c][1], h0 = f.q.h[c][0], h1 = f.q.h[c][1], e2 = f.q.e2[c];
  double s = 0.0;
  if (i < h0 && j < h1) {
    const int is = L0 - 1 - i, js = L1 - 1 - j;
    const bool ci = is == i, cj = js == j;
    const View v = f.in[c];  // r's view; x, P, Q share its layout (node-shaped vectors)
    T *X = F32 ? (T *)a.f.d.c[c] : (T *)a.x.c[c], *R = F32 ? (T *)a.f.r.c[c] : (T *)a.r.c[c];
    const T *P = F32 ? (const T *)a.f.Pnew.c[c] : (const T *)a.Pnew.c[c], *Q = F32 ? (const T *)a.f.Q.c[c] : (const T *)a.Q.c[c];
    const size_t base = (size_t)((const T *)v.p - R);
    const size_t o[4] = {base + i + (size_t)v.s1 * j, base + is + (size_t)v.s1 * j, base + i + (size_t)v.s1 * js,
                         base + is + (size_t)v.s1 * js};
    const bool w[4] = {true, !ci, !cj, !ci && !cj};
    const size_t qp = (size_t)i + (size_t)h0 * j, qs = (size_t)h0 * h1;
    float *q00 = f.out + f.q.off[c][0] + qp, *q01 = f.out + f.q.off[c][1] + qp;
    float *q10 = f.out + f.q.off[c][2] + qp, *q11 = f.out + f.q.off[c][3] + qp;
    for (int k = kg; k < e2; k += 4) {
      const size_t zk = (size_t)v.s2 * k;
      double xv[4], pv[4], rv[4], qv[4];
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        xv[t] = X[o[t] + zk];
        pv[t] = P[o[t] + zk];
        rv[t] = R[o[t] + zk];
        qv[t] = Q[o[t] + zk];
      }
      double rn[4];
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        rn[t] = rv[t] - al * qv[t];
        if (F32) rn[t] = (double)(T)rn[t];  // the value stored, in the sums and the fold too
        if (!w[t]) continue;
        X[o[t] + zk] = (F32 && a.q == 0) ? (T)(al * pv[t]) : (T)(xv[t] + al * pv[t]);
        R[o[t] + zk] = (T)rn[t];
        s += rn[t] * rn[t];
      }
      const double sj = ci ? rn[0] : rn[0] + rn[1], dj = ci ? 0.0 : rn[0] - rn[1];
      const double ss = ci ? rn[2] : rn[2] + rn[3], ds = ci ? 0.0 : rn[2] - rn[3];
      const size_t qk = qs * k;
      q00[qk] = (float)(cj ? sj : sj + ss);
      q01[qk] = (float)(cj ? 0.0 : sj - ss);
      q10[qk] = (float)(cj ? dj : dj + ds);
      q11[qk] = (float)(cj ? 0.0 : dj - ds);
    }
  }
  const double r0 = block_sum_t(s, sh, tid);
  if (tid == 0) {
    double *pr = region_w(a.d, PART_RR);
    pr[2 * bid] = r0;
    pr[2 * bid + 1] = 0.0;
  }
}

hipError_t launch_pcg_xr_fold(const PcgArgs &a, const FoldPass &f, hipStream_t s) {
  if (g_rec) return hipErrorNotSupported;
  if (fold_blocks(f.q) > MAX_PART) return hipErrorInvalidValue;
  TDGL_LAUNCH(k_pcg_xr_fold<double>, fold_grid(f.q), dim3(256), 0, s, a, f);
  return hipGetLastError();
}
hipError_t launch_pcg32_xr_fold(const PcgArgs &a, const FoldPass &f, hipStream_t s) {
  if (g_rec) return hipErrorNotSupported;
  if (fold_blocks(f.q) > MAX_PART) return hipErrorInvalidValue;
  TDGL_LAUNCH(k_pcg_xr_fold<float>, fold_grid(f.q), dim3(256), 0, s, a, f);
  return hipGetLastError();
}

// The check of a refinement round (RefFold) and the fold of the inner solve's Float32 residual.
__global__ void __launch_bounds__(256) k_ref_fold(RefFold r) {
  __shared__ double sh[8];
  const int tid = threadIdx.x;
  const int bid = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
  int *fl = r.d.fl;
  if (!*r.gate) {
    if (bid == 0 && tid == 0) {
      fl[FL_OACT] = 0;
      fl[FL_ACT0] = 0;
      fl[FL_ACT1] = 0;
    }
    return;
  }
  const double nr = sqrt(sum_parts(region(r.d, PART_RR), r.nparts, 0, sh, tid));
  const double tol = r.d.sc[SC_TOL], nb = r.d.sc[SC_NB];
  const int active = (nr > tol * nb && fl[FL_IT] < fl[FL_MAXIT]) ? 1 : 0;
  if (bid == 0 && tid == 0) {
    fl[FL_OACT] = active;
    fl[FL_ACT0] = active;
    fl[FL_ACT1] = 0;
    r.d.sc[SC_NBI] = nr;
  }
  if (!active || !r.fold) return;
  const FoldPass &f = r.f;
  int c, i, j, kg;
  fold_index(f.q, c, i, j, kg);
  const int L0 = f.q.L[c][0], L1 = f.q.L[c][1], h0 = f.q.h[c][0], h1 = f.q.h[c][1], e2 = f.q.e2[c];
  if (i >= h0 || j >= h1) return;
  const int is = L0 - 1 - i, js = L1 - 1 - j;
  const bool ci = is == i, cj = js == j;
  const View v = f.in[c];
  const float *base = (const float *)v.p;
  const size_t o_ij = (size_t)i + (size_t)v.s1 * j, o_sj = (size_t)is + (size_t)v.s1 * j;
  const size_t o_is = (size_t)i + (size_t)v.s1 * js, o_ss = (size_t)is + (size_t)v.s1 * js;
  const size_t qp = (size_t)i + (size_t)h0 * j, qs = (size_t)h0 * h1;
  float *q00 = f.out + f.q.off[c][0] + qp, *q01 = f.out + f.q.off[c][1] + qp;
  float *q10 = f.out + f.q.off[c][2] + qp, *q11 = f.out + f.q.off[c][3] + qp;
  for (int k = kg; k < e2; k += 4) {
    const float *pk = base + (size_t)v.s2 * k;
    const double a = pk[o_ij], b = pk[o_sj], cc = pk[o_is], d = pk[o_ss];
    const double sj = ci ? a : a + b, dj = ci ? 0.0 : a - b;
    const double ss = ci ? cc : cc + d, ds = ci ? 0.0 : cc - d;
    const size_t qk = qs * k;
    q00[qk] = (float)(cj ? sj : sj + ss);
    q01[qk] = (float)(cj ? 0.0 : sj - ss);
    q10[qk] = (float)(cj ? dj : dj + ds);
    q11[qk] = (float)(cj ? 0.0 : dj - ds);
  }
}
hipError_t launch_ref_fold(const RefFold &r, hipStream_t s) {
  if (g_rec) return hipErrorNotSupported;
  if (fold_blocks(r.f.q) > MAX_PART) return hipErrorInvalidValue;
  TDGL_LAUNCH(k_ref_fold, fold_grid(r.f.q), dim3(256), 0, s, r);
  return hipGetLastError();
}

hipError_t launch_fold(const FoldPass &f, const int *gate, hipStream_t s) {
  if (fold_blocks(f.q) > MAX_PART) return hipErrorInvalidValue;
  if (g_rec) return (f.out64 && !gate) ? record(BK_FOLD, Geo{}, &f, sizeof f, 0, true, fold_grid(f.q), 0) : hipErrorNotSupported;
  if (f.out64) {
    if (gate) return hipErrorInvalidValue;  // the FP64 fold is never gated
    TDGL_LAUNCH(k_fold64, fold_grid(f.q), dim3(256), 0, s, f);
  } else {
    TDGL_LAUNCH(k_fold, fold_grid(f.q), dim3(256), 0, s, f, gate);
  }
  return hipGetLastError();
}

// The inverse of the fold: the element at each of the four mirror images is the sum of the four
// quarters, the odd quarter of an axis with the sign of the image (-1 on the mirrored side); the
// products r·z and r·zold of the values written, partials per block.
// F32: u.f32 (out, r and zold Float32); ADDTH: u.addth (out = th + the value); fixed per instance
template <class TQ, bool F32, bool ADDTH>
__device__ __forceinline__ void unfold_body(const UnfoldPass &u, const TQ *in, const Blk &b, int bid) {
  __shared__ double sh[8];
  const int tid = threadIdx.x;
  int c, i, j, kg;
  fold_index(b, c, i, j, kg);
  const int L0 = u.q.L[c][0], L1 = u.q.L[c][1], h0 = u.q.h[c][0], h1 = u.q.h[c][1], e2 = u.q.e2[c];
  double d0 = 0.0, d1 = 0.0;
  if (i < h0 && j < h1) {
    const int is = L0 - 1 - i, js = L1 - 1 - j;
    const bool ci = is == i, cj = js == j;
    const size_t qp = (size_t)i + (size_t)h0 * j, qs = (size_t)h0 * h1;
    const TQ *q00 = in + u.q.off[c][0] + qp, *q01 = in + u.q.off[c][1] + qp;
    const TQ *q10 = in + u.q.off[c][2] + qp, *q11 = in + u.q.off[c][3] + qp;
    const View vo = u.out[c], vt = u.th[c];
    const size_t o[4] = {(size_t)i + (size_t)vo.s1 * j, (size_t)is + (size_t)vo.s1 * j, (size_t)i + (size_t)vo.s1 * js,
                         (size_t)is + (size_t)vo.s1 * js};
    const size_t ot[4] = {(size_t)i + (size_t)vt.s1 * j, (size_t)is + (size_t)vt.s1 * j, (size_t)i + (size_t)vt.s1 * js,
                          (size_t)is + (size_t)vt.s1 * js};
    const bool w[4] = {true, !ci, !cj, !ci && !cj};
    for (int k = kg; k < e2; k += 4) {
      const size_t qk = qs * k;
      const double r00 = q00[qk], r01 = q01[qk], r10 = q10[qk], r11 = q11[qk];
      const double val[4] = {(r00 + r10) + (r01 + r11), (r00 - r10) + (r01 - r11), (r00 + r10) - (r01 + r11),
                             (r00 - r10) - (r01 - r11)};
      const size_t zk = (size_t)vo.s2 * k;
      double rv[4] = {0, 0, 0, 0}, zv[4] = {0, 0, 0, 0}, tv[4] = {0, 0, 0, 0};
      if (u.dot)
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          if (!w[t]) continue;
          rv[t] = F32 ? (double)((const float *)u.r[c].p)[o[t] + zk] : ((const double *)u.r[c].p)[o[t] + zk];
          if (u.dot == 2)
            zv[t] = F32 ? (double)((const float *)u.zold[c].p)[o[t] + zk] : ((const double *)u.zold[c].p)[o[t] + zk];
        }
      if (ADDTH)
#pragma unroll
        for (int t = 0; t < 4; ++t)
          if (w[t]) tv[t] = ((const double *)vt.p)[ot[t] + (size_t)vt.s2 * k];
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        if (!w[t]) continue;
        // Float32: the value as written (rounded), in the dots too
        const double vw = F32 ? (double)(float)val[t] : val[t];
        if (F32)
          ((float *)vo.p)[o[t] + zk] = (float)val[t];
        else if (ADDTH)
          ((double *)vo.p)[o[t] + zk] = tv[t] + val[t];
        else
          ((double *)vo.p)[o[t] + zk] = val[t];
        if (u.dot) {
          d0 += rv[t] * vw;
          if (u.dot == 2) d1 += rv[t] * zv[t];
        }
      }
    }
  }
  if (u.dot) {
    block_sum2_t(d0, d1, sh, tid);
    if (tid == 0) {
      double *pr = region_w(u.d, PART_RZ);
      pr[2 * bid] = d0;
      pr[2 * bid + 1] = d1;
    }
  }
}
// Float32 quarters (the preconditioner's) and FP64 ones (the separable solve's) in kernels of their
// own: one kernel holding both would take the registers of the larger
template <bool F32>
__global__ void __launch_bounds__(256) k_unfold(UnfoldPass u, const int *gate) {
  if (gate && !*gate) return;
  const int bid = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
  unfold_body<float, F32, false>(u, u.in, hw_blk(), bid);
}
template <bool ADDTH>
__global__ void __launch_bounds__(256) k_unfold64(UnfoldPass u) {
  const int bid = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
  unfold_body<double, false, ADDTH>(u, u.in64, hw_blk(), bid);
}
// batched: the FP64 unfold of each run's UnfoldPass (no dots)
__global__ void __launch_bounds__(256) k_unfold_b(BatchTab t) {
  const BatchIdx q = batch_idx(t);
  const UnfoldPass &u = ((const UnfoldPass *)t.args)[q.r];
  unfold_body<double, false, true>(u, u.in64, q.b, 0);
}

hipError_t launch_unfold(const UnfoldPass &u, const int *gate, hipStream_t s) {
  if (fold_blocks(u.q) > MAX_PART) return hipErrorInvalidValue;
  if (g_rec)
    return (u.in64 && u.addth && !gate && !u.dot) ? record(BK_UNFOLD, Geo{}, &u, sizeof u, 0, true, fold_grid(u.q), 0)
                                                  : hipErrorNotSupported;
  if (u.in64) {
    if (gate || u.f32) return hipErrorInvalidValue;  // the FP64 unfold is never gated, its output FP64
    if (u.addth)
      TDGL_LAUNCH(k_unfold64<true>, fold_grid(u.q), dim3(256), 0, s, u);
    else
      TDGL_LAUNCH(k_unfold64<false>, fold_grid(u.q), dim3(256), 0, s, u);
  } else {
    if (u.addth) return hipErrorInvalidValue;  // the preconditioner writes z
    if (u.f32)
      TDGL_LAUNCH(k_unfold<true>, fold_grid(u.q), dim3(256), 0, s, u, gate);
    else
      TDGL_LAUNCH(k_unfold<false>, fold_grid(u.q), dim3(256), 0, s, u, gate);
  }
  return hipGetLastError();
}

// The z pass and projector of k_zmodes_run on modes held in quarters.  A block owns 64 modes
// (mx = 2 mm + px, my = 2 qq + py, all mz) of one parity quarter (px, py), one per lane, taken in
// the order of (mm, qq) flattened along mm (so rows of the quarter share blocks): Z of those
// modes sits in Z's quarter (px, py) at (mm, qq), Y (at my - 1) in Y's quarter (px, 1 - py) at
// (mm, qq - 1 + py), X (at mx - 1) in X's quarter (1 - px, py) at (mm - 1 + px, qq), each a
// contiguous run; the results go back to the same places.  The block's G waves split the z modes
// (and, going back, the z positions) into contiguous groups of mzc: a wave's group index is
// uniform, so the bases are read with uniform addresses (scalar loads, no LDS), and the columns are
// shared through LDS.  Every global store comes after the last load of the bases.
__device__ __forceinline__ float fma_t(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fma_t(double a, double b, double c) { return __builtin_fma(a, b, c); }
// T = float with ZModesQ (the Float32 preconditioner), T = double with ZModesQ64 (the FP64
// separable solve); block (bx, bz) of the launch's grid
// LB: the four bases staged in LDS after the columns (zq_lds_bases elements each) and read there
// with wave-uniform addresses (broadcasts), instead of through the scalar cache
__host__ __device__ constexpr int zq_lds_bases(int nz) { return nz * nz + nz + 32; }
template <int G, int MZ, class T, class A, bool LB = false>
__device__ __forceinline__ void zmodes_q_body(const A &m, int bx, int bz) {
  extern __shared__ __align__(16) unsigned char zq_smem[];
  const int nx = m.nx, ny = m.ny, nz = m.nz;
  T *Cx = (T *)zq_smem, *Cy = Cx + nz * 64, *Cz = Cy + nz * 64;
  const T *zb[4] = {m.zb[0], m.zb[1], m.zb[2], m.zb[3]};
  if constexpr (LB) {
    // every base and nz + 32 elements of its slack (the last group's reads past nz, and the
    // prefetch of row nz)
    const int nb = zq_lds_bases(nz);
    T *Bs = Cz + nz * 64;
    for (int k = 0; k < 4; ++k) {
      for (int t = (int)threadIdx.x; t < nb; t += 64 * G) Bs[k * nb + t] = m.zb[k][t];
      zb[k] = Bs + k * nb;
    }
  }
  const int p = (int)(threadIdx.x & 63);
  const int g = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int mzc = (nz + G - 1) / G;  // modes (positions) of a group, at most MZ
  const int m0 = g * mzc;
  const int px = bz >> 1, py = bz & 1;
  const Quarters &q = m.q;
  const int f = bx * 64 + p;  // the mode (mm, qq) of Z's quarter, flattened along mm
  const int mm = f % q.h[2][0], qq = f / q.h[2][0];
  const int mx = 2 * mm + px, my = 2 * qq + py;
  const bool live = qq < q.h[2][1] && mx < nx && my < ny;
  const bool hx = live && mx > 0, hy = live && my > 0;
  // positions in the three quarters (clamped into each quarter; the absent values selected after)
  const int xi = min(max(mm - 1 + px, 0), q.h[0][0] - 1), xj = min(qq, q.h[0][1] - 1);
  const int yi = min(mm, q.h[1][0] - 1), yj = min(max(qq - 1 + py, 0), q.h[1][1] - 1);
  const int zi = min(mm, q.h[2][0] - 1), zj = min(qq, q.h[2][1] - 1);
  T *colx = m.modes + q.off[0][2 * (1 - px) + py] + xi + (size_t)q.h[0][0] * xj;
  T *coly = m.modes + q.off[1][2 * px + (1 - py)] + yi + (size_t)q.h[1][0] * yj;
  T *colz = m.modes + q.off[2][2 * px + py] + zi + (size_t)q.h[2][0] * zj;
  const size_t stx = (size_t)q.h[0][0] * q.h[0][1], sty = (size_t)q.h[1][0] * q.h[1][1], stz = (size_t)q.h[2][0] * q.h[2][1];
  for (int z = g; z < nz; z += G) {
    const T vx = colx[stx * z], vy = coly[sty * z], vz = colz[stz * min(z, nz - 2)];
    Cx[z * 64 + p] = hx ? vx : T(0);
    Cy[z * 64 + p] = hy ? vy : T(0);
    Cz[z * 64 + p] = (live && z < nz - 1) ? vz : T(0);
  }
  __syncthreads();
  // forward: the modes m0 .. m0 + MZ - 1 of the three columns (those past the group's mzc are the
  // next group's, or none: computed and dropped); the padded bases by rows, V(n, mz) and U(n, mz - 1)
  const T *VR = zb[0], *UR = zb[1];
  T ax[MZ], ay[MZ], az[MZ];
#pragma unroll
  for (int r = 0; r < MZ; ++r) ax[r] = ay[r] = az[r] = T(0);
  // the next row of the bases is requested before this row's products (row nz, read past the last,
  // lies in the bases' slack), so the uniform loads' latency overlaps the arithmetic
  T vn[MZ], un[MZ];
#pragma unroll
  for (int r = 0; r < MZ; ++r) {
    vn[r] = VR[m0 + r];
    un[r] = UR[m0 + r];
  }
  for (int n = 0; n < nz; ++n) {
    const T cx = Cx[n * 64 + p], cy = Cy[n * 64 + p], cz = Cz[n * 64 + p];
    T v[MZ], u[MZ];
#pragma unroll
    for (int r = 0; r < MZ; ++r) {
      v[r] = vn[r];
      u[r] = un[r];
    }
    const T *vq = VR + (size_t)(n + 1) * nz + m0, *uq = UR + (size_t)(n + 1) * nz + m0;
#pragma unroll
    for (int r = 0; r < MZ; ++r) {
      vn[r] = vq[r];
      un[r] = uq[r];
    }
#pragma unroll
    for (int r = 0; r < MZ; ++r) {
      ax[r] = fma_t(cx, v[r], ax[r]);
      ay[r] = fma_t(cy, v[r], ay[r]);
      az[r] = fma_t(cz, u[r], az[r]);
    }
  }
  __syncthreads();
  {
    const T k1 = live ? m.sx[mx] : T(0), k2 = live ? m.sy[my] : T(0);
#pragma unroll
    for (int r = 0; r < MZ; ++r) {
      const int mz = m0 + r;
      if (r < mzc && mz < nz) {
        const T k3 = m.sz[mz];
        const bool hz = mz > 0;
        T x1 = hx ? ax[r] : T(0), y1 = hy ? ay[r] : T(0), z1 = (live && hz) ? az[r] : T(0);
        const T k2s = k1 * k1 + k2 * k2 + k3 * k3;
        if (live && k2s != 0) project<T>(k1, k2, k3, (T)m.a, x1, y1, z1, hx, hy, hz);
        Cx[mz * 64 + p] = x1;
        Cy[mz * 64 + p] = y1;
        Cz[mz * 64 + p] = z1;
      }
    }
  }
  __syncthreads();
  // back: the positions m0 .. m0 + MZ - 1; the padded bases by columns, V(n, qm) and U(n, qm - 1)
  const T *VC = zb[2], *UC = zb[3];
  T ox[MZ], oy[MZ], oz[MZ];
#pragma unroll
  for (int r = 0; r < MZ; ++r) ox[r] = oy[r] = oz[r] = T(0);
#pragma unroll
  for (int r = 0; r < MZ; ++r) {
    vn[r] = VC[m0 + r];
    un[r] = UC[m0 + r];
  }
  for (int qm = 0; qm < nz; ++qm) {
    const T cx = Cx[qm * 64 + p], cy = Cy[qm * 64 + p], cz = Cz[qm * 64 + p];
    T v[MZ], u[MZ];
#pragma unroll
    for (int r = 0; r < MZ; ++r) {
      v[r] = vn[r];
      u[r] = un[r];
    }
    const T *vq = VC + (size_t)(qm + 1) * nz + m0, *uq = UC + (size_t)(qm + 1) * nz + m0;
#pragma unroll
    for (int r = 0; r < MZ; ++r) {
      vn[r] = vq[r];
      un[r] = uq[r];
    }
#pragma unroll
    for (int r = 0; r < MZ; ++r) {
      ox[r] = fma_t(cx, v[r], ox[r]);
      oy[r] = fma_t(cy, v[r], oy[r]);
      oz[r] = fma_t(cz, u[r], oz[r]);
    }
  }
  if (!live) return;
#pragma unroll
  for (int r = 0; r < MZ; ++r) {
    const int n = m0 + r;
    if (r < mzc && n < nz) {
      if (hx) colx[stx * n] = ox[r];
      if (hy) coly[sty * n] = oy[r];
      if (n < nz - 1) colz[stz * n] = oz[r];
    }
  }
}
template <int G, int MZ>
__global__ void __launch_bounds__(64 * G) k_zmodes_q(ZModesQ m, const int *gate) {
  if (gate && !*gate) return;
  zmodes_q_body<G, MZ, float, ZModesQ>(m, (int)blockIdx.x, (int)blockIdx.z);
}
template <int G, int MZ>
__global__ void __launch_bounds__(64 * G) k_zmodes_q64(ZModesQ64 m) {
  zmodes_q_body<G, MZ, double, ZModesQ64, true>(m, (int)blockIdx.x, (int)blockIdx.z);
}
// batched: run r's blocks take its own (mode block, quarter) grid
template <int G, int MZ>
__global__ void __launch_bounds__(64 * G) k_zmodes_q64_b(BatchTab t) {
  const BatchIdx q = batch_idx(t);
  zmodes_q_body<G, MZ, double, ZModesQ64, true>(((const ZModesQ64 *)t.args)[q.r], q.b.x, q.b.z);
}

template <int G>
static hipError_t zmodes_q_g(const ZModesQ &m, const int *gate, hipStream_t s) {
  const int mzc = (m.nz + G - 1) / G;
  const dim3 grid(cdiv(m.q.h[2][0] * m.q.h[2][1], 64), 1, 4), blk(64 * G);
  const size_t ldr = 3 * (size_t)m.nz * 64 * sizeof(float);
  switch (mzc <= 2 ? 2 : mzc <= 4 ? 4 : mzc <= 6 ? 6 : mzc <= 8 ? 8 : mzc <= 12 ? 12 : 16) {
    case 2: TDGL_LAUNCH((k_zmodes_q<G, 2>), grid, blk, ldr, s, m, gate); break;
    case 4: TDGL_LAUNCH((k_zmodes_q<G, 4>), grid, blk, ldr, s, m, gate); break;
    case 6: TDGL_LAUNCH((k_zmodes_q<G, 6>), grid, blk, ldr, s, m, gate); break;
    case 8: TDGL_LAUNCH((k_zmodes_q<G, 8>), grid, blk, ldr, s, m, gate); break;
    case 12: TDGL_LAUNCH((k_zmodes_q<G, 12>), grid, blk, ldr, s, m, gate); break;
    default: TDGL_LAUNCH((k_zmodes_q<G, 16>), grid, blk, ldr, s, m, gate); break;
  }
  return hipGetLastError();
}

// The z pass and projector of k_zmodes_q on v_mfma_f32_16x16x4f32.  A block of 4 waves owns 64
// columns of one parity quarter (flattened along mm as in k_zmodes_q); wave w takes the columns
// 16 w .. 16 w + 15.  With NZP = 16 NT >= nz (zero rows and columns beyond nz):
//   forward  X~(mz, col) = Σ_n V(n, mz) X(n, col)  (and Y with V, Z with U(n, mz − 1)),
//   inverse  x(n, col)   = Σ_mz V(n, mz) X~(mz, col),
// each a 16 x 16 tile per (row tile, this wave's columns) with A the basis, from LDS, k contiguous:
// the instruction u of a 16-k group g takes k = 16 g + 4 lk + u (lane l = 16 lk + lr), so one
// 16-byte LDS read gives a lane its four k-steps.  The B operands never pass through LDS: the
// forward's are the lane's own loads of its column at n = 16 g + 4 lk + u, and a forward tile
// leaves in the lane exactly the modes mz = 16 g + 4 lk + u the inverse's instruction (g, u) takes
// as B, so the projected tiles are the inverse's operands where they lie.  The forward tiles of X,
// Y and Z of the same (mz, col) sit in the same lane and register, where the projector runs.  The
// block copies the bases once from their padded image (m.zmf: Vt, Ut, V, U, each NZP rows of
// NZP + 4 words, 4 mod 8: 16 lanes' 16-byte reads are bank-conflict free) while its columns load.
// The sums are grouped four products at a time (as the plane kernels'), not in k_zmodes_q's order:
// the results differ from it by Float32 rounding.
template <int NT>
__global__ void __launch_bounds__(256) k_zmodes_mf(ZModesQ m, const int *gate) {
  if (gate && !*gate) return;
  typedef v4f V4;
  constexpr int NZP = 16 * NT, S = NZP + 4;
  extern __shared__ __align__(16) unsigned char zm_smem[];
  float *B4 = (float *)zm_smem;  // 4 bases [row][k]: Vt (mz, n), Ut (mz, n), V (n, mz), U (n, mz)
  const int nx = m.nx, ny = m.ny, nz = m.nz, tid = threadIdx.x;
  const Quarters &q = m.q;
  const int px = (int)(blockIdx.z >> 1), py = (int)(blockIdx.z & 1);
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6), l = tid & 63, lr = l & 15, lk = l >> 4;
  // this lane's column (lanes 16 consecutive columns: runs along mm) and its three component columns
  const int f = blockIdx.x * 64 + 16 * w + lr;
  const int mm = f % q.h[2][0], qq = f / q.h[2][0];
  const int mx = 2 * mm + px, my = 2 * qq + py;
  const bool live = qq < q.h[2][1] && mx < nx && my < ny;
  const bool hx = live && mx > 0, hy = live && my > 0;
  const int xi = min(max(mm - 1 + px, 0), q.h[0][0] - 1), xj = min(qq, q.h[0][1] - 1);
  const int yi = min(mm, q.h[1][0] - 1), yj = min(max(qq - 1 + py, 0), q.h[1][1] - 1);
  const int zi = min(mm, q.h[2][0] - 1), zj = min(qq, q.h[2][1] - 1);
  float *colx = m.modes + q.off[0][2 * (1 - px) + py] + xi + (size_t)q.h[0][0] * xj;
  float *coly = m.modes + q.off[1][2 * px + (1 - py)] + yi + (size_t)q.h[1][0] * yj;
  float *colz = m.modes + q.off[2][2 * px + py] + zi + (size_t)q.h[2][0] * zj;
  const size_t stx = (size_t)q.h[0][0] * q.h[0][1], sty = (size_t)q.h[1][0] * q.h[1][1], stz = (size_t)q.h[2][0] * q.h[2][1];
  // the forward's B operands: n = 16 g + 4 lk + u of the three columns (clamped reads, zero where absent)
  V4 Bx[NT], By[NT], Bz[NT];
#pragma unroll
  for (int g = 0; g < NT; ++g)
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int n = 16 * g + 4 * lk + u, nn = min(n, nz - 1);
      const float vx = colx[stx * nn], vy = coly[sty * nn], vz = colz[stz * min(n, nz - 2)];
      Bx[g][u] = (hx && n < nz) ? vx : 0.0f;
      By[g][u] = (hy && n < nz) ? vy : 0.0f;
      Bz[g][u] = (live && n < nz - 1) ? vz : 0.0f;
    }
  // the bases: a straight copy of the padded image
  for (int e = tid; e < NZP * S; e += 256) ((V4 *)B4)[e] = ((const V4 *)m.zmf)[e];
  __syncthreads();
  const float *Vt = B4, *Ut = B4 + NZP * S, *Vr = B4 + 2 * NZP * S, *Ur = B4 + 3 * NZP * S;
  // forward: acc[c][t] is the tile of rows mz = 16 t .. 16 t + 15 and this wave's 16 columns
  V4 acc[3][NT];
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[c][t] = V4{0, 0, 0, 0};
#pragma unroll
  for (int g = 0; g < NT; ++g) {
    V4 Av[NT], Au[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      Av[t] = *(const V4 *)(Vt + (16 * t + lr) * S + 16 * g + 4 * lk);
      Au[t] = *(const V4 *)(Ut + (16 * t + lr) * S + 16 * g + 4 * lk);
    }
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        acc[0][t] = __builtin_amdgcn_mfma_f32_16x16x4f32(Av[t][u], Bx[g][u], acc[0][t], 0, 0, 0);
        acc[1][t] = __builtin_amdgcn_mfma_f32_16x16x4f32(Av[t][u], By[g][u], acc[1][t], 0, 0, 0);
        acc[2][t] = __builtin_amdgcn_mfma_f32_16x16x4f32(Au[t][u], Bz[g][u], acc[2][t], 0, 0, 0);
      }
  }
  // the projector on this lane's modes: element r of tile t is mz = 16 t + 4 lk + r, column 16 w + lr
  {
    const float k1 = live ? m.sx[mx] : 0.0f, k2 = live ? m.sy[my] : 0.0f;
#pragma unroll
    for (int t = 0; t < NT; ++t)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int mz = 16 * t + 4 * lk + r;
        if (mz >= nz) continue;
        const float k3 = m.sz[mz];
        const bool hz = mz > 0;
        float x1 = hx ? acc[0][t][r] : 0.0f, y1 = hy ? acc[1][t][r] : 0.0f, z1 = (live && hz) ? acc[2][t][r] : 0.0f;
        const float k2s = k1 * k1 + k2 * k2 + k3 * k3;
        if (live && k2s != 0) project<float>(k1, k2, k3, (float)m.a, x1, y1, z1, hx, hy, hz);
        acc[0][t][r] = x1;
        acc[1][t][r] = y1;
        acc[2][t][r] = z1;
      }
  }
  // inverse: out[c][t] is the tile of rows n = 16 t .. 16 t + 15; the instruction (g, u) takes the
  // modes mz = 16 g + 4 lk + u, element u of the lane's projected tile g
  V4 out[3][NT];
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int t = 0; t < NT; ++t) out[c][t] = V4{0, 0, 0, 0};
#pragma unroll
  for (int g = 0; g < NT; ++g) {
    V4 Av[NT], Au[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      Av[t] = *(const V4 *)(Vr + (16 * t + lr) * S + 16 * g + 4 * lk);
      Au[t] = *(const V4 *)(Ur + (16 * t + lr) * S + 16 * g + 4 * lk);
    }
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        out[0][t] = __builtin_amdgcn_mfma_f32_16x16x4f32(Av[t][u], acc[0][g][u], out[0][t], 0, 0, 0);
        out[1][t] = __builtin_amdgcn_mfma_f32_16x16x4f32(Av[t][u], acc[1][g][u], out[1][t], 0, 0, 0);
        out[2][t] = __builtin_amdgcn_mfma_f32_16x16x4f32(Au[t][u], acc[2][g][u], out[2][t], 0, 0, 0);
      }
  }
  // the positions back into the quarters: element r of tile t is n = 16 t + 4 lk + r
  if (!live) return;
#pragma unroll
  for (int t = 0; t < NT; ++t)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int n = 16 * t + 4 * lk + r;
      if (n >= nz) continue;
      if (hx) colx[stx * n] = out[0][t][r];
      if (hy) coly[sty * n] = out[1][t][r];
      if (n < nz - 1) colz[stz * n] = out[2][t][r];
    }
}

// the floats of the padded image k_zmodes_mf copies into LDS (four bases of NZP rows of NZP + 4)
int zmodes_mf_image(int nz) {
  const int nzp = 16 * ((nz + 15) / 16);
  return 4 * nzp * (nzp + 4);
}

static hipError_t launch_zmodes_mf(const ZModesQ &m, const int *gate, hipStream_t s) {
  const int NT = (m.nz + 15) / 16;
  if (!m.zmf) return hipErrorInvalidValue;
  const dim3 grid(cdiv(m.q.h[2][0] * m.q.h[2][1], 64), 1, 4), blk(256);
  const size_t lds = (size_t)zmodes_mf_image(m.nz) * sizeof(float);
  switch (NT) {
    case 1: TDGL_LAUNCH((k_zmodes_mf<1>), grid, blk, lds, s, m, gate); break;
    case 2: TDGL_LAUNCH((k_zmodes_mf<2>), grid, blk, lds, s, m, gate); break;
    case 3: TDGL_LAUNCH((k_zmodes_mf<3>), grid, blk, lds, s, m, gate); break;
    case 4: TDGL_LAUNCH((k_zmodes_mf<4>), grid, blk, lds, s, m, gate); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

static int zq_mz(int mzc) { return mzc <= 2 ? 2 : mzc <= 4 ? 4 : mzc <= 6 ? 6 : mzc <= 8 ? 8 : mzc <= 12 ? 12 : 16; }
template <int G>
static hipError_t zmodes_q64_g(const ZModesQ64 &m, hipStream_t s) {
  const int mz = zq_mz((m.nz + G - 1) / G);
  const dim3 grid(cdiv(m.q.h[2][0] * m.q.h[2][1], 64), 1, 4), blk(64 * G);
  const size_t ldr = (3 * (size_t)m.nz * 64 + 4 * (size_t)zq_lds_bases(m.nz)) * sizeof(double);
  if (g_rec) return record(BK_ZMODESQ, Geo{}, &m, sizeof m, 100 * G + mz, true, grid, ldr);
  switch (mz) {
    case 2: TDGL_LAUNCH((k_zmodes_q64<G, 2>), grid, blk, ldr, s, m); break;
    case 4: TDGL_LAUNCH((k_zmodes_q64<G, 4>), grid, blk, ldr, s, m); break;
    case 6: TDGL_LAUNCH((k_zmodes_q64<G, 6>), grid, blk, ldr, s, m); break;
    case 8: TDGL_LAUNCH((k_zmodes_q64<G, 8>), grid, blk, ldr, s, m); break;
    case 12: TDGL_LAUNCH((k_zmodes_q64<G, 12>), grid, blk, ldr, s, m); break;
    default: TDGL_LAUNCH((k_zmodes_q64<G, 16>), grid, blk, ldr, s, m); break;
  }
  return hipGetLastError();
}
hipError_t launch_zmodes_q64(const ZModesQ64 &m, hipStream_t s) {
  if (m.nz < 2 || m.nz > ZQ64_NZ_MAX || m.nx < 1 || m.ny < 1) return hipErrorInvalidValue;
  if (m.waves == 8) return zmodes_q64_g<8>(m, s);
  if (m.waves == 4) return zmodes_q64_g<4>(m, s);
  return hipErrorInvalidValue;
}

hipError_t launch_zmodes_q(const ZModesQ &m, const int *gate, hipStream_t s) {
  if (g_rec) return hipErrorNotSupported;
  if (m.nz < 2 || m.nz > ZQ_NZ_MAX || m.nx < 1 || m.ny < 1) return hipErrorInvalidValue;
  if (m.waves == 0) return launch_zmodes_mf(m, gate, s);
  const int waves = m.waves;
  if (waves == 8) return zmodes_q_g<8>(m, gate, s);
  if (waves == 4) return zmodes_q_g<4>(m, gate, s);
  return hipErrorInvalidValue;
}

constexpr int ZM_LB_MAX = 64;  // z bases in LDS up to this nz
template <class T, bool LB>
__global__ void __launch_bounds__(256) k_zmodes(ZModes m, const int *gate) { zmodes_body<T, LB>(m, gate, hw_blk()); }
template <bool LB>
__global__ void __launch_bounds__(256) k_zmodes_b(BatchTab t) {
  const BatchIdx q = batch_idx(t);
  zmodes_body<double, LB>(((const ZModes *)t.args)[q.r], nullptr, q.b);
}

template <class T>
hipError_t launch_zmodes(const ZModes &m, const int *gate, hipStream_t s) {
  if (m.nz < 2 || m.nz > 256 || m.nx < 1 || m.ny < 1) return hipErrorInvalidValue;
  const int C = 256 / m.nz;
  const bool lb = m.nz <= ZM_LB_MAX;
  const size_t lds = (6 * 256 + (lb ? 2 * (size_t)m.nz * m.nz + 2 * (size_t)(m.nz - 1) * (m.nz - 1) : 0)) * sizeof(T);
  if (g_rec)
    return (sizeof(T) != 8 || gate) ? hipErrorNotSupported
                                     : record(BK_ZMODES, Geo{}, &m, sizeof m, lb ? 1 : 0, true, dim3(cdiv(m.nx * m.ny, C)), lds);
  if (m.nz <= ZT_NZ_MAX && !m.columns) {
    // x-runs of 64 modes (32 in FP64, or where 64 would leave fewer than two blocks per CU)
    const bool wide = sizeof(T) == 4 && cdiv(m.nx, 64) * m.ny >= 512;
    const int TX = wide ? 64 : 32;
    const size_t ldr = (3 * (size_t)m.nz * TX + (size_t)m.nz * m.nz + (size_t)(m.nz - 1) * (m.nz - 1)) * sizeof(T);
    if (wide)
      TDGL_LAUNCH((k_zmodes_run<T, 64>), dim3(cdiv(m.nx, 64), m.ny), dim3(256), ldr, s, m, gate);
    else
      TDGL_LAUNCH((k_zmodes_run<T, 32>), dim3(cdiv(m.nx, 32), m.ny), dim3(256), ldr, s, m, gate);
    return hipGetLastError();
  }
  if (lb)
    TDGL_LAUNCH((k_zmodes<T, true>), dim3(cdiv(m.nx * m.ny, C)), dim3(256), lds, s, m, gate);
  else
    TDGL_LAUNCH((k_zmodes<T, false>), dim3(cdiv(m.nx * m.ny, C)), dim3(256), lds, s, m, gate);
  return hipGetLastError();
}
template hipError_t launch_zmodes<float>(const ZModes &, const int *, hipStream_t);
template hipError_t launch_zmodes<double>(const ZModes &, const int *, hipStream_t);

// ============================================================================ normal fluids
// C_µ at a block node (_conv, exchange.jl:316-337, no drag): the τ-masked divergence of the band's
// link supercurrents and its Josephson exchange, in the reference's order of additions
__device__ __forceinline__ double qp_conv1(const Geo &g, const QpConv &a, int mu, int c, int sy, int sz) {
  const cplx *P = a.G[mu];
  const cplx p = cconj(P[c]);
  const double gx = a.wa[0][mu], gy = a.wa[1][mu], gz = a.wa[2][mu];
  const double *tx = a.tau[0], *ty = a.tau[1], *tz = a.tau[2];
  const cplx *Vx = a.V[0], *Vy = a.V[1], *Vz = a.V[2];
  double v = ((gx * (tx[c] * cmul(cmul(p, Vx[c]), P[c + 1]).y + tx[c - 1] * cmul(cmul(p, cconj(Vx[c - 1])), P[c - 1]).y) *
                   a.rx2 +
               gy * (ty[c] * cmul(cmul(p, Vy[c]), P[c + sy]).y + ty[c - sy] * cmul(cmul(p, cconj(Vy[c - sy])), P[c - sy]).y) *
                   a.ry2) +
              gz * (tz[c] * cmul(cmul(p, Vz[c]), P[c + sz]).y + tz[c - sz] * cmul(cmul(p, cconj(Vz[c - sz])), P[c - sz]).y) *
                  a.rz2);
  for (int nu = 0; nu < a.nb; ++nu) {
    if (nu == mu) continue;
    v += a.gam[mu][nu] * cmul(p, a.G[nu][c]).y;
  }
  return v;
}

// One thread per unknown of the region; block nodes (faces included) get the source, the rest 0
__global__ void __launch_bounds__(256) k_qp_conv(Geo g, QpConv a) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int n = a.n1 * a.nj * a.nk;
  if (t >= n) return;
  const int ix = t % a.n1, jy = (t / a.n1) % a.nj, kz = t / (a.n1 * a.nj);
  const int i = ix + 2, j = a.j0 + 1 + jy, k = a.k0 + 1 + kz;
  double src[TDGL_MAX_BANDS] = {0.0, 0.0, 0.0, 0.0};
  if (i >= g.x1 && i <= g.x2 + 1 && j >= g.y1 && j <= g.y2 + 1 && k >= g.z1 && k <= g.z2 + 1) {
    const int c = g.at(i, j, k), sy = g.px, sz = g.px * g.py;
    if (a.mode == TDGL_QP_TWO) {
      const double c1 = qp_conv1(g, a, 0, c, sy, sz), c2 = qp_conv1(g, a, 1, c, sy, sz);
      src[0] = c2 / a.s[1] - c1 / a.s[0];
    } else {
      double cv[TDGL_MAX_BANDS], C = 0.0, R[TDGL_MAX_BANDS];
      for (int mu = 0; mu < a.nb; ++mu) cv[mu] = qp_conv1(g, a, mu, c, sy, sz);
      for (int mu = 0; mu < a.nb; ++mu) C += cv[mu];
      for (int mu = 0; mu < a.nb; ++mu) R[mu] = cv[mu] / a.s[mu] - C;
      for (int kk = 0; kk < a.K; ++kk) {
        double m = 0.0;
        for (int mu = 0; mu < a.nb; ++mu) m += a.Pinv[kk][mu] * R[mu];
        src[kk] = m;
      }
    }
  }
  for (int kk = 0; kk < a.K; ++kk) a.S[kk][t] = src[kk];
}

hipError_t launch_qp_conv(const Geo &g, const QpConv &a, hipStream_t s) {
  if (g_rec) return hipErrorNotSupported;
  const int n = a.n1 * a.nj * a.nk;
  TDGL_LAUNCH(k_qp_conv, dim3(cdiv(n, 256)), dim3(256), 0, s, g, a);
  return hipGetLastError();
}

// Thread (column tc, z index kz) of mode blockIdx.y: forward z with W_z, the division, then the
// inverse z of the block's columns from LDS with V_z (sums in increasing index)
__global__ void __launch_bounds__(256) k_qp_zcol(QpZ a) {
  __shared__ double Tm[256];
  const int nk = a.nk, C = 256 / nk;
  const int t = threadIdx.x, tc = t / nk, kz = t % nk;
  const int col = blockIdx.x * C + tc;
  const size_t n = (size_t)a.n1 * a.nj * nk;
  const int mode = blockIdx.y;
  const bool live = tc < C && col < a.n1 * a.nj;
  const int ix = live ? col % a.n1 : 0, jy = live ? col / a.n1 : 0;
  const size_t cs = (size_t)a.n1 * a.nj;
  const double *in = a.in + mode * n + ix + (size_t)a.n1 * jy;
  if (live) {
    double acc = 0.0;
#pragma unroll 8
    for (int q = 0; q < nk; ++q) acc += a.Wz[kz + (size_t)nk * q] * in[cs * q];
    acc /= a.sigma * (a.ex[ix] + a.ey[jy] + a.ez[kz]) - a.lam[mode];
    Tm[t] = acc;
  }
  __syncthreads();
  if (!live) return;
  double acc = 0.0;
#pragma unroll 8
  for (int q = 0; q < nk; ++q) acc += a.Vz[kz + (size_t)nk * q] * Tm[tc * nk + q];
  a.out[mode * n + ix + (size_t)a.n1 * jy + cs * kz] = acc;
}

hipError_t launch_qp_zcol(const QpZ &a, hipStream_t s) {
  if (g_rec) return hipErrorNotSupported;
  if (a.nk < 1 || a.nk > 256 || a.K < 1 || a.K > TDGL_MAX_BANDS) return hipErrorInvalidValue;
  const int C = 256 / a.nk;
  TDGL_LAUNCH(k_qp_zcol, dim3(cdiv(a.n1 * a.nj, C), a.K), dim3(256), 0, s, a);
  return hipGetLastError();
}

// ============================================================================ PCG
// Kernels over the box i, j, k in 2..N (the union of the three ranges) with 1D blocks of 64 x 4
// nodes and kc levels of k each; vectors node-shaped, zero outside their ranges (the direction P
// is read at neighbours with no test: implicit.jl:300, 406).  Every consumer of a reduction sums
// all the producer's block partials itself, so the block count enters squared: on large grids a
// block takes several levels (box_kc), a thread's levels in increasing k.
struct BoxIdx {
  int i, j, k, c, bid;
  int nk;  // levels of the block: k .. k + nk - 1
};
// With `xcd`, the hardware's block order is taken through xcd_run: each XCD receives a contiguous
// run of the (x, y, level) order, so that the neighbours whose halos a stencil block shares mostly
// sit behind the same L2.  The block's partial slot is its place in that order either way.
__device__ __forceinline__ BoxIdx box_idx(const Geo &g, int kc, int xcd) {
  const int gx = cdiv_d(g.Nx - 1, 64), gy = cdiv_d(g.Ny - 1, 4);
  const int b = xcd ? xcd_run((int)blockIdx.x, (int)gridDim.x) : (int)blockIdx.x;
  BoxIdx r;
  r.bid = b;
  const int bx = b % gx, by = (b / gx) % gy, bz = b / (gx * gy);
  r.i = 2 + bx * 64 + (int)(threadIdx.x & 63);
  r.j = 2 + by * 4 + (int)(threadIdx.x >> 6);
  r.k = 2 + bz * kc;
  r.nk = min(kc, g.Nz - 1 - bz * kc);
  r.c = g.at(r.i, r.j, r.k);
  return r;
}
// level q of a block (its k + q)
__device__ __forceinline__ BoxIdx box_level(const Geo &g, BoxIdx b, int q) {
  b.k += q;
  b.c += q * g.px * g.py;
  return b;
}
int pcg_box_blocks(const Geo &g, int kc) { return cdiv(g.Nx - 1, 64) * cdiv(g.Ny - 1, 4) * cdiv(g.Nz - 1, kc); }
// levels per block: one while the box has at most BOX_PARTS blocks of one level, else the least
// power of two that brings it there
constexpr int BOX_PARTS = 2048;
int box_kc(const Geo &g) {
  const int nxy = cdiv(g.Nx - 1, 64) * cdiv(g.Ny - 1, 4), nz = g.Nz - 1;
  int kc = 1;
  while (kc < nz && nxy * cdiv(nz, kc) > BOX_PARTS) kc *= 2;
  return kc;
}

// Q = (D + aK) P of the three components at box node (i, j, k) (_apply_A!, implicit.jl:216-234):
// h = curl P on the plaquettes of its ranges (zero elsewhere), then D p + s (curl h).  The 24
// values of P the stencil reads are gathered unconditionally (a box node's neighbours lie inside
// the node-shaped arrays, and P is zero outside its ranges), so they are in flight together; the
// ranges then select by value.
enum {
  P0C, P0ZP, P0ZM, P0XMZP, P0XM, P0YP, P0YM, P0XMYP,
  P1C, P1ZP, P1ZM, P1YMZP, P1YM, P1XP, P1YMXP, P1XM,
  P2C, P2YP, P2ZMYP, P2ZM, P2YM, P2XP, P2ZMXP, P2XM, NSTEN
};
__device__ __forceinline__ void gather24(const double *P0, const double *P1, const double *P2, int c, int sy, int sz,
                                         double v[NSTEN]) {
  v[P0C] = P0[c], v[P0ZP] = P0[c + sz], v[P0ZM] = P0[c - sz], v[P0XMZP] = P0[c - 1 + sz], v[P0XM] = P0[c - 1];
  v[P0YP] = P0[c + sy], v[P0YM] = P0[c - sy], v[P0XMYP] = P0[c - 1 + sy];
  v[P1C] = P1[c], v[P1ZP] = P1[c + sz], v[P1ZM] = P1[c - sz], v[P1YMZP] = P1[c - sy + sz], v[P1YM] = P1[c - sy];
  v[P1XP] = P1[c + 1], v[P1YMXP] = P1[c - sy + 1], v[P1XM] = P1[c - 1];
  v[P2C] = P2[c], v[P2YP] = P2[c + sy], v[P2ZMYP] = P2[c - sz + sy], v[P2ZM] = P2[c - sz], v[P2YM] = P2[c - sy];
  v[P2XP] = P2[c + 1], v[P2ZMXP] = P2[c - sz + 1], v[P2XM] = P2[c - 1];
}
// The same gather as raw buffer loads: one 32-bit lane offset (c) shared by the 24 loads, the
// stencil offsets in scalar registers, the base K elements before the arrays so every offset is
// non-negative (K = sz + sy + 1; c - sz >= 0 for the box's nodes).  Node arrays of n elements.
__device__ __forceinline__ void gather24b(const double *P0, const double *P1, const double *P2, int c, int sy, int sz,
                                          int n, double v[NSTEN]) {
  const int K = sz + sy + 1, vo = c * 8, bytes = (n + K) * 8;
  const __amdgpu_buffer_rsrc_t r0 = buf_rsrc(P0 - K, bytes), r1 = buf_rsrc(P1 - K, bytes), r2 = buf_rsrc(P2 - K, bytes);
  auto L = [&](__amdgpu_buffer_rsrc_t r, int d) { return buf_load_s<double>(r, vo, (K + d) * 8); };
  v[P0C] = L(r0, 0), v[P0ZP] = L(r0, sz), v[P0ZM] = L(r0, -sz), v[P0XMZP] = L(r0, -1 + sz), v[P0XM] = L(r0, -1);
  v[P0YP] = L(r0, sy), v[P0YM] = L(r0, -sy), v[P0XMYP] = L(r0, -1 + sy);
  v[P1C] = L(r1, 0), v[P1ZP] = L(r1, sz), v[P1ZM] = L(r1, -sz), v[P1YMZP] = L(r1, -sy + sz), v[P1YM] = L(r1, -sy);
  v[P1XP] = L(r1, 1), v[P1YMXP] = L(r1, -sy + 1), v[P1XM] = L(r1, -1);
  v[P2C] = L(r2, 0), v[P2YP] = L(r2, sy), v[P2ZMYP] = L(r2, -sz + sy), v[P2ZM] = L(r2, -sz), v[P2YM] = L(r2, -sy);
  v[P2XP] = L(r2, 1), v[P2ZMXP] = L(r2, -sz + 1), v[P2XM] = L(r2, -1);
}
template <class T>
__device__ __forceinline__ void apply3v(const PcgArgs &a, const T v[NSTEN], const T dv[3], int i, int j, int k, T q[3]) {
  const int Nx = a.g.Nx, Ny = a.g.Ny, Nz = a.g.Nz;
  const T dydz = (T)a.dydz, dzdx = (T)a.dzdx, dxdy = (T)a.dxdy;
  const T sxy = (T)a.sxy, sxz = (T)a.sxz, syx = (T)a.syx, syz = (T)a.syz, szx = (T)a.szx, szy = (T)a.szy;
  // the ranges of hx (2:Nx, 2:Ny-1, 2:Nz-1), hy (2:Nx-1, 2:Ny, 2:Nz-1), hz (2:Nx-1, 2:Ny-1, 2:Nz)
  auto okx = [&](int ii, int jj, int kk) { return ii >= 2 && ii <= Nx && jj >= 2 && jj <= Ny - 1 && kk >= 2 && kk <= Nz - 1; };
  auto oky = [&](int ii, int jj, int kk) { return ii >= 2 && ii <= Nx - 1 && jj >= 2 && jj <= Ny && kk >= 2 && kk <= Nz - 1; };
  auto okz = [&](int ii, int jj, int kk) { return ii >= 2 && ii <= Nx - 1 && jj >= 2 && jj <= Ny - 1 && kk >= 2 && kk <= Nz; };
  const T hx_c = okx(i, j, k) ? ((v[P2YP] - v[P2C]) - v[P1ZP] + v[P1C]) / dydz : T(0);
  const T hx_zm = okx(i, j, k - 1) ? ((v[P2ZMYP] - v[P2ZM]) - v[P1C] + v[P1ZM]) / dydz : T(0);
  const T hx_ym = okx(i, j - 1, k) ? ((v[P2C] - v[P2YM]) - v[P1YMZP] + v[P1YM]) / dydz : T(0);
  const T hy_c = oky(i, j, k) ? ((v[P0ZP] - v[P0C]) - v[P2XP] + v[P2C]) / dzdx : T(0);
  const T hy_zm = oky(i, j, k - 1) ? ((v[P0C] - v[P0ZM]) - v[P2ZMXP] + v[P2ZM]) / dzdx : T(0);
  const T hy_xm = oky(i - 1, j, k) ? ((v[P0XMZP] - v[P0XM]) - v[P2C] + v[P2XM]) / dzdx : T(0);
  const T hz_c = okz(i, j, k) ? ((v[P1XP] - v[P1C]) - v[P0YP] + v[P0C]) / dxdy : T(0);
  const T hz_ym = okz(i, j - 1, k) ? ((v[P1YMXP] - v[P1YM]) - v[P0C] + v[P0YM]) / dxdy : T(0);
  const T hz_xm = okz(i - 1, j, k) ? ((v[P1C] - v[P1XM]) - v[P0XMYP] + v[P0XM]) / dxdy : T(0);
  q[0] = dv[0] * v[P0C] + sxy * (hz_c - hz_ym) - sxz * (hy_c - hy_zm);
  q[1] = dv[1] * v[P1C] + syz * (hx_c - hx_zm) - syx * (hz_c - hz_xm);
  q[2] = dv[2] * v[P2C] + szx * (hy_c - hy_xm) - szy * (hx_c - hx_ym);
}
__device__ __forceinline__ void apply3(const PcgArgs &a, const double *P0, const double *P1, const double *P2, int i,
                                       int j, int k, int c, double q[3], double pc[3]) {
  double v[NSTEN];
  gather24(P0, P1, P2, c, a.g.px, a.g.px * a.g.py, v);
  const double dv[3] = {a.D[0][c], a.D[1][c], a.D[2][c]};
  apply3v<double>(a, v, dv, i, j, k, q);
  pc[0] = v[P0C];
  pc[1] = v[P1C];
  pc[2] = v[P2C];
}

// a box thread inside the node array (the box's blocks overhang it along x and y)
__device__ __forceinline__ bool box_live(const Geo &g, const BoxIdx &b) { return b.i <= g.Nx && b.j <= g.Ny; }
// the index a thread loads from: its node, or the box's first node for an overhanging thread, so
// every load is unconditional (a load under a divergent condition waits for memory)
__device__ __forceinline__ int box_load_c(const Geo &g, const BoxIdx &b) { return box_live(g, b) ? b.c : g.at(2, 2, 2); }

// _pcg_zero_update! (implicit.jl:251-271) when the source is zero, else the warm start of
// _implicit_curl_fast! (:391-402); nb from the field kernel's partials
__global__ void __launch_bounds__(RED_T) k_pcg_start(PcgArgs a) {
  __shared__ double sh[8];
  const int tid = threadIdx.x;
  const BoxIdx b0 = box_idx(a.g, a.kc, a.xcd);
  const bool live = box_live(a.g, b0);
  // the first level's values are requested before the partials are summed
  const int cl = box_load_c(a.g, b0);
  double xv[3], pv[3];
  for (int comp = 0; comp < 3; ++comp) {
    xv[comp] = a.x.c[comp][cl];
    pv[comp] = a.xprev.c[comp][cl];
  }
  const double nb = sqrt(sum_parts(region(a.d, PART_NB), a.nparts_in, 0, sh, tid));
  if (b0.bid == 0 && tid == 0) {
    a.d.sc[SC_NB] = nb;
    a.d.fl[FL_NZ] = nb > 0 ? 1 : 0;
    a.d.fl[FL_IT] = 0;
  }
  if (!live) return;
  for (int q = 0; q < b0.nk; ++q) {
    const BoxIdx b = box_level(a.g, b0, q);
    if (q > 0)
      for (int comp = 0; comp < 3; ++comp) {
        xv[comp] = a.x.c[comp][b.c];
        pv[comp] = a.xprev.c[comp][b.c];
      }
    for (int comp = 0; comp < 3; ++comp) {
      if (!a.rg.in(comp, b.i, b.j, b.k)) continue;
      double *x = a.x.c[comp], *xp = a.xprev.c[comp];
      const int c = b.c;
      if (nb == 0) {
        if (a.warm && a.extrap) xp[c] = xv[comp];
        x[c] = 0.0;
      } else if (a.warm && a.extrap) {
        if (a.have2) {
          x[c] = 2 * xv[comp] - pv[comp];
          xp[c] = xv[comp];
        } else {
          xp[c] = xv[comp];
        }
      }
    }
  }
}

// the start of the iterations: Q = A x, r = b − Q and the partials of r·r (implicit.jl:407-410)
__global__ void __launch_bounds__(RED_T) k_pcg_resid(PcgArgs a) {
  if (!a.d.fl[FL_NZ]) return;
  __shared__ double sh[8];
  const int tid = threadIdx.x;
  const BoxIdx b0 = box_idx(a.g, a.kc, a.xcd);
  double s = 0.0;
  if (box_live(a.g, b0))
    for (int q = 0; q < b0.nk; ++q) {
      const BoxIdx b = box_level(a.g, b0, q);
      double qv[3], pc[3], bv[3];
      for (int comp = 0; comp < 3; ++comp) bv[comp] = a.b.c[comp][b.c];
      apply3(a, a.x.c[0], a.x.c[1], a.x.c[2], b.i, b.j, b.k, b.c, qv, pc);
      for (int comp = 0; comp < 3; ++comp) {
        if (!a.rg.in(comp, b.i, b.j, b.k)) continue;
        a.Q.c[comp][b.c] = qv[comp];
        const double rv = bv[comp] - qv[comp];
        a.r.c[comp][b.c] = rv;
        s += rv * rv;
      }
    }
  const double r0 = block_sum_t(s, sh, tid);
  if (tid == 0) {
    double *pr = region_w(a.d, PART_RR);
    pr[2 * b0.bid] = r0;
    pr[2 * b0.bid + 1] = 0.0;
  }
}

// iteration q: rz (and β) from the partials of the preconditioner's last pass, P = z + βP
// (P = z on the first) and zold = z (implicit.jl:412-413, 428-436)
__global__ void __launch_bounds__(RED_T) k_pcg_p(PcgArgs a) {
  if (!a.d.fl[FL_ACT0 + (a.q & 1)]) return;
  __shared__ double sh[8];
  const int tid = threadIdx.x;
  const BoxIdx b0 = box_idx(a.g, a.kc, a.xcd);
  const bool live = box_live(a.g, b0);
  double rz2, rzo;
  sum_parts2(region(a.d, PART_RZ), a.nparts_in, rz2, rzo, sh, tid);
  double beta = 0.0;
  if (a.q > 0) {
    const double rz = a.d.sc[SC_RZ0 + ((a.q + 1) & 1)];
    beta = a.flexible ? (rz2 - rzo) / rz : rz2 / rz;
  }
  if (b0.bid == 0 && tid == 0) a.d.sc[SC_RZ0 + (a.q & 1)] = rz2;
  if (!live) return;
  for (int q = 0; q < b0.nk; ++q) {
    const BoxIdx b = box_level(a.g, b0, q);
    double zv[3], pv[3] = {};
    for (int comp = 0; comp < 3; ++comp) {
      zv[comp] = a.z.c[comp][b.c];
      if (a.q > 0) pv[comp] = a.Pold.c[comp][b.c];
    }
    for (int comp = 0; comp < 3; ++comp) {
      if (!a.rg.in(comp, b.i, b.j, b.k)) continue;
      const int c = b.c;
      a.Pnew.c[comp][c] = a.q == 0 ? zv[comp] : zv[comp] + beta * pv[comp];
      if (a.flexible) a.zold.c[comp][c] = zv[comp];
    }
  }
}

// iteration q: Q = A P and the partials of P·Q (implicit.jl:418-420)
__global__ void __launch_bounds__(RED_T) k_pcg_ap(PcgArgs a) {
  if (!a.d.fl[FL_ACT0 + (a.q & 1)]) return;
  __shared__ double sh[8];
  const int tid = threadIdx.x;
  const BoxIdx b0 = box_idx(a.g, a.kc, a.xcd);
  double s = 0.0;
  if (box_live(a.g, b0))
    for (int q = 0; q < b0.nk; ++q) {
      const BoxIdx b = box_level(a.g, b0, q);
      double qv[3], pc[3];
      apply3(a, a.Pnew.c[0], a.Pnew.c[1], a.Pnew.c[2], b.i, b.j, b.k, b.c, qv, pc);
      for (int comp = 0; comp < 3; ++comp) {
        if (!a.rg.in(comp, b.i, b.j, b.k)) continue;
        a.Q.c[comp][b.c] = qv[comp];
        s += pc[comp] * qv[comp];
      }
    }
  const double r0 = block_sum_t(s, sh, tid);
  if (tid == 0) {
    double *pr = region_w(a.d, PART_PQ);
    pr[2 * b0.bid] = r0;
    pr[2 * b0.bid + 1] = 0.0;
  }
}

// iteration q in one launch: β from the partials of the preconditioner's last pass, the direction
// P = z + βPold (P = z on the first) at the node and at every neighbour the operator reads, then
// Q = A P and the partials of P·Q; the node's P and zold = z are stored (implicit.jl:412-420,
// 428-436).  z and Pold are zero outside the ranges, so P is too.
__global__ void __launch_bounds__(RED_T) k_pcg_pap(PcgArgs a) {
  if (!a.d.fl[FL_ACT0 + (a.q & 1)]) return;
  __shared__ double sh[8];
  const int tid = threadIdx.x;
  const BoxIdx b0 = box_idx(a.g, a.kc, a.xcd);
  const bool live = box_live(a.g, b0);
  const int sy = a.g.px, sz = a.g.px * a.g.py;
  const int cl = box_load_c(a.g, b0);
  // the first partials of r·z and r·zold are requested before the first level's stencil loads
  // (two per thread: the registers of the 48 stencil values in flight leave room for two), then
  // summed exactly as sum_parts2 sums them
  const int np = a.nparts_in;
  const __amdgpu_buffer_rsrc_t rp = buf_rsrc(region(a.d, PART_RZ), 16 * np);
  double2 pv[2];
#pragma unroll
  for (int u = 0; u < 2; ++u) pv[u] = buf_load_d2(rp, 16 * min(tid + u * RED_T, np - 1));
  double vz[NSTEN], vp[NSTEN], dv[3];
  gather24b(a.z.c[0], a.z.c[1], a.z.c[2], cl, sy, sz, a.g.n, vz);
  if (a.q > 0) gather24b(a.Pold.c[0], a.Pold.c[1], a.Pold.c[2], cl, sy, sz, a.g.n, vp);
  for (int comp = 0; comp < 3; ++comp) dv[comp] = a.D[comp][cl];
  double rz2 = 0.0, rzo = 0.0;
#pragma unroll
  for (int u = 0; u < 2; ++u) {
    const bool in = tid + u * RED_T < np;
    rz2 += in ? pv[u].x : 0.0;
    rzo += in ? pv[u].y : 0.0;
  }
  for (int b1 = 2 * RED_T; b1 < np; b1 += 2 * RED_T)
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const int e = b1 + tid + u * RED_T;
      const double2 v = buf_load_d2(rp, 16 * min(e, np - 1));
      rz2 += e < np ? v.x : 0.0;
      rzo += e < np ? v.y : 0.0;
    }
  block_sum2_t(rz2, rzo, sh, tid);
  double beta = 0.0;
  if (a.q > 0) {
    const double rz = a.d.sc[SC_RZ0 + ((a.q + 1) & 1)];
    beta = a.flexible ? (rz2 - rzo) / rz : rz2 / rz;
  }
  if (b0.bid == 0 && tid == 0) a.d.sc[SC_RZ0 + (a.q & 1)] = rz2;
  double s = 0.0;
  if (live)
    for (int q = 0; q < b0.nk; ++q) {
      const BoxIdx b = box_level(a.g, b0, q);
      if (q > 0) {
        gather24b(a.z.c[0], a.z.c[1], a.z.c[2], b.c, sy, sz, a.g.n, vz);
        if (a.q > 0) gather24b(a.Pold.c[0], a.Pold.c[1], a.Pold.c[2], b.c, sy, sz, a.g.n, vp);
        for (int comp = 0; comp < 3; ++comp) dv[comp] = a.D[comp][b.c];
      }
      double v[NSTEN], qv[3];
      for (int n = 0; n < NSTEN; ++n) v[n] = a.q == 0 ? vz[n] : vz[n] + beta * vp[n];
      apply3v<double>(a, v, dv, b.i, b.j, b.k, qv);
      const int cc[3] = {P0C, P1C, P2C};
      for (int comp = 0; comp < 3; ++comp) {
        if (!a.rg.in(comp, b.i, b.j, b.k)) continue;
        const double pn = v[cc[comp]];
        a.Pnew.c[comp][b.c] = pn;
        if (a.flexible && a.zcopy) a.zold.c[comp][b.c] = vz[cc[comp]];
        a.Q.c[comp][b.c] = qv[comp];
        s += pn * qv[comp];
      }
    }
  const double r0 = block_sum_t(s, sh, tid);
  if (tid == 0) {
    double *pr = region_w(a.d, PART_PQ);
    pr[2 * b0.bid] = r0;
    pr[2 * b0.bid + 1] = 0.0;
  }
}

// The same iteration as k_pcg_pap over an LDS tile: the block's 64 x 4 nodes of one k level and
// their halo (66 x 6) hold P = z + βPold of the three components for the levels k-1, k, k+1, so
// that z and Pold are read once per node (and halo node) instead of once per stencil use; the next
// level's plane is requested before the current level's operator is evaluated.  P at each node
// is the value every old-kernel thread computed for it, the operator and the block's partials
// take the same values in the same order: the results are the same bits.
constexpr int PT_W = 66, PT_H = 6, PT_N = PT_W * PT_H;  // the tile with its halo
// T = float: the refinement's inner iteration on a.f (its P, z, Q and D in Float32; β and the
// partials in FP64)
template <class T>
__global__ void __launch_bounds__(RED_T) k_pcg_pap_tile(PcgArgs a) {
  constexpr bool F32 = sizeof(T) == 4;
  if (!a.d.fl[FL_ACT0 + (a.q & 1)]) return;
  __shared__ double sh[8];
  __shared__ T Pl[3][3][PT_N];  // [component][level slot][row * 66 + col]
  const T *const *Zc = F32 ? (const T *const *)a.f.z.c : (const T *const *)a.z.c;
  const T *const *Pc = F32 ? (const T *const *)a.f.Pold.c : (const T *const *)a.Pold.c;
  T *const *Nc = F32 ? (T *const *)a.f.Pnew.c : (T *const *)a.Pnew.c;
  T *const *Qc = F32 ? (T *const *)a.f.Q.c : (T *const *)a.Q.c;
  const T *const *Dc = F32 ? (const T *const *)a.f.D : (const T *const *)a.D;
  const int tid = threadIdx.x;
  const BoxIdx b0 = box_idx(a.g, a.kc, a.xcd);
  const bool live = box_live(a.g, b0);
  const int i0 = b0.i - (tid & 63) - 1, j0 = b0.j - (tid >> 6) - 1;  // the tile's first (halo) node
  const bool pold = a.q > 0;
  // this thread's two tile points (p = tid, tid + 256 when below 396) as node indices of level 1
  int pidx[2];
  bool pok[2];
#pragma unroll
  for (int u = 0; u < 2; ++u) {
    const int p = tid + u * RED_T;
    pok[u] = p < PT_N;
    const int col = p % PT_W, row = p / PT_W;
    const int ii = min(max(i0 + col, 1), a.g.Nx + 1), jj = min(max(j0 + row, 1), a.g.Ny + 1);
    pidx[u] = a.g.at(ii, jj, 1);
  }
  const int sz = a.g.px * a.g.py;
  T rz_[2][3], rp_[2][3];  // a plane in flight: z and Pold of the two points, three components
  auto fetch = [&](int kk) {
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const int c = pidx[u] + sz * (kk - 1);
#pragma unroll
      for (int comp = 0; comp < 3; ++comp) {  // clamped addresses: every load in bounds
        rz_[u][comp] = Zc[comp][c];
        rp_[u][comp] = pold ? Pc[comp][c] : T(0);
      }
    }
  };
  auto stash = [&](int kk, T beta) {
    const int slot = kk % 3;
#pragma unroll
    for (int u = 0; u < 2; ++u)
      if (pok[u])
#pragma unroll
        for (int comp = 0; comp < 3; ++comp)
          Pl[comp][slot][tid + u * RED_T] = pold ? rz_[u][comp] + beta * rp_[u][comp] : rz_[u][comp];
  };
  // β from the partials of r·z and r·zold, summed as k_pcg_pap sums them, while the first plane
  // is in flight
  const int kf = b0.k;  // the block's first level
  fetch(kf - 1);
  const int np = a.nparts_in;
  const __amdgpu_buffer_rsrc_t rp = buf_rsrc(region(a.d, PART_RZ), 16 * np);
  double rz2 = 0.0, rzo = 0.0;
  for (int b1 = 0; b1 < np; b1 += 2 * RED_T)
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const int e = b1 + tid + u * RED_T;
      const double2 v = buf_load_d2(rp, 16 * min(e, np - 1));
      rz2 += e < np ? v.x : 0.0;
      rzo += e < np ? v.y : 0.0;
    }
  block_sum2_t(rz2, rzo, sh, tid);
  double beta = 0.0;
  if (pold) {
    const double rz = a.d.sc[SC_RZ0 + ((a.q + 1) & 1)];
    beta = a.flexible ? (rz2 - rzo) / rz : rz2 / rz;
  }
  if (b0.bid == 0 && tid == 0) a.d.sc[SC_RZ0 + (a.q & 1)] = rz2;
  const T betaT = (T)beta;
  stash(kf - 1, betaT);
  fetch(kf);
  stash(kf, betaT);
  fetch(kf + 1);
  stash(kf + 1, betaT);
  const int tx = (tid & 63) + 1, ty = (tid >> 6) + 1;
  auto L = [&](int comp, int kk, int row, int col) { return Pl[comp][kk % 3][row * PT_W + col]; };
  // D of the node: the palette in LDS and the level's code requested a level ahead, or the three
  // arrays a level ahead
  __shared__ T palL[3 * DPAL_MAX];
  const bool pal = a.Dcode != nullptr;
  if (pal && tid < 3 * DPAL_MAX) palL[tid] = F32 ? (T)a.Dpal32[tid] : (T)a.Dpal[tid];
  const int cnode = box_load_c(a.g, b0), slev = a.g.px * a.g.py;
  uint32_t code_n = pal ? a.Dcode[cnode] : 0u;
  T dn[3] = {T(0), T(0), T(0)};
  if (!pal)
#pragma unroll
    for (int comp = 0; comp < 3; ++comp) dn[comp] = Dc[comp][cnode];
  double s = 0.0;
  for (int q = 0; q < b0.nk; ++q) {
    const int k = kf + q;
    __syncthreads();
    const bool more = q + 1 < b0.nk;
    if (more) fetch(k + 2);
    const uint32_t code = code_n;
    const T dcur[3] = {dn[0], dn[1], dn[2]};
    if (more) {
      const int cn = cnode + slev * (q + 1);
      if (pal)
        code_n = a.Dcode[cn];
      else
#pragma unroll
        for (int comp = 0; comp < 3; ++comp) dn[comp] = Dc[comp][cn];
    }
    if (live) {
      const BoxIdx b = box_level(a.g, b0, q);
      T v[NSTEN], qv[3];
      v[P0C] = L(0, k, ty, tx), v[P0ZP] = L(0, k + 1, ty, tx), v[P0ZM] = L(0, k - 1, ty, tx);
      v[P0XMZP] = L(0, k + 1, ty, tx - 1), v[P0XM] = L(0, k, ty, tx - 1), v[P0YP] = L(0, k, ty + 1, tx);
      v[P0YM] = L(0, k, ty - 1, tx), v[P0XMYP] = L(0, k, ty + 1, tx - 1);
      v[P1C] = L(1, k, ty, tx), v[P1ZP] = L(1, k + 1, ty, tx), v[P1ZM] = L(1, k - 1, ty, tx);
      v[P1YMZP] = L(1, k + 1, ty - 1, tx), v[P1YM] = L(1, k, ty - 1, tx), v[P1XP] = L(1, k, ty, tx + 1);
      v[P1YMXP] = L(1, k, ty - 1, tx + 1), v[P1XM] = L(1, k, ty, tx - 1);
      v[P2C] = L(2, k, ty, tx), v[P2YP] = L(2, k, ty + 1, tx), v[P2ZMYP] = L(2, k - 1, ty + 1, tx);
      v[P2ZM] = L(2, k - 1, ty, tx), v[P2YM] = L(2, k, ty - 1, tx), v[P2XP] = L(2, k, ty, tx + 1);
      v[P2ZMXP] = L(2, k - 1, ty, tx + 1), v[P2XM] = L(2, k, ty, tx - 1);
      T dv[3];
#pragma unroll
      for (int comp = 0; comp < 3; ++comp)
        dv[comp] = pal ? palL[DPAL_MAX * comp + (int)((code >> (8 * comp)) & 255u)] : dcur[comp];
      apply3v<T>(a, v, dv, b.i, b.j, b.k, qv);
      const int cc[3] = {P0C, P1C, P2C};
      for (int comp = 0; comp < 3; ++comp) {
        if (!a.rg.in(comp, b.i, b.j, b.k)) continue;
        const T pn = v[cc[comp]];
        Nc[comp][b.c] = pn;
        if (!F32 && a.flexible && a.zcopy) a.zold.c[comp][b.c] = a.z.c[comp][b.c];
        Qc[comp][b.c] = qv[comp];
        s += (double)pn * (double)qv[comp];
      }
    }
    __syncthreads();  // level k-1 is read: its slot takes level k+2
    if (more) stash(k + 2, betaT);
  }
  const double r0 = block_sum_t(s, sh, tid);
  if (tid == 0) {
    double *pr = region_w(a.d, PART_PQ);
    pr[2 * b0.bid] = r0;
    pr[2 * b0.bid + 1] = 0.0;
  }
}

// iteration q: α = rz/pAp, x += αP, r −= αQ and the partials of r·r; IT += 1
__global__ void __launch_bounds__(RED_T) k_pcg_xr(PcgArgs a) {
  if (!a.d.fl[FL_ACT0 + (a.q & 1)]) return;
  __shared__ double sh[8];
  const int tid = threadIdx.x;
  const BoxIdx b0 = box_idx(a.g, a.kc, a.xcd);
  const bool live = box_live(a.g, b0);
  const int cl = box_load_c(a.g, b0);
  double xv[3], pv[3], rv[3], qv[3];
  for (int comp = 0; comp < 3; ++comp) {
    xv[comp] = a.x.c[comp][cl];
    pv[comp] = a.Pnew.c[comp][cl];
    rv[comp] = a.r.c[comp][cl];
    qv[comp] = a.Q.c[comp][cl];
  }
  const double pAp = sum_parts(region(a.d, PART_PQ), a.nparts_in, 0, sh, tid);
  const double al = a.d.sc[SC_RZ0 + (a.q & 1)] / pAp;
  if (b0.bid == 0 && tid == 0) a.d.fl[FL_IT] += 1;
  double s = 0.0;
  if (live)
    for (int q = 0; q < b0.nk; ++q) {
      const BoxIdx b = box_level(a.g, b0, q);
      // the next level's values are requested before this level's are used
      double xn[3], pn[3], rn[3], qn[3];
      const bool more = q + 1 < b0.nk;
      if (more) {
        const int cn = b.c + a.g.px * a.g.py;
        for (int comp = 0; comp < 3; ++comp) {
          xn[comp] = a.x.c[comp][cn];
          pn[comp] = a.Pnew.c[comp][cn];
          rn[comp] = a.r.c[comp][cn];
          qn[comp] = a.Q.c[comp][cn];
        }
      }
      for (int comp = 0; comp < 3; ++comp) {
        if (!a.rg.in(comp, b.i, b.j, b.k)) continue;
        const int c = b.c;
        a.x.c[comp][c] = xv[comp] + al * pv[comp];
        const double r1 = rv[comp] - al * qv[comp];
        a.r.c[comp][c] = r1;
        s += r1 * r1;
      }
      if (more)
        for (int comp = 0; comp < 3; ++comp) {
          xv[comp] = xn[comp];
          pv[comp] = pn[comp];
          rv[comp] = rn[comp];
          qv[comp] = qn[comp];
        }
    }
  const double r0 = block_sum_t(s, sh, tid);
  if (tid == 0) {
    double *pr = region_w(a.d, PART_RR);
    pr[2 * b0.bid] = r0;
    pr[2 * b0.bid + 1] = 0.0;
  }
}

// φ = θ + x on the ranges (implicit.jl:440-444) and the cycle's counts of the solve
// A refinement round's residual: r32 = b − A x with the FP64 operator (rounded to Float32 when
// stored) and the partials of ||b − A x||² in FP64 (only those with norm_only); gated by
// fl[gate_fl]
__global__ void __launch_bounds__(RED_T) k_pcg_resid32(PcgArgs a) {
  if (!a.d.fl[a.gate_fl]) return;
  __shared__ double sh[8];
  const int tid = threadIdx.x;
  const BoxIdx b0 = box_idx(a.g, a.kc, a.xcd);
  double s = 0.0;
  if (box_live(a.g, b0))
    for (int q = 0; q < b0.nk; ++q) {
      const BoxIdx b = box_level(a.g, b0, q);
      double qv[3], pc[3], bv[3];
      for (int comp = 0; comp < 3; ++comp) bv[comp] = a.b.c[comp][b.c];
      apply3(a, a.x.c[0], a.x.c[1], a.x.c[2], b.i, b.j, b.k, b.c, qv, pc);
      for (int comp = 0; comp < 3; ++comp) {
        if (!a.rg.in(comp, b.i, b.j, b.k)) continue;
        const double rv = bv[comp] - qv[comp];
        if (!a.norm_only) a.f.r.c[comp][b.c] = (float)rv;
        s += rv * rv;
      }
    }
  const double r0 = block_sum_t(s, sh, tid);
  if (tid == 0) {
    double *pr = region_w(a.d, PART_RR);
    pr[2 * b0.bid] = r0;
    pr[2 * b0.bid + 1] = 0.0;
  }
}
// the end of a refinement round: x += d (FP64 += Float32), gated by fl[gate_fl]
__global__ void __launch_bounds__(RED_T) k_pcg_xupdate(PcgArgs a) {
  if (!a.d.fl[a.gate_fl]) return;
  const BoxIdx b0 = box_idx(a.g, a.kc, a.xcd);
  if (!box_live(a.g, b0)) return;
  for (int q = 0; q < b0.nk; ++q) {
    const BoxIdx b = box_level(a.g, b0, q);
    for (int comp = 0; comp < 3; ++comp)
      if (a.rg.in(comp, b.i, b.j, b.k)) a.x.c[comp][b.c] += (double)a.f.d.c[comp][b.c];
  }
}
__global__ void __launch_bounds__(RED_T) k_pcg_finish(PcgArgs a) {
  const BoxIdx b0 = box_idx(a.g, a.kc, a.xcd);
  if (box_live(a.g, b0))
    for (int q = 0; q < b0.nk; ++q) {
      const BoxIdx b = box_level(a.g, b0, q);
      for (int comp = 0; comp < 3; ++comp)
        if (a.rg.in(comp, b.i, b.j, b.k)) a.ph[comp][b.c] = a.th[comp][b.c] + a.x.c[comp][b.c];
    }
  if (b0.bid == 0 && threadIdx.x == 0) {
    int *fl = a.d.fl;
    if (fl[FL_NZ] && fl[a.ovf_fl ? a.ovf_fl : FL_ACT0 + (a.act_slot & 1)]) fl[FL_OVF] = 1;
    fl[FL_ITSUM] += fl[FL_IT];
    if (fl[FL_IT] > fl[FL_ITMAX]) fl[FL_ITMAX] = fl[FL_IT];
    const int c = a.cls >= 0 && a.cls < PCG_NCLS ? a.cls : 0;
    if (fl[FL_IT] > fl[FL_ITMAXC + c]) fl[FL_ITMAXC + c] = fl[FL_IT];
  }
}

static hipError_t launch_box(void (*k)(PcgArgs), const PcgArgs &a, hipStream_t s) {
  if (g_rec) return hipErrorNotSupported;
  if (a.kc < 1 || a.kc > a.g.Nz - 1) return hipErrorInvalidValue;
  const int nb = pcg_box_blocks(a.g, a.kc);
  if (nb > MAX_PART) return hipErrorInvalidValue;
  TDGL_LAUNCH(k, dim3(nb), dim3(RED_T), 0, s, a);
  return hipGetLastError();
}
hipError_t launch_pcg_start(const PcgArgs &a, hipStream_t s) { return launch_box(k_pcg_start, a, s); }
hipError_t launch_pcg_resid(const PcgArgs &a, hipStream_t s) { return launch_box(k_pcg_resid, a, s); }
hipError_t launch_pcg_p(const PcgArgs &a, hipStream_t s) { return launch_box(k_pcg_p, a, s); }
hipError_t launch_pcg_ap(const PcgArgs &a, hipStream_t s) { return launch_box(k_pcg_ap, a, s); }
hipError_t launch_pcg_pap(const PcgArgs &a, hipStream_t s) {
  return launch_box(a.tile ? k_pcg_pap_tile<double> : k_pcg_pap, a, s);
}
hipError_t launch_pcg32_pap(const PcgArgs &a, hipStream_t s) { return launch_box(k_pcg_pap_tile<float>, a, s); }
hipError_t launch_pcg_xr(const PcgArgs &a, hipStream_t s) { return launch_box(k_pcg_xr, a, s); }
hipError_t launch_pcg_finish(const PcgArgs &a, hipStream_t s) { return launch_box(k_pcg_finish, a, s); }
hipError_t launch_pcg_resid32(const PcgArgs &a, hipStream_t s) { return launch_box(k_pcg_resid32, a, s); }
hipError_t launch_pcg_xupdate(const PcgArgs &a, hipStream_t s) { return launch_box(k_pcg_xupdate, a, s); }

__global__ void k_pcg_cycle_reset(PcgDev d) {
  if (threadIdx.x == 0) {
    d.fl[FL_OVF] = 0;
    d.fl[FL_ITSUM] = 0;
    d.fl[FL_ITMAX] = 0;
    for (int c = 0; c < PCG_NCLS; ++c) d.fl[FL_ITMAXC + c] = 0;
  }
}

hipError_t launch_pcg_cycle_reset(PcgDev d, hipStream_t s) {
  if (g_rec) return hipErrorNotSupported;
  TDGL_LAUNCH(k_pcg_cycle_reset, dim3(1), dim3(64), 0, s, d);
  return hipGetLastError();
}

// ============================================================================ batched launches
size_t batch_args_size(int kind) {
  switch (kind) {
    case BK_STAGE: return sizeof(StageArgs);
    case BK_FIELD: return sizeof(FieldArgs);
    case BK_TX: return sizeof(TxPass);
    case BK_TX2: return sizeof(Tx2Modes);
    case BK_CURL: return sizeof(CurlArgs);
    case BK_GHOSTS: return sizeof(GhostArgs);
    case BK_METER: return sizeof(MeterArgs);
    case BK_METERFIN: return sizeof(MeterFinArgs);
    case BK_PLANE: return sizeof(PlanePass);
    case BK_ZMODES: return sizeof(ZModes);
    case BK_FOLD: return sizeof(FoldPass);
    case BK_UNFOLD: return sizeof(UnfoldPass);
    case BK_ZMODESQ: return sizeof(ZModesQ64);
    case BK_QPLANE: return sizeof(QPass64);
    default: return 0;
  }
}
// the grid of a record's single-run launch (as its launcher computes it)
dim3 batch_grid(int kind, const Geo &g, const void *args) {
  switch (kind) {
    case BK_STAGE: {
      const int ni = g.x2 - g.x1, nj = g.y2 - g.y1, nk = g.z2 - g.z1;
      if (ni <= 0 || nj <= 0 || nk <= 0) return dim3(0, 0, 0);
      return dim3(cdiv(ni, 64), cdiv(nj, 4), nk);
    }
    case BK_FIELD:
    case BK_CURL: return dim3(cdiv(g.Nx + 1, 64), cdiv(g.Ny + 1, 4), g.Nz + 1);
    case BK_GHOSTS: {
      const int ni = g.x2 - g.x1 + 2, nj = g.y2 - g.y1 + 2, nk = g.z2 - g.z1 + 2;
      return dim3(cdiv(2 * (nj * nk + ni * nk + ni * nj), 256));
    }
    case BK_TX: return tx_grid(*(const TxPass *)args);
    case BK_TX2: {
      const Tx2Modes &m = *(const Tx2Modes *)args;
      return dim3(cdiv(m.nx * m.ny * m.nz, 256));
    }
    default: return meter_batch_grid(kind, args);
  }
}
hipError_t launch_batch(int kind, int variant, size_t lds, const BatchTab &t, int blocks, hipStream_t s) {
  if (blocks <= 0) return hipSuccess;
  switch (kind) {
    case BK_STAGE: TDGL_LAUNCH(k_stage_b, dim3(blocks), dim3(64, 4, 1), 0, s, t); break;
    case BK_FIELD: TDGL_LAUNCH(k_field_b, dim3(blocks), dim3(64, 4, 1), 0, s, t); break;
    case BK_CURL: TDGL_LAUNCH(k_curl_b, dim3(blocks), dim3(64, 4, 1), 0, s, t); break;
    case BK_GHOSTS: TDGL_LAUNCH(k_ghosts_b, dim3(blocks), dim3(256), 0, s, t); break;
    case BK_TX: TDGL_LAUNCH(k_tx_b, dim3(blocks), dim3(256), 0, s, t); break;
    case BK_TX2: TDGL_LAUNCH(k_tx2_modes_b, dim3(blocks), dim3(256), 0, s, t); break;
    case BK_ZMODES:
      if (variant)
        TDGL_LAUNCH(k_zmodes_b<true>, dim3(blocks), dim3(256), lds, s, t);
      else
        TDGL_LAUNCH(k_zmodes_b<false>, dim3(blocks), dim3(256), lds, s, t);
      break;
    case BK_QPLANE: {
      const int jt = variant / 100, waves = variant % 100;
      if (waves < 4 || waves > QP64_MT) return hipErrorInvalidValue;
      switch (jt) {
        case 1: TDGL_LAUNCH((k_qplane64_b<1>), dim3(blocks), dim3(64 * waves), lds, s, t); break;
        case 2: TDGL_LAUNCH((k_qplane64_b<2>), dim3(blocks), dim3(64 * waves), lds, s, t); break;
        case 3: TDGL_LAUNCH((k_qplane64_b<3>), dim3(blocks), dim3(64 * waves), lds, s, t); break;
        case 4: TDGL_LAUNCH((k_qplane64_b<4>), dim3(blocks), dim3(64 * waves), lds, s, t); break;
        case 5: TDGL_LAUNCH((k_qplane64_b<5>), dim3(blocks), dim3(64 * waves), lds, s, t); break;
        case 6: TDGL_LAUNCH((k_qplane64_b<6>), dim3(blocks), dim3(64 * waves), lds, s, t); break;
        default: return hipErrorInvalidValue;
      }
      break;
    }
    case BK_FOLD: TDGL_LAUNCH(k_fold_b, dim3(blocks), dim3(256), 0, s, t); break;
    case BK_UNFOLD: TDGL_LAUNCH(k_unfold_b, dim3(blocks), dim3(256), 0, s, t); break;
    case BK_ZMODESQ:
      switch (variant) {
#define TDGL_ZQB_CASE(g, mz) \
  case 100 * g + mz: TDGL_LAUNCH((k_zmodes_q64_b<g, mz>), dim3(blocks), dim3(64 * g), lds, s, t); break;
        TDGL_ZQB_CASE(8, 2) TDGL_ZQB_CASE(8, 4) TDGL_ZQB_CASE(8, 6) TDGL_ZQB_CASE(8, 8) TDGL_ZQB_CASE(8, 12)
        TDGL_ZQB_CASE(8, 16) TDGL_ZQB_CASE(4, 2) TDGL_ZQB_CASE(4, 4) TDGL_ZQB_CASE(4, 6) TDGL_ZQB_CASE(4, 8)
        TDGL_ZQB_CASE(4, 12) TDGL_ZQB_CASE(4, 16)
#undef TDGL_ZQB_CASE
        default: return hipErrorInvalidValue;
      }
      break;
    case BK_PLANE:
      if (variant >= 1000) {
        switch (variant - 1000) {
#define TDGL_PLANE1B_CASE(x, tt, m)                                                                 \
  case (x * 4 + tt) * 4 + m:                                                                        \
    TDGL_LAUNCH((k_plane_mf1_b<x, tt, m>), dim3(blocks), dim3(PLANE1_T), lds, s, t);               \
    break;
          TDGL_PLANE1B_CASE(2, 2, 2) TDGL_PLANE1B_CASE(2, 3, 2) TDGL_PLANE1B_CASE(2, 2, 3) TDGL_PLANE1B_CASE(2, 3, 3)
          TDGL_PLANE1B_CASE(3, 2, 2) TDGL_PLANE1B_CASE(3, 3, 2) TDGL_PLANE1B_CASE(3, 2, 3) TDGL_PLANE1B_CASE(3, 3, 3)
          TDGL_PLANE1B_CASE(4, 2, 2) TDGL_PLANE1B_CASE(4, 3, 2) TDGL_PLANE1B_CASE(4, 2, 3) TDGL_PLANE1B_CASE(4, 3, 3)
          TDGL_PLANE1B_CASE(5, 2, 2) TDGL_PLANE1B_CASE(5, 3, 2) TDGL_PLANE1B_CASE(5, 2, 3) TDGL_PLANE1B_CASE(5, 3, 3)
          TDGL_PLANE1B_CASE(5, 3, 1)
          TDGL_PLANE1_SMALL(TDGL_PLANE1B_CASE)
#undef TDGL_PLANE1B_CASE
          default: return hipErrorInvalidValue;
        }
      } else {
        switch (variant) {
#define TDGL_PLANEB_CASE(tj) \
  case tj: TDGL_LAUNCH(k_plane_b<tj>, dim3(blocks), dim3(256), lds, s, t); break;
          TDGL_PLANEB_CASE(1) TDGL_PLANEB_CASE(2) TDGL_PLANEB_CASE(3) TDGL_PLANEB_CASE(4)
          TDGL_PLANEB_CASE(5) TDGL_PLANEB_CASE(6) TDGL_PLANEB_CASE(7) TDGL_PLANEB_CASE(8)
#undef TDGL_PLANEB_CASE
          default: return hipErrorInvalidValue;
        }
      }
      break;
    default: return launch_batch_meter(kind, t, blocks, s);
  }
  return hipGetLastError();
}

}  // namespace tdgl

