// kernels.hip — the HIP kernels of one release-2.0 time step on gfx950.
//
// Every kernel keeps the reference's order of operations per node (gtdgl/src/bands.jl,
// field.jl, implicit.jl); the file is compiled with -ffp-contract=off, so products and sums
// round as the CPU step's do.  Transcendentals (sincos, atan2, sqrt) come from the device math
// library.  Node-shaped layout and ranges: common.h.
#include "kernels.h"

namespace tdgl {

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

// One (sub)stage of band µ on the interior nodes, fused: the band pass (_band_row!,
// bands.jl:71-97) with KWT (:135-140), the rotating-frame rate or term (_frame_rate!
// :213-234 / _frame_term! :274-286), the RKL/RKC combination (_combine! :425-430) and the
// turn back of the frame (_frame_turn! :321-327), in that order, as the reference applies them.
// Stage ghosts: from the second (sub)stage on, the face values of S are the covariant Neumann
// ghosts of S with this stage's links (ghosts!, :524-539), which for the nodes a stencil reads are
// the link times the stencil's own centre: computed here instead of stored.
__global__ void __launch_bounds__(256) k_stage(Geo g, StageArgs a) {
  const int i = g.x1 + 1 + blockIdx.x * blockDim.x + threadIdx.x;
  const int j = g.y1 + 1 + blockIdx.y * blockDim.y + threadIdx.y;
  const int k = g.z1 + 1 + blockIdx.z;
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
    p = a.gamma2 > 0 ? kwt_update(gc, csub(acc, gc), a.gamma2) : acc;
  }
  double om = 0.0;
  if (a.frame == FRAME_RATE) {
    const double a2 = abs2(gc);
    om = a2 > 1e-3 ? jclamp(cmul(cconj(gc), csub(p, gc)).y / (a2 * a.h), -a.wmax, a.wmax) : 0.0;
    a.Om[c] = om;
    p = csub(p, imx_mul(om * a.h, gc));
  } else if (a.frame == FRAME_TERM) {
    om = a.Om[c];
    p = csub(p, imx_mul(om * a.h, gc));
  } else if (a.turn) {
    om = a.Om[c];
  }
  if (a.combine) p = cadd(rmul(a.cmu, p), rmul(a.cnu, a.A[c]));
  if (a.turn) p = cmul(p, dcis(om * a.dt));
  a.D[c] = p;
}

hipError_t launch_stage(const Geo &g, const StageArgs &a, hipStream_t s) {
  const int ni = g.x2 - g.x1, nj = g.y2 - g.y1, nk = g.z2 - g.z1;
  if (ni <= 0 || nj <= 0 || nk <= 0) return hipSuccess;
  dim3 b(64, 4, 1), gr(cdiv(ni, 64), cdiv(nj, 4), nk);
  hipLaunchKernelGGL(k_stage, gr, b, 0, s, g, a);
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

__global__ void __launch_bounds__(256) k_field(Geo g, FieldArgs a) {
  __shared__ double sh[256];
  const int tid = threadIdx.x + blockDim.x * threadIdx.y;
  const int i = 1 + blockIdx.x * blockDim.x + threadIdx.x;
  const int j = 1 + blockIdx.y * blockDim.y + threadIdx.y;
  const int k = 1 + blockIdx.z;
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
      const int bid = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
      a.part[2 * bid] = r;
      a.part[2 * bid + 1] = 0.0;
    }
  }
}

int field_blocks(const Geo &g) { return cdiv(g.Nx + 1, 64) * cdiv(g.Ny + 1, 4) * (g.Nz + 1); }

hipError_t launch_field(const Geo &g, const FieldArgs &a, hipStream_t s) {
  if (a.part && field_blocks(g) > MAX_PART) return hipErrorInvalidValue;
  dim3 b(64, 4, 1), gr(cdiv(g.Nx + 1, 64), cdiv(g.Ny + 1, 4), g.Nz + 1);
  hipLaunchKernelGGL(k_field, gr, b, 0, s, g, a);
  return hipGetLastError();
}

// Links U = e^{−iφ} (links!, field.jl:105-121), field h = curl φ (curl_h!, :124-137) and the
// field boundary (field_bc!, :189-228: applied H, slab or bar self-field; the later assignment
// of the reference wins where two cover one element).
__global__ void __launch_bounds__(256) k_curl(Geo g, CurlArgs a) {
  const int i = 1 + blockIdx.x * blockDim.x + threadIdx.x;
  const int j = 1 + blockIdx.y * blockDim.y + threadIdx.y;
  const int k = 1 + blockIdx.z;
  if (i > g.Nx + 1 || j > g.Ny + 1 || k > g.Nz + 1) return;
  const int Nx = g.Nx, Ny = g.Ny, Nz = g.Nz;
  const int c = g.at(i, j, k);
  const int sy = g.px, sz = g.px * g.py;
  const double *phx = a.ph[0], *phy = a.ph[1], *phz = a.ph[2];
  const double H = a.hp[0], HJ = a.hp[1];
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
    if (set) a.h[2][c] = v;
  }
}

hipError_t launch_curl(const Geo &g, const CurlArgs &a, hipStream_t s) {
  dim3 b(64, 4, 1), gr(cdiv(g.Nx + 1, 64), cdiv(g.Ny + 1, 4), g.Nz + 1);
  hipLaunchKernelGGL(k_curl, gr, b, 0, s, g, a);
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
__global__ void __launch_bounds__(256) k_ghosts(Geo g, GhostArgs a) {
  const int ni = g.x2 - g.x1 + 2, nj = g.y2 - g.y1 + 2, nk = g.z2 - g.z1 + 2;
  int t = blockIdx.x * blockDim.x + threadIdx.x;
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

hipError_t launch_ghosts(const Geo &g, const GhostArgs &a, hipStream_t s) {
  const int ni = g.x2 - g.x1 + 2, nj = g.y2 - g.y1 + 2, nk = g.z2 - g.z1 + 2;
  const int n = 2 * (nj * nk + ni * nk + ni * nj);
  hipLaunchKernelGGL(k_ghosts, dim3(cdiv(n, 256)), dim3(256), 0, s, g, a);
  return hipGetLastError();
}

__global__ void __launch_bounds__(256) k_copies(CopyList c) {
  const int a = blockIdx.y;
  if (a >= c.cnt) return;
  const double *src = c.src[a];
  double *dst = c.dst[a];
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < c.n[a]; t += gridDim.x * blockDim.x) dst[t] = src[t];
}

hipError_t launch_copies(const CopyList &c, hipStream_t s) {
  if (c.cnt <= 0) return hipSuccess;
  int m = 0;
  for (int a = 0; a < c.cnt; ++a) m = c.n[a] > m ? c.n[a] : m;
  const int bx = cdiv(m, 256) < 1024 ? cdiv(m, 256) : 1024;
  hipLaunchKernelGGL(k_copies, dim3(bx, c.cnt), dim3(256), 0, s, c);
  return hipGetLastError();
}

// ============================================================================ reductions
// Deterministic sums: a fixed assignment of elements to threads, a fixed LDS tree per block,
// block partials in block order.  A consumer sums a producer's partials itself (every block the
// same value) after the producer's kernel has ended.
constexpr int RED_T = 256;

__device__ __forceinline__ double block_sum_t(double v, double *sh, int tid) {
  sh[tid] = v;
  __syncthreads();
  for (int w = RED_T / 2; w > 0; w >>= 1) {
    if (tid < w) sh[tid] += sh[tid + w];
    __syncthreads();
  }
  const double r = sh[0];
  __syncthreads();
  return r;
}

// Σ over the partials [slot][which] of a region, identical in every block
__device__ __forceinline__ double sum_parts(const double *part, int n, int which, double *sh, int tid) {
  double s = 0.0;
  for (int b = tid; b < n; b += RED_T) s += part[2 * b + which];
  return block_sum_t(s, sh, tid);
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
__global__ void __launch_bounds__(256) k_tx(TxPass p, const int *gate) {
  constexpr int TXK = TxDepth<TA>::K;
  __shared__ TA As[TXK][TXM + 1];
  __shared__ TA Is[TXK][TXO + 1];
  __shared__ double sh[RED_T];
  const int tid = threadIdx.x;
  const int bid = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
  if (p.conv.mode != CONV_NONE) {
    // the convergence check that decides whether this solve goes on (implicit.jl:305, 319, 426)
    if (!*p.conv.in) {
      if (bid == 0 && tid == 0) *p.conv.out = 0;
      return;
    }
    const double rr = sum_parts(region(p.d, PART_RR), p.conv.nparts, 0, sh, tid);
    const double tol = p.d.sc[SC_TOL], nb = p.d.sc[SC_NB];
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
  const int comp = blockIdx.z;
  const int e0 = p.e[comp][0], e1 = p.e[comp][1], e2 = p.e[comp][2];
  const int axis = p.axis;
  const int L = axis == 0 ? e0 : (axis == 1 ? e1 : e2);
  const int O = (e0 * e1 * e2) / L;
  const int m0 = blockIdx.y * TXM, o0 = blockIdx.x * TXO;
  double d0 = 0.0, d1 = 0.0;
  if (m0 < L && o0 < O) {
    const TI *in = (const TI *)p.in[comp].p;
    const TA *A = (const TA *)p.A[comp];
    const int tm = (tid / 16) * 2, to = (tid % 16) * 2;
    TA acc[2][2] = {{0, 0}, {0, 0}};
    for (int n0 = 0; n0 < L; n0 += TXK) {
      for (int t = tid; t < TXK * TXM; t += 256) {
        const int kk = t / TXM, mm = t % TXM;
        const int n = n0 + kk, m = m0 + mm;
        As[kk][mm] = (n < L && m < L) ? A[n * L + m] : TA(0);
      }
      for (int t = tid; t < TXK * TXO; t += 256) {
        int kk, oo;
        if (axis == 0) {
          kk = t % TXK;
          oo = t / TXK;
        } else {
          kk = t / TXO;
          oo = t % TXO;
        }
        const int n = n0 + kk, o = o0 + oo;
        TA v = TA(0);
        if (n < L && o < O) {
          int i0, i1, i2;
          tx_coords(axis, e0, e1, n, o, i0, i1, i2);
          v = (TA)in[voff(p.in[comp], i0, i1, i2)];
        }
        Is[kk][oo] = v;
      }
      __syncthreads();
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
    for (int dm = 0; dm < 2; ++dm)
      for (int dout = 0; dout < 2; ++dout) {
        const int m = m0 + tm + dm, o = o0 + to + dout;
        if (m >= L || o >= O) continue;
        int i0, i1, i2;
        tx_coords(axis, e0, e1, m, o, i0, i1, i2);
        const int q = voff(p.out[comp], i0, i1, i2);
        if (p.addth) {
          const double v = ((const double *)p.th[comp].p)[voff(p.th[comp], i0, i1, i2)] + (double)acc[dm][dout];
          ((double *)p.out[comp].p)[q] = v;
        } else {
          const TO v = (TO)acc[dm][dout];
          ((TO *)p.out[comp].p)[q] = v;
          if (p.dot) {
            const double rv = ((const double *)p.r[comp].p)[voff(p.r[comp], i0, i1, i2)];
            d0 += rv * (double)v;
            if (p.dot == 2) d1 += rv * ((const double *)p.zold[comp].p)[voff(p.zold[comp], i0, i1, i2)];
          }
        }
      }
  }
  if (p.dot) {
    const double r0 = block_sum_t(d0, sh, tid);
    const double r1 = block_sum_t(d1, sh, tid);
    if (tid == 0) {
      double *pr = region_w(p.d, PART_RZ);
      pr[2 * bid] = r0;
      pr[2 * bid + 1] = r1;
    }
  }
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
  hipLaunchKernelGGL((k_tx<TI, TA, TO>), tx_grid(p), dim3(256), 0, s, p, gate);
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
__global__ void __launch_bounds__(256) k_tx2_modes(Tx2Modes m, const int *gate) {
  if (gate && !*gate) return;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int nx = m.nx, ny = m.ny, nz = m.nz;
  if (t >= nx * ny * nz) return;
  const int mx = t % nx, my = (t / nx) % ny, mz = t / (nx * ny);
  const T *in = (const T *)m.in;
  T *out = (T *)m.out;
  const T *AV = (const T *)m.Az[0], *AU = (const T *)m.Az[1];
  const T k1 = ((const T *)m.sx)[mx], k2 = ((const T *)m.sy)[my], k3 = ((const T *)m.sz)[mz];
  const bool hx = mx > 0, hy = my > 0, hz = mz > 0;
  // X: extents (nx-1, ny, nz); Y: (nx, ny-1, nz); Z: (nx, ny, nz-1)
  T x1 = 0, y1 = 0, z1 = 0;
  if (hx) {
    const T *col = in + m.off[0] + (mx - 1) + (nx - 1) * my;
    const int st = (nx - 1) * ny;
    T acc = 0;
#pragma unroll 8
    for (int n = 0; n < nz; ++n) acc += col[st * n] * AV[n * nz + mz];
    x1 = acc;
  }
  if (hy) {
    const T *col = in + m.off[1] + mx + nx * (my - 1);
    const int st = nx * (ny - 1);
    T acc = 0;
#pragma unroll 8
    for (int n = 0; n < nz; ++n) acc += col[st * n] * AV[n * nz + mz];
    y1 = acc;
  }
  if (hz) {
    const T *col = in + m.off[2] + mx + nx * my;
    const int st = nx * ny;
    T acc = 0;
#pragma unroll 8
    for (int n = 0; n < nz - 1; ++n) acc += col[st * n] * AU[n * (nz - 1) + (mz - 1)];
    z1 = acc;
  }
  const T k2s = k1 * k1 + k2 * k2 + k3 * k3;
  if (k2s != 0) project<T>(k1, k2, k3, (T)m.a, x1, y1, z1, hx, hy, hz);
  if (hx) out[m.off[0] + (mx - 1) + (nx - 1) * (my + ny * mz)] = x1;
  if (hy) out[m.off[1] + mx + nx * ((my - 1) + (ny - 1) * mz)] = y1;
  if (hz) out[m.off[2] + mx + nx * (my + ny * (mz - 1))] = z1;
}

template <class T>
hipError_t launch_tx2_modes(const Tx2Modes &m, const int *gate, hipStream_t s) {
  hipLaunchKernelGGL((k_tx2_modes<T>), dim3(cdiv(m.nx * m.ny * m.nz, 256)), dim3(256), 0, s, m, gate);
  return hipGetLastError();
}
template hipError_t launch_tx2_modes<double>(const Tx2Modes &, const int *, hipStream_t);
template hipError_t launch_tx2_modes<float>(const Tx2Modes &, const int *, hipStream_t);

// ============================================================================ PCG
// Kernels over the box i, j, k in 2..N (the union of the three ranges) with 1D blocks of 64 x 4
// nodes; vectors node-shaped, zero outside their ranges (the direction P is read at neighbours
// with no test: implicit.jl:300, 406).
struct BoxIdx {
  int i, j, k, c, bid;
};
__device__ __forceinline__ BoxIdx box_idx(const Geo &g) {
  const int gx = cdiv_d(g.Nx - 1, 64), gy = cdiv_d(g.Ny - 1, 4);
  const int b = blockIdx.x;
  BoxIdx r;
  r.bid = b;
  const int bx = b % gx, by = (b / gx) % gy, bz = b / (gx * gy);
  r.i = 2 + bx * 64 + (int)(threadIdx.x & 63);
  r.j = 2 + by * 4 + (int)(threadIdx.x >> 6);
  r.k = 2 + bz;
  r.c = g.at(r.i, r.j, r.k);
  return r;
}
int pcg_box_blocks(const Geo &g) { return cdiv(g.Nx - 1, 64) * cdiv(g.Ny - 1, 4) * (g.Nz - 1); }

// h of the direction on the ranges of _apply_A! (implicit.jl:216-224), zero elsewhere; P is the
// direction of this kernel: z + β Pold (or a vector itself when Pold is null)
struct PDir {
  const double *z[3], *Po[3];
  double beta;
  int first;
  __device__ __forceinline__ double at(int comp, int c) const {
    if (!Po[0]) return z[comp][c];
    return first ? z[comp][c] : z[comp][c] + beta * Po[comp][c];
  }
};
__device__ __forceinline__ double hx_of(const Geo &g, const PDir &P, int i, int j, int k, int c, double dydz) {
  if (!(i >= 2 && i <= g.Nx && j >= 2 && j <= g.Ny - 1 && k >= 2 && k <= g.Nz - 1)) return 0.0;
  const int sy = g.px, sz = g.px * g.py;
  return ((P.at(2, c + sy) - P.at(2, c)) - P.at(1, c + sz) + P.at(1, c)) / dydz;
}
__device__ __forceinline__ double hy_of(const Geo &g, const PDir &P, int i, int j, int k, int c, double dzdx) {
  if (!(i >= 2 && i <= g.Nx - 1 && j >= 2 && j <= g.Ny && k >= 2 && k <= g.Nz - 1)) return 0.0;
  const int sz = g.px * g.py;
  return ((P.at(0, c + sz) - P.at(0, c)) - P.at(2, c + 1) + P.at(2, c)) / dzdx;
}
__device__ __forceinline__ double hz_of(const Geo &g, const PDir &P, int i, int j, int k, int c, double dxdy) {
  if (!(i >= 2 && i <= g.Nx - 1 && j >= 2 && j <= g.Ny - 1 && k >= 2 && k <= g.Nz)) return 0.0;
  const int sy = g.px;
  return ((P.at(1, c + 1) - P.at(1, c)) - P.at(0, c + sy) + P.at(0, c)) / dxdy;
}

// q = (D + aK) p of one component at (i, j, k) (_apply_A!, implicit.jl:226-234)
__device__ __forceinline__ double apply_one(const PcgArgs &a, const PDir &P, int comp, int i, int j, int k, int c,
                                            double p) {
  const Geo &g = a.g;
  const int sy = g.px, sz = g.px * g.py;
  if (comp == 0)
    return a.D[0][c] * p + a.sxy * (hz_of(g, P, i, j, k, c, a.dxdy) - hz_of(g, P, i, j - 1, k, c - sy, a.dxdy)) -
           a.sxz * (hy_of(g, P, i, j, k, c, a.dzdx) - hy_of(g, P, i, j, k - 1, c - sz, a.dzdx));
  if (comp == 1)
    return a.D[1][c] * p + a.syz * (hx_of(g, P, i, j, k, c, a.dydz) - hx_of(g, P, i, j, k - 1, c - sz, a.dydz)) -
           a.syx * (hz_of(g, P, i, j, k, c, a.dxdy) - hz_of(g, P, i - 1, j, k, c - 1, a.dxdy));
  return a.D[2][c] * p + a.szx * (hy_of(g, P, i, j, k, c, a.dzdx) - hy_of(g, P, i - 1, j, k, c - 1, a.dzdx)) -
         a.szy * (hx_of(g, P, i, j, k, c, a.dydz) - hx_of(g, P, i, j - 1, k, c - sy, a.dydz));
}

// _pcg_zero_update! (implicit.jl:251-271) when the source is zero, else the warm start of
// _implicit_curl_fast! (:391-402); nb from the field kernel's partials
__global__ void __launch_bounds__(RED_T) k_pcg_start(PcgArgs a) {
  __shared__ double sh[RED_T];
  const int tid = threadIdx.x;
  const double nb = sqrt(sum_parts(region(a.d, PART_NB), a.nparts_in, 0, sh, tid));
  const BoxIdx b = box_idx(a.g);
  if (b.bid == 0 && tid == 0) {
    a.d.sc[SC_NB] = nb;
    a.d.fl[FL_NZ] = nb > 0 ? 1 : 0;
    a.d.fl[FL_IT] = 0;
  }
  for (int comp = 0; comp < 3; ++comp) {
    if (!a.rg.in(comp, b.i, b.j, b.k)) continue;
    double *x = a.x.c[comp], *xp = a.xprev.c[comp];
    const int c = b.c;
    if (nb == 0) {
      if (a.warm && a.extrap) xp[c] = x[c];
      x[c] = 0.0;
    } else if (a.warm && a.extrap) {
      if (a.have2) {
        const double v = x[c];
        x[c] = 2 * v - xp[c];
        xp[c] = v;
      } else {
        xp[c] = x[c];
      }
    }
  }
}

// the start of the iterations: Q = A x, r = b − Q and the partials of r·r (implicit.jl:407-410)
__global__ void __launch_bounds__(RED_T) k_pcg_resid(PcgArgs a) {
  if (!a.d.fl[FL_NZ]) return;
  __shared__ double sh[RED_T];
  const int tid = threadIdx.x;
  const BoxIdx b = box_idx(a.g);
  PDir P{};
  for (int comp = 0; comp < 3; ++comp) P.z[comp] = a.x.c[comp];
  double s = 0.0;
  for (int comp = 0; comp < 3; ++comp) {
    if (!a.rg.in(comp, b.i, b.j, b.k)) continue;
    const double Aq = apply_one(a, P, comp, b.i, b.j, b.k, b.c, a.x.c[comp][b.c]);
    a.Q.c[comp][b.c] = Aq;
    const double rv = a.b.c[comp][b.c] - Aq;
    a.r.c[comp][b.c] = rv;
    s += rv * rv;
  }
  const double r0 = block_sum_t(s, sh, tid);
  if (tid == 0) {
    double *pr = region_w(a.d, PART_RR);
    pr[2 * b.bid] = r0;
    pr[2 * b.bid + 1] = 0.0;
  }
}

// iteration q: rz (and β) from the partials of the preconditioner's last pass, P = z + βP
// (P = z on the first) and zold = z (implicit.jl:412-413, 428-436)
__global__ void __launch_bounds__(RED_T) k_pcg_p(PcgArgs a) {
  if (!a.d.fl[FL_ACT0 + (a.q & 1)]) return;
  __shared__ double sh[RED_T];
  const int tid = threadIdx.x;
  const BoxIdx b = box_idx(a.g);
  const double rz2 = sum_parts(region(a.d, PART_RZ), a.nparts_in, 0, sh, tid);
  double beta = 0.0;
  if (a.q > 0) {
    const double rzo = a.flexible ? sum_parts(region(a.d, PART_RZ), a.nparts_in, 1, sh, tid) : 0.0;
    const double rz = a.d.sc[SC_RZ0 + ((a.q + 1) & 1)];
    beta = a.flexible ? (rz2 - rzo) / rz : rz2 / rz;
  }
  if (b.bid == 0 && tid == 0) a.d.sc[SC_RZ0 + (a.q & 1)] = rz2;
  for (int comp = 0; comp < 3; ++comp) {
    if (!a.rg.in(comp, b.i, b.j, b.k)) continue;
    const int c = b.c;
    const double zv = a.z.c[comp][c];
    a.Pnew.c[comp][c] = a.q == 0 ? zv : zv + beta * a.Pold.c[comp][c];
    if (a.flexible) a.zold.c[comp][c] = zv;
  }
}

// iteration q: Q = A P and the partials of P·Q (implicit.jl:418-420)
__global__ void __launch_bounds__(RED_T) k_pcg_ap(PcgArgs a) {
  if (!a.d.fl[FL_ACT0 + (a.q & 1)]) return;
  __shared__ double sh[RED_T];
  const int tid = threadIdx.x;
  const BoxIdx b = box_idx(a.g);
  PDir P{};
  for (int comp = 0; comp < 3; ++comp) P.z[comp] = a.Pnew.c[comp];
  double s = 0.0;
  for (int comp = 0; comp < 3; ++comp) {
    if (!a.rg.in(comp, b.i, b.j, b.k)) continue;
    const double p = a.Pnew.c[comp][b.c];
    const double Aq = apply_one(a, P, comp, b.i, b.j, b.k, b.c, p);
    a.Q.c[comp][b.c] = Aq;
    s += p * Aq;
  }
  const double r0 = block_sum_t(s, sh, tid);
  if (tid == 0) {
    double *pr = region_w(a.d, PART_PQ);
    pr[2 * b.bid] = r0;
    pr[2 * b.bid + 1] = 0.0;
  }
}

// iteration q: α = rz/pAp, x += αP, r −= αQ and the partials of r·r; IT += 1
__global__ void __launch_bounds__(RED_T) k_pcg_xr(PcgArgs a) {
  if (!a.d.fl[FL_ACT0 + (a.q & 1)]) return;
  __shared__ double sh[RED_T];
  const int tid = threadIdx.x;
  const BoxIdx b = box_idx(a.g);
  const double pAp = sum_parts(region(a.d, PART_PQ), a.nparts_in, 0, sh, tid);
  const double al = a.d.sc[SC_RZ0 + (a.q & 1)] / pAp;
  if (b.bid == 0 && tid == 0) a.d.fl[FL_IT] += 1;
  double s = 0.0;
  for (int comp = 0; comp < 3; ++comp) {
    if (!a.rg.in(comp, b.i, b.j, b.k)) continue;
    const int c = b.c;
    a.x.c[comp][c] += al * a.Pnew.c[comp][c];
    const double rv = a.r.c[comp][c] - al * a.Q.c[comp][c];
    a.r.c[comp][c] = rv;
    s += rv * rv;
  }
  const double r0 = block_sum_t(s, sh, tid);
  if (tid == 0) {
    double *pr = region_w(a.d, PART_RR);
    pr[2 * b.bid] = r0;
    pr[2 * b.bid + 1] = 0.0;
  }
}

// φ = θ + x on the ranges (implicit.jl:440-444) and the cycle's counts of the solve
__global__ void __launch_bounds__(RED_T) k_pcg_finish(PcgArgs a) {
  const BoxIdx b = box_idx(a.g);
  for (int comp = 0; comp < 3; ++comp)
    if (a.rg.in(comp, b.i, b.j, b.k)) a.ph[comp][b.c] = a.th[comp][b.c] + a.x.c[comp][b.c];
  if (b.bid == 0 && threadIdx.x == 0) {
    int *fl = a.d.fl;
    if (fl[FL_NZ] && fl[FL_ACT0 + (a.act_slot & 1)]) fl[FL_OVF] = 1;
    fl[FL_ITSUM] += fl[FL_IT];
    if (fl[FL_IT] > fl[FL_ITMAX]) fl[FL_ITMAX] = fl[FL_IT];
  }
}

static hipError_t launch_box(void (*k)(PcgArgs), const PcgArgs &a, hipStream_t s) {
  const int nb = pcg_box_blocks(a.g);
  if (nb > MAX_PART) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k, dim3(nb), dim3(RED_T), 0, s, a);
  return hipGetLastError();
}
hipError_t launch_pcg_start(const PcgArgs &a, hipStream_t s) { return launch_box(k_pcg_start, a, s); }
hipError_t launch_pcg_resid(const PcgArgs &a, hipStream_t s) { return launch_box(k_pcg_resid, a, s); }
hipError_t launch_pcg_p(const PcgArgs &a, hipStream_t s) { return launch_box(k_pcg_p, a, s); }
hipError_t launch_pcg_ap(const PcgArgs &a, hipStream_t s) { return launch_box(k_pcg_ap, a, s); }
hipError_t launch_pcg_xr(const PcgArgs &a, hipStream_t s) { return launch_box(k_pcg_xr, a, s); }
hipError_t launch_pcg_finish(const PcgArgs &a, hipStream_t s) { return launch_box(k_pcg_finish, a, s); }

__global__ void k_pcg_cycle_reset(PcgDev d) {
  if (threadIdx.x == 0) {
    d.fl[FL_OVF] = 0;
    d.fl[FL_ITSUM] = 0;
    d.fl[FL_ITMAX] = 0;
  }
}

hipError_t launch_pcg_cycle_reset(PcgDev d, hipStream_t s) {
  hipLaunchKernelGGL(k_pcg_cycle_reset, dim3(1), dim3(64), 0, s, d);
  return hipGetLastError();
}

}  // namespace tdgl
