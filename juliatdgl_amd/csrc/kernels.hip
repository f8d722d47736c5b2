// kernels.hip — the HIP kernels of one release-2.0 time step on gfx950.
//
// Every kernel keeps the reference's order of operations per node (gtdgl/src/bands.jl,
// field.jl, implicit.jl); the file is compiled with -ffp-contract=off, so products and sums
// round as the CPU step's do.  Transcendentals (sincos, atan2, sqrt) come from the device math
// library.  Node-shaped layout and ranges: common.h.
#include "kernels.h"

namespace tdgl {

// ============================================================================ band stage
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
    const cplx lxp = a.Lx[c], lyp = a.Ly[c], lzp = a.Lz[c];
    const cplx lxm = a.Lx[c - sx], lym = a.Ly[c - sy], lzm = a.Lz[c - sz];
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

// ============================================================================ stage links
// Rate of the frame at a block node: the faces copy the interior (precedence x, y, z), so the
// value at any block node is the one of the nearest interior node (bands.jl:224-232).
__device__ __forceinline__ double om_at(const Geo &g, const double *Om, int i, int j, int k) {
  i = i < g.x1 + 1 ? g.x1 + 1 : (i > g.x2 ? g.x2 : i);
  j = j < g.y1 + 1 ? g.y1 + 1 : (j > g.y2 ? g.y2 : j);
  k = k < g.z1 + 1 ? g.z1 + 1 : (k > g.z2 ? g.z2 : k);
  return Om[g.at(i, j, k)];
}

__device__ __forceinline__ void link_one(const Geo &g, const LinkArgs &a, int ax, int c, int i, int j, int k) {
  const int di = ax == 0, dj = ax == 1, dk = ax == 2;
  switch (a.mode) {
    case LINK_ROTFRAC: {
      const cplx v = a.V[ax][c];
      const cplx q = cmul(v, cconj(a.U[ax][c]));
      const double sf = -q.y;
      if (q.x > 0 && fabs(sf) <= 0.05) {
        const double t = (sf + sf * sf * sf / 6) * a.c;
        const double t2 = t * t;
        a.L[ax][c] = cmul(v, mk(1 - t2 / 2, -t * (1 - t2 / 6)));
      } else {
        a.L[ax][c] = cmul(v, dcis(dangle(q) * a.c));
      }
      break;
    }
    case LINK_ROTLINKS:
      a.L[ax][c] = cmul(a.V[ax][c], scis(a.R[ax][c] * a.c));
      break;
    case LINK_RATE: {
      const cplx q = cmul(a.V[ax][c], cconj(a.U[ax][c]));
      const double sf = -q.y;
      const double t = q.x > 0 && fabs(sf) <= 0.05 ? -(sf + sf * sf * sf / 6) : dangle(q);
      a.R[ax][c] = t + (om_at(g, a.Om, i + di, j + dj, k + dk) - om_at(g, a.Om, i, j, k)) * a.dt;
      break;
    }
    case LINK_ROOT: {
      const cplx v = a.V[ax][c];
      a.L[ax][c] = v;
      const cplx q = cmul(v, cconj(a.U[ax][c]));
      const double sf = -q.y;
      if (q.x > 0 && fabs(sf) <= 0.05) {
        const double t = (sf + sf * sf * sf / 6) / a.n;
        const double t2 = t * t;
        a.Q[ax][c] = mk(1 - t2 / 2, -t * (1 - t2 / 6));
      } else {
        a.Q[ax][c] = dcis(dangle(q) / a.n);
      }
      break;
    }
    case LINK_ADVANCE: {
      cplx l = cmul(a.L[ax][c], a.Q[ax][c]);
      if (a.frame) l = cmul(l, scis((om_at(g, a.Om, i + di, j + dj, k + dk) - om_at(g, a.Om, i, j, k)) * a.c));
      a.L[ax][c] = l;
      break;
    }
  }
}

// Over the block: x links on _rx = (x1:x2, y1:y2+1, z1:z2+1), y links on _ry, z links on _rz
// (bands.jl:349-351).
__global__ void __launch_bounds__(256) k_links(Geo g, LinkArgs a) {
  const int i = g.x1 + blockIdx.x * blockDim.x + threadIdx.x;
  const int j = g.y1 + blockIdx.y * blockDim.y + threadIdx.y;
  const int k = g.z1 + blockIdx.z;
  if (i > g.x2 + 1 || j > g.y2 + 1 || k > g.z2 + 1) return;
  const int c = g.at(i, j, k);
  if (i <= g.x2) link_one(g, a, 0, c, i, j, k);
  if (j <= g.y2) link_one(g, a, 1, c, i, j, k);
  if (k <= g.z2) link_one(g, a, 2, c, i, j, k);
}

hipError_t launch_links(const Geo &g, const LinkArgs &a, hipStream_t s) {
  const int ni = g.x2 - g.x1 + 2, nj = g.y2 - g.y1 + 2, nk = g.z2 - g.z1 + 2;
  dim3 b(64, 4, 1), gr(cdiv(ni, 64), cdiv(nj, 4), nk);
  hipLaunchKernelGGL(k_links, gr, b, 0, s, g, a);
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
__global__ void __launch_bounds__(256) k_field(Geo g, FieldArgs a) {
  const int i = 1 + blockIdx.x * blockDim.x + threadIdx.x;
  const int j = 1 + blockIdx.y * blockDim.y + threadIdx.y;
  const int k = 1 + blockIdx.z;
  if (i > g.Nx + 1 || j > g.Ny + 1 || k > g.Nz + 1) return;
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
      if (a.gather && a.rg.in(0, i, j, k)) a.B[a.rg.idx(0, i, j, k)] = ph - a.th[0][c];
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
      if (a.gather && a.rg.in(1, i, j, k)) a.B[a.rg.idx(1, i, j, k)] = ph - a.th[1][c];
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
      if (a.gather && a.rg.in(2, i, j, k)) a.B[a.rg.idx(2, i, j, k)] = ph - a.th[2][c];
    } else {
      ph = a.ph[2][c];
    }
    if (a.scale) ph = a.th[2][c] + (ph - a.th[2][c]) / a.D[2][c];
    if (inr || a.scale) a.ph[2][c] = ph;
  }
}

hipError_t launch_field(const Geo &g, const FieldArgs &a, hipStream_t s) {
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
        v = -a.HJ;
        set = true;
      }
      if (k == Nz) {
        v = +a.HJ;
        set = true;
      }
    } else {
      if (k == 1 || k == Nz) {
        v = a.HJ * a.bar_hy_z[(j - 1) + (Ny + 1) * (k == 1 ? 0 : 1)];
        if (k == 1 && k == Nz) v = a.HJ * a.bar_hy_z[(j - 1) + (Ny + 1)];
        set = true;
      }
      if (j == 1 || j == Ny + 1) {
        v = a.HJ * a.bar_hy_y[(j == 1 ? 0 : 1) + 2 * (k - 1)];
        set = true;
      }
      if (i == 1 || i == Nx) {
        v = a.HJ * a.bar_hy_x[(j - 1) + (Ny + 1) * (k - 1)];
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
      v = a.H;
      set = true;
    }
    if (a.bar) {
      if (k == 1 || k == Nz + 1) v = a.H + a.HJ * a.bar_hz_z[(j - 1) + Ny * (k == 1 ? 0 : 1)];
      if (j == 1 || j == Ny) v = a.H + a.HJ * a.bar_hz_y[(j == 1 ? 0 : 1) + 2 * (k - 1)];
      if (i == 1 || i == Nx) v = a.H + a.HJ * a.bar_hz_x[(j - 1) + Ny * (k - 1)];
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

__global__ void __launch_bounds__(256) k_scatter(Geo g, Ranges rg, const double *B, const double *th0,
                                                 const double *th1, const double *th2, double *ph0, double *ph1,
                                                 double *ph2) {
  const int i = 2 + blockIdx.x * blockDim.x + threadIdx.x;
  const int j = 2 + blockIdx.y * blockDim.y + threadIdx.y;
  const int k = 2 + blockIdx.z;
  const int c = g.at(i, j, k);
  if (rg.in(0, i, j, k)) ph0[c] = th0[c] + B[rg.idx(0, i, j, k)];
  if (rg.in(1, i, j, k)) ph1[c] = th1[c] + B[rg.idx(1, i, j, k)];
  if (rg.in(2, i, j, k)) ph2[c] = th2[c] + B[rg.idx(2, i, j, k)];
}

hipError_t launch_scatter(const Geo &g, const Ranges &rg, const double *B, const double *const th[3],
                          double *const ph[3], hipStream_t s) {
  dim3 b(64, 4, 1), gr(cdiv(g.Nx - 1, 64), cdiv(g.Ny - 1, 4), g.Nz - 1);
  hipLaunchKernelGGL(k_scatter, gr, b, 0, s, g, rg, B, th[0], th[1], th[2], ph[0], ph[1], ph[2]);
  return hipGetLastError();
}

// ============================================================================ separable solve
// out[m along axis] = Σ_n A[n*L + m] in[n along axis] for the three components (blockIdx.y).
// Each thread one output element; the sum runs over n in increasing order.
template <class TI, class TA, class TO>
__global__ void __launch_bounds__(256) k_tx(TxPass p, const int *gate) {
  if (gate && !*gate) return;
  const int comp = blockIdx.y;
  const int e0 = p.e[comp][0], e1 = p.e[comp][1], e2 = p.e[comp][2];
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= e0 * e1 * e2) return;
  const TI *in = (const TI *)p.in[comp];
  TO *out = (TO *)p.out[comp];
  const TA *A = (const TA *)p.A[comp];
  const int i0 = t % e0;
  const int r = t / e0;
  const int i1 = r % e1;
  const int i2 = r / e1;
  TA acc = 0;
  if (p.axis == 0) {
    const TI *line = in + e0 * (i1 + e1 * i2);
    for (int n = 0; n < e0; ++n) acc += A[n * e0 + i0] * (TA)line[n];
  } else if (p.axis == 1) {
    const TI *line = in + i0 + e0 * e1 * i2;
    for (int n = 0; n < e1; ++n) acc += (TA)line[e0 * n] * A[n * e1 + i1];
  } else {
    const TI *line = in + i0 + e0 * i1;
    for (int n = 0; n < e2; ++n) acc += (TA)line[e0 * e1 * n] * A[n * e2 + i2];
  }
  out[t] = (TO)acc;
}

static int tx_blocks(const TxPass &p) {
  int m = 0;
  for (int c = 0; c < 3; ++c) {
    const int n = p.e[c][0] * p.e[c][1] * p.e[c][2];
    m = n > m ? n : m;
  }
  return cdiv(m, 256);
}

template <class T>
hipError_t launch_tx(const TxPass &p, const int *gate, hipStream_t s) {
  hipLaunchKernelGGL((k_tx<T, T, T>), dim3(tx_blocks(p), 3), dim3(256), 0, s, p, gate);
  return hipGetLastError();
}
template hipError_t launch_tx<double>(const TxPass &, const int *, hipStream_t);
template hipError_t launch_tx<float>(const TxPass &, const int *, hipStream_t);

hipError_t launch_tx_d2f(const TxPass &p, const int *gate, hipStream_t s) {
  // the input is rounded to Float32 first (BX .= r[1], implicit.jl:357)
  hipLaunchKernelGGL((k_tx<double, float, float>), dim3(tx_blocks(p), 3), dim3(256), 0, s, p, gate);
  return hipGetLastError();
}
hipError_t launch_tx_f2d(const TxPass &p, const int *gate, hipStream_t s) {
  hipLaunchKernelGGL((k_tx<float, float, double>), dim3(tx_blocks(p), 3), dim3(256), 0, s, p, gate);
  return hipGetLastError();
}

// The per-wavevector projector (I + aK)^{-1} = P_L + P_T/(1 + a|k|²) (implicit.jl:67-78).
template <class T>
__global__ void __launch_bounds__(256) k_modes(ModeArgs m, const int *gate) {
  if (gate && !*gate) return;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int nx = m.nx, ny = m.ny, nz = m.nz;
  if (t >= nx * ny * nz) return;
  const int mx = t % nx, my = (t / nx) % ny, mz = t / (nx * ny);
  const T *sx = (const T *)m.sx, *sy = (const T *)m.sy, *sz = (const T *)m.sz;
  T *X = (T *)m.X, *Y = (T *)m.Y, *Z = (T *)m.Z;
  const T k1 = sx[mx], k2 = sy[my], k3 = sz[mz];
  const T k2s = k1 * k1 + k2 * k2 + k3 * k3;
  if (k2s == 0) return;
  const int ix = (mx - 1) + (nx - 1) * (my + ny * mz);
  const int iy = mx + nx * ((my - 1) + (ny - 1) * mz);
  const int iz = mx + nx * (my + ny * (mz - 1));
  const T x1 = mx > 0 ? X[ix] : T(0);
  const T y1 = my > 0 ? Y[iy] : T(0);
  const T z1 = mz > 0 ? Z[iz] : T(0);
  const T d = (k1 * x1 + k2 * y1 + k3 * z1) / k2s;
  const T f = T(1) / (T(1) + (T)m.a * k2s);
  if (mx > 0) X[ix] = k1 * d + f * (x1 - k1 * d);
  if (my > 0) Y[iy] = k2 * d + f * (y1 - k2 * d);
  if (mz > 0) Z[iz] = k3 * d + f * (z1 - k3 * d);
}

template <class T>
hipError_t launch_modes(const ModeArgs &m, const int *gate, hipStream_t s) {
  hipLaunchKernelGGL((k_modes<T>), dim3(cdiv(m.nx * m.ny * m.nz, 256)), dim3(256), 0, s, m, gate);
  return hipGetLastError();
}
template hipError_t launch_modes<double>(const ModeArgs &, const int *, hipStream_t);
template hipError_t launch_modes<float>(const ModeArgs &, const int *, hipStream_t);

// ============================================================================ PCG
// P outside its ranges is zero (implicit.jl:300, 406)
__device__ __forceinline__ double pv(const Ranges &rg, const double *P, int comp, int i, int j, int k) {
  return rg.in(comp, i, j, k) ? P[rg.idx(comp, i, j, k)] : 0.0;
}
// h of the search direction on the ranges of _apply_A! (implicit.jl:216-224), zero elsewhere
__device__ __forceinline__ double hx_of(const Geo &g, const Ranges &rg, const double *P, int i, int j, int k,
                                        double dydz) {
  if (!(i >= 2 && i <= g.Nx && j >= 2 && j <= g.Ny - 1 && k >= 2 && k <= g.Nz - 1)) return 0.0;
  return ((pv(rg, P, 2, i, j + 1, k) - pv(rg, P, 2, i, j, k)) - pv(rg, P, 1, i, j, k + 1) + pv(rg, P, 1, i, j, k)) / dydz;
}
__device__ __forceinline__ double hy_of(const Geo &g, const Ranges &rg, const double *P, int i, int j, int k,
                                        double dzdx) {
  if (!(i >= 2 && i <= g.Nx - 1 && j >= 2 && j <= g.Ny && k >= 2 && k <= g.Nz - 1)) return 0.0;
  return ((pv(rg, P, 0, i, j, k + 1) - pv(rg, P, 0, i, j, k)) - pv(rg, P, 2, i + 1, j, k) + pv(rg, P, 2, i, j, k)) / dzdx;
}
__device__ __forceinline__ double hz_of(const Geo &g, const Ranges &rg, const double *P, int i, int j, int k,
                                        double dxdy) {
  if (!(i >= 2 && i <= g.Nx - 1 && j >= 2 && j <= g.Ny - 1 && k >= 2 && k <= g.Nz)) return 0.0;
  return ((pv(rg, P, 1, i + 1, j, k) - pv(rg, P, 1, i, j, k)) - pv(rg, P, 0, i, j + 1, k) + pv(rg, P, 0, i, j, k)) / dxdy;
}

struct ApplyArgs {
  const double *P;
  double *Q;
  const double *D0, *D1, *D2;
  double sxy, sxz, syx, syz, szx, szy, dydz, dzdx, dxdy;
};

// q = (D + aK) p on the ranges (_apply_A!, implicit.jl:212-236), h recomputed per use
__global__ void __launch_bounds__(256) k_apply_A(Geo g, Ranges rg, ApplyArgs a, const int *gate) {
  if (gate && !*gate) return;
  const int i = 2 + blockIdx.x * blockDim.x + threadIdx.x;
  const int j = 2 + blockIdx.y * blockDim.y + threadIdx.y;
  const int k = 2 + blockIdx.z;
  const int c = g.at(i, j, k);
  const double *P = a.P;
  if (rg.in(0, i, j, k)) {
    const double p = P[rg.idx(0, i, j, k)];
    a.Q[rg.idx(0, i, j, k)] = a.D0[c] * p + a.sxy * (hz_of(g, rg, P, i, j, k, a.dxdy) - hz_of(g, rg, P, i, j - 1, k, a.dxdy)) -
                              a.sxz * (hy_of(g, rg, P, i, j, k, a.dzdx) - hy_of(g, rg, P, i, j, k - 1, a.dzdx));
  }
  if (rg.in(1, i, j, k)) {
    const double p = P[rg.idx(1, i, j, k)];
    a.Q[rg.idx(1, i, j, k)] = a.D1[c] * p + a.syz * (hx_of(g, rg, P, i, j, k, a.dydz) - hx_of(g, rg, P, i, j, k - 1, a.dydz)) -
                              a.syx * (hz_of(g, rg, P, i, j, k, a.dxdy) - hz_of(g, rg, P, i - 1, j, k, a.dxdy));
  }
  if (rg.in(2, i, j, k)) {
    const double p = P[rg.idx(2, i, j, k)];
    a.Q[rg.idx(2, i, j, k)] = a.D2[c] * p + a.szx * (hy_of(g, rg, P, i, j, k, a.dzdx) - hy_of(g, rg, P, i - 1, j, k, a.dzdx)) -
                              a.szy * (hx_of(g, rg, P, i, j, k, a.dydz) - hx_of(g, rg, P, i, j - 1, k, a.dydz));
  }
}

hipError_t launch_apply_A(const Geo &g, const Ranges &rg, const double *P, double *Q, const double *const D[3],
                          double sxy, double sxz, double syx, double syz, double szx, double szy, double dx, double dy,
                          double dz, const int *gate, hipStream_t s) {
  ApplyArgs a{P, Q, D[0], D[1], D[2], sxy, sxz, syx, syz, szx, szy, dy * dz, dz * dx, dx * dy};
  dim3 b(64, 4, 1), gr(cdiv(g.Nx - 1, 64), cdiv(g.Ny - 1, 4), g.Nz - 1);
  hipLaunchKernelGGL(k_apply_A, gr, b, 0, s, g, rg, a, gate);
  return hipGetLastError();
}

// Deterministic dot products: a fixed grid of NPART blocks, each a grid-stride sum in a fixed
// order and a fixed tree in LDS; the partials are summed by one block in k_finish.
constexpr int NPART = 512;
constexpr int DOT_T = 256;
int dot_blocks() { return NPART; }

__device__ __forceinline__ double block_sum(double v, double *sh) {
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  const double r = sh[0];
  __syncthreads();
  return r;
}

__global__ void __launch_bounds__(DOT_T) k_dot(const double *a, const double *b, const double *c, int n, double *part,
                                              const int *gate) {
  if (gate && !*gate) return;
  __shared__ double sh[DOT_T];
  double s0 = 0.0, s1 = 0.0;
  for (int t = blockIdx.x * DOT_T + threadIdx.x; t < n; t += NPART * DOT_T) {
    s0 += a[t] * b[t];
    if (c) s1 += a[t] * c[t];
  }
  const double r0 = block_sum(s0, sh);
  const double r1 = c ? block_sum(s1, sh) : 0.0;
  if (threadIdx.x == 0) {
    part[blockIdx.x] = r0;
    part[NPART + blockIdx.x] = r1;
  }
}

hipError_t launch_dot(const double *a, const double *b, const double *c, int n, double *part, const int *gate,
                      hipStream_t s) {
  hipLaunchKernelGGL(k_dot, dim3(NPART), dim3(DOT_T), 0, s, a, b, c, n, part, gate);
  return hipGetLastError();
}

__global__ void __launch_bounds__(NPART) k_finish(int what, PcgDev d, int flexible, const int *gate) {
  if (gate && !*gate) return;
  __shared__ double sh[NPART];
  const double s0 = block_sum(d.part[threadIdx.x], sh);
  const double s1 = (what == FIN_BETA && flexible) ? block_sum(d.part[NPART + threadIdx.x], sh) : 0.0;
  if (threadIdx.x != 0) return;
  double *sc = d.sc;
  int *fl = d.fl;
  switch (what) {
    case FIN_NB:
      sc[SC_NB] = sqrt(s0);
      fl[FL_ACTIVE] = sc[SC_NB] > 0 ? 1 : 0;
      fl[FL_COLD] = fl[FL_ACTIVE];
      fl[FL_IT] = 0;
      break;
    case FIN_RR0:
      fl[FL_ACTIVE] = (sc[SC_NB] > 0 && sqrt(s0) > sc[SC_TOLNB] * sc[SC_NB]) ? 1 : 0;
      break;
    case FIN_RZ0:
      sc[SC_RZ] = s0;
      break;
    case FIN_ALPHA:
      fl[FL_IT] += 1;
      sc[SC_PAP] = s0;
      sc[SC_ALPHA] = sc[SC_RZ] / s0;
      break;
    case FIN_RR:
      sc[SC_RR] = s0;
      if (sqrt(s0) <= sc[SC_TOLNB] * sc[SC_NB] || fl[FL_IT] >= fl[FL_MAXIT]) fl[FL_ACTIVE] = 0;
      break;
    case FIN_BETA: {
      const double rz2 = s0;
      sc[SC_BETA] = flexible ? (rz2 - s1) / sc[SC_RZ] : rz2 / sc[SC_RZ];
      sc[SC_RZ] = rz2;
      break;
    }
  }
}

hipError_t launch_finish(int what, PcgDev d, int flexible, const int *gate, hipStream_t s) {
  hipLaunchKernelGGL(k_finish, dim3(1), dim3(NPART), 0, s, what, d, flexible, gate);
  return hipGetLastError();
}

__global__ void __launch_bounds__(256) k_vec(VecArgs v, const int *gate) {
  if (gate && !*gate) return;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= v.n) return;
  switch (v.op) {
    case V_COPY:
      v.P[t] = v.z[t];
      if (v.flexible) v.zold[t] = v.z[t];
      break;
    case V_XR: {
      const double al = v.sc[SC_ALPHA];
      v.x[t] += al * v.P[t];
      v.r[t] -= al * v.Q[t];
      break;
    }
    case V_P: {
      const double be = v.sc[SC_BETA];
      v.P[t] = v.z[t] + be * v.P[t];
      if (v.flexible) v.zold[t] = v.z[t];
      break;
    }
    case V_RESID:
      v.r[t] = v.b[t] - v.Q[t];
      break;
    case V_START: {
      // _pcg_zero_update! (implicit.jl:251-271) when the source is zero, else the warm start
      // of _implicit_curl_fast! (:391-402)
      if (v.sc[SC_NB] == 0) {
        if (v.warm && v.extrap) v.xprev[t] = v.x[t];
        v.x[t] = 0.0;
      } else if (v.warm && v.extrap) {
        if (v.have2) {
          const double x = v.x[t];
          v.x[t] = 2 * x - v.xprev[t];
          v.xprev[t] = x;
        } else {
          v.xprev[t] = v.x[t];
        }
      }
      break;
    }
  }
}

hipError_t launch_vec(const VecArgs &v, const int *gate, hipStream_t s) {
  hipLaunchKernelGGL(k_vec, dim3(cdiv(v.n, 256)), dim3(256), 0, s, v, gate);
  return hipGetLastError();
}

}  // namespace tdgl
