// oracle.cpp — CPU restatement of the release-2.0 time step (gtdgl/src/step.jl:13-35).
//
// TEST INFRASTRUCTURE ONLY.  This is the parity oracle and the CPU baseline of the HIP stepper:
// tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it; the product
// (juliatdgl_amd) never does.
//
// It restates, in the reference's loop order and order of additions, the options the HIP
// stepper runs: N bands; Euler with substeps (extrapolated link roots), RKL1 and damped RKC1
// stages; KWT; the rotating frame; supercurrent; field increment; the implicit curl-curl by the
// separable solve or by PCG (exact, fast, Float32 preconditioner); explicit and D-scaled
// explicit field; links; curl; slab and bar field boundaries; Neumann ghosts; advance; two and N
// normal fluids with a finite exchange (uniform, separable: exchange.jl); thermal noise
// (noise.jl: Euler–Maruyama and polar KWT kicks, SK-ROCK, the field kick) with the KWT log-polar
// stepping and the low-passed frame that come with it (bands.jl).
// Arrays are Julia-shaped (column-major, 1-based accessors) so every index reads as in the
// reference.  Complex arithmetic is written out with the formulas of Julia's Base (no fused
// multiply-add: build with -ffp-contract=off).
//
// Pinned by: the golden states of gtdgl/test/golden (k2_R13_explicit, k2_R4_prod, k25_R4_prod:
// Euler + substeps + extrapolated links, separable and explicit field, slab BC); the closed
// forms of gtdgl/test/test_physics.jl and TDGLGPU/test/test_invariants.jl (KWT 2x2 solve,
// amplitude relaxation rates, uniform equilibrium, pure gauge); RKC polynomial identities; the
// separable solve and the PCG (exact, fast, Float32 preconditioner) against (I + aK) and
// (D + aK) assembled in the tests from the circulation of the links around the plaquettes (the
// true residual within the tolerance, and a sparse direct solve); the PCG at D = 1 equal to the
// separable solve; the zero-source rule and its checkpoint continuation of
// gtdgl/test/pcg_zero_fixture.jl; the 93 manufactured rotations of gtdgl/test/rotation_fixture.jl
// with the bounds of test_link_rotations.jl (_rotfrac!, _rot! roots, _link_rate!, _cis); the
// anisotropic masses by the Euler factor of discrete Neumann plane waves along each axis; the
// noise by the equilibria of gtdgl/test/test_physics.jl:475-535, 697-735 (exp(−F/T̃) of decoupled
// nodes with and without KWT, the phase and amplitude rates, equipartition of the field's modes)
// and its normals by the Philox4x32-10 known-answer vectors of Random123.  The tests that do so
// are tests/test_oracle_*.py.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../include/tdglhip.h"

namespace {

struct C {
  double re, im;
};
// Julia Base.complex arithmetic (complex.jl), no contraction
inline C cmul(C a, C b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
inline C cadd(C a, C b) { return {a.re + b.re, a.im + b.im}; }
inline C csub(C a, C b) { return {a.re - b.re, a.im - b.im}; }
inline C rmul(double x, C a) { return {x * a.re, x * a.im}; }
inline C conj(C a) { return {a.re, -a.im}; }
inline double abs2(C a) { return a.re * a.re + a.im * a.im; }
inline C cis(double x) { return {std::cos(x), std::sin(x)}; }
inline double angle(C a) { return std::atan2(a.im, a.re); }
// im * x * y with x real (the frame term): Complex(±0, x) * y
inline C imx_mul(double x, C y) { return {0.0 * y.re - x * y.im, 0.0 * y.im + x * y.re}; }

template <class T>
struct A3 {
  int n1 = 0, n2 = 0, n3 = 0;
  std::vector<T> v;
  A3() {}
  A3(int a, int b, int c) : n1(a), n2(b), n3(c), v((size_t)a * b * c, T{}) {}
  inline T &operator()(int i, int j, int k) { return v[(size_t)(i - 1) + (size_t)n1 * ((j - 1) + (size_t)n2 * (k - 1))]; }
  inline const T &operator()(int i, int j, int k) const {
    return v[(size_t)(i - 1) + (size_t)n1 * ((j - 1) + (size_t)n2 * (k - 1))];
  }
  size_t size() const { return v.size(); }
};
using A3c = A3<C>;
using A3r = A3<double>;

template <class T>
void load(A3<T> &a, const double *p) {
  std::memcpy(a.v.data(), p, a.v.size() * sizeof(T));
}
template <class T>
void save(const A3<T> &a, double *p) {
  std::memcpy(p, a.v.data(), a.v.size() * sizeof(T));
}

const double ROT_MIN2 = 1e-3;  // bands.jl:208
const double ROT_MAXDT = 0.5;  // bands.jl:209

inline double jclamp(double x, double lo, double hi) { return x > hi ? hi : (x < lo ? lo : x); }

// kwt_update, bands.jl:135-140
inline C kwt_update(C g, C d, double g2) {
  double a2 = abs2(g);
  double q = 1 + g2 * a2;
  C dp = rmul(std::sqrt(q), d);
  double s = cmul(conj(g), dp).re / q;
  return cadd(g, csub(dp, rmul(s, rmul(g2, g))));
}

// kwt_update_rot, bands.jl:150-158: the increment's part across ψ as a rotation
inline C kwt_update_rot(C g, C d, double g2) {
  double a2 = abs2(g);
  double q = 1 + g2 * a2;
  C dp = rmul(std::sqrt(q), d);
  C Dl = csub(dp, rmul(cmul(conj(g), dp).re / q, rmul(g2, g)));
  if (!(a2 > 1e-24)) return cadd(g, Dl);
  double a = std::sqrt(a2);
  C u = {g.re / a, g.im / a};
  C w = cmul(conj(u), Dl);
  return cmul(u, rmul(a + w.re, cis(w.im / a)));
}

// _cis, bands.jl:289
inline C scis(double a) {
  if (std::fabs(a) <= 0.05) {
    double a2 = a * a;
    return {1 - a2 / 2, a * (1 - a2 / 6)};
  }
  return cis(a);
}

// 1D bases of the separable solve: V on nn nodes (cosine, DCT-II), U on nn-1 edges (sine,
// DST-I), s the singular values of the forward difference in ascending order
// (implicit.jl:15-22; the SVD there gives these bases up to signs that cancel in the solve).
struct Basis {
  int nn;
  std::vector<double> V, U, s;  // V (nn x nn), U (nn-1 x nn-1), column-major
};
Basis make_basis(int nn) {
  Basis b;
  b.nn = nn;
  b.V.assign((size_t)nn * nn, 0.0);
  b.U.assign((size_t)(nn - 1) * (nn - 1), 0.0);
  b.s.assign(nn, 0.0);
  const double pi = 3.14159265358979323846;
  for (int k = 0; k < nn; ++k) {
    b.s[k] = k == 0 ? 0.0 : 2.0 * std::sin(pi * k / (2.0 * nn));
    for (int i = 0; i < nn; ++i)
      b.V[i + (size_t)nn * k] = k == 0 ? 1.0 / std::sqrt((double)nn)
                                       : std::sqrt(2.0 / nn) * std::cos(pi * k * (i + 0.5) / nn);
  }
  for (int k = 1; k < nn; ++k)
    for (int e = 0; e < nn - 1; ++e)
      b.U[e + (size_t)(nn - 1) * (k - 1)] = -std::sqrt(2.0 / nn) * std::sin(pi * k * (e + 1.0) / nn);
  return b;
}

// out = A transformed by M1' along dim 1, M2' along dim 2, M3' along dim 3 (_tx!, implicit.jl:30-37);
// M_d is (n_d x n_d) column-major; `tr` transposes M_d (the inverse transform).  Every output
// element is the sum over n in increasing order; the loops run the contiguous index innermost
// (vectorised across independent outputs, which keeps each element's order of additions).
template <class R>
void tx(std::vector<R> &out, const std::vector<R> &A, int n1, int n2, int n3, const std::vector<R> &M1,
        const std::vector<R> &M2, const std::vector<R> &M3, bool tr) {
  // Mt[d][n*L + m] = (M_d or M_d')(n, m): the row of the n-th input across the outputs m
  auto rows = [&](const std::vector<R> &M, int L) {
    std::vector<R> t((size_t)L * L);
    for (int n = 0; n < L; ++n)
      for (int m = 0; m < L; ++m) t[(size_t)n * L + m] = tr ? M[m + (size_t)L * n] : M[n + (size_t)L * m];
    return t;
  };
  const std::vector<R> R1 = rows(M1, n1), R2 = rows(M2, n2), R3 = rows(M3, n3);
  std::vector<R> T1((size_t)n1 * n2 * n3), T2((size_t)n1 * n2 * n3);
  // T1[m, j, k] = sum_n M1[n, m] A[n, j, k]
#pragma omp parallel for schedule(static)
  for (int k = 0; k < n3; ++k)
    for (int j = 0; j < n2; ++j) {
      R *o = &T1[(size_t)n1 * (j + (size_t)n2 * k)];
      const R *a = &A[(size_t)n1 * (j + (size_t)n2 * k)];
      for (int m = 0; m < n1; ++m) o[m] = 0;
      for (int n = 0; n < n1; ++n) {
        const R an = a[n];
        const R *r = &R1[(size_t)n * n1];
        for (int m = 0; m < n1; ++m) o[m] += r[m] * an;
      }
    }
  // T2[i, m, k] = sum_n T1[i, n, k] M2[n, m]
#pragma omp parallel for schedule(static)
  for (int k = 0; k < n3; ++k)
    for (int m = 0; m < n2; ++m) {
      R *o = &T2[(size_t)n1 * (m + (size_t)n2 * k)];
      for (int i = 0; i < n1; ++i) o[i] = 0;
      for (int n = 0; n < n2; ++n) {
        const R w = R2[(size_t)n * n2 + m];
        const R *t = &T1[(size_t)n1 * (n + (size_t)n2 * k)];
        for (int i = 0; i < n1; ++i) o[i] += t[i] * w;
      }
    }
  // out[i, j, m] = sum_n T2[i, j, n] M3[n, m]
  out.assign((size_t)n1 * n2 * n3, 0);
#pragma omp parallel for schedule(static)
  for (int m = 0; m < n3; ++m) {
    R *o = &out[(size_t)n1 * n2 * m];
    for (int n = 0; n < n3; ++n) {
      const R w = R3[(size_t)n * n3 + m];
      const R *t = &T2[(size_t)n1 * n2 * n];
      for (size_t q = 0; q < (size_t)n1 * n2; ++q) o[q] += t[q] * w;
    }
  }
}

struct Sep {
  int Nx, Ny, Nz;
  Basis bx, by, bz;
  double a;
  int ex[3], ey[3], ez[3];  // extents of rx, ry, rz
  std::vector<float> V32[3], U32[3], s32[3];
};

// The mode projector of _sep_solve! (implicit.jl:67-78), in place on X, Y, Z.
template <class R>
void modes(std::vector<R> &X, std::vector<R> &Y, std::vector<R> &Z, const Sep &S, const R *sx, const R *sy,
           const R *sz, R a) {
  int nx = S.bx.nn, ny = S.by.nn, nz = S.bz.nn;
  // X: (nx-1, ny, nz), Y: (nx, ny-1, nz), Z: (nx, ny, nz-1)
  auto iX = [&](int mx, int my, int mz) { return (size_t)mx + (size_t)(nx - 1) * (my + (size_t)ny * mz); };
  auto iY = [&](int mx, int my, int mz) { return (size_t)mx + (size_t)nx * (my + (size_t)(ny - 1) * mz); };
  auto iZ = [&](int mx, int my, int mz) { return (size_t)mx + (size_t)nx * (my + (size_t)ny * mz); };
  for (int mz = 0; mz < nz; ++mz)
    for (int my = 0; my < ny; ++my)
      for (int mx = 0; mx < nx; ++mx) {
        R k1 = sx[mx], k2 = sy[my], k3 = sz[mz];
        R k2s = k1 * k1 + k2 * k2 + k3 * k3;
        if (k2s == 0) continue;
        R x1 = mx > 0 ? X[iX(mx - 1, my, mz)] : R(0);
        R y1 = my > 0 ? Y[iY(mx, my - 1, mz)] : R(0);
        R z1 = mz > 0 ? Z[iZ(mx, my, mz - 1)] : R(0);
        R d = (k1 * x1 + k2 * y1 + k3 * z1) / k2s;
        R f = R(1) / (R(1) + a * k2s);
        if (mx > 0) X[iX(mx - 1, my, mz)] = k1 * d + f * (x1 - k1 * d);
        if (my > 0) Y[iY(mx, my - 1, mz)] = k2 * d + f * (y1 - k2 * d);
        if (mz > 0) Z[iZ(mx, my, mz - 1)] = k3 * d + f * (z1 - k3 * d);
      }
}

// (I + aK)^{-1} applied in place (_sep_solve!, implicit.jl:62-83)
void sep_solve(const Sep &S, std::vector<double> &BX, std::vector<double> &BY, std::vector<double> &BZ) {
  std::vector<double> X, Y, Z;
  tx(X, BX, S.ex[0], S.ex[1], S.ex[2], S.bx.U, S.by.V, S.bz.V, false);
  tx(Y, BY, S.ey[0], S.ey[1], S.ey[2], S.bx.V, S.by.U, S.bz.V, false);
  tx(Z, BZ, S.ez[0], S.ez[1], S.ez[2], S.bx.V, S.by.V, S.bz.U, false);
  modes<double>(X, Y, Z, S, S.bx.s.data(), S.by.s.data(), S.bz.s.data(), S.a);
  tx(BX, X, S.ex[0], S.ex[1], S.ex[2], S.bx.U, S.by.V, S.bz.V, true);
  tx(BY, Y, S.ey[0], S.ey[1], S.ey[2], S.bx.V, S.by.U, S.bz.V, true);
  tx(BZ, Z, S.ez[0], S.ez[1], S.ez[2], S.bx.V, S.by.V, S.bz.U, true);
}

// The Float32 separable inverse (_precond32!, implicit.jl:353-373)
void sep_solve32(const Sep &S, std::vector<double> *z, const std::vector<double> *r) {
  std::vector<float> B[3], T[3];
  for (int c = 0; c < 3; ++c) {
    B[c].resize(r[c].size());
    for (size_t i = 0; i < r[c].size(); ++i) B[c][i] = (float)r[c][i];
  }
  tx(T[0], B[0], S.ex[0], S.ex[1], S.ex[2], S.U32[0], S.V32[1], S.V32[2], false);
  tx(T[1], B[1], S.ey[0], S.ey[1], S.ey[2], S.V32[0], S.U32[1], S.V32[2], false);
  tx(T[2], B[2], S.ez[0], S.ez[1], S.ez[2], S.V32[0], S.V32[1], S.U32[2], false);
  modes<float>(T[0], T[1], T[2], S, S.s32[0].data(), S.s32[1].data(), S.s32[2].data(), (float)S.a);
  tx(B[0], T[0], S.ex[0], S.ex[1], S.ex[2], S.U32[0], S.V32[1], S.V32[2], true);
  tx(B[1], T[1], S.ey[0], S.ey[1], S.ey[2], S.V32[0], S.U32[1], S.V32[2], true);
  tx(B[2], T[2], S.ez[0], S.ez[1], S.ez[2], S.V32[0], S.V32[1], S.U32[2], true);
  for (int c = 0; c < 3; ++c)
    for (size_t i = 0; i < B[c].size(); ++i) z[c][i] = (double)B[c][i];
}

double dot(const std::vector<double> &a, const std::vector<double> &b) {
  double s = 0;
  for (size_t i = 0; i < a.size(); ++i) s += a[i] * b[i];
  return s;
}

}  // namespace

struct orc {
  tdgl_desc d;
  int N;
  int Nx, Ny, Nz, x1, x2, y1, y2, z1, z2;
  // masks
  A3r f, tx, ty, tz;
  A3r Dx, Dy, Dz;
  std::vector<double> bar[6];
  // state (state.jl:51-60); the swaps of advance! swap these objects
  std::vector<A3c> G, F;
  A3r thx, thy, thz, phx, phy, phz;
  A3c Ux, Uy, Uz, Vx, Vy, Vz;
  A3r Jx, Jy, Jz, hx, hy, hz;
  // workspace (scheme.jl:156-165)
  A3c W, W2, L[6];
  A3r Om, R[3];
  // separable solve / PCG
  Sep sep;
  bool have_sep = false;
  std::vector<double> BX, BY, BZ;
  A3r P[3], Q[3], H[3];
  std::vector<double> x[3], r[3], z[3], b[3], xprev[3], zold[3];
  bool warm = false, have2 = false;
  int iters = 0;
  long long total_iters = 0;
  // normal fluids (exchange.jl): the bases of the region, and the solved modes (region-shaped)
  int qp = TDGL_QP_NONE, qK = 0, qj0 = 0, qnj = 0, qk0 = 0, qnk = 0;
  std::vector<double> qVx, qWx, qex, qVy, qWy, qey, qVz, qWz, qez;
  std::vector<double> qM[TDGL_MAX_BANDS], qT1, qT2;
  std::vector<double> qSrc[TDGL_MAX_BANDS];  // the sources of the last solve (before it)
  // open boundary (openbc.jl): the clock, δJ and θ_δ on the three link lattices, h_δ of the last
  // two recomputations (h-shaped)
  int ocount = -1;
  A3r odJ[3], oth[3], oold[3], onew[3];
  // thermal noise (noise.jl): the stream position, the draws of the current step (per band at the
  // nodes, per family at the links), SK-ROCK's phase kicks per band, the caller's draws
  uint64_t nctr = 0;
  std::vector<A3c> Xi;
  A3r Nl[3];
  std::vector<A3r> Qth;
  A3c Sp;
  A3r Om2;
  std::vector<double> hdraws;
  int hslots = 0;
  uint64_t hbase = 0;
  bool noise = false, sk_any = false;
};

namespace {

// Link ranges the band stencil and its ghosts read (bands.jl:349-351)
struct Rng {
  int i0, i1, j0, j1, k0, k1;
};
Rng rx(const orc &o) { return {o.x1, o.x2, o.y1, o.y2 + 1, o.z1, o.z2 + 1}; }
Rng ry(const orc &o) { return {o.x1, o.x2 + 1, o.y1, o.y2, o.z1, o.z2 + 1}; }
Rng rz(const orc &o) { return {o.x1, o.x2 + 1, o.y1, o.y2 + 1, o.z1, o.z2}; }
template <class F>
void forr(const Rng &r, F fn) {
  for (int k = r.k0; k <= r.k1; ++k)
    for (int j = r.j0; j <= r.j1; ++j)
      for (int i = r.i0; i <= r.i1; ++i) fn(i, j, k);
}

// _band_row! over the interior (bands.jl:71-97, one material, no drag, no heat field)
void band_pass(orc &o, A3c &D, const A3c &S, int mu, const A3c &Lx, const A3c &Ly, const A3c &Lz,
               const tdgl_bandcoef &cf) {
  const int N = o.N;
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = o.z1 + 1; k <= o.z2; ++k)
    for (int j = o.y1 + 1; j <= o.y2; ++j)
      for (int i = o.x1 + 1; i <= o.x2; ++i) {
        C g = S(i, j, k);
        double fm = o.f(i, j, k);
        if (o.d.hard && fm == 0.0) {
          D(i, j, k) = {0.0, 0.0};
          continue;
        }
        C vxm = conj(Lx(i - 1, j, k)), vym = conj(Ly(i, j - 1, k)), vzm = conj(Lz(i, j, k - 1));
        C tx_ = rmul(cf.kx, cadd(csub(cmul(Lx(i, j, k), S(i + 1, j, k)), rmul(2.0, g)), cmul(vxm, S(i - 1, j, k))));
        C ty_ = rmul(cf.ky, cadd(csub(cmul(Ly(i, j, k), S(i, j + 1, k)), rmul(2.0, g)), cmul(vym, S(i, j - 1, k))));
        C tz_ = rmul(cf.kz, cadd(csub(cmul(Lz(i, j, k), S(i, j, k + 1)), rmul(2.0, g)), cmul(vzm, S(i, j, k - 1))));
        C nl = rmul(fm * cf.eps - cf.beta * abs2(g), rmul(cf.c, g));
        C acc = cadd(cadd(cadd(cadd(g, tx_), ty_), tz_), nl);
        for (int nu = 0; nu < N; ++nu) {
          if (nu == mu) continue;
          acc = cadd(acc, rmul(cf.cpl[nu], o.G[nu](i, j, k)));
        }
        D(i, j, k) = cf.gamma2 > 0 ? (o.noise ? kwt_update_rot(g, csub(acc, g), cf.gamma2)
                                             : kwt_update(g, csub(acc, g), cf.gamma2))
                                  : acc;
      }
}

// _frame_rate! without noise (bands.jl:213-234)
void frame_rate(orc &o, A3r &Om, A3c &Cc, const A3c &Y, double h, double dt) {
  double wmax = ROT_MAXDT / dt;
  for (int k = o.z1 + 1; k <= o.z2; ++k)
    for (int j = o.y1 + 1; j <= o.y2; ++j)
      for (int i = o.x1 + 1; i <= o.x2; ++i) {
        C y = Y(i, j, k);
        double a2 = abs2(y);
        double w = a2 > ROT_MIN2 ? jclamp(cmul(conj(y), csub(Cc(i, j, k), y)).im / (a2 * h), -wmax, wmax) : 0.0;
        Om(i, j, k) = w;
        Cc(i, j, k) = csub(Cc(i, j, k), imx_mul(w * h, y));
      }
  for (int k = o.z1 + 1; k <= o.z2; ++k)
    for (int j = o.y1 + 1; j <= o.y2; ++j) {
      Om(o.x1, j, k) = Om(o.x1 + 1, j, k);
      Om(o.x2 + 1, j, k) = Om(o.x2, j, k);
    }
  for (int k = o.z1 + 1; k <= o.z2; ++k)
    for (int i = o.x1; i <= o.x2 + 1; ++i) {
      Om(i, o.y1, k) = Om(i, o.y1 + 1, k);
      Om(i, o.y2 + 1, k) = Om(i, o.y2, k);
    }
  for (int j = o.y1; j <= o.y2 + 1; ++j)
    for (int i = o.x1; i <= o.x2 + 1; ++i) {
      Om(i, j, o.z1) = Om(i, j, o.z1 + 1);
      Om(i, j, o.z2 + 1) = Om(i, j, o.z2);
    }
}

// _frame_term! without noise (bands.jl:274-286)
void frame_term(orc &o, A3c &Cc, const A3c &S, const A3r &Om, double h) {
  for (int k = o.z1 + 1; k <= o.z2; ++k)
    for (int j = o.y1 + 1; j <= o.y2; ++j)
      for (int i = o.x1 + 1; i <= o.x2; ++i) Cc(i, j, k) = csub(Cc(i, j, k), imx_mul(Om(i, j, k) * h, S(i, j, k)));
}

// _frame_turn! (bands.jl:321-327)
void frame_turn(orc &o, A3c &F, const A3r &Om, double dt) {
  for (int k = o.z1 + 1; k <= o.z2; ++k)
    for (int j = o.y1 + 1; j <= o.y2; ++j)
      for (int i = o.x1 + 1; i <= o.x2; ++i) F(i, j, k) = cmul(F(i, j, k), cis(Om(i, j, k) * dt));
}

// _frame_links! (bands.jl:292-297)
void frame_links(orc &o, A3c &Lx, A3c &Ly, A3c &Lz, const A3r &Om, double tau) {
  forr(rx(o), [&](int i, int j, int k) { Lx(i, j, k) = cmul(Lx(i, j, k), scis((Om(i + 1, j, k) - Om(i, j, k)) * tau)); });
  forr(ry(o), [&](int i, int j, int k) { Ly(i, j, k) = cmul(Ly(i, j, k), scis((Om(i, j + 1, k) - Om(i, j, k)) * tau)); });
  forr(rz(o), [&](int i, int j, int k) { Lz(i, j, k) = cmul(Lz(i, j, k), scis((Om(i, j, k + 1) - Om(i, j, k)) * tau)); });
}

// _link_rate! (bands.jl:303-310)
void link_rate(orc &o, A3r &Rr, const A3c &V, const A3c &U, const A3r &Om, double dt, const Rng &rg, int oi, int oj,
               int ok) {
  forr(rg, [&](int i, int j, int k) {
    C q = cmul(V(i, j, k), conj(U(i, j, k)));
    double sf = -q.im;
    double a = q.re > 0 && std::fabs(sf) <= 0.05 ? -(sf + sf * sf * sf / 6) : angle(q);
    Rr(i, j, k) = a + (Om(i + oi, j + oj, k + ok) - Om(i, j, k)) * dt;
  });
}

// _rotlinks! (bands.jl:311-316)
void rotlinks(A3c &L, const A3c &V, const A3r &Rr, double c, const Rng &rg) {
  forr(rg, [&](int i, int j, int k) { L(i, j, k) = cmul(V(i, j, k), scis(Rr(i, j, k) * c)); });
}

// _rot! (bands.jl:355-369)
void rot(A3c &L, A3c &Rc, const A3c &V, const A3c &U, int n, const Rng &rg) {
  forr(rg, [&](int i, int j, int k) {
    C v = V(i, j, k);
    L(i, j, k) = v;
    C q = cmul(v, conj(U(i, j, k)));
    double sf = -q.im;
    if (q.re > 0 && std::fabs(sf) <= 0.05) {
      double a = (sf + sf * sf * sf / 6) / n;
      double a2 = a * a;
      Rc(i, j, k) = {1 - a2 / 2, -a * (1 - a2 / 6)};
    } else {
      Rc(i, j, k) = cis(angle(q) / n);
    }
  });
}

// _advance_links! (bands.jl:371-376)
void advance_links(orc &o, A3c &Lx, A3c &Ly, A3c &Lz, const A3c &Rx_, const A3c &Ry_, const A3c &Rz_) {
  forr(rx(o), [&](int i, int j, int k) { Lx(i, j, k) = cmul(Lx(i, j, k), Rx_(i, j, k)); });
  forr(ry(o), [&](int i, int j, int k) { Ly(i, j, k) = cmul(Ly(i, j, k), Ry_(i, j, k)); });
  forr(rz(o), [&](int i, int j, int k) { Lz(i, j, k) = cmul(Lz(i, j, k), Rz_(i, j, k)); });
}

// _rotfrac! (bands.jl:409-423)
void rotfrac(A3c &L, const A3c &V, const A3c &U, double c, const Rng &rg) {
  forr(rg, [&](int i, int j, int k) {
    C v = V(i, j, k);
    C q = cmul(v, conj(U(i, j, k)));
    double sf = -q.im;
    if (q.re > 0 && std::fabs(sf) <= 0.05) {
      double a = (sf + sf * sf * sf / 6) * c;
      double a2 = a * a;
      L(i, j, k) = cmul(v, C{1 - a2 / 2, -a * (1 - a2 / 6)});
    } else {
      L(i, j, k) = cmul(v, cis(angle(q) * c));
    }
  });
}

// _combine! (bands.jl:425-430)
void combine(orc &o, A3c &Cc, const A3c &A, double mu, double nu) {
  for (int k = o.z1 + 1; k <= o.z2; ++k)
    for (int j = o.y1 + 1; j <= o.y2; ++j)
      for (int i = o.x1 + 1; i <= o.x2; ++i) Cc(i, j, k) = cadd(rmul(mu, Cc(i, j, k)), rmul(nu, A(i, j, k)));
}

// faces!(Ω) of _frame_rate_smooth! (bands.jl:250-255): the faces copy the interior, x, y, z
void om_faces(orc &o, A3r &Om) {
  for (int k = o.z1 + 1; k <= o.z2; ++k)
    for (int j = o.y1 + 1; j <= o.y2; ++j) {
      Om(o.x1, j, k) = Om(o.x1 + 1, j, k);
      Om(o.x2 + 1, j, k) = Om(o.x2, j, k);
    }
  for (int k = o.z1 + 1; k <= o.z2; ++k)
    for (int i = o.x1; i <= o.x2 + 1; ++i) {
      Om(i, o.y1, k) = Om(i, o.y1 + 1, k);
      Om(i, o.y2 + 1, k) = Om(i, o.y2, k);
    }
  for (int j = o.y1; j <= o.y2 + 1; ++j)
    for (int i = o.x1; i <= o.x2 + 1; ++i) {
      Om(i, j, o.z1) = Om(i, j, o.z1 + 1);
      Om(i, j, o.z2 + 1) = Om(i, j, o.z2);
    }
}

// _frame_rate_smooth! (bands.jl:243-270), with noise: the rate, two low-pass sweeps
// Ω ← Ω/2 + (sum of the 6 neighbours)/12, and the frame term as the rotation C ← C e^{−iω̂h}
void frame_rate_smooth(orc &o, A3r &Om, A3c &Cc, const A3c &Y, double h, double dt, int passes) {
  const double wmax = ROT_MAXDT / dt;
  for (int k = o.z1 + 1; k <= o.z2; ++k)
    for (int j = o.y1 + 1; j <= o.y2; ++j)
      for (int i = o.x1 + 1; i <= o.x2; ++i) {
        C y = Y(i, j, k);
        double a2 = abs2(y);
        Om(i, j, k) = a2 > ROT_MIN2 ? jclamp(cmul(conj(y), csub(Cc(i, j, k), y)).im / (a2 * h), -wmax, wmax) : 0.0;
      }
  om_faces(o, Om);
  for (int p = 0; p < passes; ++p) {
    o.Om2 = Om;
    const A3r &T = o.Om2;
    for (int k = o.z1 + 1; k <= o.z2; ++k)
      for (int j = o.y1 + 1; j <= o.y2; ++j)
        for (int i = o.x1 + 1; i <= o.x2; ++i) {
          double nb = (T(i + 1, j, k) + T(i - 1, j, k)) + (T(i, j + 1, k) + T(i, j - 1, k)) + (T(i, j, k + 1) + T(i, j, k - 1));
          Om(i, j, k) = 0.5 * T(i, j, k) + nb / 12;
        }
    om_faces(o, Om);
  }
  for (int k = o.z1 + 1; k <= o.z2; ++k)
    for (int j = o.y1 + 1; j <= o.y2; ++j)
      for (int i = o.x1 + 1; i <= o.x2; ++i) Cc(i, j, k) = cmul(Cc(i, j, k), cis(-(Om(i, j, k) * h)));
}

// _frame_term!(...; turn = true) (bands.jl:276-281): with noise the frame term is a rotation
void frame_term_turn(orc &o, A3c &Cc, const A3r &Om, double h) {
  for (int k = o.z1 + 1; k <= o.z2; ++k)
    for (int j = o.y1 + 1; j <= o.y2; ++j)
      for (int i = o.x1 + 1; i <= o.x2; ++i) Cc(i, j, k) = cmul(Cc(i, j, k), cis(-(Om(i, j, k) * h)));
}

// complex quotient, log and exp (z/w by the textbook formula; log z = log|z| + i arg z with
// |z| = hypot; e^z = e^{Re z} e^{i Im z}): the device computes them alike
inline C cdiv(C a, C b) {
  double den = b.re * b.re + b.im * b.im;
  return {(a.re * b.re + a.im * b.im) / den, (a.im * b.re - a.re * b.im) / den};
}
inline C clog(C z) { return {std::log(std::hypot(z.re, z.im)), std::atan2(z.im, z.re)}; }
inline C cexp(C z) { return rmul(std::exp(z.re), cis(z.im)); }

// _combine_log! (bands.jl:439-445): the stages of a KWT band with noise combined in log-polar form
void combine_log(orc &o, A3c &Cc, const A3c &A, double mu) {
  for (int k = o.z1 + 1; k <= o.z2; ++k)
    for (int j = o.y1 + 1; j <= o.y2; ++j)
      for (int i = o.x1 + 1; i <= o.x2; ++i) {
        C a = A(i, j, k), c = Cc(i, j, k);
        Cc(i, j, k) = abs2(a) > ROT_MIN2 && abs2(c) > ROT_MIN2 ? cmul(a, cexp(rmul(mu, clog(cdiv(c, a)))))
                                                              : cadd(rmul(mu, c), rmul(1 - mu, a));
      }
}

// _sk_turn! (bands.jl:447-453): D ← X e^{i c δθ} on the interior
void sk_turn(orc &o, A3c &D, const A3c &X, const A3r &Qt, double c) {
  for (int k = o.z1 + 1; k <= o.z2; ++k)
    for (int j = o.y1 + 1; j <= o.y2; ++j)
      for (int i = o.x1 + 1; i <= o.x2; ++i) {
        double q = Qt(i, j, k);
        D(i, j, k) = q == 0.0 ? X(i, j, k) : cmul(X(i, j, k), cis(c * q));
      }
}

// ghosts! (bands.jl:524-539)
void ghosts(orc &o, A3c &X, const A3c &Vx, const A3c &Vy, const A3c &Vz) {
  for (int k = o.z1; k <= o.z2 + 1; ++k)
    for (int j = o.y1; j <= o.y2 + 1; ++j) {
      X(o.x1, j, k) = cmul(Vx(o.x1, j, k), X(o.x1 + 1, j, k));
      X(o.x2 + 1, j, k) = cmul(conj(Vx(o.x2, j, k)), X(o.x2, j, k));
    }
  for (int k = o.z1; k <= o.z2 + 1; ++k)
    for (int i = o.x1; i <= o.x2 + 1; ++i) {
      X(i, o.y1, k) = cmul(Vy(i, o.y1, k), X(i, o.y1 + 1, k));
      X(i, o.y2 + 1, k) = cmul(conj(Vy(i, o.y2, k)), X(i, o.y2, k));
    }
  for (int j = o.y1; j <= o.y2 + 1; ++j)
    for (int i = o.x1; i <= o.x2 + 1; ++i) {
      X(i, j, o.z1) = cmul(Vz(i, j, o.z1), X(i, j, o.z1 + 1));
      X(i, j, o.z2 + 1) = cmul(conj(Vz(i, j, o.z2)), X(i, j, o.z2));
    }
}

// _band_substeps! (bands.jl:378-406)
void band_substeps(orc &o, int mu, int n) {
  const tdgl_bandcoef &cf = o.d.bcs[mu];
  A3c &F = o.F[mu];
  A3c &W = o.W;
  A3c *Lx, *Ly, *Lz;
  if (o.d.extrap) {
    Lx = &o.L[0];
    Ly = &o.L[1];
    Lz = &o.L[2];
    rot(o.L[0], o.L[3], o.Vx, o.Ux, n, rx(o));
    rot(o.L[1], o.L[4], o.Vy, o.Uy, n, ry(o));
    rot(o.L[2], o.L[5], o.Vz, o.Uz, n, rz(o));
  } else {
    Lx = &o.Vx;
    Ly = &o.Vy;
    Lz = &o.Vz;
  }
  const A3c *src = &o.G[mu];
  bool rt = o.d.rotframe;
  double h = o.d.dt / n;
  for (int m = 1; m <= n; ++m) {
    A3c *dst = ((n - m) % 2 == 0) ? &F : &W;
    band_pass(o, *dst, *src, mu, *Lx, *Ly, *Lz, cf);
    if (rt) {
      if (m == 1) {
        if (o.noise)
          frame_rate_smooth(o, o.Om, *dst, *src, h, o.d.dt, 2);
        else
          frame_rate(o, o.Om, *dst, *src, h, o.d.dt);
      } else if (o.noise) {
        frame_term_turn(o, *dst, o.Om, h);
      } else {
        frame_term(o, *dst, *src, o.Om, h);
      }
    }
    if (m < n) {
      if (o.d.extrap) advance_links(o, *Lx, *Ly, *Lz, o.L[3], o.L[4], o.L[5]);
      if (rt) frame_links(o, *Lx, *Ly, *Lz, o.Om, h);
      ghosts(o, *dst, *Lx, *Ly, *Lz);
    }
    src = dst;
  }
  if (rt) frame_turn(o, F, o.Om, o.d.dt);
}

// _band_rkl! (bands.jl:464-521)
void band_rkl(orc &o, int mu, int ns) {
  const tdgl_bandcoef &cf = o.d.bcs[mu];
  bool kc = o.d.rkc[mu] != 0;
  A3c *Lx, *Ly, *Lz;
  if (o.d.extrap) {
    o.L[0] = o.Vx;
    o.L[1] = o.Vy;
    o.L[2] = o.Vz;
    Lx = &o.L[0];
    Ly = &o.L[1];
    Lz = &o.L[2];
  } else {
    Lx = &o.Vx;
    Ly = &o.Vy;
    Lz = &o.Vz;
  }
  A3c *buf[3] = {&o.W, &o.W2, &o.F[mu]};
  bool rt = o.d.rotframe;
  // h = dt/q, q = ns(ns+1)/2 (RKL1) or RKCCoef.q (bands.jl:468, 477)
  double qd = kc ? o.d.divisor[mu] : (double)(ns * (ns + 1) / 2);
  double h = o.d.dt / qd;
  const A3c *A = &o.G[mu];
  const A3c *B = &o.G[mu];
  A3c *Cc = buf[0];
  const bool lg = o.noise && o.d.bcs[mu].gamma2 > 0;  // _logcomb (bands.jl:455)
  const bool skq = o.noise && o.d.noise_sk[mu];       // _skrock (bands.jl:446)
  for (int j = 1; j <= ns; ++j) {
    Cc = buf[(j - 1) % 3];
    if (j == 1 && skq) {
      // SK-ROCK's first stage: K₁ = E(Ψⁿ e^{iν₁δθ}) e^{i(κ₁−ν₁)δθ} (bands.jl:482-487)
      const double nu1 = o.d.noise_sk_w[mu][0], ka1 = o.d.noise_sk_w[mu][1];
      sk_turn(o, o.Sp, *B, o.Qth[mu], nu1);
      ghosts(o, o.Sp, *Lx, *Ly, *Lz);
      band_pass(o, *Cc, o.Sp, mu, *Lx, *Ly, *Lz, cf);
      if (rt) frame_rate_smooth(o, o.Om, *Cc, o.Sp, h, o.d.dt, 2);
      sk_turn(o, *Cc, *Cc, o.Qth[mu], ka1 - nu1);
    } else {
      band_pass(o, *Cc, *B, mu, *Lx, *Ly, *Lz, cf);
      if (rt) {
        if (j == 1) {
          if (o.noise)
            frame_rate_smooth(o, o.Om, *Cc, *B, h, o.d.dt, 2);
          else
            frame_rate(o, o.Om, *Cc, *B, h, o.d.dt);
        } else if (o.noise) {
          frame_term_turn(o, *Cc, o.Om, h);
        } else {
          frame_term(o, *Cc, *B, o.Om, h);
        }
      }
    }
    if (rt && j == 1 && o.d.extrap && ns > 1) {
      link_rate(o, o.R[0], o.Vx, o.Ux, o.Om, o.d.dt, rx(o), 1, 0, 0);
      link_rate(o, o.R[1], o.Vy, o.Uy, o.Om, o.d.dt, ry(o), 0, 1, 0);
      link_rate(o, o.R[2], o.Vz, o.Uz, o.Om, o.d.dt, rz(o), 0, 0, 1);
    }
    if (j >= 2) {
      double mj, nj;
      if (kc) {
        mj = o.d.rkc_mu[mu][j - 1];
        nj = o.d.rkc_nu[mu][j - 1];
      } else {
        mj = (2.0 * j - 1) / j;
        nj = (1.0 - j) / j;
      }
      if (lg)
        combine_log(o, *Cc, *A, mj);
      else
        combine(o, *Cc, *A, mj, nj);
    }
    if (j < ns) {
      if (o.d.extrap) {
        double c = kc ? o.d.rkc_c[mu][j - 1] : (double)(j * (j + 1)) / (double)(ns * (ns + 1));
        if (rt) {
          rotlinks(*Lx, o.Vx, o.R[0], c, rx(o));
          rotlinks(*Ly, o.Vy, o.R[1], c, ry(o));
          rotlinks(*Lz, o.Vz, o.R[2], c, rz(o));
        } else {
          rotfrac(*Lx, o.Vx, o.Ux, c, rx(o));
          rotfrac(*Ly, o.Vy, o.Uy, c, ry(o));
          rotfrac(*Lz, o.Vz, o.Uz, c, rz(o));
        }
      }
      ghosts(o, *Cc, *Lx, *Ly, *Lz);
    }
    A = B;
    B = Cc;
  }
  if (Cc != &o.F[mu]) o.F[mu] = *Cc;
  if (rt) frame_turn(o, o.F[mu], o.Om, o.d.dt);
}

// update_bands! (bands.jl:330-346)
void update_bands(orc &o) {
  for (int mu = 0; mu < o.N; ++mu) {
    int n = o.d.nsub[mu];
    if (o.d.stages[mu] > 0) {
      band_rkl(o, mu, o.d.stages[mu]);
    } else if (n == 1) {
      band_pass(o, o.F[mu], o.G[mu], mu, o.Vx, o.Vy, o.Vz, o.d.bc[mu]);
      if (o.d.rotframe) {
        if (o.noise)
          frame_rate_smooth(o, o.Om, o.F[mu], o.G[mu], o.d.dt, o.d.dt, 2);
        else
          frame_rate(o, o.Om, o.F[mu], o.G[mu], o.d.dt, o.d.dt);
        frame_turn(o, o.F[mu], o.Om, o.d.dt);
      }
    } else {
      band_substeps(o, mu, n);
    }
  }
}

// supercurrent! (field.jl:19-46), one material, no drag
inline double jlink(const orc &o, int ax, C v, int i, int j, int k, int i2, int j2, int k2) {
  double acc = o.d.wa[ax][0] * cmul(cmul(conj(o.G[0](i, j, k)), v), o.G[0](i2, j2, k2)).im;
  for (int mu = 1; mu < o.N; ++mu) acc += o.d.wa[ax][mu] * cmul(cmul(conj(o.G[mu](i, j, k)), v), o.G[mu](i2, j2, k2)).im;
  return acc;
}
void supercurrent(orc &o) {
  const int Nx = o.Nx, Ny = o.Ny, Nz = o.Nz;
#pragma omp parallel for schedule(static)
  for (int k = 2; k <= Nz; ++k) {
    for (int j = 2; j <= Ny; ++j)
      for (int i = 1; i <= Nx; ++i) o.Jx(i, j, k) = jlink(o, 0, o.Vx(i, j, k), i, j, k, i + 1, j, k);
    for (int j = 1; j <= Ny; ++j)
      for (int i = 2; i <= Nx; ++i) o.Jy(i, j, k) = jlink(o, 1, o.Vy(i, j, k), i, j, k, i, j + 1, k);
  }
#pragma omp parallel for schedule(static)
  for (int k = 1; k <= Nz; ++k)
    for (int j = 2; j <= Ny; ++j)
      for (int i = 2; i <= Nx; ++i) o.Jz(i, j, k) = jlink(o, 2, o.Vz(i, j, k), i, j, k, i, j, k + 1);
}

// field_increment! (field.jl:84-102)
void field_increment(orc &o) {
  const int Nx = o.Nx, Ny = o.Ny, Nz = o.Nz;
  const tdgl_desc &d = o.d;
#pragma omp parallel for schedule(static)
  for (int k = 2; k <= Nz; ++k)
    for (int j = 2; j <= Ny; ++j)
      for (int i = 1; i <= Nx; ++i)
        o.phx(i, j, k) = ((o.thx(i, j, k) + (d.d1 * o.tx(i, j, k)) * o.Jx(i, j, k)) - d.sxy * (o.hz(i, j, k) - o.hz(i, j - 1, k))) +
                         d.sxz * (o.hy(i, j, k) - o.hy(i, j, k - 1));
#pragma omp parallel for schedule(static)
  for (int k = 2; k <= Nz; ++k)
    for (int j = 1; j <= Ny; ++j)
      for (int i = 2; i <= Nx; ++i)
        o.phy(i, j, k) = ((o.thy(i, j, k) + (d.d1 * o.ty(i, j, k)) * o.Jy(i, j, k)) - d.syz * (o.hx(i, j, k) - o.hx(i, j, k - 1))) +
                         d.syx * (o.hz(i, j, k) - o.hz(i - 1, j, k));
#pragma omp parallel for schedule(static)
  for (int k = 1; k <= Nz; ++k)
    for (int j = 2; j <= Ny; ++j)
      for (int i = 2; i <= Nx; ++i)
        o.phz(i, j, k) = ((o.thz(i, j, k) + (d.d1 * o.tz(i, j, k)) * o.Jz(i, j, k)) - d.szx * (o.hy(i, j, k) - o.hy(i - 1, j, k))) +
                         d.szy * (o.hx(i, j, k) - o.hx(i, j - 1, k));
}

// gather / scatter of the ranges of the separable solve: rx = (2:Nx-1, 2:Ny, 2:Nz), ...
void gather_diff(std::vector<double> &B, const A3r &ph, const A3r &th, int i0, int i1, int j0, int j1, int k0, int k1) {
  size_t p = 0;
  B.resize((size_t)(i1 - i0 + 1) * (j1 - j0 + 1) * (k1 - k0 + 1));
  for (int k = k0; k <= k1; ++k)
    for (int j = j0; j <= j1; ++j)
      for (int i = i0; i <= i1; ++i) B[p++] = ph(i, j, k) - th(i, j, k);
}
void scatter_add(A3r &ph, const A3r &th, const std::vector<double> &B, int i0, int i1, int j0, int j1, int k0, int k1) {
  size_t p = 0;
  for (int k = k0; k <= k1; ++k)
    for (int j = j0; j <= j1; ++j)
      for (int i = i0; i <= i1; ++i) ph(i, j, k) = th(i, j, k) + B[p++];
}

// implicit_curl!(s, ::SepSolver) (implicit.jl:91-104)
void implicit_sep(orc &o) {
  const int Nx = o.Nx, Ny = o.Ny, Nz = o.Nz;
  gather_diff(o.BX, o.phx, o.thx, 2, Nx - 1, 2, Ny, 2, Nz);
  gather_diff(o.BY, o.phy, o.thy, 2, Nx, 2, Ny - 1, 2, Nz);
  gather_diff(o.BZ, o.phz, o.thz, 2, Nx, 2, Ny, 2, Nz - 1);
  sep_solve(o.sep, o.BX, o.BY, o.BZ);
  scatter_add(o.phx, o.thx, o.BX, 2, Nx - 1, 2, Ny, 2, Nz);
  scatter_add(o.phy, o.thy, o.BY, 2, Nx, 2, Ny - 1, 2, Nz);
  scatter_add(o.phz, o.thz, o.BZ, 2, Nx, 2, Ny, 2, Nz - 1);
}

// _apply_A! (implicit.jl:212-236)
void apply_A(orc &o) {
  const int Nx = o.Nx, Ny = o.Ny, Nz = o.Nz;
  const tdgl_desc &d = o.d;
  A3r &px = o.P[0], &py = o.P[1], &pz = o.P[2];
  A3r &qx = o.Q[0], &qy = o.Q[1], &qz = o.Q[2];
  A3r &hx = o.H[0], &hy = o.H[1], &hz = o.H[2];
  for (int k = 2; k <= Nz - 1; ++k)
    for (int j = 2; j <= Ny - 1; ++j)
      for (int i = 2; i <= Nx; ++i)
        hx(i, j, k) = ((pz(i, j + 1, k) - pz(i, j, k)) - py(i, j, k + 1) + py(i, j, k)) / (d.dy * d.dz);
  for (int k = 2; k <= Nz - 1; ++k)
    for (int j = 2; j <= Ny; ++j)
      for (int i = 2; i <= Nx - 1; ++i)
        hy(i, j, k) = ((px(i, j, k + 1) - px(i, j, k)) - pz(i + 1, j, k) + pz(i, j, k)) / (d.dz * d.dx);
  for (int k = 2; k <= Nz; ++k)
    for (int j = 2; j <= Ny - 1; ++j)
      for (int i = 2; i <= Nx - 1; ++i)
        hz(i, j, k) = ((py(i + 1, j, k) - py(i, j, k)) - px(i, j + 1, k) + px(i, j, k)) / (d.dx * d.dy);
  for (int k = 2; k <= Nz; ++k)
    for (int j = 2; j <= Ny; ++j)
      for (int i = 2; i <= Nx - 1; ++i)
        qx(i, j, k) = o.Dx(i, j, k) * px(i, j, k) + d.sxy * (hz(i, j, k) - hz(i, j - 1, k)) - d.sxz * (hy(i, j, k) - hy(i, j, k - 1));
  for (int k = 2; k <= Nz; ++k)
    for (int j = 2; j <= Ny - 1; ++j)
      for (int i = 2; i <= Nx; ++i)
        qy(i, j, k) = o.Dy(i, j, k) * py(i, j, k) + d.syz * (hx(i, j, k) - hx(i, j, k - 1)) - d.syx * (hz(i, j, k) - hz(i - 1, j, k));
  for (int k = 2; k <= Nz - 1; ++k)
    for (int j = 2; j <= Ny; ++j)
      for (int i = 2; i <= Nx; ++i)
        qz(i, j, k) = o.Dz(i, j, k) * pz(i, j, k) + d.szx * (hy(i, j, k) - hy(i - 1, j, k)) - d.szy * (hx(i, j, k) - hx(i, j - 1, k));
}

struct R3 {
  int i0, i1, j0, j1, k0, k1;
};
R3 prange(const orc &o, int c) {
  if (c == 0) return {2, o.Nx - 1, 2, o.Ny, 2, o.Nz};
  if (c == 1) return {2, o.Nx, 2, o.Ny - 1, 2, o.Nz};
  return {2, o.Nx, 2, o.Ny, 2, o.Nz - 1};
}
void scatter_full(A3r &full, const std::vector<double> &v, const R3 &r) {
  std::fill(full.v.begin(), full.v.end(), 0.0);
  size_t p = 0;
  for (int k = r.k0; k <= r.k1; ++k)
    for (int j = r.j0; j <= r.j1; ++j)
      for (int i = r.i0; i <= r.i1; ++i) full(i, j, k) = v[p++];
}
template <class F>
void over_range(const R3 &r, F fn) {
  size_t p = 0;
  for (int k = r.k0; k <= r.k1; ++k)
    for (int j = r.j0; j <= r.j1; ++j)
      for (int i = r.i0; i <= r.i1; ++i) fn(p++, i, j, k);
}
double dot3(const std::vector<double> *a, const std::vector<double> *b) { return dot(a[0], b[0]) + dot(a[1], b[1]) + dot(a[2], b[2]); }

void precond(orc &o, std::vector<double> *zz, const std::vector<double> *rr, bool f32) {
  if (f32) {
    sep_solve32(o.sep, zz, rr);
    return;
  }
  o.BX = rr[0];
  o.BY = rr[1];
  o.BZ = rr[2];
  sep_solve(o.sep, o.BX, o.BY, o.BZ);
  zz[0] = o.BX;
  zz[1] = o.BY;
  zz[2] = o.BZ;
}

// _pcg_zero_update! (implicit.jl:251-271)
void pcg_zero_update(orc &o) {
  if (o.d.pcg_fast || o.d.pcg_f32) {
    if (o.warm && (o.d.pcg_extrap && o.d.pcg_fast)) {
      for (int c = 0; c < 3; ++c) o.xprev[c] = o.x[c];
      o.have2 = true;
    } else {
      o.have2 = false;
    }
  }
  for (int c = 0; c < 3; ++c) std::fill(o.x[c].begin(), o.x[c].end(), 0.0);
  o.warm = true;
  o.iters = 0;
  const int Nx = o.Nx, Ny = o.Ny, Nz = o.Nz;
  for (int k = 2; k <= Nz; ++k)
    for (int j = 2; j <= Ny; ++j)
      for (int i = 2; i <= Nx - 1; ++i) o.phx(i, j, k) = o.thx(i, j, k);
  for (int k = 2; k <= Nz; ++k)
    for (int j = 2; j <= Ny - 1; ++j)
      for (int i = 2; i <= Nx; ++i) o.phy(i, j, k) = o.thy(i, j, k);
  for (int k = 2; k <= Nz - 1; ++k)
    for (int j = 2; j <= Ny; ++j)
      for (int i = 2; i <= Nx; ++i) o.phz(i, j, k) = o.thz(i, j, k);
}

// implicit_curl!(s, ::PCGSolver) and _implicit_curl_fast! (implicit.jl:279-334, 376-446)
void implicit_pcg(orc &o) {
  const tdgl_desc &d = o.d;
  const bool fast = d.pcg_fast || d.pcg_f32;  // PCGFast present
  const bool f32 = d.pcg_f32 != 0;
  const bool extrap = d.pcg_extrap && d.pcg_fast;
  R3 rg[3] = {prange(o, 0), prange(o, 1), prange(o, 2)};
  A3r *ph[3] = {&o.phx, &o.phy, &o.phz};
  A3r *th[3] = {&o.thx, &o.thy, &o.thz};
  A3r *Dd[3] = {&o.Dx, &o.Dy, &o.Dz};
  for (int c = 0; c < 3; ++c) {
    o.b[c].resize((size_t)(rg[c].i1 - rg[c].i0 + 1) * (rg[c].j1 - rg[c].j0 + 1) * (rg[c].k1 - rg[c].k0 + 1));
    over_range(rg[c], [&](size_t p, int i, int j, int k) { o.b[c][p] = (*ph[c])(i, j, k) - (*th[c])(i, j, k); });
  }
  for (int c = 0; c < 3; ++c)
    for (size_t i = 0; i < ph[c]->v.size(); ++i) ph[c]->v[i] = th[c]->v[i] + (ph[c]->v[i] - th[c]->v[i]) / Dd[c]->v[i];
  double nb = std::sqrt(dot3(o.b, o.b));
  if (nb == 0) {
    pcg_zero_update(o);
    return;
  }
  if (!fast) {
    if (!o.warm) precond(o, o.x, o.b, false);
  } else {
    if (!o.warm) {
      precond(o, o.x, o.b, false);
      o.have2 = false;
    } else if (!extrap) {
    } else if (o.have2) {
      for (int c = 0; c < 3; ++c)
        for (size_t i = 0; i < o.x[c].size(); ++i) {
          double v = o.x[c][i];
          o.x[c][i] = 2 * v - o.xprev[c][i];
          o.xprev[c][i] = v;
        }
    } else {
      for (int c = 0; c < 3; ++c) o.xprev[c] = o.x[c];
      o.have2 = true;
    }
  }
  o.warm = true;
  for (int c = 0; c < 3; ++c) scatter_full(o.P[c], o.x[c], rg[c]);
  apply_A(o);
  for (int c = 0; c < 3; ++c) {
    o.r[c].resize(o.b[c].size());
    over_range(rg[c], [&](size_t p, int i, int j, int k) { o.r[c][p] = o.b[c][p] - o.Q[c](i, j, k); });
  }
  int it = 0;
  if (nb > 0 && std::sqrt(dot3(o.r, o.r)) > d.pcg_tol * nb) {
    for (int c = 0; c < 3; ++c) o.z[c].resize(o.r[c].size());
    precond(o, o.z, o.r, fast && f32);
    for (int c = 0; c < 3; ++c) scatter_full(o.P[c], o.z[c], rg[c]);
    if (fast && f32)
      for (int c = 0; c < 3; ++c) o.zold[c] = o.z[c];
    double rz = dot3(o.r, o.z);
    while (it < d.pcg_maxit) {
      ++it;
      apply_A(o);
      double pAp = 0.0;
      for (int c = 0; c < 3; ++c) {
        double s = 0;
        over_range(rg[c], [&](size_t, int i, int j, int k) { s += o.P[c](i, j, k) * o.Q[c](i, j, k); });
        pAp += s;
      }
      double alpha = rz / pAp;
      for (int c = 0; c < 3; ++c)
        over_range(rg[c], [&](size_t p, int i, int j, int k) {
          o.x[c][p] += alpha * o.P[c](i, j, k);
          o.r[c][p] -= alpha * o.Q[c](i, j, k);
        });
      if (std::sqrt(dot3(o.r, o.r)) <= d.pcg_tol * nb) break;
      precond(o, o.z, o.r, fast && f32);
      double rz2 = dot3(o.r, o.z);
      double beta;
      if (fast && f32) {
        beta = (rz2 - dot3(o.r, o.zold)) / rz;
        for (int c = 0; c < 3; ++c) o.zold[c] = o.z[c];
      } else {
        beta = rz2 / rz;
      }
      rz = rz2;
      for (int c = 0; c < 3; ++c)
        over_range(rg[c], [&](size_t p, int i, int j, int k) { o.P[c](i, j, k) = o.z[c][p] + beta * o.P[c](i, j, k); });
    }
  }
  o.iters = it;
  o.total_iters += it;
  for (int c = 0; c < 3; ++c)
    over_range(rg[c], [&](size_t p, int i, int j, int k) { (*ph[c])(i, j, k) = (*th[c])(i, j, k) + o.x[c][p]; });
}

// scale_increment! (implicit.jl:337-342)
void scale_increment(orc &o) {
  A3r *ph[3] = {&o.phx, &o.phy, &o.phz};
  A3r *th[3] = {&o.thx, &o.thy, &o.thz};
  A3r *Dd[3] = {&o.Dx, &o.Dy, &o.Dz};
  for (int c = 0; c < 3; ++c)
    for (size_t i = 0; i < ph[c]->v.size(); ++i) ph[c]->v[i] = th[c]->v[i] + (ph[c]->v[i] - th[c]->v[i]) / Dd[c]->v[i];
}

// links! (field.jl:105-121)
void links(orc &o) {
  const int Nx = o.Nx, Ny = o.Ny, Nz = o.Nz;
#pragma omp parallel for schedule(static)
  for (int k = 1; k <= Nz; ++k) {
    if (2 <= k && k <= Nz) {
      for (int j = 2; j <= Ny; ++j)
        for (int i = 1; i <= Nx; ++i) o.Ux(i, j, k) = cis(-o.phx(i, j, k));
      for (int j = 1; j <= Ny; ++j)
        for (int i = 2; i <= Nx; ++i) o.Uy(i, j, k) = cis(-o.phy(i, j, k));
    }
    for (int j = 2; j <= Ny; ++j)
      for (int i = 2; i <= Nx; ++i) o.Uz(i, j, k) = cis(-o.phz(i, j, k));
  }
}

// curl_h! (field.jl:124-137)
void curl_h(orc &o) {
  const int Nx = o.Nx, Ny = o.Ny, Nz = o.Nz;
  const double dx = o.d.dx, dy = o.d.dy, dz = o.d.dz;
  for (int k = 2; k <= Nz - 1; ++k)
    for (int j = 2; j <= Ny - 1; ++j)
      for (int i = 2; i <= Nx; ++i)
        o.hx(i, j, k) = ((o.phz(i, j + 1, k) - o.phz(i, j, k)) - o.phy(i, j, k + 1) + o.phy(i, j, k)) / (dy * dz);
  for (int k = 2; k <= Nz - 1; ++k)
    for (int j = 2; j <= Ny; ++j)
      for (int i = 2; i <= Nx - 1; ++i)
        o.hy(i, j, k) = ((o.phx(i, j, k + 1) - o.phx(i, j, k)) - o.phz(i + 1, j, k) + o.phz(i, j, k)) / (dz * dx);
  for (int k = 2; k <= Nz; ++k)
    for (int j = 2; j <= Ny - 1; ++j)
      for (int i = 2; i <= Nx - 1; ++i)
        o.hz(i, j, k) = ((o.phy(i + 1, j, k) - o.phy(i, j, k)) - o.phx(i, j + 1, k) + o.phx(i, j, k)) / (dx * dy);
}

// field_bc! (field.jl:189-228), no tilt, no open boundary
void open_add(orc &o);

void field_bc(orc &o, double H, double HJ) {
  const int Nx = o.Nx, Ny = o.Ny, Nz = o.Nz;
  A3r &hx = o.hx, &hy = o.hy, &hz = o.hz;
  for (int j = 1; j <= Ny; ++j)
    for (int i = 1; i <= Nx; ++i) hz(i, j, 1) = hz(i, j, Nz + 1) = H;
  for (int k = 1; k <= Nz + 1; ++k)
    for (int j = 1; j <= Ny; ++j) hz(1, j, k) = hz(Nx, j, k) = H;
  for (int k = 1; k <= Nz + 1; ++k)
    for (int i = 1; i <= Nx; ++i) hz(i, 1, k) = hz(i, Ny, k) = H;
  for (int k = 1; k <= Nz; ++k)
    for (int j = 1; j <= Ny; ++j) hx(1, j, k) = hx(Nx + 1, j, k) = 0.0;
  for (int k = 1; k <= Nz; ++k)
    for (int i = 1; i <= Nx + 1; ++i) hx(i, 1, k) = hx(i, Ny, k) = 0.0;
  for (int k = 1; k <= Nz; ++k)
    for (int i = 1; i <= Nx; ++i) hy(i, 1, k) = hy(i, Ny + 1, k) = 0.0;
  if (o.d.selffield_bc == TDGL_BC_SLAB) {
    for (int j = 1; j <= Ny + 1; ++j)
      for (int i = 1; i <= Nx; ++i) {
        hy(i, j, 1) = -HJ;
        hy(i, j, Nz) = +HJ;
      }
    return;
  }
  const std::vector<double> &hy_z = o.bar[0], &hy_y = o.bar[1], &hz_z = o.bar[2], &hz_y = o.bar[3], &hz_x = o.bar[4],
                            &hy_x = o.bar[5];
  // hy_z (Ny+1, 2)
  for (int c = 0; c < 2; ++c) {
    int k = c == 0 ? 1 : Nz;
    for (int j = 1; j <= Ny + 1; ++j)
      for (int i = 1; i <= Nx; ++i) hy(i, j, k) = HJ * hy_z[(j - 1) + (size_t)(Ny + 1) * c];
  }
  // hy_y (2, Nz)
  for (int c = 0; c < 2; ++c) {
    int j = c == 0 ? 1 : Ny + 1;
    for (int k = 1; k <= Nz; ++k)
      for (int i = 1; i <= Nx; ++i) hy(i, j, k) = HJ * hy_y[c + 2 * (size_t)(k - 1)];
  }
  // hz_z (Ny, 2)
  for (int c = 0; c < 2; ++c) {
    int k = c == 0 ? 1 : Nz + 1;
    for (int j = 1; j <= Ny; ++j)
      for (int i = 1; i <= Nx; ++i) hz(i, j, k) = H + HJ * hz_z[(j - 1) + (size_t)Ny * c];
  }
  // hz_y (2, Nz+1)
  for (int c = 0; c < 2; ++c) {
    int j = c == 0 ? 1 : Ny;
    for (int k = 1; k <= Nz + 1; ++k)
      for (int i = 1; i <= Nx; ++i) hz(i, j, k) = H + HJ * hz_y[c + 2 * (size_t)(k - 1)];
  }
  // hz_x (Ny, Nz+1)
  for (int c = 0; c < 2; ++c) {
    int i = c == 0 ? 1 : Nx;
    for (int k = 1; k <= Nz + 1; ++k)
      for (int j = 1; j <= Ny; ++j) hz(i, j, k) = H + HJ * hz_x[(j - 1) + (size_t)Ny * (k - 1)];
  }
  // hy_x (Ny+1, Nz)
  for (int c = 0; c < 2; ++c) {
    int i = c == 0 ? 1 : Nx;
    for (int k = 1; k <= Nz; ++k)
      for (int j = 1; j <= Ny + 1; ++j) hy(i, j, k) = HJ * hy_x[(j - 1) + (size_t)(Ny + 1) * (k - 1)];
  }
  if (o.d.selffield_bc == TDGL_BC_OPEN) open_add(o);
}

// ---------------------------------------------------------------- open boundary (openbc.jl)
// G(i, j, k) of −Δ_h: the table within R, 1/4πr with the cubic correction beyond (_lgf, :101-102)
double lgf(const orc &o, int i, int j, int k) {
  i = std::abs(i);
  j = std::abs(j);
  k = std::abs(k);
  const int R = o.d.open_R;
  if (i <= R && j <= R && k <= R) return o.d.open_lgf[i + (size_t)(R + 1) * (j + (size_t)(R + 1) * k)];
  const double r2 = (double)(i * i + j * j + k * k), r = std::sqrt(r2);
  const double i4 = (double)i * i * i * i, j4 = (double)j * j * j * j, k4 = (double)k * k * k * k;
  return (1 + (5 * (i4 + j4 + k4) / (r2 * r2) - 3) / (8 * r2)) / (4 * M_PI * r);
}

// δJ = κ² ∇_h × h − HJ·J_bar on the links of the box, the face-normal links left out
// (_delta_current!, openbc.jl:143-160)
void open_delta_current(orc &o, double HJ) {
  const int Nx = o.Nx, Ny = o.Ny, Nz = o.Nz;
  const double k2 = o.d.open_kappa2, dx = o.d.dx, dy = o.d.dy, dz = o.d.dz;
  A3r &bx = o.odJ[0], &by = o.odJ[1], &bz = o.odJ[2];
  std::fill(bx.v.begin(), bx.v.end(), 0.0);
  std::fill(by.v.begin(), by.v.end(), 0.0);
  std::fill(bz.v.begin(), bz.v.end(), 0.0);
  const A3r &hx = o.hx, &hy = o.hy, &hz = o.hz;
  for (int k = 2; k <= Nz; ++k)
    for (int j = 2; j <= Ny; ++j)
      for (int i = 2; i <= Nx - 1; ++i)
        bx(i, j, k) = k2 * ((hz(i, j, k) - hz(i, j - 1, k)) / dy - (hy(i, j, k) - hy(i, j, k - 1)) / dz) -
                      HJ * o.d.open_Jbar[(j - 1) + (size_t)(Ny + 1) * (k - 1)];
  for (int k = 2; k <= Nz; ++k)
    for (int j = 2; j <= Ny - 1; ++j)
      for (int i = 2; i <= Nx; ++i) by(i, j, k) = k2 * ((hx(i, j, k) - hx(i, j, k - 1)) / dz - (hz(i, j, k) - hz(i - 1, j, k)) / dx);
  for (int k = 2; k <= Nz - 1; ++k)
    for (int j = 2; j <= Ny; ++j)
      for (int i = 2; i <= Nx; ++i) bz(i, j, k) = k2 * ((hy(i, j, k) - hy(i - 1, j, k)) / dx - (hx(i, j, k) - hx(i, j - 1, k)) / dy);
}

// θ_δ = (dx³/κ²) G ∗ δJ on every link of the box (_convolve!, openbc.jl:163-172): the linear
// convolution that the reference computes by zero padding and FFT, here as the direct sum
void open_convolve(orc &o) {
  for (int c = 0; c < 3; ++c) {
    const A3r &b = o.odJ[c];
    A3r &t = o.oth[c];
    std::vector<int> si, sj, sk;
    std::vector<double> sv;
    for (int k = 1; k <= b.n3; ++k)
      for (int j = 1; j <= b.n2; ++j)
        for (int i = 1; i <= b.n1; ++i)
          if (b(i, j, k) != 0.0) {
            si.push_back(i);
            sj.push_back(j);
            sk.push_back(k);
            sv.push_back(b(i, j, k));
          }
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 1; k <= t.n3; ++k)
      for (int j = 1; j <= t.n2; ++j)
        for (int i = 1; i <= t.n1; ++i) {
          double acc = 0.0;
          for (size_t q = 0; q < sv.size(); ++q) acc += (o.d.open_fac * lgf(o, i - si[q], j - sj[q], k - sk[q])) * sv[q];
          t(i, j, k) = acc;
        }
  }
}

// h_δ = ∇_h × θ_δ / area on the tangential plaquettes of the six faces (_face_field!, :175-188)
void open_face_field(orc &o, A3r *h) {
  const int Nx = o.Nx, Ny = o.Ny, Nz = o.Nz;
  const double dx = o.d.dx, dy = o.d.dy, dz = o.d.dz;
  const A3r &px = o.oth[0], &py = o.oth[1], &pz = o.oth[2];
  auto hzv = [&](int i, int j, int k) { return ((py(i + 1, j, k) - py(i, j, k)) - px(i, j + 1, k) + px(i, j, k)) / (dx * dy); };
  auto hxv = [&](int i, int j, int k) { return ((pz(i, j + 1, k) - pz(i, j, k)) - py(i, j, k + 1) + py(i, j, k)) / (dy * dz); };
  auto hyv = [&](int i, int j, int k) { return ((px(i, j, k + 1) - px(i, j, k)) - pz(i + 1, j, k) + pz(i, j, k)) / (dz * dx); };
  A3r &hx = h[0], &hy = h[1], &hz = h[2];
  for (int j : {1, Ny})
    for (int k = 1; k <= Nz + 1; ++k)
      for (int i = 1; i <= Nx; ++i) hz(i, j, k) = hzv(i, j, k);
  for (int j : {1, Ny})
    for (int k = 1; k <= Nz; ++k)
      for (int i = 1; i <= Nx + 1; ++i) hx(i, j, k) = hxv(i, j, k);
  for (int k : {1, Nz})
    for (int j = 1; j <= Ny + 1; ++j)
      for (int i = 1; i <= Nx; ++i) hy(i, j, k) = hyv(i, j, k);
  for (int k : {1, Nz})
    for (int j = 1; j <= Ny; ++j)
      for (int i = 1; i <= Nx + 1; ++i) hx(i, j, k) = hxv(i, j, k);
  for (int i : {1, Nx})
    for (int k = 1; k <= Nz + 1; ++k)
      for (int j = 1; j <= Ny; ++j) hz(i, j, k) = hzv(i, j, k);
  for (int i : {1, Nx})
    for (int k = 1; k <= Nz; ++k)
      for (int j = 1; j <= Ny + 1; ++j) hy(i, j, k) = hyv(i, j, k);
}

// open_update! (openbc.jl:196-215): the clock; every `every` steps (and the first time) h_δ anew
void open_update(orc &o, double HJ) {
  const bool first = o.ocount < 0;
  o.ocount += 1;
  if (o.ocount >= o.d.open_every || first) {
    for (int c = 0; c < 3; ++c) o.oold[c] = o.onew[c];
    open_delta_current(o, HJ);
    open_convolve(o);
    open_face_field(o, o.onew);
    if (first)
      for (int c = 0; c < 3; ++c) o.oold[c] = o.onew[c];
    o.ocount = 0;
  }
}

// h_δ at the current point of the ramp, on the six faces after the bar field (_open_add!, :218-231)
void open_add(orc &o) {
  const int Nx = o.Nx, Ny = o.Ny, Nz = o.Nz;
  const double f = std::min((double)(o.ocount + 1) / (double)o.d.open_every, 1.0);
  const A3r *ol = o.oold, *nw = o.onew;
  auto rampv = [&](int c, int i, int j, int k) { return ol[c](i, j, k) + f * (nw[c](i, j, k) - ol[c](i, j, k)); };
  for (int j : {1, Ny})
    for (int k = 1; k <= Nz + 1; ++k)
      for (int i = 2; i <= Nx - 1; ++i) o.hz(i, j, k) += rampv(2, i, j, k);
  for (int j : {1, Ny})
    for (int k = 1; k <= Nz; ++k)
      for (int i = 2; i <= Nx; ++i) o.hx(i, j, k) = rampv(0, i, j, k);
  for (int k : {1, Nz})
    for (int j = 1; j <= Ny + 1; ++j)
      for (int i = 2; i <= Nx - 1; ++i) o.hy(i, j, k) += rampv(1, i, j, k);
  for (int k : {1, Nz})
    for (int j = 1; j <= Ny; ++j)
      for (int i = 2; i <= Nx; ++i) o.hx(i, j, k) = rampv(0, i, j, k);
  for (int i : {1, Nx})
    for (int k = 1; k <= Nz + 1; ++k)
      for (int j = 1; j <= Ny; ++j) o.hz(i, j, k) += rampv(2, i, j, k);
  for (int i : {1, Nx})
    for (int k = 1; k <= Nz; ++k)
      for (int j = 1; j <= Ny + 1; ++j) o.hy(i, j, k) += rampv(1, i, j, k);
}

// neumann! (field.jl:231-245) and advance! (field.jl:248-263)
void neumann(orc &o) {
  for (int mu = 0; mu < o.N; ++mu) ghosts(o, o.F[mu], o.Ux, o.Uy, o.Uz);
}
void advance(orc &o) {
  for (int mu = 0; mu < o.N; ++mu)
    for (int k = o.z1; k <= o.z2 + 1; ++k)
      for (int j = o.y1; j <= o.y2 + 1; ++j)
        for (int i = o.x1; i <= o.x2 + 1; ++i) o.G[mu](i, j, k) = o.F[mu](i, j, k);
  std::swap(o.thx, o.phx);
  std::swap(o.thy, o.phy);
  std::swap(o.thz, o.phz);
  std::swap(o.Vx, o.Ux);
  std::swap(o.Vy, o.Uy);
  std::swap(o.Vz, o.Uz);
}

// ---------------------------------------------------------------- normal fluids (exchange.jl)
// C_µ on a block node: the divergence of band µ's τ-masked link supercurrents plus the
// Josephson exchange with the other bands (_conv, exchange.jl:316-337; no drag)
double qp_conv(const orc &o, int mu, int i, int j, int k) {
  const A3c &P = o.G[mu];
  const C p = conj(P(i, j, k));
  const double gx = o.d.wa[0][mu], gy = o.d.wa[1][mu], gz = o.d.wa[2][mu];
  const double rx2 = 1 / (o.d.dx * o.d.dx), ry2 = 1 / (o.d.dy * o.d.dy), rz2 = 1 / (o.d.dz * o.d.dz);
  double c = ((gx * (o.tx(i, j, k) * cmul(cmul(p, o.Vx(i, j, k)), P(i + 1, j, k)).im +
                     o.tx(i - 1, j, k) * cmul(cmul(p, conj(o.Vx(i - 1, j, k))), P(i - 1, j, k)).im) * rx2 +
               gy * (o.ty(i, j, k) * cmul(cmul(p, o.Vy(i, j, k)), P(i, j + 1, k)).im +
                     o.ty(i, j - 1, k) * cmul(cmul(p, conj(o.Vy(i, j - 1, k))), P(i, j - 1, k)).im) * ry2) +
              gz * (o.tz(i, j, k) * cmul(cmul(p, o.Vz(i, j, k)), P(i, j, k + 1)).im +
                    o.tz(i, j, k - 1) * cmul(cmul(p, conj(o.Vz(i, j, k - 1))), P(i, j, k - 1)).im) * rz2);
  for (int nu = 0; nu < o.N; ++nu) {
    if (nu == mu) continue;
    c += o.d.qp_gamma[mu][nu] * cmul(p, o.G[nu](i, j, k)).im;
  }
  return c;
}

// B = A transformed by Mx along x, My along y, Mz along z (_tx3!, exchange.jl:304-311); M
// column-major n x n, B(i, ..) = Σ_a M(i, a) A(a, ..) with the sum in increasing a
void qp_tx3(std::vector<double> &B, const std::vector<double> &A, std::vector<double> &T, int n1, int n2, int n3,
            const std::vector<double> &Mx, const std::vector<double> &My, const std::vector<double> &Mz) {
  auto at = [&](int i, int j, int k) { return (size_t)i + (size_t)n1 * (j + (size_t)n2 * k); };
  for (int k = 0; k < n3; ++k)
    for (int j = 0; j < n2; ++j)
      for (int i = 0; i < n1; ++i) {
        double acc = 0.0;
        for (int a = 0; a < n1; ++a) acc += Mx[i + (size_t)n1 * a] * A[at(a, j, k)];
        T[at(i, j, k)] = acc;
      }
  for (int k = 0; k < n3; ++k)
    for (int j = 0; j < n2; ++j)
      for (int i = 0; i < n1; ++i) {
        double acc = 0.0;
        for (int b = 0; b < n2; ++b) acc += T[at(i, b, k)] * My[j + (size_t)n2 * b];
        B[at(i, j, k)] = acc;
      }
  for (int k = 0; k < n3; ++k)
    for (int j = 0; j < n2; ++j)
      for (int i = 0; i < n1; ++i) {
        double acc = 0.0;
        for (int c = 0; c < n3; ++c) acc += B[at(i, j, c)] * Mz[k + (size_t)n3 * c];
        T[at(i, j, k)] = acc;
      }
  B = T;
}

// the modes, divided, back (_qp_sep!, exchange.jl:295-302)
void qp_sep(orc &o, std::vector<double> &X, double lam) {
  const int n1 = o.Nx - 1, n2 = o.qnj, n3 = o.qnk;
  qp_tx3(o.qT2, X, o.qT1, n1, n2, n3, o.qWx, o.qWy, o.qWz);
  for (int k = 0; k < n3; ++k)
    for (int j = 0; j < n2; ++j)
      for (int i = 0; i < n1; ++i) {
        double &t = o.qT2[(size_t)i + (size_t)n1 * (j + (size_t)n2 * k)];
        t /= o.d.qp_sigma * (o.qex[i] + o.qey[j] + o.qez[k]) - lam;
      }
  qp_tx3(X, o.qT2, o.qT1, n1, n2, n3, o.qVx, o.qVy, o.qVz);
}

// the potentials from the state at the start of the step (qp_potential!, exchange.jl:344-389 for
// two bands, :477-516 for N): the source on every block node, the separable solves
void qp_potential(orc &o) {
  const int n1 = o.Nx - 1, nj = o.qnj, nk = o.qnk, N = o.N;
  const size_t n = (size_t)n1 * nj * nk;
  for (int q = 0; q < o.qK; ++q) o.qM[q].assign(n, 0.0);
  std::vector<double> R[TDGL_MAX_BANDS];
  if (o.qp == TDGL_QP_N)
    for (int mu = 0; mu < N; ++mu) R[mu].assign(n, 0.0);
  for (int k = o.z1; k <= o.z2 + 1; ++k)
    for (int j = o.y1; j <= o.y2 + 1; ++j)
      for (int i = o.x1; i <= o.x2 + 1; ++i) {
        const size_t q = (size_t)(i - 2) + (size_t)n1 * ((j - o.qj0 - 1) + (size_t)nj * (k - o.qk0 - 1));
        if (o.qp == TDGL_QP_TWO) {
          const double c1 = qp_conv(o, 0, i, j, k), c2 = qp_conv(o, 1, i, j, k);
          o.qM[0][q] = c2 / o.d.qp_s[1] - c1 / o.d.qp_s[0];
        } else {
          double c[TDGL_MAX_BANDS], C = 0.0;
          for (int mu = 0; mu < N; ++mu) c[mu] = qp_conv(o, mu, i, j, k);
          for (int mu = 0; mu < N; ++mu) C += c[mu];
          for (int mu = 0; mu < N; ++mu) R[mu][q] = c[mu] / o.d.qp_s[mu] - C;
        }
      }
  if (o.qp == TDGL_QP_N)
    for (int kk = 0; kk < o.qK; ++kk)
      for (size_t q = 0; q < n; ++q) {
        double m = 0.0;
        for (int mu = 0; mu < N; ++mu) m += o.d.qp_Pinv[kk][mu] * R[mu][q];
        o.qM[kk][q] = m;
      }
  for (int kk = 0; kk < o.qK; ++kk) o.qSrc[kk] = o.qM[kk];
  for (int kk = 0; kk < o.qK; ++kk) qp_sep(o, o.qM[kk], o.d.qp_lambda[kk]);
}

// every band turns by its potential on the nodes the band update writes (qp_turn!,
// exchange.jl:400-412 and :533-543)
void qp_turn(orc &o) {
  const int n1 = o.Nx - 1, nj = o.qnj;
  for (int mu = 0; mu < o.N; ++mu)
    for (int k = o.z1 + 1; k <= o.z2; ++k)
      for (int j = o.y1 + 1; j <= o.y2; ++j)
        for (int i = o.x1 + 1; i <= o.x2; ++i) {
          const size_t q = (size_t)(i - 2) + (size_t)n1 * ((j - o.qj0 - 1) + (size_t)nj * (k - o.qk0 - 1));
          double a;
          if (o.qp == TDGL_QP_TWO) {
            a = -o.d.qp_turn[mu] * o.qM[0][q];
          } else {
            double pot = 0.0;
            for (int kk = 0; kk < o.qK; ++kk) pot += o.d.qp_P[mu][kk] * o.qM[kk][q];
            a = -o.d.dt * pot;
          }
          o.F[mu](i, j, k) = cmul(o.F[mu](i, j, k), cis(a));
        }
}

// ---------------------------------------------------------------------------- thermal noise
// Philox4x32-10 (Salmon, Moraes, Dror & Shaw, "Parallel random numbers: as easy as 1, 2, 3",
// SC'11): ten rounds of two 32x32->64 multiplications on the counter, the key bumped by the
// Weyl constants between rounds.
void philox(uint32_t c[4], uint32_t k0, uint32_t k1) {
  for (int r = 0; r < 10; ++r) {
    if (r) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    const uint64_t p0 = (uint64_t)0xD2511F53u * c[0], p1 = (uint64_t)0xCD9E8D57u * c[2];
    const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0, hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    const uint32_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
    c[0] = n0;
    c[1] = lo1;
    c[2] = n2;
    c[3] = lo0;
  }
}

// Two standard normals from one Philox block (Box & Muller, Ann. Math. Stat. 29, 610 (1958)):
// counter (q, step low, step high, tag), key = the seed; u in (0, 1) and v in [0, 1) from 52
// bits each, r = √(−2 log u), (r cos 2πv, r sin 2πv).
void normal2(uint64_t seed, uint64_t t, uint32_t tag, uint32_t q, double &n1, double &n2) {
  uint32_t c[4] = {q, (uint32_t)t, (uint32_t)(t >> 32), tag};
  philox(c, (uint32_t)seed, (uint32_t)(seed >> 32));
  const double e52 = 2.220446049250313e-16;  // 2^-52
  const double u = ((double)(((uint64_t)(c[0] >> 6) << 26) | (uint64_t)(c[1] >> 6)) + 0.5) * e52;
  const double v = (double)(((uint64_t)(c[2] >> 6) << 26) | (uint64_t)(c[3] >> 6)) * e52;
  const double r = std::sqrt(-2.0 * std::log(u));
  const double a = 6.283185307179586 * v;
  n1 = r * std::cos(a);
  n2 = r * std::sin(a);
}

inline bool kicked(const orc &o, int i, int j, int k) { return !(o.d.hard && o.f(i, j, k) == 0.0); }

// Every draw of the step: ξ per band at the nodes the kick visits (noise.jl:47-52), n at the
// links field_kick! visits (:188-196); from the generator or from the caller's draws.
void noise_fill(orc &o) {
  const int Nx = o.Nx, Ny = o.Ny, Nz = o.Nz;
  const size_t nn = (size_t)(Nx + 1) * (Ny + 1) * (Nz + 1);
  const size_t nl[3] = {(size_t)Nx * (Ny + 1) * (Nz + 1), (size_t)(Nx + 1) * Ny * (Nz + 1), (size_t)(Nx + 1) * (Ny + 1) * Nz};
  if (o.d.noise_host) {
    const size_t stride = (size_t)o.N * 2 * nn + (o.d.noise_field ? nl[0] + nl[1] + nl[2] : 0);
    const uint64_t slot = o.nctr - o.hbase;
    if (o.nctr < o.hbase || slot >= (uint64_t)o.hslots) std::abort();
    const double *p = o.hdraws.data() + slot * stride;
    for (int mu = 0; mu < o.N; ++mu, p += 2 * nn) load(o.Xi[mu], p);
    if (o.d.noise_field)
      for (int c = 0; c < 3; p += nl[c], ++c) load(o.Nl[c], p);
    return;
  }
  const uint64_t seed = o.d.noise_seed, t = o.nctr;
  for (int mu = 0; mu < o.N; ++mu)
    for (int k = o.z1 + 1; k <= o.z2; ++k)
      for (int j = o.y1 + 1; j <= o.y2; ++j)
        for (int i = o.x1 + 1; i <= o.x2; ++i) {
          if (!kicked(o, i, j, k)) continue;
          const uint32_t q = (uint32_t)((i - 1) + (size_t)(Nx + 1) * ((j - 1) + (size_t)(Ny + 1) * (k - 1)));
          double a, b;
          normal2(seed, t, (uint32_t)mu, q, a, b);
          o.Xi[mu](i, j, k) = {a, b};
        }
  if (!o.d.noise_field) return;
  const int e[3][3] = {{Nx, Ny + 1, Nz + 1}, {Nx + 1, Ny, Nz + 1}, {Nx + 1, Ny + 1, Nz}};
  const int lo[3][3] = {{1, 2, 2}, {2, 1, 2}, {2, 2, 1}};
  const int hi[3][3] = {{Nx, Ny, Nz}, {Nx, Ny, Nz}, {Nx, Ny, Nz}};
  for (int c = 0; c < 3; ++c)
    for (int k = lo[c][2]; k <= hi[c][2]; ++k)
      for (int j = lo[c][1]; j <= hi[c][1]; ++j)
        for (int i = lo[c][0]; i <= hi[c][0]; ++i) {
          const uint32_t q = (uint32_t)((i - 1) + (size_t)e[c][0] * ((j - 1) + (size_t)e[c][1] * (k - 1)));
          double a, b;
          normal2(seed, t, 16u + (uint32_t)c, q, a, b);
          o.Nl[c](i, j, k) = a;
        }
}

// _kwt_kick (noise.jl:75-87): the polar kick of a KWT node at ψ = Ψⁿ, the Cartesian one where
// |ψ|² or |F|² is under _ROT_MIN2
C kwt_kick(C F, C psi, C xi, double a, double c, double dr, double g2, double Gm, double d0) {
  const double r2 = abs2(psi), sk = std::sqrt(1 + g2 * r2);
  if (!(r2 > 1e-24)) return cadd(F, rmul(a, xi));
  const double sr = std::sqrt(r2);
  const C u = {psi.re / sr, psi.im / sr};
  const C w = cmul(conj(u), xi);
  if (r2 > ROT_MIN2 && abs2(F) > ROT_MIN2) {
    const double rho = std::sqrt(r2), aF = std::hypot(F.re, F.im);
    const double drho = c * w.re / std::sqrt(Gm * sk) + d0 / (rho * (sk * sk * sk));
    const double dth = c * w.im * std::sqrt(sk / Gm) / rho;
    return cmul(C{F.re / aF, F.im / aF}, rmul(aF + drho, cis(dth)));
  }
  const C t = cmul(rmul(c, u), C{w.re / std::sqrt(Gm * sk), w.im * std::sqrt(sk / Gm)});
  return csub(cadd(F, t), rmul(dr * (1 + sk * sk) / (sk * sk * sk), psi));
}

// noise_draw! (noise.jl:99-124): δθ of every SK-ROCK band from Ψⁿ and its draws
void noise_draw(orc &o) {
  const double dV = o.d.dx * o.d.dy * o.d.dz;
  const double c = std::sqrt(o.d.noise_T * o.d.dt / dV);
  for (int mu = 0; mu < o.N; ++mu) {
    A3r &Qt = o.Qth[mu];
    std::fill(Qt.v.begin(), Qt.v.end(), 0.0);
    if (!o.d.noise_sk[mu]) continue;
    const double g2 = o.d.bcs[mu].gamma2, Gm = o.d.noise_Gamma[mu];
    for (int k = o.z1 + 1; k <= o.z2; ++k)
      for (int j = o.y1 + 1; j <= o.y2; ++j)
        for (int i = o.x1 + 1; i <= o.x2; ++i) {
          if (!kicked(o, i, j, k)) continue;
          const C psi = o.G[mu](i, j, k), xi = o.Xi[mu](i, j, k);
          const double r2 = abs2(psi);
          if (!(g2 > 0 && r2 > ROT_MIN2)) continue;
          const double sk = std::sqrt(1 + g2 * r2), rho = std::sqrt(r2);
          Qt(i, j, k) = c * cmul(conj(C{psi.re / rho, psi.im / rho}), xi).im * std::sqrt(sk / Gm) / rho;
        }
  }
}

// noise_kick! (noise.jl:44-65) or, after SK-ROCK stages, noise_apply! (:127-153): the kick of the
// step on every band, after its update
void noise_kick(orc &o) {
  const double dV = o.d.dx * o.d.dy * o.d.dz, Tdt = o.d.noise_T * o.d.dt;
  const double c = std::sqrt(Tdt / dV);
  for (int mu = 0; mu < o.N; ++mu) {
    const double Gm = o.d.noise_Gamma[mu], g2 = o.d.bcs[mu].gamma2;
    const double a = std::sqrt(Tdt / (Gm * dV));
    const double dr = Tdt / (2 * dV) * g2 / Gm, d0 = Tdt / (2 * dV * Gm);
    A3c &F = o.F[mu];
    for (int k = o.z1 + 1; k <= o.z2; ++k)
      for (int j = o.y1 + 1; j <= o.y2; ++j)
        for (int i = o.x1 + 1; i <= o.x2; ++i) {
          if (!kicked(o, i, j, k)) continue;
          const C xi = o.Xi[mu](i, j, k);
          if (g2 == 0) {
            F(i, j, k) = cadd(F(i, j, k), rmul(a, xi));
          } else if (o.sk_any && o.Qth[mu](i, j, k) != 0.0) {
            // the phase went in with the stages: the amplitude here
            const C psi = o.G[mu](i, j, k);
            const double rho = std::hypot(psi.re, psi.im), sk = std::sqrt(1 + g2 * (rho * rho));
            const C u = {psi.re / rho, psi.im / rho};
            const double drho = c * cmul(conj(u), xi).re / std::sqrt(Gm * sk) + Tdt / (2 * dV * Gm) / (rho * (sk * sk * sk));
            const C f = F(i, j, k);
            const double aF = std::hypot(f.re, f.im);
            F(i, j, k) = aF > std::sqrt(ROT_MIN2) ? rmul((aF + drho) / aF, f) : cadd(f, rmul(drho, u));
          } else {
            F(i, j, k) = kwt_kick(F(i, j, k), o.G[mu](i, j, k), xi, a, c, dr, g2, Gm, d0);
          }
        }
  }
}

// field_kick! (noise.jl:182-198): φ_l += d_l √(T̃ dt D_l/(σn dV)) n on the links the field step evolves
void field_kick(orc &o) {
  const int Nx = o.Nx, Ny = o.Ny, Nz = o.Nz;
  const double c = std::sqrt(o.d.noise_T * o.d.dt / (o.d.noise_sigma_n * o.d.dx * o.d.dy * o.d.dz));
  const double ax = c * o.d.dx, ay = c * o.d.dy, az = c * o.d.dz;
  const bool hasD = o.d.Dx != nullptr;
  for (int k = 2; k <= Nz; ++k)
    for (int j = 2; j <= Ny; ++j)
      for (int i = 1; i <= Nx; ++i) o.phx(i, j, k) += (hasD ? ax * std::sqrt(o.Dx(i, j, k)) : ax) * o.Nl[0](i, j, k);
  for (int k = 2; k <= Nz; ++k)
    for (int j = 1; j <= Ny; ++j)
      for (int i = 2; i <= Nx; ++i) o.phy(i, j, k) += (hasD ? ay * std::sqrt(o.Dy(i, j, k)) : ay) * o.Nl[1](i, j, k);
  for (int k = 1; k <= Nz; ++k)
    for (int j = 2; j <= Ny; ++j)
      for (int i = 2; i <= Nx; ++i) o.phz(i, j, k) += (hasD ? az * std::sqrt(o.Dz(i, j, k)) : az) * o.Nl[2](i, j, k);
}

void step(orc &o, double H, double HJ) {
  if (o.d.selffield_bc == TDGL_BC_OPEN) open_update(o, HJ);
  if (o.qp != TDGL_QP_NONE) qp_potential(o);
  if (o.noise) noise_fill(o);
  if (o.sk_any) noise_draw(o);
  update_bands(o);
  if (o.qp != TDGL_QP_NONE) qp_turn(o);
  if (o.noise) noise_kick(o);
  supercurrent(o);
  field_increment(o);
  if (o.noise && o.d.noise_field) field_kick(o);
  switch (o.d.field_mode) {
    case TDGL_FIELD_SEPARABLE: implicit_sep(o); break;
    case TDGL_FIELD_PCG: implicit_pcg(o); break;
    case TDGL_FIELD_EXPLICIT_SCALED: scale_increment(o); break;
    default: break;
  }
  links(o);
  curl_h(o);
  field_bc(o, H, HJ);
  neumann(o);
  advance(o);
  if (o.noise) o.nctr += 1;
}


}  // namespace

extern "C" {

orc *orc_create(const tdgl_desc *d) {
  orc *o = new orc();
  o->d = *d;
  o->N = d->nbands;
  o->Nx = d->Nx;
  o->Ny = d->Ny;
  o->Nz = d->Nz;
  o->x1 = d->x1;
  o->x2 = d->x2;
  o->y1 = d->y1;
  o->y2 = d->y2;
  o->z1 = d->z1;
  o->z2 = d->z2;
  const int Nx = d->Nx, Ny = d->Ny, Nz = d->Nz;
  o->f = A3r(Nx + 1, Ny + 1, Nz + 1);
  o->tx = A3r(Nx + 1, Ny + 1, Nz + 1);
  o->ty = A3r(Nx + 1, Ny + 1, Nz + 1);
  o->tz = A3r(Nx + 1, Ny + 1, Nz + 1);
  load(o->f, d->f);
  load(o->tx, d->taux);
  load(o->ty, d->tauy);
  load(o->tz, d->tauz);
  if (d->Dx) {
    o->Dx = A3r(Nx, Ny + 1, Nz + 1);
    o->Dy = A3r(Nx + 1, Ny, Nz + 1);
    o->Dz = A3r(Nx + 1, Ny + 1, Nz);
    load(o->Dx, d->Dx);
    load(o->Dy, d->Dy);
    load(o->Dz, d->Dz);
  }
  if (d->selffield_bc == TDGL_BC_BAR || d->selffield_bc == TDGL_BC_OPEN) {
    size_t n[6] = {(size_t)(Ny + 1) * 2, (size_t)2 * Nz, (size_t)Ny * 2, (size_t)2 * (Nz + 1), (size_t)Ny * (Nz + 1),
                   (size_t)(Ny + 1) * Nz};
    const double *src[6] = {d->bar_hy_z, d->bar_hy_y, d->bar_hz_z, d->bar_hz_y, d->bar_hz_x, d->bar_hy_x};
    for (int q = 0; q < 6; ++q) o->bar[q].assign(src[q], src[q] + n[q]);
  }
  for (int mu = 0; mu < o->N; ++mu) {
    o->G.emplace_back(Nx + 1, Ny + 1, Nz + 1);
    o->F.emplace_back(Nx + 1, Ny + 1, Nz + 1);
  }
  o->thx = o->phx = o->Jx = A3r(Nx, Ny + 1, Nz + 1);
  o->thy = o->phy = o->Jy = A3r(Nx + 1, Ny, Nz + 1);
  o->thz = o->phz = o->Jz = A3r(Nx + 1, Ny + 1, Nz);
  o->Ux = o->Vx = A3c(Nx, Ny + 1, Nz + 1);
  o->Uy = o->Vy = A3c(Nx + 1, Ny, Nz + 1);
  o->Uz = o->Vz = A3c(Nx + 1, Ny + 1, Nz);
  o->hx = A3r(Nx + 1, Ny, Nz);
  o->hy = A3r(Nx, Ny + 1, Nz);
  o->hz = A3r(Nx, Ny, Nz + 1);
  o->W = o->W2 = A3c(Nx + 1, Ny + 1, Nz + 1);
  o->L[0] = o->L[3] = A3c(Nx, Ny + 1, Nz + 1);
  o->L[1] = o->L[4] = A3c(Nx + 1, Ny, Nz + 1);
  o->L[2] = o->L[5] = A3c(Nx + 1, Ny + 1, Nz);
  o->Om = A3r(Nx + 1, Ny + 1, Nz + 1);
  o->R[0] = A3r(Nx, Ny + 1, Nz + 1);
  o->R[1] = A3r(Nx + 1, Ny, Nz + 1);
  o->R[2] = A3r(Nx + 1, Ny + 1, Nz);
  if (d->field_mode == TDGL_FIELD_SEPARABLE || d->field_mode == TDGL_FIELD_PCG) {
    Sep &S = o->sep;
    S.Nx = Nx;
    S.Ny = Ny;
    S.Nz = Nz;
    S.bx = make_basis(Nx - 1);
    S.by = make_basis(Ny - 1);
    S.bz = make_basis(Nz - 1);
    S.a = d->a;
    int ex[3] = {Nx - 2, Ny - 1, Nz - 1}, ey[3] = {Nx - 1, Ny - 2, Nz - 1}, ez[3] = {Nx - 1, Ny - 1, Nz - 2};
    std::memcpy(S.ex, ex, sizeof ex);
    std::memcpy(S.ey, ey, sizeof ey);
    std::memcpy(S.ez, ez, sizeof ez);
    const Basis *bb[3] = {&S.bx, &S.by, &S.bz};
    for (int c = 0; c < 3; ++c) {
      S.V32[c].assign(bb[c]->V.begin(), bb[c]->V.end());
      S.U32[c].assign(bb[c]->U.begin(), bb[c]->U.end());
      S.s32[c].assign(bb[c]->s.begin(), bb[c]->s.end());
    }
    o->have_sep = true;
  }
  if (d->selffield_bc == TDGL_BC_OPEN) {
    o->odJ[0] = o->oth[0] = A3r(Nx, Ny + 1, Nz + 1);
    o->odJ[1] = o->oth[1] = A3r(Nx + 1, Ny, Nz + 1);
    o->odJ[2] = o->oth[2] = A3r(Nx + 1, Ny + 1, Nz);
    o->oold[0] = o->onew[0] = A3r(Nx + 1, Ny, Nz);
    o->oold[1] = o->onew[1] = A3r(Nx, Ny + 1, Nz);
    o->oold[2] = o->onew[2] = A3r(Nx, Ny, Nz + 1);
  }
  if (d->qp_mode != TDGL_QP_NONE) {
    o->qp = d->qp_mode;
    o->qK = d->qp_nmodes;
    o->qj0 = d->qp_j0;
    o->qnj = d->qp_nj;
    o->qk0 = d->qp_k0;
    o->qnk = d->qp_nk;
    const int n1 = Nx - 1, nj = d->qp_nj, nk = d->qp_nk;
    o->qVx.assign(d->qp_Vx, d->qp_Vx + (size_t)n1 * n1);
    o->qWx.assign(d->qp_Wx, d->qp_Wx + (size_t)n1 * n1);
    o->qex.assign(d->qp_ex, d->qp_ex + n1);
    o->qVy.assign(d->qp_Vy, d->qp_Vy + (size_t)nj * nj);
    o->qWy.assign(d->qp_Wy, d->qp_Wy + (size_t)nj * nj);
    o->qey.assign(d->qp_ey, d->qp_ey + nj);
    o->qVz.assign(d->qp_Vz, d->qp_Vz + (size_t)nk * nk);
    o->qWz.assign(d->qp_Wz, d->qp_Wz + (size_t)nk * nk);
    o->qez.assign(d->qp_ez, d->qp_ez + nk);
    o->qT1.assign((size_t)n1 * nj * nk, 0.0);
    o->qT2.assign((size_t)n1 * nj * nk, 0.0);
  }
  if (d->noise) {
    o->noise = true;
    for (int mu = 0; mu < o->N; ++mu) {
      o->Xi.emplace_back(Nx + 1, Ny + 1, Nz + 1);
      o->Qth.emplace_back(Nx + 1, Ny + 1, Nz + 1);
      o->sk_any = o->sk_any || d->noise_sk[mu];
    }
    o->Nl[0] = A3r(Nx, Ny + 1, Nz + 1);
    o->Nl[1] = A3r(Nx + 1, Ny, Nz + 1);
    o->Nl[2] = A3r(Nx + 1, Ny + 1, Nz);
    o->Sp = A3c(Nx + 1, Ny + 1, Nz + 1);
    o->Om2 = A3r(Nx + 1, Ny + 1, Nz + 1);
  }
  if (d->field_mode == TDGL_FIELD_PCG) {
    o->P[0] = o->Q[0] = A3r(Nx, Ny + 1, Nz + 1);
    o->P[1] = o->Q[1] = A3r(Nx + 1, Ny, Nz + 1);
    o->P[2] = o->Q[2] = A3r(Nx + 1, Ny + 1, Nz);
    o->H[0] = A3r(Nx + 1, Ny, Nz);
    o->H[1] = A3r(Nx, Ny + 1, Nz);
    o->H[2] = A3r(Nx, Ny, Nz + 1);
    for (int c = 0; c < 3; ++c) {
      R3 r = prange(*o, c);
      size_t n = (size_t)(r.i1 - r.i0 + 1) * (r.j1 - r.j0 + 1) * (r.k1 - r.k0 + 1);
      o->x[c].assign(n, 0.0);
      o->xprev[c].assign(n, 0.0);
      o->zold[c].assign(n, 0.0);
    }
  }
  return o;
}

void orc_destroy(orc *o) { delete o; }

// the noise stream: its position, the caller's draws (the layout of tdgl_noise_draws), and the
// generator itself for its known-answer vectors
uint64_t orc_noise_get(orc *o) { return o->nctr; }
void orc_noise_set(orc *o, uint64_t c) { o->nctr = c; }
void orc_noise_draws(orc *o, int nsteps, const double *draws) {
  const size_t nn = (size_t)(o->Nx + 1) * (o->Ny + 1) * (o->Nz + 1);
  const size_t nl = (size_t)o->Nx * (o->Ny + 1) * (o->Nz + 1) + (size_t)(o->Nx + 1) * o->Ny * (o->Nz + 1) +
                    (size_t)(o->Nx + 1) * (o->Ny + 1) * o->Nz;
  const size_t stride = (size_t)o->N * 2 * nn + (o->d.noise_field ? nl : 0);
  o->hdraws.assign(draws, draws + stride * (size_t)nsteps);
  o->hslots = nsteps;
  o->hbase = o->nctr;
}
void orc_philox(const uint32_t *ctr, const uint32_t *key, uint32_t *out) {
  uint32_t c[4] = {ctr[0], ctr[1], ctr[2], ctr[3]};
  philox(c, key[0], key[1]);
  for (int q = 0; q < 4; ++q) out[q] = c[q];
}
void orc_normal2(uint64_t seed, uint64_t t, uint32_t tag, uint32_t q, double *out) {
  normal2(seed, t, tag, q, out[0], out[1]);
}

// the solved modes of the last normal-fluid potential (δ for two bands), region-shaped
void orc_qp_modes(orc *o, int k, double *out) {
  if (k >= 0 && k < o->qK) std::memcpy(out, o->qM[k].data(), o->qM[k].size() * sizeof(double));
}
void orc_qp_source(orc *o, int k, double *out) {
  if (k >= 0 && k < o->qK) std::memcpy(out, o->qSrc[k].data(), o->qSrc[k].size() * sizeof(double));
}
// the open boundary's clock and h_δ of its last two recomputations (h-shaped, hx hy hz each)
int orc_open_state(orc *o, double *hold, double *hnew) {
  if (o->d.selffield_bc != TDGL_BC_OPEN) return -1;
  size_t q = 0;
  for (int c = 0; c < 3; ++c) {
    std::memcpy(hold + q, o->oold[c].v.data(), o->oold[c].size() * sizeof(double));
    std::memcpy(hnew + q, o->onew[c].v.data(), o->onew[c].size() * sizeof(double));
    q += o->oold[c].size();
  }
  return o->ocount;
}
// the clock and ramp of the open boundary (restore_open!, io.jl:163-171)
void orc_set_open(orc *o, int count, const double *hold, const double *hnew) {
  if (o->d.selffield_bc != TDGL_BC_OPEN) return;
  size_t q = 0;
  for (int c = 0; c < 3; ++c) {
    std::memcpy(o->oold[c].v.data(), hold + q, o->oold[c].size() * sizeof(double));
    std::memcpy(o->onew[c].v.data(), hnew + q, o->onew[c].size() * sizeof(double));
    q += o->oold[c].size();
  }
  o->ocount = count;
}
// θ_δ and δJ of the last recomputation (link-shaped x, y, z each)
void orc_open_fields(orc *o, double *theta, double *dj) {
  size_t q = 0;
  for (int c = 0; c < 3; ++c) {
    std::memcpy(theta + q, o->oth[c].v.data(), o->oth[c].size() * sizeof(double));
    std::memcpy(dj + q, o->odJ[c].v.data(), o->odJ[c].size() * sizeof(double));
    q += o->oth[c].size();
  }
}
// the potentials of the state as it is now (qp_potential! without a step)
void orc_qp_potential(orc *o) {
  if (o->qp != TDGL_QP_NONE) qp_potential(*o);
}

void orc_set_state(orc *o, const tdgl_state *s) {
  for (int mu = 0; mu < o->N; ++mu) {
    load(o->G[mu], s->G[mu]);
    load(o->F[mu], s->F[mu]);
  }
  load(o->thx, s->thx);
  load(o->thy, s->thy);
  load(o->thz, s->thz);
  load(o->phx, s->phx);
  load(o->phy, s->phy);
  load(o->phz, s->phz);
  load(o->Ux, s->Ux);
  load(o->Uy, s->Uy);
  load(o->Uz, s->Uz);
  load(o->Vx, s->Vx);
  load(o->Vy, s->Vy);
  load(o->Vz, s->Vz);
  load(o->Jx, s->Jx);
  load(o->Jy, s->Jy);
  load(o->Jz, s->Jz);
  load(o->hx, s->hx);
  load(o->hy, s->hy);
  load(o->hz, s->hz);
}

void orc_get_state(orc *o, tdgl_state *s) {
  for (int mu = 0; mu < o->N; ++mu) {
    save(o->G[mu], s->G[mu]);
    save(o->F[mu], s->F[mu]);
  }
  save(o->thx, s->thx);
  save(o->thy, s->thy);
  save(o->thz, s->thz);
  save(o->phx, s->phx);
  save(o->phy, s->phy);
  save(o->phz, s->phz);
  save(o->Ux, s->Ux);
  save(o->Uy, s->Uy);
  save(o->Uz, s->Uz);
  save(o->Vx, s->Vx);
  save(o->Vy, s->Vy);
  save(o->Vz, s->Vz);
  save(o->Jx, s->Jx);
  save(o->Jy, s->Jy);
  save(o->Jz, s->Jz);
  save(o->hx, s->hx);
  save(o->hy, s->hy);
  save(o->hz, s->hz);
}

void orc_step(orc *o, int n, double H, double HJ) {
  for (int q = 0; q < n; ++q) step(*o, H, HJ);
}

void orc_get_pcg(orc *o, tdgl_pcg *p) {
  for (int c = 0; c < 3; ++c) {
    if (p->x[c]) std::memcpy(p->x[c], o->x[c].data(), o->x[c].size() * 8);
    if (p->xprev[c]) std::memcpy(p->xprev[c], o->xprev[c].data(), o->xprev[c].size() * 8);
  }
  p->warm = o->warm;
  p->have2 = o->have2;
  p->iters = o->iters;
}

void orc_set_pcg(orc *o, const tdgl_pcg *p) {
  for (int c = 0; c < 3; ++c) {
    if (p->x[c]) std::memcpy(o->x[c].data(), p->x[c], o->x[c].size() * 8);
    if (p->xprev[c]) std::memcpy(o->xprev[c].data(), p->xprev[c], o->xprev[c].size() * 8);
  }
  o->warm = p->warm != 0;
  o->have2 = p->have2 != 0;
  o->iters = p->iters;
}

long long orc_total_iters(orc *o) { return o->total_iters; }

// The separable inverse alone, in place on range arrays (tests).
void orc_sep_solve(orc *o, double *BX, double *BY, double *BZ) {
  std::vector<double> X(BX, BX + o->BX.size()), Y, Z;
  const Sep &S = o->sep;
  size_t nx = (size_t)S.ex[0] * S.ex[1] * S.ex[2], ny = (size_t)S.ey[0] * S.ey[1] * S.ey[2],
         nz = (size_t)S.ez[0] * S.ez[1] * S.ez[2];
  X.assign(BX, BX + nx);
  Y.assign(BY, BY + ny);
  Z.assign(BZ, BZ + nz);
  sep_solve(S, X, Y, Z);
  std::memcpy(BX, X.data(), nx * 8);
  std::memcpy(BY, Y.data(), ny * 8);
  std::memcpy(BZ, Z.data(), nz * 8);
}

// The 1D bases (tests): V (nn x nn), U (nn-1 x nn-1), s (nn)
void orc_basis(int nn, double *V, double *U, double *s) {
  Basis b = make_basis(nn);
  std::memcpy(V, b.V.data(), b.V.size() * 8);
  std::memcpy(U, b.U.data(), b.U.size() * 8);
  std::memcpy(s, b.s.data(), b.s.size() * 8);
}

// One KWT node update (tests: the 2x2 system of test_physics.jl:51-59)
void orc_kwt_update(const double *g, const double *d, double g2, double *out) {
  C r = kwt_update(C{g[0], g[1]}, C{d[0], d[1]}, g2);
  out[0] = r.re;
  out[1] = r.im;
}

// The field solve of the step alone (implicit_curl! or scale_increment!, step.jl:25-29) on the
// state as set: φ holds φₑ, θ the links of the step (tests: the solve against an operator built
// independently, the zero-source rule of pcg_zero_fixture.jl).
void orc_field_solve(orc *o) {
  switch (o->d.field_mode) {
    case TDGL_FIELD_SEPARABLE: implicit_sep(*o); break;
    case TDGL_FIELD_PCG: implicit_pcg(*o); break;
    case TDGL_FIELD_EXPLICIT_SCALED: scale_increment(*o); break;
    default: break;
  }
}

// _link_rate! of one link (bands.jl:303-310): phase of V Ū plus (ω₁ − ω₀) dt
double orc_link_rate(const double *v, const double *u, double om0, double om1, double dt) {
  A3c V(1, 1, 1), U(1, 1, 1);
  A3r Rr(2, 1, 1), Om(2, 1, 1);
  V(1, 1, 1) = C{v[0], v[1]};
  U(1, 1, 1) = C{u[0], u[1]};
  Om(1, 1, 1) = om0;
  Om(2, 1, 1) = om1;
  orc o0;
  link_rate(o0, Rr, V, U, Om, dt, Rng{1, 1, 1, 1, 1, 1}, 1, 0, 0);
  return Rr(1, 1, 1);
}

// _cis (bands.jl:288-289): the short series below |a| = 0.05
void orc_scis(double a, double *out) {
  C r = scis(a);
  out[0] = r.re;
  out[1] = r.im;
}

// One temporal link rotation (tests: rotation_fixture.jl semantics): mode 0 _rotfrac!, 1 _rot! root
void orc_rotation(int mode, const double *v, const double *u, double c, int n, double *out) {
  A3c L(1, 1, 1), R(1, 1, 1), V(1, 1, 1), U(1, 1, 1);
  V(1, 1, 1) = C{v[0], v[1]};
  U(1, 1, 1) = C{u[0], u[1]};
  Rng r{1, 1, 1, 1, 1, 1};
  if (mode == 0) {
    rotfrac(L, V, U, c, r);
    out[0] = L(1, 1, 1).re;
    out[1] = L(1, 1, 1).im;
  } else {
    rot(L, R, V, U, n, r);
    out[0] = R(1, 1, 1).re;
    out[1] = R(1, 1, 1).im;
  }
}

}  // extern "C"
