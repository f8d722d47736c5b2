// common.h — shared device helpers of the HIP stepper: complex arithmetic in the order of
// Julia's Base (gtdgl's reference arithmetic; built with -ffp-contract=off so that no product
// and sum fuse), the node-shaped grid geometry, and launch helpers.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/tdglhip.h"

#define TDGL_HD __host__ __device__ __forceinline__

namespace tdgl {

typedef double2 cplx;

TDGL_HD cplx mk(double re, double im) { return make_double2(re, im); }
TDGL_HD cplx cmul(cplx a, cplx b) { return mk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
TDGL_HD cplx cadd(cplx a, cplx b) { return mk(a.x + b.x, a.y + b.y); }
TDGL_HD cplx csub(cplx a, cplx b) { return mk(a.x - b.x, a.y - b.y); }
TDGL_HD cplx rmul(double s, cplx a) { return mk(s * a.x, s * a.y); }
TDGL_HD cplx cconj(cplx a) { return mk(a.x, -a.y); }
TDGL_HD double abs2(cplx a) { return a.x * a.x + a.y * a.y; }
// Complex(±0, x) * y: the frame term im·(ω h)·y
TDGL_HD cplx imx_mul(double x, cplx y) { return mk(0.0 * y.x - x * y.y, 0.0 * y.y + x * y.x); }

__device__ __forceinline__ cplx dcis(double x) {
  double s, c;
  sincos(x, &s, &c);
  return mk(c, s);
}
__device__ __forceinline__ double dangle(cplx a) { return atan2(a.y, a.x); }

// e^{ia}: short series while |a| <= 0.05, exactly above (bands.jl:289)
__device__ __forceinline__ cplx scis(double a) {
  if (fabs(a) <= 0.05) {
    double a2 = a * a;
    return mk(1 - a2 / 2, a * (1 - a2 / 6));
  }
  return dcis(a);
}

// KWT step of one node (bands.jl:135-140)
__device__ __forceinline__ cplx kwt_update(cplx g, cplx d, double g2) {
  double a2 = abs2(g);
  double q = 1 + g2 * a2;
  cplx dp = rmul(sqrt(q), d);
  double s = cmul(cconj(g), dp).x / q;
  return cadd(g, csub(dp, rmul(s, rmul(g2, g))));
}

// KWT step of one node with its tangential part as a rotation (kwt_update_rot, bands.jl:150-158)
__device__ __forceinline__ cplx kwt_update_rot(cplx g, cplx d, double g2) {
  double a2 = abs2(g);
  double q = 1 + g2 * a2;
  cplx dp = rmul(sqrt(q), d);
  cplx dl = csub(dp, rmul(cmul(cconj(g), dp).x / q, rmul(g2, g)));
  if (!(a2 > 1e-24)) return cadd(g, dl);
  double a = sqrt(a2);
  cplx u = mk(g.x / a, g.y / a);
  cplx w = cmul(cconj(u), dl);
  return cmul(u, rmul(a + w.x, dcis(w.y / a)));
}

// complex quotient (textbook formula), log (log|z| with |z| = hypot, arg z) and exp
__device__ __forceinline__ cplx cdivz(cplx a, cplx b) {
  double den = b.x * b.x + b.y * b.y;
  return mk((a.x * b.x + a.y * b.y) / den, (a.y * b.x - a.x * b.y) / den);
}
__device__ __forceinline__ cplx clogz(cplx z) { return mk(log(hypot(z.x, z.y)), atan2(z.y, z.x)); }
__device__ __forceinline__ cplx cexpz(cplx z) { return rmul(exp(z.x), dcis(z.y)); }

TDGL_HD double jclamp(double x, double lo, double hi) { return x > hi ? hi : (x < lo ? lo : x); }

// Node-shaped geometry: every 3D array of a run is stored with the strides of the node array
// (Nx+1, Ny+1, Nz+1), the x pitch padded to a multiple of 8 elements; 1-based (i, j, k) as in
// the reference.  Link and plaquette arrays use the leading part of the same layout.
struct Geo {
  int Nx, Ny, Nz;
  int x1, x2, y1, y2, z1, z2;
  int px, py;  // pitch along x (elements), Ny+1
  int n;       // elements per array: px*py*(Nz+1)
  TDGL_HD int at(int i, int j, int k) const { return (i - 1) + px * ((j - 1) + py * (k - 1)); }
};

// Compact layout of the ranges of the implicit solve (implicit.jl:52):
// rx = (2:Nx-1, 2:Ny, 2:Nz), ry = (2:Nx, 2:Ny-1, 2:Nz), rz = (2:Nx, 2:Ny, 2:Nz-1),
// component c at offset off[c] of a vector of length total, i fastest.
struct Ranges {
  int e[3][3];   // extents per component
  int off[3];    // offsets of the components
  int total;
  TDGL_HD bool in(int c, int i, int j, int k) const {
    return i >= 2 && j >= 2 && k >= 2 && i - 2 < e[c][0] && j - 2 < e[c][1] && k - 2 < e[c][2];
  }
  TDGL_HD int idx(int c, int i, int j, int k) const { return off[c] + (i - 2) + e[c][0] * ((j - 2) + e[c][1] * (k - 2)); }
};

inline int cdiv(int a, int b) { return (a + b - 1) / b; }
__device__ __forceinline__ int cdiv_d(int a, int b) { return (a + b - 1) / b; }

}  // namespace tdgl
