// openbc.hip — the open boundary of the field on the device (gtdgl/src/openbc.jl): δJ, the
// convolution with the lattice Green function by zero-padded FFTs, h_δ on the faces, and the
// clock that decides on which steps they run.
//
// The convolution is Hockney's: a padded size M >= 2n - 1 per axis (a product of 2, 3 and 5) and
// the Green function placed cyclically, so the first n outputs are the linear convolution, the
// same as the reference's 2n padding computes.  G is even along every axis, so its transform is
// real.  The padding is never stored: the x pass reads only the box and zero-extends, the z pass
// (forward, times Ĝ, inverse) keeps only the box's planes, the inverse passes write only the box.
#include "kernels.h"

namespace tdgl {

// ---------------------------------------------------------------- the clock (open_update!, :196-215)
__global__ void k_open_clock(int *ofl, int every) {
  if (threadIdx.x != 0) return;
  const int first = ofl[OF_COUNT] < 0;
  const int c = ofl[OF_COUNT] + 1;
  const int re = (c >= every || first) ? 1 : 0;
  ofl[OF_RECOMP] = re;
  ofl[OF_FIRST] = first;
  ofl[OF_COUNT] = re ? 0 : c;
}
hipError_t launch_open_clock(int *ofl, int every, hipStream_t s) {
  if (g_rec) return hipErrorNotSupported;
  TDGL_LAUNCH(k_open_clock, dim3(1), dim3(64), 0, s, ofl, every);
  return hipGetLastError();
}

// ---------------------------------------------------------------- δJ (_delta_current!, :143-160)
// δJ on the links of lattice c (blockIdx.y), compact (n1, n2, n3), zero on the face-normal links
__global__ void __launch_bounds__(256) k_open_dj(Geo g, OpenDJ a) {
  if (!a.ofl[OF_RECOMP]) return;
  const int c = blockIdx.y;
  const int n1 = a.n[c][0], n2 = a.n[c][1], n3 = a.n[c][2];
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n1 * n2 * n3) return;
  const int i = 1 + t % n1, j = 1 + (t / n1) % n2, k = 1 + t / (n1 * n2);
  const int Nx = g.Nx, Ny = g.Ny, Nz = g.Nz;
  const double *hx = a.h[0], *hy = a.h[1], *hz = a.h[2];
  const int sy = g.px, sz = g.px * g.py;
  double v = 0.0;
  if (c == 0) {
    if (k >= 2 && k <= Nz && j >= 2 && j <= Ny && i >= 2 && i <= Nx - 1) {
      const int q = g.at(i, j, k);
      v = a.k2 * ((hz[q] - hz[q - sy]) / a.dy - (hy[q] - hy[q - sz]) / a.dz) - a.HJp[0] * a.Jbar[(j - 1) + (Ny + 1) * (k - 1)];
    }
  } else if (c == 1) {
    if (k >= 2 && k <= Nz && j >= 2 && j <= Ny - 1 && i >= 2 && i <= Nx) {
      const int q = g.at(i, j, k);
      v = a.k2 * ((hx[q] - hx[q - sz]) / a.dz - (hz[q] - hz[q - 1]) / a.dx);
    }
  } else {
    if (k >= 2 && k <= Nz - 1 && j >= 2 && j <= Ny && i >= 2 && i <= Nx) {
      const int q = g.at(i, j, k);
      v = a.k2 * ((hy[q] - hy[q - 1]) / a.dx - (hx[q] - hx[q - sy]) / a.dy);
    }
  }
  a.dj[c][t] = v;
}
hipError_t launch_open_dj(const Geo &g, const OpenDJ &a, hipStream_t s) {
  if (g_rec) return hipErrorNotSupported;
  int m = 0;
  for (int c = 0; c < 3; ++c) m = std::max(m, a.n[c][0] * a.n[c][1] * a.n[c][2]);
  TDGL_LAUNCH(k_open_dj, dim3(cdiv(m, 256), 3), dim3(256), 0, s, g, a);
  return hipGetLastError();
}

// ---------------------------------------------------------------- line FFTs
__device__ __forceinline__ cplx imul_s(cplx v, int s) { return s > 0 ? mk(-v.y, v.x) : mk(v.y, -v.x); }  // (s i) v

// the R-point DFT with exponent sign s, in place
__device__ __forceinline__ void dft_small(cplx *v, int R, int s) {
  if (R == 2) {
    const cplx a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  } else if (R == 4) {
    const cplx t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]), t2 = cadd(v[1], v[3]), t3 = imul_s(csub(v[1], v[3]), s);
    v[0] = cadd(t0, t2);
    v[1] = cadd(t1, t3);
    v[2] = csub(t0, t2);
    v[3] = csub(t1, t3);
  } else if (R == 3) {
    const double h3 = 0.86602540378443864676;  // √3/2
    const cplx t1 = cadd(v[1], v[2]), t2 = imul_s(rmul(h3, csub(v[1], v[2])), s);
    const cplx m = csub(v[0], rmul(0.5, t1));
    v[0] = cadd(v[0], t1);
    v[1] = cadd(m, t2);
    v[2] = csub(m, t2);
  } else {  // 5
    const double c1 = 0.30901699437494742410, c2 = -0.80901699437494742410;   // cos 2π/5, cos 4π/5
    const double s1 = 0.95105651629515357212, s2 = 0.58778525229247312917;    // sin 2π/5, sin 4π/5
    const cplx a1 = cadd(v[1], v[4]), b1 = csub(v[1], v[4]), a2 = cadd(v[2], v[3]), b2 = csub(v[2], v[3]);
    const cplx p1 = cadd(v[0], cadd(rmul(c1, a1), rmul(c2, a2)));
    const cplx p2 = cadd(v[0], cadd(rmul(c2, a1), rmul(c1, a2)));
    const cplx q1 = imul_s(cadd(rmul(s1, b1), rmul(s2, b2)), s);
    const cplx q2 = imul_s(csub(rmul(s2, b1), rmul(s1, b2)), s);
    v[0] = cadd(v[0], cadd(a1, a2));
    v[1] = cadd(p1, q1);
    v[4] = csub(p1, q1);
    v[2] = cadd(p2, q2);
    v[3] = csub(p2, q2);
  }
}

// Stockham autosort passes over `lines` lines of length M in LDS; returns the buffer holding the
// result (natural order).  Pass with radix R: butterfly j reads src[j + r M/R], twiddles r by
// ω^(r (j mod Ns) M/(Ns R)), writes dst[(j/Ns) Ns R + (j mod Ns) + r Ns].
__device__ cplx *fft_lines(cplx *A, cplx *B, int lines, const FftAxis &ax, int dir) {
  const int M = ax.M;
  cplx *src = A, *dst = B;
  int Ns = 1;
  for (int s = 0; s < ax.nrad; ++s) {
    const int R = ax.rad[s], MR = M / R, step = M / (Ns * R);
    for (int item = threadIdx.x; item < lines * MR; item += blockDim.x) {
      const int l = item / MR, j = item % MR, k = j % Ns;
      const cplx *S = src + l * M;
      cplx *D = dst + l * M;
      cplx v[5];
      for (int r = 0; r < R; ++r) v[r] = S[j + r * MR];
      for (int r = 1; r < R; ++r) {
        cplx w = ax.tw[r * k * step];
        if (dir > 0) w.y = -w.y;
        v[r] = cmul(v[r], w);
      }
      dft_small(v, R, dir);
      const int base = (j / Ns) * Ns * R + k;
      for (int r = 0; r < R; ++r) D[base + r * Ns] = v[r];
    }
    __syncthreads();
    cplx *t = src;
    src = dst;
    dst = t;
    Ns *= R;
  }
  return src;
}

constexpr int FFT_LDS = 2048;  // complex values per LDS buffer (two buffers)

__global__ void __launch_bounds__(256) k_fft(FftPass p, const int *gate) {
  if (gate && !*gate) return;
  extern __shared__ __align__(16) unsigned char fft_smem[];
  cplx *A = (cplx *)fft_smem, *B = A + FFT_LDS;
  const int M = p.ax.M, LPB = p.lpb;
  const int nlines = p.nl0 * p.nl1;
  const int L0 = blockIdx.x * LPB;
  const int tot = LPB * M;
  // load (zero beyond nin and beyond the last line); clamped addresses, selected at the store
  for (int t0 = threadIdx.x; t0 < tot; t0 += 4 * blockDim.x) {
    cplx v[4];
    bool ok[4];
    int dsti[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int t = t0 + u * blockDim.x;
      int l, e;
      if (p.lines_fast) {
        l = t % LPB;
        e = t / LPB;
      } else {
        l = t / M;
        e = t % M;
      }
      const int L = L0 + l;
      ok[u] = t < tot && L < nlines && e < p.nin;
      dsti[u] = t < tot ? l * M + e : -1;
      const int Lc = L < nlines ? L : nlines - 1, ec = e < p.nin ? e : p.nin - 1;
      const long long idx = (long long)(Lc % p.nl0) * p.in_l0 + (long long)(Lc / p.nl0) * p.in_l1 + (long long)ec * p.in_s;
      v[u] = p.in_real ? mk(((const double *)p.in)[idx], 0.0) : ((const cplx *)p.in)[idx];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u)
      if (dsti[u] >= 0) A[dsti[u]] = ok[u] ? v[u] : mk(0.0, 0.0);
  }
  __syncthreads();
  cplx *R = fft_lines(A, B, LPB, p.ax, p.dir);
  if (p.ghat) {
    // times Ĝ (real), then back (the z pass of a convolution)
    for (int t = threadIdx.x; t < tot; t += blockDim.x) {
      const int l = t / M, e = t % M;
      const int L = L0 + l;
      if (L >= nlines) continue;
      const double gv = p.ghat[(long long)(L % p.nl0) * p.g_l0 + (long long)(L / p.nl0) * p.g_l1 + (long long)e * p.g_s];
      R[t] = rmul(gv, R[t]);
    }
    __syncthreads();
    cplx *other = R == A ? B : A;
    R = fft_lines(R, other, LPB, p.ax, -p.dir);
  }
  // store the first nout of every line
  for (int t = threadIdx.x; t < LPB * p.nout; t += blockDim.x) {
    int l, e;
    if (p.lines_fast) {
      l = t % LPB;
      e = t / LPB;
    } else {
      l = t / p.nout;
      e = t % p.nout;
    }
    const int L = L0 + l;
    if (L >= nlines) continue;
    const cplx v = R[l * M + e];
    const long long idx = (long long)(L % p.nl0) * p.out_l0 + (long long)(L / p.nl0) * p.out_l1 + (long long)e * p.out_s;
    if (p.out_real)
      ((double *)p.out)[idx] = v.x * p.scale;
    else
      ((cplx *)p.out)[idx] = rmul(p.scale, v);
  }
}

int fft_lpb(int M) { return std::max(1, std::min(64, FFT_LDS / M)); }

hipError_t launch_fft(FftPass p, const int *gate, hipStream_t s) {
  if (g_rec) return hipErrorNotSupported;
  if (p.ax.M < 1 || p.ax.M > FFT_LDS || p.nin < 1 || p.nin > p.ax.M || p.nout < 1 || p.nout > p.ax.M)
    return hipErrorInvalidValue;
  p.lpb = fft_lpb(p.ax.M);
  const int nlines = p.nl0 * p.nl1;
  TDGL_LAUNCH(k_fft, dim3(cdiv(nlines, p.lpb)), dim3(256), 2 * FFT_LDS * sizeof(cplx), s, p, gate);
  return hipGetLastError();
}

// ---------------------------------------------------------------- the padded Green function
// fac·G(d) at the padded index a (offset a for a <= M/2, a - M above: even, so Ĝ is real)
__device__ __forceinline__ double lgf_at(const OpenGreen &a, int i, int j, int k) {
  i = abs(i);
  j = abs(j);
  k = abs(k);
  const int R = a.R;
  if (i <= R && j <= R && k <= R) return a.lgf[i + (R + 1) * (j + (R + 1) * k)];
  const double r2 = (double)i * i + (double)j * j + (double)k * k, r = sqrt(r2);
  const double i4 = (double)i * i * i * i, j4 = (double)j * j * j * j, k4 = (double)k * k * k * k;
  return (1 + (5 * (i4 + j4 + k4) / (r2 * r2) - 3) / (8 * r2)) / (4 * 3.14159265358979323846 * r);
}
__global__ void __launch_bounds__(256) k_open_green(OpenGreen a) {
  const long long n = (long long)a.M[0] * a.M[1] * a.M[2];
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int x = (int)(t % a.M[0]), y = (int)((t / a.M[0]) % a.M[1]), z = (int)(t / ((long long)a.M[0] * a.M[1]));
  const int ox = x <= a.M[0] / 2 ? x : x - a.M[0], oy = y <= a.M[1] / 2 ? y : y - a.M[1], oz = z <= a.M[2] / 2 ? z : z - a.M[2];
  a.out[t] = a.fac * lgf_at(a, ox, oy, oz);
}
hipError_t launch_open_green(const OpenGreen &a, hipStream_t s) {
  const long long n = (long long)a.M[0] * a.M[1] * a.M[2];
  TDGL_LAUNCH(k_open_green, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a);
  return hipGetLastError();
}

// ---------------------------------------------------------------- h_δ on the faces (_face_field!, :175-188)
__global__ void __launch_bounds__(256) k_open_face(Geo g, OpenFace a) {
  if (!a.ofl[OF_RECOMP]) return;
  const int i = 1 + blockIdx.x * blockDim.x + threadIdx.x;
  const int j = 1 + blockIdx.y;
  const int k = 1 + blockIdx.z;
  const int Nx = g.Nx, Ny = g.Ny, Nz = g.Nz;
  if (i > Nx + 1) return;
  // θ_δ on lattice c at link (i, j, k), compact (n1, n2, n3)
  auto th = [&](int c, int ii, int jj, int kk) {
    return a.theta[c][(ii - 1) + a.n[c][0] * ((jj - 1) + a.n[c][1] * (kk - 1))];
  };
  const int q = g.at(i, j, k);
  // hz (Nx, Ny, Nz+1): faces j ∈ {1, Ny} and i ∈ {1, Nx}
  if (i <= Nx && j <= Ny && (j == 1 || j == Ny || i == 1 || i == Nx))
    a.hnew[2][q] = ((th(1, i + 1, j, k) - th(1, i, j, k)) - th(0, i, j + 1, k) + th(0, i, j, k)) / (a.dx * a.dy);
  // hx (Nx+1, Ny, Nz): faces j ∈ {1, Ny} and k ∈ {1, Nz}
  if (j <= Ny && k <= Nz && (j == 1 || j == Ny || k == 1 || k == Nz))
    a.hnew[0][q] = ((th(2, i, j + 1, k) - th(2, i, j, k)) - th(1, i, j, k + 1) + th(1, i, j, k)) / (a.dy * a.dz);
  // hy (Nx, Ny+1, Nz): faces k ∈ {1, Nz} and i ∈ {1, Nx}
  if (i <= Nx && k <= Nz && (k == 1 || k == Nz || i == 1 || i == Nx))
    a.hnew[1][q] = ((th(0, i, j, k + 1) - th(0, i, j, k)) - th(2, i + 1, j, k) + th(2, i, j, k)) / (a.dz * a.dx);
}
hipError_t launch_open_face(const Geo &g, const OpenFace &a, hipStream_t s) {
  if (g_rec) return hipErrorNotSupported;
  TDGL_LAUNCH(k_open_face, dim3(cdiv(g.Nx + 1, 64), g.Ny + 1, g.Nz + 1), dim3(64), 0, s, g, a);
  return hipGetLastError();
}

}  // namespace tdgl
