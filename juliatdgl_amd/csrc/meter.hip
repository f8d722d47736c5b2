// meter.hip — a ladder's per-cycle measurements on the device (gtdgl/src/protocol.jl:416-464): the
// voltmeter plane sums and θ line integral, the unwrapped path phase of every band across the
// voltmeter, the current through the mid plane, and the largest change of |Ψ| and of hz since the
// last cycle.  The host reads a record of a few doubles per cycle instead of the whole state.
//
// One launch of k_meter does all the reductions, block by block in four kinds of blocks (lines,
// block nodes, hz, the mid plane), each writing its partials; k_meter_finish sums them in block
// order (the same record for the same state, bit for bit) and keeps the rung's running dmax/nbig.
#include "kernels.h"

namespace tdgl {

namespace {

constexpr double TWO_PI = 6.283185307179586;

// Julia's max: NaN if either is NaN
__device__ __forceinline__ double jmax(double a, double b) { return (a != a || b != b) ? a + b : (a > b ? a : b); }

__device__ __forceinline__ double wave_sum_m(double v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o);
  return v;
}
__device__ __forceinline__ double wave_max_m(double v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v = jmax(v, __shfl_xor(v, o));
  return v;
}

// d − 2π round(d / 2π), round half away from zero (oround_f, geometry.jl:13)
__device__ __forceinline__ double wrap2pi(double d) { return d - TWO_PI * round(d / TWO_PI); }

}  // namespace

// line blocks: one wavefront per line (j, k), 4 per block
__device__ void meter_lines(const MeterArgs &a, double *part, int blk) {
  __shared__ double sl[4][METER_Q];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int nl = a.nl1 * a.nl2, L = blk * 4 + w;
  double q[METER_Q];
#pragma unroll
  for (int t = 0; t < METER_Q; ++t) q[t] = 0.0;
  if (L < nl) {
    const int jj = L % a.nl1, kk = L / a.nl1;
    const int j = a.g.y1 + jj, k = a.g.z1 + kk;
    double D[TDGL_MAX_BANDS];
    for (int mu = 0; mu < a.N; ++mu) {
      double s = 0.0;
      for (int ii = a.vi + lane; ii < a.vo; ii += 64) {
        const int c = a.g.at(ii, j, k);
        s += dangle(cmul(cmul(cconj(a.G[mu][c]), a.Vx[c]), a.G[mu][c + 1]));  // (conj(G) V) G[+x]
      }
      D[mu] = wave_sum_m(s);
    }
    double ta = 0.0;
    for (int ii = a.vi + lane; ii < a.vo; ii += 64) ta += a.phx[a.g.at(ii, j, k)];
    ta = wave_sum_m(ta);
    if (lane == 0) {
      const bool ok = a.lok[L] != 0, f1 = a.lf1[L] != 0;
      if (a.Min[L]) q[METER_PIN] = a.phx[a.g.at(a.vi, j, k)];
      if (a.Mou[L]) q[METER_POUT] = a.phx[a.g.at(a.vo, j, k)];
      if (ok) q[METER_TA] = ta;
      if (a.first) {
        for (int mu = 0; mu < a.N; ++mu) {
          a.Dprev[mu][L] = D[mu];
          a.T[mu][L] = 0.0;
        }
      } else {
        const double d1 = wrap2pi(D[0] - a.Dprev[0][L]);
        if (ok) {
          q[METER_DMAX] = fabs(d1);
          q[METER_NBIG] = fabs(d1) > 2.0 ? 1.0 : 0.0;
        }
        for (int mu = 0; mu < a.N; ++mu) {
          const double t = a.T[mu][L] + wrap2pi(D[mu] - a.Dprev[mu][L]);
          a.T[mu][L] = t;
          a.Dprev[mu][L] = D[mu];
          if (ok) {
            q[METER_T + mu] = t;
            q[METER_TABS + mu] = fabs(t);
          }
        }
        if (f1) q[METER_TF1] = a.T[0][L];
      }
    }
  }
  if (lane == 0)
#pragma unroll
    for (int t = 0; t < METER_Q; ++t) sl[w][t] = q[t];
  __syncthreads();
  if (threadIdx.x < METER_Q) {
    const int t = threadIdx.x;
    double v;
    if (t == METER_DMAX)
      v = jmax(jmax(sl[0][t], sl[1][t]), jmax(sl[2][t], sl[3][t]));
    else
      v = (sl[0][t] + sl[1][t]) + (sl[2][t] + sl[3][t]);
    part[(size_t)blk * METER_Q + t] = v;
  }
}

// block-node blocks: |Ψ_µ| against the last cycle's (or, at the start, only taken)
__device__ void meter_nodes(const MeterArgs &a, double *part, int blk) {
  __shared__ double sm[4][TDGL_MAX_BANDS + 1];
  const int nbx = a.g.x2 - a.g.x1 + 2, nby = a.g.y2 - a.g.y1 + 2, nbz = a.g.z2 - a.g.z1 + 2;
  const int nb = nbx * nby * nbz;
  const int t = blk * 256 + threadIdx.x;
  double dif[TDGL_MAX_BANDS] = {0.0, 0.0, 0.0, 0.0}, amax = 0.0;
  if (t < nb) {
    const int i = a.g.x1 + t % nbx, j = a.g.y1 + (t / nbx) % nby, k = a.g.z1 + t / (nbx * nby);
    const int c = a.g.at(i, j, k);
    for (int mu = 0; mu < a.N; ++mu) {
      const cplx v = a.G[mu][c];
      const double m = hypot(v.x, v.y);  // abs(::ComplexF64)
      if (!a.init) {
        dif[mu] = fabs(m - a.cooper[mu][t]);
        amax = jmax(amax, m);
      }
      a.cooper[mu][t] = m;
    }
  }
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int mu = 0; mu < TDGL_MAX_BANDS; ++mu) dif[mu] = wave_max_m(dif[mu]);
  amax = wave_max_m(amax);
  if (lane == 0) {
    for (int mu = 0; mu < TDGL_MAX_BANDS; ++mu) sm[w][mu] = dif[mu];
    sm[w][TDGL_MAX_BANDS] = amax;
  }
  __syncthreads();
  if (threadIdx.x <= TDGL_MAX_BANDS) {
    const int q = threadIdx.x;
    part[(size_t)blk * METER_Q + q] = jmax(jmax(sm[0][q], sm[1][q]), jmax(sm[2][q], sm[3][q]));
  }
}

// hz blocks: the largest change of hz (Nx, Ny, Nz+1) since the last cycle
__device__ void meter_hz(const MeterArgs &a, double *part, int blk) {
  __shared__ double sm[4];
  const int Nx = a.g.Nx, Ny = a.g.Ny, Nz = a.g.Nz;
  const int n = Nx * Ny * (Nz + 1);
  const int t = blk * 256 + threadIdx.x;
  double d = 0.0;
  if (t < n) {
    const int c = a.g.at(1 + t % Nx, 1 + (t / Nx) % Ny, 1 + t / (Nx * Ny));
    const double v = a.hz[c];
    if (!a.init) d = fabs(v - a.hold[c]);
    a.hold[c] = v;
  }
  d = wave_max_m(d);
  if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = d;
  __syncthreads();
  if (threadIdx.x == 0) part[(size_t)blk * METER_Q] = jmax(jmax(sm[0], sm[1]), jmax(sm[2], sm[3]));
}

// the mid-plane block (midplane_current, observables.jl:102-114): one block, fixed order
__device__ void meter_mid(const MeterArgs &a, double *part) {
  __shared__ double sm[2][4];
  const int Ny = a.g.Ny, Nz = a.g.Nz, ii = a.mid;
  const int n = (Ny - 1) * (Nz - 1);
  double it = 0.0, is = 0.0;
  for (int t = threadIdx.x; t < n; t += 256) {
    const int jj = 2 + t % (Ny - 1), kk = 2 + t / (Ny - 1);
    const int c = a.g.at(ii, jj, kk);
    const int sy = a.g.px, sz = a.g.px * a.g.py;
    const double jx = a.ks2 * ((a.hz[c] - a.hz[c - sy]) / a.dy - (a.hy[c] - a.hy[c - sz]) / a.dz);
    it += jx;
    if (a.g.y1 <= jj && jj <= a.g.y2 + 1 && a.g.z1 <= kk && kk <= a.g.z2 + 1) is += jx;
  }
  it = wave_sum_m(it);
  is = wave_sum_m(is);
  if ((threadIdx.x & 63) == 0) {
    sm[0][threadIdx.x >> 6] = it;
    sm[1][threadIdx.x >> 6] = is;
  }
  __syncthreads();
  if (threadIdx.x < 2) {
    const int q = threadIdx.x;
    part[q] = ((sm[q][0] + sm[q][1]) + (sm[q][2] + sm[q][3])) * a.dy * a.dz;
  }
}

__device__ __forceinline__ void meter_body(const MeterArgs &a, int b) {
  if (b < a.nlb) {
    if (!a.init) meter_lines(a, a.part, b);
    return;
  }
  b -= a.nlb;
  if (b < a.nnb) {
    meter_nodes(a, a.part + (size_t)a.nlb * METER_Q, b);
    return;
  }
  b -= a.nnb;
  if (b < a.nhb) {
    meter_hz(a, a.part + (size_t)(a.nlb + a.nnb) * METER_Q, b);
    return;
  }
  if (!a.init) meter_mid(a, a.part + (size_t)(a.nlb + a.nnb + a.nhb) * METER_Q);
}

__global__ void __launch_bounds__(256) k_meter(MeterArgs a) { meter_body(a, blockIdx.x); }

// The record from the partials, each quantity summed by one thread in block order; the rung's
// dmax and nbig carried in rung[0..1] (zero at its first cycle).
__device__ __forceinline__ void meter_finish_body(const MeterArgs &a, double *rec, double *rung) {
  const int t = threadIdx.x;
  const double *pl = a.part, *pn = a.part + (size_t)a.nlb * METER_Q, *ph = a.part + (size_t)(a.nlb + a.nnb) * METER_Q,
               *pm = a.part + (size_t)(a.nlb + a.nnb + a.nhb) * METER_Q;
  __shared__ double q[METER_Q];
  if (t < METER_Q) {
    double v = 0.0;
    for (int b = 0; b < a.nlb; ++b) v = t == METER_DMAX ? jmax(v, pl[(size_t)b * METER_Q + t]) : v + pl[(size_t)b * METER_Q + t];
    q[t] = v;
  }
  __syncthreads();
  if (t == 0) {
    const double nin = a.nin, nou = a.nou, nlin = a.nlin;
    rec[MR_V] = q[METER_POUT] / nou - q[METER_PIN] / nin;
    rec[MR_TA] = a.nlin > 0 ? q[METER_TA] / nlin : 0.0;
    rec[MR_TF1] = a.nlin_f1 > 0 ? q[METER_TF1] / (double)a.nlin_f1 : 0.0;
    for (int mu = 0; mu < TDGL_MAX_BANDS; ++mu) {
      rec[MR_T + mu] = mu < a.N ? q[METER_T + mu] / nlin : 0.0;
      rec[MR_TABS + mu] = mu < a.N ? q[METER_TABS + mu] : 0.0;
    }
    if (a.first) {
      rung[0] = 0.0;
      rung[1] = 0.0;
    } else {
      rung[0] = jmax(rung[0], q[METER_DMAX]);
      rung[1] += q[METER_NBIG];
    }
    rec[MR_DMAX] = rung[0];
    rec[MR_NBIG] = rung[1];
    rec[MR_IT] = pm[0];
    rec[MR_IS] = pm[1];
    double dif[TDGL_MAX_BANDS] = {0.0, 0.0, 0.0, 0.0}, amax = 0.0, dh = 0.0;
    for (int b = 0; b < a.nnb; ++b) {
      for (int mu = 0; mu < TDGL_MAX_BANDS; ++mu) dif[mu] = jmax(dif[mu], pn[(size_t)b * METER_Q + mu]);
      amax = jmax(amax, pn[(size_t)b * METER_Q + TDGL_MAX_BANDS]);
    }
    for (int b = 0; b < a.nhb; ++b) dh = jmax(dh, ph[(size_t)b * METER_Q]);
    for (int mu = 0; mu < TDGL_MAX_BANDS; ++mu) rec[MR_DIF + mu] = dif[mu];
    rec[MR_AMAX] = amax;
    rec[MR_DIFFH] = dh;
  }
}

__global__ void k_meter_finish(MeterArgs a, double *rec, double *rung) { meter_finish_body(a, rec, rung); }

// batched: run r owns blocks off[r] .. off[r+1]-1 of k_meter's grid, and block r of the finish
__global__ void __launch_bounds__(256) k_meter_b(BatchTab t) {
  const int b = blockIdx.x;
  int lo = 0, hi = t.n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (t.off[mid] <= b)
      lo = mid;
    else
      hi = mid - 1;
  }
  meter_body(((const MeterArgs *)t.args)[lo], b - t.off[lo]);
}
__global__ void k_meter_finish_b(BatchTab t) {
  const MeterFinArgs &f = ((const MeterFinArgs *)t.args)[blockIdx.x];
  meter_finish_body(f.a, f.rec, f.rung);
}

int meter_blocks(const MeterArgs &a) { return a.nlb + a.nnb + a.nhb + 1; }

hipError_t launch_meter(const MeterArgs &a, hipStream_t s) {
  if (g_rec) return hipErrorNotSupported;
  TDGL_LAUNCH(k_meter, dim3(meter_blocks(a)), dim3(256), 0, s, a);
  return hipGetLastError();
}
hipError_t launch_meter_finish(const MeterArgs &a, double *rec, double *rung, hipStream_t s) {
  if (g_rec) return hipErrorNotSupported;
  TDGL_LAUNCH(k_meter_finish, dim3(1), dim3(64), 0, s, a, rec, rung);
  return hipGetLastError();
}

dim3 meter_batch_grid(int kind, const void *args) {
  if (kind == BK_METER) return dim3(meter_blocks(*(const MeterArgs *)args));
  return dim3(1);
}
hipError_t launch_batch_meter(int kind, const BatchTab &t, int blocks, hipStream_t s) {
  if (kind == BK_METER) {
    TDGL_LAUNCH(k_meter_b, dim3(blocks), dim3(256), 0, s, t);
  } else if (kind == BK_METERFIN) {
    TDGL_LAUNCH(k_meter_finish_b, dim3(blocks), dim3(64), 0, s, t);
  } else {
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace tdgl

