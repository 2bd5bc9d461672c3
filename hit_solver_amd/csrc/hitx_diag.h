// HIT diagnostics and snapshot fields on the device, in FP64 whatever the
// storage type of the state:
//   OOP/hit2d.py:471-548 compute_diagnostics (means, fluctuation moments,
//     turbulent Mach number, divergence/vorticity rms),
//   OOP/hit2d.py:146-155 + 459-468 spectral derivatives, divergence_and_vorticity,
//   OOP/hit2d.py:429-453 large_scale_kinetic_fraction,
//   OOP/hit2d.py:370-399 + 452-456 source_power(large_scale_drag_source),
//   OOP/hit2d.py:571-596 save_snapshot fields (rho, u, v, p, divergence, vorticity).
//
// Spectral derivative.  numpy's ifft2(1j*k*fft2(f)).real along one axis of
// length n equals, in exact arithmetic, the circular convolution
//   (df)_j = sum_l d((j - l) mod n) f_l,
//   d(0) = 0,  d(m) = (2 pi / (n h)) * (-1)^m / 2 * cot(pi m / n)   (n even),
//              d(m) = (2 pi / (n h)) * (-1)^m / 2 * csc(pi m / n)   (n odd):
// the k-sum of i k e^{ikx} with the Nyquist mode dropped (for real f its
// term is purely imaginary and .real discards it).  The convolution is
// evaluated directly, tiled through LDS (n^2 per line; the diagnostics run
// every few tens of steps).
//
// Low-pass band (0 < |k| <= kmax, OOP/hit2d.py:352-367) by explicit
// projection onto the band's modes: a row DFT along x for |kx| <= kDiagMaxK
// in the first pass, the y sums per mode in a second, pointwise synthesis
// in the moment pass.  Means are sum / N, as numpy's mean.
#pragma once

#include "hitx_kernels.h"

namespace hitx {

constexpr int kDiagMaxK = 4;       // largest |kx|, |ky| of a low-pass mode
constexpr int kDiagMaxModes = 64;  // modes of the low-pass band
constexpr int kDiagF = 4;          // projected fields: u, v, rho*u, rho*v
constexpr int kDiagS = 6;          // first-pass sums: rho, u, v, p, c, u^2+v^2
constexpr int kDiagRowW = kDiagS + kDiagF * (kDiagMaxK + 1) * 2;
constexpr int kDiagMomW = 9;       // moment-pass sums (HITX_DIAG_* 6..14)
constexpr int kDiagPad = 32;       // offset of the extended derivative tables

struct DiagModes {
  int n;
  int kxmax;  // largest |kx| in the list
  int kymax;
  int8_t kx[kDiagMaxModes], ky[kDiagMaxModes];
};

// Pass 1, one block per row y: primitives to the field planes, the row's
// sums and its DFT along x, exp(-2 pi i kx x / nx), of u, v, q1, q2.
template <class TQ>
__global__ void __launch_bounds__(kBlock)
    k_diag_fields(Params<double> P, const TQ* __restrict__ q, double* __restrict__ rho_o, double* __restrict__ u_o,
                  double* __restrict__ v_o, double* __restrict__ p_o, double* __restrict__ rowp, int kxmax) {
  const int nx = P.nx, N = nx * P.ny, y = blockIdx.x;
  double acc[kDiagRowW];
#pragma unroll
  for (int k = 0; k < kDiagRowW; ++k) acc[k] = 0.0;
  for (int x = threadIdx.x; x < nx; x += kBlock) {
    const int idx = y * nx + x;
    const double a = (double)q[idx], b = (double)q[N + idx], c = (double)q[2 * N + idx],
                 d = (double)q[3 * N + idx];
    const Prim<double> pr = prim(P, a, b, c, d);
    rho_o[idx] = pr.rho;
    u_o[idx] = pr.u;
    v_o[idx] = pr.v;
    p_o[idx] = pr.p;
    acc[0] += pr.rho;
    acc[1] += pr.u;
    acc[2] += pr.v;
    acc[3] += pr.p;
    acc[4] += sqrt(P.gamma * pr.p / pr.rho);
    acc[5] += pr.u * pr.u + pr.v * pr.v;
    const double f[kDiagF] = {pr.u, pr.v, b, c};
#pragma unroll
    for (int k = 0; k <= kDiagMaxK; ++k) {
      if (k > kxmax) break;
      double s, co;
      sincospi(2.0 * (double)((k * x) % nx) / (double)nx, &s, &co);
#pragma unroll
      for (int j = 0; j < kDiagF; ++j) {
        acc[kDiagS + (j * (kDiagMaxK + 1) + k) * 2] += f[j] * co;
        acc[kDiagS + (j * (kDiagMaxK + 1) + k) * 2 + 1] -= f[j] * s;
      }
    }
  }
  block_reduce_write<kDiagRowW, 0>(acc, rowp + (size_t)y * kDiagRowW);
}

// Pass 2: block j = field * M.n + mode sums the rows' x-DFT times
// exp(-2 pi i ky y / ny) over y (a negative kx takes the conjugate row value).
__global__ void __launch_bounds__(kBlock)
    k_diag_modes(const double* __restrict__ rowp, int ny, DiagModes M, double* __restrict__ modes) {
  const int j = blockIdx.x, f = j / M.n, m = j - f * M.n;
  const int kx = M.kx[m], ky = M.ky[m], ax = kx < 0 ? -kx : kx;
  const int col = kDiagS + (f * (kDiagMaxK + 1) + ax) * 2;
  double v[2] = {0.0, 0.0};
  for (int y = threadIdx.x; y < ny; y += kBlock) {
    const double rr = rowp[(size_t)y * kDiagRowW + col];
    const double ri = kx < 0 ? -rowp[(size_t)y * kDiagRowW + col + 1] : rowp[(size_t)y * kDiagRowW + col + 1];
    const int r = (int)((((long)ky * y) % ny + ny) % ny);
    double s, co;
    sincospi(2.0 * (double)r / (double)ny, &s, &co);
    v[0] += rr * co + ri * s;
    v[1] += ri * co - rr * s;
  }
  block_reduce_write<2, 0>(v, modes + 2 * j);
}

// Spectral derivative along AXIS (0: x, lines are rows; 1: y, lines are
// columns) of two planes a and b.  ACC = false writes oa = da, ob = db;
// ACC = true (AXIS 1 only) updates oa = oa + da/dy, ob = ob - db/dy
// (divergence du/dx + dv/dy and vorticity dv/dx - du/dy with a = v, b = u).
// dk is the extended table dk[t] = d((t - kDiagPad) mod n), t < 2 n + 2 kDiagPad.
constexpr int kDerLines = 64, kDerPos = 32, kDerPer = kDerPos / kWaves;
constexpr int kDerLds = (kDerLines * (kDerPos + 1) > kDerPos * (kDerLines + 1)) ? kDerLines * (kDerPos + 1)
                                                                              : kDerPos * (kDerLines + 1);
template <int AXIS, bool ACC>
__global__ void __launch_bounds__(kBlock)
    k_diag_deriv(int nx, int ny, const double* __restrict__ dk, const double* __restrict__ a,
                 const double* __restrict__ b, double* __restrict__ oa, double* __restrict__ ob) {
  __shared__ double sa[kDerLds], sb[kDerLds];
  const int n = AXIS ? ny : nx, nl = AXIS ? nx : ny;
  const int l0 = blockIdx.x * kDerLines, p0 = blockIdx.y * kDerPos;
  const int tl = threadIdx.x & 63;
  const int g = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int pg = p0 + g * kDerPer;  // first output position of this wave
  double ca[kDerPer], cb[kDerPer];
#pragma unroll
  for (int i = 0; i < kDerPer; ++i) ca[i] = cb[i] = 0.0;
  for (int c0 = 0; c0 < n; c0 += kDerPos) {
    for (int e = threadIdx.x; e < kDerPos * kDerLines; e += kBlock) {
      const int pi = AXIS ? e / kDerLines : e % kDerPos;
      const int li = AXIS ? e % kDerLines : e / kDerPos;
      const int gl = l0 + li, gp = c0 + pi;
      double va = 0.0, vb = 0.0;
      if (gl < nl && gp < n) {
        const size_t idx = AXIS ? (size_t)gp * nx + gl : (size_t)gl * nx + gp;
        va = a[idx];
        vb = b[idx];
      }
      sa[pi * (kDerLines + 1) + li] = va;
      sb[pi * (kDerLines + 1) + li] = vb;
    }
    __syncthreads();
#pragma unroll 4
    for (int j = 0; j < kDerPos; ++j) {
      const double va = sa[j * (kDerLines + 1) + tl], vb = sb[j * (kDerLines + 1) + tl];
      // d((pg + i) - (c0 + j)), uniform across the wave
      const double* dp = dk + (pg - c0 - j + n + kDiagPad);
#pragma unroll
      for (int i = 0; i < kDerPer; ++i) {
        ca[i] += dp[i] * va;
        cb[i] += dp[i] * vb;
      }
    }
    __syncthreads();
  }
  const int gl = l0 + tl;
  if (AXIS) {
    if (gl < nl) {
#pragma unroll
      for (int i = 0; i < kDerPer; ++i) {
        const int p = pg + i;
        if (p < n) {
          const size_t idx = (size_t)p * nx + gl;
          if (ACC) {
            oa[idx] = oa[idx] + ca[i];
            ob[idx] = ob[idx] - cb[i];
          } else {
            oa[idx] = ca[i];
            ob[idx] = cb[i];
          }
        }
      }
    }
  } else {
    // stage through LDS so the row stores are contiguous
#pragma unroll
    for (int i = 0; i < kDerPer; ++i) {
      sa[tl * (kDerPos + 1) + g * kDerPer + i] = ca[i];
      sb[tl * (kDerPos + 1) + g * kDerPer + i] = cb[i];
    }
    __syncthreads();
    for (int e = threadIdx.x; e < kDerPos * kDerLines; e += kBlock) {
      const int pi = e % kDerPos, li = e / kDerPos;
      const int y = l0 + li, x = p0 + pi;
      if (y < nl && x < n) {
        oa[(size_t)y * nx + x] = sa[li * (kDerPos + 1) + pi];
        ob[(size_t)y * nx + x] = sb[li * (kDerPos + 1) + pi];
      }
    }
  }
}

// Pass 3, one block per row y: fluctuation moments about the means, the
// divergence and vorticity squares, the low-pass kinetic energy, the drag
// energy source u*s1 + v*s2 with s = -c * lowpass(momentum), and the count
// of mask nodes.  sums[1], sums[2] are the first pass's sums of u and v.
__global__ void __launch_bounds__(kBlock)
    k_diag_moments(int nx, int ny, const double* __restrict__ rho_i,
                   const double* __restrict__ u_i, const double* __restrict__ v_i,
                   const double* __restrict__ div_i, const double* __restrict__ vort_i,
                   const uint8_t* __restrict__ mask, const double* __restrict__ sums,
                   const double* __restrict__ modes, DiagModes M, double neg_coef, double* __restrict__ rowm) {
  __shared__ double fm[kDiagF][kDiagMaxModes][2];
  __shared__ double ey[2 * kDiagMaxK + 1][2];
  const int N = nx * ny, y = blockIdx.x;
  for (int i = threadIdx.x; i < kDiagF * M.n; i += kBlock) {
    const int f = i / M.n, m = i - f * M.n;
    fm[f][m][0] = modes[2 * i];
    fm[f][m][1] = modes[2 * i + 1];
  }
  if (threadIdx.x <= 2 * kDiagMaxK) {
    const int k = threadIdx.x - kDiagMaxK;
    const int r = (int)((((long)k * y) % ny + ny) % ny);
    double s, co;
    sincospi(2.0 * (double)r / (double)ny, &s, &co);
    ey[threadIdx.x][0] = co;
    ey[threadIdx.x][1] = s;
  }
  __syncthreads();
  const double ub = sums[1] / (double)N, vb = sums[2] / (double)N;
  double acc[kDiagMomW];
#pragma unroll
  for (int k = 0; k < kDiagMomW; ++k) acc[k] = 0.0;
  for (int x = threadIdx.x; x < nx; x += kBlock) {
    const int idx = y * nx + x;
    const double rho = rho_i[idx], u = u_i[idx], v = v_i[idx];
    const double uf = u - ub, vf = v - vb;
    acc[0] += rho * (uf * uf);
    acc[1] += rho * (vf * vf);
    acc[2] += (rho * uf) * vf;
    acc[3] += rho * (uf * uf + vf * vf);
    const double dv_ = div_i[idx], w = vort_i[idx];
    acc[4] += dv_ * dv_;
    acc[5] += w * w;
    if (M.n > 0) {
      double ex[kDiagMaxK + 1][2];
#pragma unroll
      for (int k = 0; k <= kDiagMaxK; ++k) {
        if (k > M.kxmax) break;
        sincospi(2.0 * (double)((k * x) % nx) / (double)nx, &ex[k][1], &ex[k][0]);
      }
      double lo[kDiagF] = {0.0, 0.0, 0.0, 0.0};
      for (int m = 0; m < M.n; ++m) {
        const int kx = M.kx[m], ky = M.ky[m], ax = kx < 0 ? -kx : kx;
        double xr = 0.0, xi = 0.0;
#pragma unroll
        for (int k = 0; k <= kDiagMaxK; ++k)
          if (k == ax) {
            xr = ex[k][0];
            xi = ex[k][1];
          }
        if (kx < 0) xi = -xi;
        const double yr = ey[ky + kDiagMaxK][0], yi = ey[ky + kDiagMaxK][1];
        const double pr = xr * yr - xi * yi, pi = xr * yi + xi * yr;  // exp(+i k.x)
#pragma unroll
        for (int f = 0; f < kDiagF; ++f) lo[f] += fm[f][m][0] * pr - fm[f][m][1] * pi;
      }
#pragma unroll
      for (int f = 0; f < kDiagF; ++f) lo[f] /= (double)N;
      acc[6] += rho * (lo[0] * lo[0] + lo[1] * lo[1]);
      const double s1 = neg_coef * lo[2], s2 = neg_coef * lo[3];
      acc[7] += u * s1 + v * s2;
    }
    if (mask) acc[8] += (double)mask[idx];
  }
  block_reduce_write<kDiagMomW, 0>(acc, rowm + (size_t)y * kDiagMomW);
}

// ---------------------------------------------------------------- Riemann
// Run diagnostics and output fields of a (non-periodic) Riemann state:
//   OOP/problems/riemann_config3.py:432-452 validate_raw_physical_state (raw
//     rho = q0 and p from q without floors: finiteness and positivity),
//   :459-463 vorticity_z = dv/dx - du/dy of the floored primitives by the
//     finite differences of :268-275 (central inside, one-sided at the ends),
//   :629-650 _diagnostics (min/max of raw rho and p, max |vorticity|, WENO
//     mask count), :497-519 the save_riemann_npz planes rho, u, v, p, omega.
// min and max propagate NaN, as numpy's min and max do.
constexpr int kRDiagW = 10;  // HITX_RDIAG_* slots

__device__ __forceinline__ double nan_min(double a, double b) { return (a != a) ? a : ((b != b) ? b : fmin(a, b)); }
__device__ __forceinline__ double nan_max(double a, double b) { return (a != a) ? a : ((b != b) ? b : fmax(a, b)); }

// slot k: 0 and 2 minima, 1, 3 and 4 maxima, the rest sums
__device__ __forceinline__ double rdiag_op(int k, double a, double b) {
  return (k == 0 || k == 2) ? nan_min(a, b) : (k <= 4 ? nan_max(a, b) : a + b);
}

__device__ __forceinline__ void rdiag_block_write(double (&v)[kRDiagW], double* out) {
  __shared__ double sh[kWaves][kRDiagW];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < kRDiagW; ++k) {
    double a = v[k];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) a = rdiag_op(k, a, __shfl_xor(a, off, 64));
    if (lane == 0) sh[w][k] = a;
  }
  __syncthreads();
  if (threadIdx.x < kRDiagW) {
    const int k = threadIdx.x;
    double a = sh[0][k];
    for (int j = 1; j < kWaves; ++j) a = rdiag_op(k, a, sh[j][k]);
    out[k] = a;
  }
}

__device__ __forceinline__ void rdiag_init(double (&v)[kRDiagW]) {
  v[0] = v[2] = __builtin_inf();
  v[1] = v[3] = -__builtin_inf();
  v[4] = 0.0;  // |omega| >= 0
#pragma unroll
  for (int k = 5; k < kRDiagW; ++k) v[k] = 0.0;
}

template <class TQ>
__device__ __forceinline__ void rd_uv(const Params<double>& P, const TQ* __restrict__ q, int N, int idx, double& u,
                                      double& v) {
  const double rho = npmax((double)q[idx], P.rho_floor);
  u = (double)q[N + idx] / rho;
  v = (double)q[2 * N + idx] / rho;
}

template <class TQ>
__global__ void __launch_bounds__(kBlock)
    k_riemann_diag(Params<double> P, const TQ* __restrict__ q, const uint8_t* __restrict__ mask,
                   double* __restrict__ fields, double* __restrict__ part) {
  const int nx = P.nx, ny = P.ny, N = nx * ny;
  double acc[kRDiagW];
  rdiag_init(acc);
  for (int idx = blockIdx.x * kBlock + threadIdx.x; idx < N; idx += gridDim.x * kBlock) {
    const int y = idx / nx, x = idx - y * nx;
    const double a = (double)q[idx], b = (double)q[N + idx], c = (double)q[2 * N + idx], d = (double)q[3 * N + idx];
    // raw density and pressure (no floors)
    const double ur = b / a, vr = c / a;
    const double pr = P.gm1 * (d - (0.5 * a) * (ur * ur + vr * vr));
    acc[0] = nan_min(acc[0], a);
    acc[1] = nan_max(acc[1], a);
    acc[2] = nan_min(acc[2], pr);
    acc[3] = nan_max(acc[3], pr);
    const bool fin_q = isfinite(a) && isfinite(b) && isfinite(c) && isfinite(d);
    acc[5] += fin_q ? 0.0 : 1.0;
    acc[6] += (isfinite(a) && isfinite(pr)) ? 0.0 : 1.0;
    acc[7] += (a > 0.0) ? 0.0 : 1.0;  // not positive (NaN included)
    acc[8] += (pr > 0.0) ? 0.0 : 1.0;
    if (mask) acc[9] += (double)mask[idx];
    // vorticity of the floored primitives
    double ul, vl, uh, vh, t;
    double dvdx, dudy;
    if (x == 0) {
      rd_uv(P, q, N, idx, t, vl);
      rd_uv(P, q, N, idx + 1, t, vh);
      dvdx = (vh - vl) / P.dx;
    } else if (x == nx - 1) {
      rd_uv(P, q, N, idx - 1, t, vl);
      rd_uv(P, q, N, idx, t, vh);
      dvdx = (vh - vl) / P.dx;
    } else {
      rd_uv(P, q, N, idx - 1, t, vl);
      rd_uv(P, q, N, idx + 1, t, vh);
      dvdx = (vh - vl) / (2.0 * P.dx);
    }
    if (y == 0) {
      rd_uv(P, q, N, idx, ul, t);
      rd_uv(P, q, N, idx + nx, uh, t);
      dudy = (uh - ul) / P.dy;
    } else if (y == ny - 1) {
      rd_uv(P, q, N, idx - nx, ul, t);
      rd_uv(P, q, N, idx, uh, t);
      dudy = (uh - ul) / P.dy;
    } else {
      rd_uv(P, q, N, idx - nx, ul, t);
      rd_uv(P, q, N, idx + nx, uh, t);
      dudy = (uh - ul) / (2.0 * P.dy);
    }
    const double om = dvdx - dudy;
    acc[4] = nan_max(acc[4], fabs(om));
    if (fields) {
      const Prim<double> pm = prim(P, a, b, c, d);
      fields[idx] = pm.rho;
      fields[N + idx] = pm.u;
      fields[2 * N + idx] = pm.v;
      fields[3 * N + idx] = pm.p;
      fields[4 * N + (size_t)idx] = om;
    }
  }
  rdiag_block_write(acc, part + (size_t)blockIdx.x * kRDiagW);
}

__global__ void __launch_bounds__(kBlock)
    k_riemann_diag_finalize(const double* __restrict__ part, int nblk, double* __restrict__ out) {
  double acc[kRDiagW];
  rdiag_init(acc);
  for (int b = threadIdx.x; b < nblk; b += kBlock)
#pragma unroll
    for (int k = 0; k < kRDiagW; ++k) acc[k] = rdiag_op(k, acc[k], part[(size_t)b * kRDiagW + k]);
  rdiag_block_write(acc, out);
}

}  // namespace hitx
