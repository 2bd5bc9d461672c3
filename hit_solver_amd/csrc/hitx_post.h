// Post-processing of HIT snapshot planes on the device: shell-averaged
// energy and enstrophy spectra (reference OOP/postprocess/spectra.py:48-149)
// and two-point auto/cross correlations (OOP/postprocess/
// two_point_correlation.py:107-139, 141-232).  All arithmetic is double.
//
// Transforms are direct discrete Fourier transforms along lines staged in
// LDS, first along x, then along y: out[k] = sum_j in[j] W[(k j) mod n],
// W[m] = exp(-+2 pi i m / n) from a per-call table, the index advanced by k
// modulo n (exact argument reduction, no accumulated phase error).  The
// reference calls numpy's FFT; a direct sum agrees with it to rounding
// (O(n eps) against O(log n eps)), which the diagnostics' tolerances cover.
#pragma once

#include "hitx_device.h"

namespace hitx {

constexpr int kDftThreads = 256;
constexpr int kPostBlock = 256;

// Twiddle table W[m] = (cos(2 pi m / n), sin(2 pi m / n)), m < n.
__global__ void __launch_bounds__(kPostBlock) k_dft_table(int n, double* __restrict__ tw) {
  const int m = blockIdx.x * kPostBlock + threadIdx.x;
  if (m >= n) return;
  double s, c;
  sincospi(2.0 * (double)m / (double)n, &s, &c);
  tw[2 * m] = c;
  tw[2 * m + 1] = s;
}

// One line of n values per block (complex, or real when REAL_IN), read with
// element stride es from in + line * ls, written complex with element stride
// eo to out + line * lo.  sign -1: forward (exp(-i)); +1: inverse (exp(+i)),
// then scaled by `scale`.  LDS: the line and the twiddle table, 32 n bytes.
template <bool REAL_IN>
__global__ void __launch_bounds__(kDftThreads) k_dft_lines(int n, int nlines, const double* __restrict__ in,
                                                           long long ls, long long es, double* __restrict__ out,
                                                           long long lo, long long eo, const double* __restrict__ tw,
                                                           int sign, double scale) {
  extern __shared__ __align__(16) double dsm[];
  double* xr = dsm;          // [n] real parts of the line
  double* xi = dsm + n;      // [n] imaginary parts
  double* wc = dsm + 2 * n;  // [n] cos
  double* ws = dsm + 3 * n;  // [n] sin (signed for the direction)
  const int line = blockIdx.x;
  if (line >= nlines) return;
  const double* src = in + (size_t)line * ls;
  for (int j = threadIdx.x; j < n; j += kDftThreads) {
    if (REAL_IN) {
      xr[j] = src[(size_t)j * es];
      xi[j] = 0.0;
    } else {
      xr[j] = src[(size_t)j * es * 2];
      xi[j] = src[(size_t)j * es * 2 + 1];
    }
    wc[j] = tw[2 * j];
    ws[j] = sign < 0 ? -tw[2 * j + 1] : tw[2 * j + 1];
  }
  __syncthreads();
  double* dst = out + (size_t)line * lo;
  for (int k = threadIdx.x; k < n; k += kDftThreads) {
    double ar = 0.0, ai = 0.0;
    int m = 0;  // (k j) mod n
    for (int j = 0; j < n; ++j) {
      const double c = wc[m], s = ws[m], a = xr[j], b = xi[j];
      ar += a * c - b * s;
      ai += a * s + b * c;
      m += k;
      m = m >= n ? m - n : m;
    }
    dst[(size_t)k * eo * 2] = ar * scale;
    dst[(size_t)k * eo * 2 + 1] = ai * scale;
  }
}

// ---------------------------------------------------------------- fields
// Per-block partial sums of the means the fluctuations need: rho, u, v,
// rho*u, rho*v, p, omega, theta (slots PM_*), a fixed tree per block.
enum { PM_RHO = 0, PM_U, PM_V, PM_RU, PM_RV, PM_P, PM_W, PM_T, PM_COUNT };

__device__ __forceinline__ void block_sum_write(double (&v)[PM_COUNT], double* __restrict__ out) {
  __shared__ double red[PM_COUNT][kPostBlock];
  for (int k = 0; k < PM_COUNT; ++k) red[k][threadIdx.x] = v[k];
  __syncthreads();
  for (int s = kPostBlock / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s)
      for (int k = 0; k < PM_COUNT; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x < PM_COUNT) out[threadIdx.x] = red[threadIdx.x][0];
}

__global__ void __launch_bounds__(kPostBlock) k_post_partials(int N, const double* __restrict__ rho,
                                                              const double* __restrict__ u,
                                                              const double* __restrict__ v,
                                                              const double* __restrict__ p,
                                                              const double* __restrict__ w,
                                                              const double* __restrict__ t,
                                                              double* __restrict__ part) {
  double acc[PM_COUNT] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int i = blockIdx.x * kPostBlock + threadIdx.x; i < N; i += gridDim.x * kPostBlock) {
    const double r = rho[i], a = u[i], b = v[i];
    acc[PM_RHO] += r;
    acc[PM_U] += a;
    acc[PM_V] += b;
    acc[PM_RU] += r * a;
    acc[PM_RV] += r * b;
    if (p) acc[PM_P] += p[i];
    if (w) acc[PM_W] += w[i];
    if (t) acc[PM_T] += t[i];
  }
  block_sum_write(acc, part + (size_t)blockIdx.x * PM_COUNT);
}

// Means from the partials (one block, fixed order), in the slots PM_*.
__global__ void __launch_bounds__(kPostBlock) k_post_means(int nblk, int N, const double* __restrict__ part,
                                                           double* __restrict__ means) {
  double acc[PM_COUNT] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int b = threadIdx.x; b < nblk; b += kPostBlock)
    for (int k = 0; k < PM_COUNT; ++k) acc[k] += part[(size_t)b * PM_COUNT + k];
  __shared__ double m[PM_COUNT];
  block_sum_write(acc, m);
  __syncthreads();
  if (threadIdx.x < PM_COUNT) means[threadIdx.x] = m[threadIdx.x] / (double)N;
}

// Planes out[8][N]: sqrt(rho) u', sqrt(rho) v', u', v', rho', p', omega',
// theta' (Reynolds or Favre velocity fluctuations, spectra.py:203-214);
// planes of absent inputs are left untouched.
__global__ void __launch_bounds__(kPostBlock) k_post_fields(int N, int favre, const double* __restrict__ rho,
                                                            const double* __restrict__ u,
                                                            const double* __restrict__ v,
                                                            const double* __restrict__ p,
                                                            const double* __restrict__ w,
                                                            const double* __restrict__ t,
                                                            const double* __restrict__ means,
                                                            double* __restrict__ out) {
  const double rm = means[PM_RHO];
  const double um = favre ? means[PM_RU] / rm : means[PM_U];
  const double vm = favre ? means[PM_RV] / rm : means[PM_V];
  for (int i = blockIdx.x * kPostBlock + threadIdx.x; i < N; i += gridDim.x * kPostBlock) {
    const double r = rho[i], uf = u[i] - um, vf = v[i] - vm, sr = sqrt(r);
    out[i] = sr * uf;
    out[(size_t)N + i] = sr * vf;
    out[2 * (size_t)N + i] = uf;
    out[3 * (size_t)N + i] = vf;
    out[4 * (size_t)N + i] = r - rm;
    if (p) out[5 * (size_t)N + i] = p[i] - means[PM_P];
    if (w) out[6 * (size_t)N + i] = w[i] - means[PM_W];
    if (t) out[7 * (size_t)N + i] = t[i] - means[PM_T];
  }
}

// Second-order central divergence and vorticity on the periodic grid
// (two_point_correlation.py:472-485), for snapshots without those planes.
__global__ void __launch_bounds__(kPostBlock) k_post_divvort(int nx, int ny, double dx, double dy,
                                                             const double* __restrict__ u,
                                                             const double* __restrict__ v,
                                                             double* __restrict__ div, double* __restrict__ vort) {
  const int N = nx * ny;
  for (int i = blockIdx.x * kPostBlock + threadIdx.x; i < N; i += gridDim.x * kPostBlock) {
    const int y = i / nx, x = i - y * nx;
    const int xr = x + 1 == nx ? 0 : x + 1, xl = x == 0 ? nx - 1 : x - 1;
    const int yr = y + 1 == ny ? 0 : y + 1, yl = y == 0 ? ny - 1 : y - 1;
    const double du_dx = (u[y * nx + xr] - u[y * nx + xl]) / (2.0 * dx);
    const double du_dy = (u[yr * nx + x] - u[yl * nx + x]) / (2.0 * dy);
    const double dv_dx = (v[y * nx + xr] - v[y * nx + xl]) / (2.0 * dx);
    const double dv_dy = (v[yr * nx + x] - v[yl * nx + x]) / (2.0 * dy);
    div[i] = du_dx + dv_dy;
    vort[i] = dv_dx - du_dy;
  }
}

// ---------------------------------------------------------------- shells
// Wavenumber of index m of an n-point transform with spacing h, as
// 2 pi * np.fft.fftfreq(n, h)[m].
__device__ __forceinline__ double fftfreq_k(int m, int n, double h) {
  const int N2 = (n - 1) / 2 + 1;
  const double f = (double)(m < N2 ? m : m - n) * (1.0 / ((double)n * h));
  return 2.0 * M_PI * f;
}

// One row (fixed ky) per block: the modes' shell index rint(|k| / dk) and
// energy / enstrophy densities in LDS, then one thread per shell sums its
// modes.  Along a row the shell index is non-decreasing over the
// non-negative frequencies (m <= nx/2) and non-increasing over the negative
// ones, so each shell's modes are two index ranges, found by binary search
// and summed in index order.  Row partials hist[row][2][nshell] (energy,
// enstrophy); shells beyond nshell - 1 are dropped (spectra.py:107-126).
__global__ void __launch_bounds__(kPostBlock) k_shell_rows(int nx, int ny, double dx, double dy, int nshell,
                                                           const double* __restrict__ hat,
                                                           double* __restrict__ hist) {
  extern __shared__ __align__(16) double ssm[];
  double* de = ssm;                                   // [nx]
  double* dz = ssm + nx;                              // [nx]
  int* sh = reinterpret_cast<int*>(ssm + 2 * nx);     // [nx]
  const int y = blockIdx.x;
  const size_t N = (size_t)nx * ny;
  const double ky = fftfreq_k(y, ny, dy);
  const double dk = fmin(2.0 * M_PI / (nx * dx), 2.0 * M_PI / (ny * dy));
  const double norm = (double)N * (double)N;
  for (int m = threadIdx.x; m < nx; m += kPostBlock) {
    const double kx = fftfreq_k(m, nx, dx);
    const size_t o = ((size_t)y * nx + m) * 2;
    const double wur = hat[o], wui = hat[o + 1];                  // sqrt(rho) u'
    const double wvr = hat[2 * N + o], wvi = hat[2 * N + o + 1];  // sqrt(rho) v'
    const double ur = hat[4 * N + o], ui = hat[4 * N + o + 1];    // u'
    const double vr = hat[6 * N + o], vi = hat[6 * N + o + 1];    // v'
    de[m] = 0.5 * ((wur * wur + wui * wui) + (wvr * wvr + wvi * wvi)) / norm;
    const double orr = kx * vr - ky * ur, oi = kx * vi - ky * ui;  // omega_hat = 1j * (kx v - ky u)
    dz[m] = 0.5 * (orr * orr + oi * oi) / norm;
    sh[m] = (int)rint(sqrt(kx * kx + ky * ky) / dk);
  }
  __syncthreads();
  const int half = nx / 2;  // indices 0 .. half: non-decreasing; half+1 .. nx-1: non-increasing
  for (int s = threadIdx.x; s < nshell; s += kPostBlock) {
    double se = 0.0, sz = 0.0;
    // first index >= s in [0, half], then advance while == s
    int lo = 0, hi = half + 1;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (sh[mid] < s) lo = mid + 1; else hi = mid;
    }
    for (int m = lo; m <= half && sh[m] == s; ++m) {
      se += de[m];
      sz += dz[m];
    }
    // [half+1, nx): first index whose shell <= s, then advance while == s
    lo = half + 1;
    hi = nx;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (sh[mid] > s) lo = mid + 1; else hi = mid;
    }
    for (int m = lo; m < nx && sh[m] == s; ++m) {
      se += de[m];
      sz += dz[m];
    }
    hist[((size_t)y * 2) * nshell + s] = se;
    hist[((size_t)y * 2 + 1) * nshell + s] = sz;
  }
}

// Sum of the row partials per shell, rows in order.
__global__ void __launch_bounds__(kPostBlock) k_shell_sum(int ny, int nshell, const double* __restrict__ hist,
                                                          double* __restrict__ energy,
                                                          double* __restrict__ enstrophy) {
  const int s = blockIdx.x * kPostBlock + threadIdx.x;
  if (s >= nshell) return;
  double e = 0.0, z = 0.0;
  for (int y = 0; y < ny; ++y) {
    e += hist[((size_t)y * 2) * nshell + s];
    z += hist[((size_t)y * 2 + 1) * nshell + s];
  }
  energy[s] = e;
  enstrophy[s] = z;
}

// ---------------------------------------------------------------- correlations
// The maps of one correlation call (at most kCorrMaxMaps): the two planes
// of each product and its normalisation -- 0 none, 1 its own zero lag, 2
// sqrt(R_aa(0) R_bb(0)) of the two maps named in ref[m].
constexpr int kCorrMaxMaps = 16;
struct CorrMaps {
  int n;
  int pair[kCorrMaxMaps][2];
  int mode[kCorrMaxMaps];
  int ref[kCorrMaxMaps][2];
};

// prod[m] = hat[a_m] * conj(hat[b_m]) for the maps (complex planes).
__global__ void __launch_bounds__(kPostBlock) k_cross_spectra(int N, CorrMaps M, const double* __restrict__ hat,
                                                              double* __restrict__ prod) {
  for (size_t i = blockIdx.x * (size_t)kPostBlock + threadIdx.x; i < (size_t)N * M.n;
       i += (size_t)gridDim.x * kPostBlock) {
    const int mp = (int)(i / N);
    const size_t o = i - (size_t)mp * N;
    const double* a = hat + (size_t)M.pair[mp][0] * N * 2;
    const double* b = hat + (size_t)M.pair[mp][1] * N * 2;
    const double ar = a[2 * o], ai = a[2 * o + 1], br = b[2 * o], bi = b[2 * o + 1];
    prod[(size_t)mp * N * 2 + 2 * o] = ar * br + ai * bi;
    prod[(size_t)mp * N * 2 + 2 * o + 1] = ai * br - ar * bi;
  }
}

// Real part of the inverse transforms divided by N (the correlation), then
// normalised: mode 1 by its own zero-lag value (or the unit impulse when that
// is within eps of zero, two_point_correlation.py:113-120), mode 2 by
// sqrt(mean(a^2) mean(b^2)), here sqrt(R_aa(0) R_bb(0)) of two maps of the
// same call (Parseval; the zero map when within eps, :130-138).
__global__ void __launch_bounds__(kPostBlock) k_corr_finish(int N, CorrMaps M, const double* __restrict__ inv,
                                                            double* __restrict__ out) {
  constexpr double eps = 2.220446049250313e-16;
  for (size_t i = blockIdx.x * (size_t)kPostBlock + threadIdx.x; i < (size_t)N * M.n;
       i += (size_t)gridDim.x * kPostBlock) {
    const int mp = (int)(i / N);
    const size_t o = i - (size_t)mp * N;
    const double r = inv[((size_t)mp * N + o) * 2] / (double)N;
    double v = r;
    if (M.mode[mp] == 1) {
      const double z = inv[(size_t)mp * N * 2] / (double)N;
      v = fabs(z) > eps ? r / z : (o == 0 ? 1.0 : 0.0);
    } else if (M.mode[mp] == 2) {
      const double za = inv[(size_t)M.ref[mp][0] * N * 2] / (double)N;
      const double zb = inv[(size_t)M.ref[mp][1] * N * 2] / (double)N;
      const double sc = sqrt(za * zb);
      v = sc > eps ? r / sc : 0.0;
    }
    out[(size_t)mp * N + o] = v;
  }
}

}  // namespace hitx
