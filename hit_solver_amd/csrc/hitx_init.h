// Initial states built on the device (setup; the reference builds them with
// NumPy on the host):
//   OOP/hit2d.py:158-204 initialize_hit_2d: psi = sum_m a_m cos(kx_m x + ky_m y
//     + phase_m) at the cell centres (modes and phases in the reference's loop
//     and draw order, drawn on the host), u = dpsi/dy, v = -dpsi/dx by the
//     spectral derivative (the periodic convolution of hitx_diag.h), mean
//     removal, scale target_mach * mean(c) / rms(u, v) with rho = 1,
//     p = 1/gamma, conservative variables with the floors of
//     OOP/equations.py:147-163;
//   OOP/problems/riemann_config3.py:60-99 (and the configurations 3 at 0.8
//     and 6) quadrant states: top = y >= y_split, right = x >= x_split at the
//     cell centres, conservative variables with the same floors.
// Everything is formed in double and stored in the plan's dtype.
#pragma once

#include "hitx_diag.h"

namespace hitx {

constexpr int kInitModeW = 4;  // kx, ky, amplitude, phase

// psi at every node, modes accumulated in list order as the reference's loop.
__global__ void __launch_bounds__(kBlock)
    k_init_psi(int nx, int ny, double dx, double dy, double x_min, double y_min, const double* __restrict__ modes,
               int n_modes, double* __restrict__ psi) {
  const int N = nx * ny;
  for (int idx = blockIdx.x * kBlock + threadIdx.x; idx < N; idx += gridDim.x * kBlock) {
    const int j = idx / nx, i = idx - j * nx;
    const double X = x_min + dx * ((double)i + 0.5);
    const double Y = y_min + dy * ((double)j + 0.5);
    double s = 0.0;
    for (int m = 0; m < n_modes; ++m) {
      const double* md = modes + (size_t)m * kInitModeW;
      s = s + md[2] * cos(md[0] * X + md[1] * Y + md[3]);
    }
    psi[idx] = s;
  }
}

// sums of u = dy and v = -dx (slots 0, 1), and of the uniform sound speed
// sqrt(gamma p0 / 1) (slot 2)
__global__ void __launch_bounds__(kBlock)
    k_init_sums(int N, const double* __restrict__ dpx, const double* __restrict__ dpy, double c0,
                double* __restrict__ part) {
  double v[3] = {0.0, 0.0, 0.0};
  for (int idx = blockIdx.x * kBlock + threadIdx.x; idx < N; idx += gridDim.x * kBlock) {
    v[0] += dpy[idx];
    v[1] += -dpx[idx];
    v[2] += c0;
  }
  block_reduce_write<3, 0>(v, part + (size_t)blockIdx.x * 3);
}

// sum of u'^2 + v'^2 with u' = u - sum_u/N, v' = v - sum_v/N
__global__ void __launch_bounds__(kBlock)
    k_init_rms(int N, const double* __restrict__ dpx, const double* __restrict__ dpy,
               const double* __restrict__ sums, double* __restrict__ part) {
  const double ub = sums[0] / (double)N, vb = sums[1] / (double)N;
  double v[1] = {0.0};
  for (int idx = blockIdx.x * kBlock + threadIdx.x; idx < N; idx += gridDim.x * kBlock) {
    const double u = dpy[idx] - ub, w = -dpx[idx] - vb;
    v[0] += u * u + w * w;
  }
  block_reduce_write<1, 0>(v, part + blockIdx.x);
}

// out[0] = rms(u', v'), out[1] = target_mach * mean(c) / rms
__global__ void k_init_scale(int N, const double* __restrict__ sums, double target_mach,
                             double* __restrict__ out) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    const double rms = sqrt(sums[3] / (double)N);
    out[0] = rms;
    out[1] = target_mach * (sums[2] / (double)N) / rms;
  }
}

template <class T>
__global__ void __launch_bounds__(kBlock)
    k_init_hit_write(int N, const double* __restrict__ dpx, const double* __restrict__ dpy,
                     const double* __restrict__ sums, const double* __restrict__ scal, double gamma, double rho0,
                     double p0, T* __restrict__ q) {
  const double ub = sums[0] / (double)N, vb = sums[1] / (double)N, sc = scal[1];
  for (int idx = blockIdx.x * kBlock + threadIdx.x; idx < N; idx += gridDim.x * kBlock) {
    const double u = (dpy[idx] - ub) * sc, v = (-dpx[idx] - vb) * sc;
    q[idx] = (T)rho0;
    q[N + idx] = (T)(rho0 * u);
    q[2 * N + idx] = (T)(rho0 * v);
    q[3 * N + idx] = (T)(p0 / (gamma - 1.0) + (0.5 * rho0) * (u * u + v * v));
  }
}

// Quadrant state: rows of quad are (rho, u, v, p) for I top-right,
// II top-left, III bottom-left, IV bottom-right.
struct Quadrants {
  double s[4][4];
};

template <class T>
__global__ void __launch_bounds__(kBlock)
    k_init_riemann(int nx, int ny, double dx, double dy, double x_min, double y_min, double x_split,
                   double y_split, Quadrants Qd, double gamma, double rho_floor, double p_floor,
                   T* __restrict__ q) {
  const int N = nx * ny;
  for (int idx = blockIdx.x * kBlock + threadIdx.x; idx < N; idx += gridDim.x * kBlock) {
    const int j = idx / nx, i = idx - j * nx;
    const double X = x_min + dx * ((double)i + 0.5);
    const double Y = y_min + dy * ((double)j + 0.5);
    const bool top = Y >= y_split, right = X >= x_split;
    const int k = top ? (right ? 0 : 1) : (right ? 3 : 2);
    const double rho = npmax(Qd.s[k][0], rho_floor), u = Qd.s[k][1], v = Qd.s[k][2];
    const double p = npmax(Qd.s[k][3], p_floor);
    q[idx] = (T)rho;
    q[N + idx] = (T)(rho * u);
    q[2 * N + idx] = (T)(rho * v);
    q[3 * N + idx] = (T)(p / (gamma - 1.0) + (0.5 * rho) * (u * u + v * v));
  }
}

}  // namespace hitx
