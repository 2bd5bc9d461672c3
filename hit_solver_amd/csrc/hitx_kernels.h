// hitx kernels: stage reductions, shock sensor, line sweeps with the compact
// solve, stage update with the HIT source closure, hyperviscosity, forcing.
//
// Layout in HBM: state q is [4][ny][nx] (x fastest), masks uint8 [ny][nx].
// The y-direction sweep runs on a transposed copy qT [4][nx][ny] so both
// sweeps stage whole lines through LDS with unit-stride global access.
#pragma once

#include "../../include/hitx.h"
#include "hitx_device.h"
#include "hitx_wave.h"

namespace hitx {

__device__ __forceinline__ int wrap(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

// Tile coordinates of this block of a 2-D tile grid in XCD-contiguous order.
struct Tile2 {
  int bx, by;
};
__device__ __forceinline__ Tile2 tile_xcd() {
  const int gx = gridDim.x, total = gx * gridDim.y;
  const int b = xcd_remap((int)(blockIdx.y * gx + blockIdx.x), total);
  return Tile2{b % gx, b / gx};
}
__device__ __forceinline__ int clampi(int i, int lo, int hi) { return i < lo ? lo : (i > hi ? hi : i); }
// g wrapped into [0, n) (periodic grids); in-range indices skip the modulo
__device__ __forceinline__ int wrapn(int g, int n) { return (unsigned)g < (unsigned)n ? g : ((g % n) + n) % n; }

template <class T>
__device__ __forceinline__ T absT(T a) { return fabs(a); }

// ================================================================ reductions
// ---------------------------------------------------------------- column walks
// A wave covers 64 adjacent columns and walks down a run of rows.  With a
// halo of one column the edge lanes only supply x-neighbours (62 columns
// owned per wave); rows are walked with the previous and next row held in
// registers, so every node's primitives are formed once (plus one halo row
// above and below the run).  Twiddles of a row are wave-uniform.
struct Walk {
  int x, xl, ya, yb;  // lane column, column loaded (wrapped or clamped), rows [ya, yb)
  bool own, live;     // lane owns column x; wave has rows
};

__device__ __forceinline__ Walk walk_setup(int nx, int ny, int periodic, int rb, int halo, int bid) {
  const int lane = threadIdx.x & 63;
  const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int gw = bid * kWaves + wid;
  const int segw = 64 - 2 * halo;
  const int nseg = (nx + segw - 1) / segw;
  const int seg = gw % nseg, rc = gw / nseg;
  Walk w;
  w.ya = rc * rb;
  w.yb = min(w.ya + rb, ny);
  w.live = w.ya < ny;
  w.x = seg * segw - halo + lane;
  w.own = w.live && lane >= halo && lane < 64 - halo && w.x < nx;
  w.xl = periodic ? ((w.x % nx) + nx) % nx : clampi(w.x, 0, nx - 1);
  return w;
}

__host__ __device__ inline int walk_blocks(int nx, int ny, int rb, int halo) {
  const int segw = 64 - 2 * halo;
  const int waves = ((nx + segw - 1) / segw) * ((ny + rb - 1) / rb);
  return (waves + kWaves - 1) / kWaves;
}

// Value of lane l-1 (dir = -1) or l+1 (dir = +1) in a wavefront: DPP
// wave_shr:1 / wave_shl:1 (one VALU op, no LDS); the end lanes receive 0.
__device__ __forceinline__ int dpp_shift(int v, int dir) {
  return dir < 0 ? __builtin_amdgcn_update_dpp(0, v, 0x138, 0xf, 0xf, false)
                 : __builtin_amdgcn_update_dpp(0, v, 0x130, 0xf, 0xf, false);
}
__device__ __forceinline__ float lane_step(float v, int dir) {
  return __int_as_float(dpp_shift(__float_as_int(v), dir));
}
__device__ __forceinline__ double lane_step(double v, int dir) {
  const long long b = __double_as_longlong(v);
  const int lo = dpp_shift((int)(b & 0xffffffffll), dir), hi = dpp_shift((int)(b >> 32), dir);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
template <class T>
__device__ __forceinline__ T lane_from(T v, int src) {
  return __shfl(v, src & 63, 64);
}

// Per-stage reductions that gate the sensor and the flux split
// (OOP/spatial_operator.py:148 eps, :271 div_rms, :279 vort_rms, :332 alpha),
// plus the HIT closure's cooling mean pressure (OOP/hit2d.py:423) and drag
// projection onto the low-k modes (OOP/hit2d.py:391-394).  Sensor inputs use
// enforce(q) when enforce_first (the operator rhs); cooling and drag use q as
// given (the RK stage state), as the reference does.
//
// Column walk (halo 1, P.rb rows per wave).  The drag projection is
// separable: each lane sums rho*u and rho*v times cos/sin(ky*y) for
// ky = 0..dky over its rows, and multiplies by exp(-i*kx*x) once at the end.
// enforce_first: 0 the state as given, 1 enforce first (OOP/spatial_operator.py
// rhs: the sensor sees enforce(q)), 2 enforce first, the state being an RK
// stage state that the kernel which wrote it already enforced.  With 2 the
// FP32 reductions take the state as given (enforcing an enforced state again
// moves a node by at most an ulp): one set of primitives per node and the
// cooling pressure shared.  FP64 enforces first as the reference does.
template <class T>
__device__ __forceinline__ int reduce_enforce_first(int enforce_first) {
  return (sizeof(T) == 4 && enforce_first == 2) ? 0 : (enforce_first != 0 ? 1 : 0);
}

template <class T, bool DRAG>
__device__ __forceinline__ void stage_reduce_body(const Params<T>& P, const T* __restrict__ q,
                                                  int enforce_first_in, double* __restrict__ part, int bid) {
  const int enforce_first = reduce_enforce_first<T>(enforce_first_in);
  constexpr int NR = DRAG ? (R_DRAG + kDragCols) : R_DRAG;
  double v[R_DRAG];
#pragma unroll
  for (int k = 0; k < R_DRAG; ++k) v[k] = 0.0;
  double bs[2][5];  // [rho u, rho v][1, cos y, sin y, cos 2y, sin 2y] column sums
#pragma unroll
  for (int f = 0; f < 2; ++f)
#pragma unroll
    for (int j = 0; j < 5; ++j) bs[f][j] = 0.0;
  const int nx = P.nx, ny = P.ny, N = nx * ny;
  const bool per = DRAG || P.periodic;  // the drag source exists on periodic grids only
  const Walk W = walk_setup(nx, ny, per, P.rb, 1, bid);
  const int lane = threadIdx.x & 63;
  auto row_of = [&](int y) { return per ? (y < 0 ? y + ny : (y >= ny ? y - ny : y)) : clampi(y, 0, ny - 1); };
  const Divisor<T> d2x = divisor(P.two_dx), d2y = divisor(P.two_dy);
  // maxima of |rho|, |p|, |e|, |u|+c, |v|+c in the state's type (widening
  // to double is exact and monotone, so the maxima are the same)
  T mx[5] = {T(0), T(0), T(0), T(0), T(0)};
  // the drag end phase's column twiddles of |kx| = 1, 2, loaded before the
  // walk so their latency is hidden behind it
  double twc[4] = {0.0, 0.0, 0.0, 0.0};
  if (DRAG) {
    const int xc = W.own ? W.x : 0;  // any column for lanes with zero sums
#pragma unroll
    for (int k = 1; k <= 2; ++k) {
      if (k <= P.dkx) {
        const double* t = P.twx + ((size_t)(k + P.kmx) * nx + xc) * 2;
        twc[2 * k - 2] = t[0];
        twc[2 * k - 1] = t[1];
      }
    }
  }
  if (W.live) {
    auto prim_e = [&](T a, T b, T c, T d) {
      if (enforce_first) enforce(P, a, b, c, d);
      return prim(P, a, b, c, d);
    };
    // Rows rotate through three register sets (primitives of rows y-1, y,
    // y+1 and the state as given of rows y, y+1, y+2; cooling and the drag
    // sums read the state as given, no reload).  The loop body handles three
    // rows with the sets' roles permuted, so no row is copied between sets.
    struct Raw {
      T a, b, c, d;
    };
    Prim<T> P0 = prim_at(P, q, N, row_of(W.ya - 1) * nx + W.xl, enforce_first), P1, P2;
    Raw R0, R1, R2;
    load4(q, N, W.ya * nx + W.xl, R1.a, R1.b, R1.c, R1.d);
    load4(q, N, row_of(W.ya + 1) * nx + W.xl, R2.a, R2.b, R2.c, R2.d);
    P1 = prim_e(R1.a, R1.b, R1.c, R1.d);
    P2 = prim_e(R2.a, R2.b, R2.c, R2.d);
    // row y with pv, cu, nn the primitives of rows y-1, y, y+1, cr the state
    // as given of row y; the loads of row y+2 go to nr (in flight during the
    // row) and its primitives to pn (the set row y-1 held)
    // the drag twiddles cos/sin(ky y) of |ky| = 1, 2 for the current row
    // (twr) and the next (twn): wave-uniform loads issued a row ahead, so
    // their latency overlaps the row instead of stalling it
    double twr[4] = {0.0, 0.0, 0.0, 0.0}, twn[4] = {0.0, 0.0, 0.0, 0.0};
    auto tw_load = [&](int y, double (&t)[4]) {
      if (!DRAG) return;
#pragma unroll
      for (int k = 1; k <= 2; ++k) {
        if (k <= P.dky) {
          const double* ty = P.twy + ((size_t)(k + P.kmy) * ny + y) * 2;
          t[2 * k - 2] = ty[0];
          t[2 * k - 1] = ty[1];
        }
      }
    };
    tw_load(W.ya, twr);
    auto row = [&](int y, Prim<T>& pv, const Prim<T>& cu, const Prim<T>& nn, const Raw& cr, Raw& nr,
                   Prim<T>& pn) {
      const bool more = y + 1 < W.yb;
      if (more) load4(q, N, row_of(y + 2) * nx + W.xl, nr.a, nr.b, nr.c, nr.d);
      if (more) tw_load(y + 1, twn);
      const T ul = lane_step(cu.u, -1), ur = lane_step(cu.u, 1);
      const T vl = lane_step(cu.v, -1), vr = lane_step(cu.v, 1);
      const T c0 = cr.a, c1 = cr.b, c2 = cr.c, c3 = cr.d;
      if (W.own) {
        const int x = W.x;
        const T cs = sqrt_alg(dv(P.gamma * cu.p, cu.rho));
        mx[0] = fmax(mx[0], absT(cu.rho));
        mx[1] = fmax(mx[1], absT(cu.p));
        const T e = dv(cu.p, cu.rho * P.gm1);
        mx[2] = fmax(mx[2], absT(e));
        mx[3] = fmax(mx[3], absT(cu.u) + cs);
        mx[4] = fmax(mx[4], absT(cu.v) + cs);
        T ux, uy, vx, vy;
        if (per) {
          ux = dv(ur - ul, d2x);
          vx = dv(vr - vl, d2x);
          uy = dv(nn.u - pv.u, d2y);
          vy = dv(nn.v - pv.v, d2y);
        } else {  // OOP/problems/riemann_config3.py:268-275 (clamped neighbours)
          const bool xin = (x > 0 && x < nx - 1), yin = (y > 0 && y < ny - 1);
          ux = dv(ur - ul, xin ? P.two_dx : P.dx);
          vx = dv(vr - vl, xin ? P.two_dx : P.dx);
          uy = dv(nn.u - pv.u, yin ? P.two_dy : P.dy);
          vy = dv(nn.v - pv.v, yin ? P.two_dy : P.dy);
        }
        // sums of squares and products accumulate with fused multiply-adds
        // (summation order is free here; the pointwise algebra above is not)
        const T div = ux + vy;
        v[R_DIV2] = fma((double)div, (double)div, v[R_DIV2]);
        if (P.shear_on) {
          const T w = vx - uy;
          v[R_VORT2] = fma((double)w, (double)w, v[R_VORT2]);
        }
        if (P.cool_on || DRAG) {
          const T a = c0, b = c1, c = c2, d = c3;
          if (P.cool_on) v[R_PSUM] += (double)(enforce_first ? prim(P, a, b, c, d).p : cu.p);
          if (DRAG) {
            const double fb = (double)b, fc = (double)c;
            bs[0][0] += fb;
            bs[1][0] += fc;
#pragma unroll
            for (int k = 1; k <= 2; ++k) {
              if (k <= P.dky) {
                const double cy = twr[2 * k - 2], sy = twr[2 * k - 1];
                bs[0][2 * k - 1] = fma(fb, cy, bs[0][2 * k - 1]);
                bs[0][2 * k] = fma(fb, sy, bs[0][2 * k]);
                bs[1][2 * k - 1] = fma(fc, cy, bs[1][2 * k - 1]);
                bs[1][2 * k] = fma(fc, sy, bs[1][2 * k]);
              }
            }
          }
        }
      }
      if (more) pn = prim_e(nr.a, nr.b, nr.c, nr.d);
#pragma unroll
      for (int k = 0; k < 4; ++k) twr[k] = twn[k];
    };
    for (int y = W.ya; y < W.yb; y += 3) {
      row(y, P0, P1, P2, R1, R0, P0);
      if (y + 1 >= W.yb) break;
      row(y + 1, P1, P2, P0, R2, R1, P1);
      if (y + 2 >= W.yb) break;
      row(y + 2, P2, P0, P1, R0, R2, P2);
    }
  }
#pragma unroll
  for (int k = 0; k < 5; ++k) v[R_MAXR + k] = (double)mx[k];
  block_reduce_write<R_DRAG, 0x1Full>(v, part + (size_t)bid * NR);
  if (DRAG) {
    // Mode (kx, ky), field f: sum over the wave's columns of
    // exp(-i kx x) * sum_y f exp(-i ky y).  The lanes' column sums and their
    // columns' twiddles of |kx| = 1, 2 go to LDS (one load per lane); lane
    // j < 2*ndrag then forms output (mode j/2, field j%2) over the 64 columns
    // (columns a wave does not own have zero sums).  cos(-a) = cos(a),
    // sin(-a) = -sin(a).
    __shared__ double cs_[kWaves][10][64];
    __shared__ double tw_[kWaves][4][64];
    __shared__ double ms_[kWaves][48];
    const int wid = threadIdx.x >> 6;
#pragma unroll
    for (int f = 0; f < 2; ++f)
#pragma unroll
      for (int j = 0; j < 5; ++j) cs_[wid][f * 5 + j][lane] = W.own ? bs[f][j] : 0.0;
#pragma unroll
    for (int k = 0; k < 4; ++k) tw_[wid][k][lane] = twc[k];
    __syncthreads();
    if (lane < 2 * P.ndrag) {
      const int m = lane >> 1, f = lane & 1;
      const int kx = P.drag_modes[2 * m], ky = P.drag_modes[2 * m + 1];
      const int ka = ky < 0 ? -ky : ky, kxa = kx < 0 ? -kx : kx;
      const int sc = f * 5 + (ka == 0 ? 0 : 2 * ka - 1);
      double re = 0.0, im = 0.0;
#pragma unroll 8
      for (int l = 0; l < 64; ++l) {
        const double cx = kxa == 0 ? 1.0 : tw_[wid][2 * kxa - 2][l];
        const double sa = kxa == 0 ? 0.0 : tw_[wid][2 * kxa - 1][l];
        const double sx = kx < 0 ? -sa : sa;
        const double br = cs_[wid][sc][l];
        const double bi = ka == 0 ? 0.0 : (ky > 0 ? -cs_[wid][sc + 1][l] : cs_[wid][sc + 1][l]);
        re += cx * br + sx * bi;
        im += cx * bi - sx * br;
      }
      ms_[wid][4 * m + 2 * f] = re;
      ms_[wid][4 * m + 2 * f + 1] = im;
    }
    __syncthreads();
    for (int j = threadIdx.x; j < 4 * P.ndrag; j += blockDim.x) {
      double a = ms_[0][j];
      for (int w = 1; w < kWaves; ++w) a += ms_[w][j];
      part[(size_t)bid * NR + R_DRAG + j] = a;
    }
  }
}

template <class T, bool DRAG>
__global__ void __launch_bounds__(kBlock) k_stage_reduce(Params<T> P, const T* __restrict__ q,
                                                         int enforce_first, double* __restrict__ part) {
  stage_reduce_body<T, DRAG>(P, q, enforce_first, part, blockIdx.x);
}

// The stage scalars from the reduced row r: thread 0 writes the scalar
// slots, the block the drag slots.
template <class T, bool DRAG>
__device__ __forceinline__ void stage_scal_write(const Params<T>& P, const double* r, double* __restrict__ scal) {
  const double N = (double)P.nx * (double)P.ny;
  if (threadIdx.x == 0) {
    scal[S_EPS_R] = 1e-14 + 1e-12 * r[R_MAXR];
    scal[S_EPS_P] = 1e-14 + 1e-12 * r[R_MAXP];
    scal[S_EPS_E] = 1e-14 + 1e-12 * r[R_MAXE];
    const double div_rms = sqrt(r[R_DIV2] / N);
    scal[S_COMP_ON] = div_rms > 1e-12 ? 1.0 : 0.0;
    scal[S_DIV_THR] = -P.comp_thr * div_rms;
    const double w_rms = sqrt(r[R_VORT2] / N);
    scal[S_SHEAR_ON] = (P.shear_on && w_rms > 1e-12) ? 1.0 : 0.0;
    scal[S_SHEAR_THR] = P.shear_thr * w_rms;
    scal[S_ALPHA_X] = r[R_AX];
    scal[S_ALPHA_Y] = r[R_AY];
    if (P.cool_on) {  // OOP/hit2d.py:422-425
      const double err = (r[R_PSUM] / N - P.cool_pref) / (P.gamma_d - 1.0);
      scal[S_COOL] = -err / P.cool_tc;
    } else {
      scal[S_COOL] = 0.0;
    }
  }
  if (DRAG)
    for (int k = threadIdx.x; k < 4 * P.ndrag; k += blockDim.x) scal[S_DRAG + k] = r[R_DRAG + k] / N;
}

// One block: reduce the partial rows in a fixed order and write the stage
// scalars (a separate launch: kernel boundaries order the partial writes).
template <class T, bool DRAG>
__device__ __forceinline__ void stage_finalize_body(const Params<T>& P, const double* __restrict__ part, int nblk,
                                                    double* __restrict__ scal) {
  constexpr int NR = DRAG ? (R_DRAG + kDragCols) : R_DRAG;
  __shared__ double r[NR];
  reduce_partials<NR, 0x1Full>(part, nblk, r);
  __syncthreads();
  stage_scal_write<T, DRAG>(P, r, scal);
}

// Per-row table of the drag low-pass (one thread per row and field): for row
// y and field f (rho u, rho v), G(kx) = sum over modes with that kx of
// A(m)*exp(i ky y), A = the projection / N, and drows[y][f] = {Re G(0),
// Re(G(1)+G(-1)), Im(G(1)-G(-1)), Re(G(2)+G(-2)), Im(G(2)-G(-2))}, so the
// low-pass field at (x, y) is H0 + sum_k (Hc_k cos(k x) - Hs_k sin(k x)) (the
// real part of the inverse transform, OOP/hit2d.py:366-367).
// dc: the drag coefficients, [mode][field][re, im] (the stage scalars from S_DRAG).
template <class T>
__device__ __forceinline__ void drows_body(const Params<T>& P, const double* __restrict__ dc,
                                           double* __restrict__ drows, int t) {
  if (t >= 2 * P.ny) return;
  const int y = t >> 1, f = t & 1;
  double h[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  for (int m = 0; m < P.ndrag; ++m) {
    const int kx = P.drag_modes[2 * m], ky = P.drag_modes[2 * m + 1];
    const double ar = dc[4 * m + 2 * f], ai = dc[4 * m + 2 * f + 1];
    const double* ty = P.twy + ((size_t)(ky + P.kmy) * P.ny + y) * 2;
    const double gr = ar * ty[0] - ai * ty[1], gi = ar * ty[1] + ai * ty[0];
    const double gs = kx > 0 ? gi : -gi;
    if (kx == 0) {
      h[0] += gr;
    } else if (kx == 1 || kx == -1) {
      h[1] += gr;
      h[2] += gs;
    } else {
      h[3] += gr;
      h[4] += gs;
    }
  }
#pragma unroll
  for (int j = 0; j < 5; ++j) drows[((size_t)y * 2 + f) * 5 + j] = h[j];
}

template <class T>
__global__ void __launch_bounds__(kBlock) k_drows(Params<T> P, const double* __restrict__ scal,
                                                  double* __restrict__ drows) {
  drows_body<T>(P, scal + S_DRAG, drows, blockIdx.x * kBlock + threadIdx.x);
}

// The drag rows with the finalize's drag part folded in: every block reduces
// the drag columns of the column-walk partial rows itself (the finalize's
// per-column order), block 0 also writes the drag slots of the stage scalars.
template <class T>
__device__ __forceinline__ void drows_fin_body(const Params<T>& P, const double* __restrict__ part, int nblk,
                                               double* __restrict__ scal, double* __restrict__ drows, int t,
                                               bool first) {
  __shared__ double r[kDragCols];
  __shared__ double dc[kDragCols];
  reduce_partial_cols<kDragCols, 0>(part, nblk, R_DRAG + kDragCols, R_DRAG, r);
  __syncthreads();
  const double N = (double)P.nx * (double)P.ny;
  for (int k = threadIdx.x; k < 4 * P.ndrag; k += blockDim.x) {
    dc[k] = r[k] / N;
    if (first) scal[S_DRAG + k] = dc[k];
  }
  __syncthreads();
  drows_body<T>(P, dc, drows, t);
}
template <class T>
__global__ void __launch_bounds__(kBlock) k_drows_fin(Params<T> P, const double* __restrict__ part, int nblk,
                                                      double* __restrict__ scal, double* __restrict__ drows) {
  drows_fin_body<T>(P, part, nblk, scal, drows, blockIdx.x * kBlock + threadIdx.x, blockIdx.x == 0);
}

template <class T, bool DRAG>
__global__ void __launch_bounds__(kBlock) k_stage_finalize(Params<T> P, const double* __restrict__ part, int nblk,
                                                           double* __restrict__ scal) {
  stage_finalize_body<T, DRAG>(P, part, nblk, scal);
}

// Generic column reduction: out[k] = sum over nblk rows of part[row*nr + k]
// (one block per slot k; fixed order).
__global__ void __launch_bounds__(kBlock) k_finalize_sum(const double* __restrict__ part, int nblk, int nr,
                                                         double* __restrict__ out) {
  const int k = blockIdx.x;
  double v[1] = {0.0};
  for (int b = threadIdx.x; b < nblk; b += blockDim.x) v[0] += part[(size_t)b * nr + k];
  __shared__ double r[1];
  block_reduce_write<1, 0>(v, r);
  __syncthreads();
  if (threadIdx.x == 0) out[k] = r[0];
}

__global__ void __launch_bounds__(kBlock) k_finalize_max0(const double* __restrict__ part, int nblk, int nr,
                                                          double* __restrict__ out) {
  // slot 0 is a max, the others sums (step reductions)
  const int k = blockIdx.x;
  double v[1] = {0.0};
  for (int b = threadIdx.x; b < nblk; b += blockDim.x) {
    const double a = part[(size_t)b * nr + k];
    v[0] = k == 0 ? fmax(v[0], a) : v[0] + a;
  }
  __shared__ double r[1];
  if (k == 0)
    block_reduce_write<1, 1>(v, r);
  else
    block_reduce_write<1, 0>(v, r);
  __syncthreads();
  if (threadIdx.x == 0) out[k] = r[0];
}

// Step-level reductions: max spectral radius (OOP/spatial_operator.py:369-377,
// OOP/problems/riemann_config3.py:371-378), sum(u^2+v^2) and sum(c) for the
// turbulent Mach number (OOP/hit2d.py:279-290).
template <class T>
__device__ __forceinline__ void step_reduce_body(const Params<T>& P, const T* __restrict__ q,
                                                        double* __restrict__ part, int bid, int nblk) {
  double v[3] = {0.0, 0.0, 0.0};
  const int N = P.nx * P.ny, stride = nblk * kBlock;
  auto node = [&](T a, T b, T cc, T d) {
    const Prim<T> c = prim(P, a, b, cc, d);
    const T cs = sqrt_alg(dv(P.gamma * c.p, c.rho));
    T sr = dv(absT(c.u) + cs, P.dx);
    sr = sr + dv(absT(c.v) + cs, P.dy);
    v[0] = fmax(v[0], (double)sr);
    v[1] += (double)(c.u * c.u + c.v * c.v);
    v[2] += (double)cs;
  };
  int idx = bid * kBlock + threadIdx.x;
  // four nodes' loads in flight together, accumulated in node order
  for (; idx + 3 * stride < N; idx += 4 * stride) {
    T a[4], b[4], c[4], d[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) load4(q, N, idx + j * stride, a[j], b[j], c[j], d[j]);
#pragma unroll
    for (int j = 0; j < 4; ++j) node(a[j], b[j], c[j], d[j]);
  }
  for (; idx < N; idx += stride) {
    T a, b, c, d;
    load4(q, N, idx, a, b, c, d);
    node(a, b, c, d);
  }
  block_reduce_write<3, 1>(v, part + (size_t)bid * 3);
}

template <class T>
__global__ void __launch_bounds__(kBlock) k_step_reduce(Params<T> P, const T* __restrict__ q,
                                                        double* __restrict__ part) {
  step_reduce_body<T>(P, q, part, blockIdx.x, gridDim.x);
}

// ================================================================ sensor
// Relative jump of f at a node along one axis: periodic form
// OOP/spatial_operator.py:147-166, non-periodic form
// OOP/problems/riemann_config3.py:128-146 (pair jumps credited to both nodes,
// clamped curvature).  fl/fr are the left/right neighbours (clamped in the
// non-periodic case), at_lo/at_hi flag the domain ends.
template <class T>
__device__ __forceinline__ T rel_jump(T fl, T f, T fr, T eps, bool periodic, bool at_lo, bool at_hi) {
  const T cv = dv(absT((fr - T(2.0) * f) + fl), (((absT(fr) + T(2.0) * absT(f)) + absT(fl)) + eps));
  if (periodic) {
    const T jr = dv(absT(fr - f), ((absT(fr) + absT(f)) + eps));
    const T jl = dv(absT(f - fl), ((absT(f) + absT(fl)) + eps));
    return npmax(npmax(jl, jr), cv);
  }
  T j = T(0);
  if (!at_hi) j = npmax(j, dv(absT(fr - f), ((absT(fr) + absT(f)) + eps)));
  if (!at_lo) j = npmax(j, dv(absT(f - fl), ((absT(f) + absT(fl)) + eps)));
  return npmax(j, cv);
}

// FP32 form of "rel_jump(...) > t": each quotient a/b (b > 0) compared as
// a > t*b, without the reciprocals (three quarter-rate ops per call); it
// differs from the quotient form only by rounding, which the FP32 mode's
// error budget covers (FP64 keeps the quotients for bit parity).
__device__ __forceinline__ bool jump_over(float fl, float f, float fr, float eps, float t, bool periodic, bool at_lo,
                                          bool at_hi) {
  const float af = fabsf(f), afl = fabsf(fl), afr = fabsf(fr);
  bool o = fabsf((fr - 2.0f * f) + fl) > t * (((afr + 2.0f * af) + afl) + eps);
  if (periodic || !at_hi) o = o || fabsf(fr - f) > t * ((afr + af) + eps);
  if (periodic || !at_lo) o = o || fabsf(f - fl) > t * ((af + afl) + eps);
  return o;
}

// Whether the relative jump of density, pressure or internal energy at a node
// exceeds the threshold: a node's value and its x and y neighbours of each
// field (OOP/spatial_operator.py:283-303, OOP/problems/riemann_config3.py:309-318).
template <class T>
struct Jn {
  T xl, c, xr, yl, yr;  // left / centre / right along x, left / right along y
};
template <class T>
__device__ __forceinline__ bool jumps_over(const Jn<T>& r, const Jn<T>& p, const Jn<T>& e, T er, T ep, T ee,
                                           double thr, bool per, bool xlo, bool xhi, bool ylo, bool yhi) {
  if constexpr (std::is_same<T, float>::value) {
    const float t = (float)thr;
    return jump_over(r.xl, r.c, r.xr, er, t, per, xlo, xhi) || jump_over(r.yl, r.c, r.yr, er, t, per, ylo, yhi) ||
           jump_over(p.xl, p.c, p.xr, ep, t, per, xlo, xhi) || jump_over(p.yl, p.c, p.yr, ep, t, per, ylo, yhi) ||
           jump_over(e.xl, e.c, e.xr, ee, t, per, xlo, xhi) || jump_over(e.yl, e.c, e.yr, ee, t, per, ylo, yhi);
  } else {
    const T jr = npmax(rel_jump(r.xl, r.c, r.xr, er, per, xlo, xhi), rel_jump(r.yl, r.c, r.yr, er, per, ylo, yhi));
    const T jp = npmax(rel_jump(p.xl, p.c, p.xr, ep, per, xlo, xhi), rel_jump(p.yl, p.c, p.yr, ep, per, ylo, yhi));
    const T je = npmax(rel_jump(e.xl, e.c, e.xr, ee, per, xlo, xhi), rel_jump(e.yl, e.c, e.yr, ee, per, ylo, yhi));
    return (double)jr > thr || (double)jp > thr || (double)je > thr;
  }
}

// Fused sensor on a 32x16 output tile: primitives of (enforced) q on the tile
// plus a (w+1) halo in LDS, undilated mask on the tile plus a w halo, the
// separable (2w+1)^2 OR dilation, mask and maskT stores, and (when qT is not
// null) the enforced state written transposed for the y sweep.
// Same per-node arithmetic as k_rawmask; out-of-domain nodes of the
// non-periodic grid are clamped copies for the neighbour gathers and false
// in the dilation, as in OOP/problems/riemann_config3.py:128-162.
constexpr int kSTX = 32, kSTY = 16;
template <class T>
__device__ __forceinline__ void sensor_body(const Params<T>& P, const T* __restrict__ q, int enforce_first,
                                                   const double* __restrict__ scal, uint8_t* __restrict__ mask,
                                                   uint8_t* __restrict__ maskT, T* __restrict__ qT, int bx, int by) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int nx = P.nx, ny = P.ny, N = nx * ny, w = P.width, H = w + 1;
  const int EX = kSTX + 2 * H, EY = kSTY + 2 * H, RX = kSTX + 2 * w, RY = kSTY + 2 * w;
  const int x0 = bx * kSTX, y0 = by * kSTY;
  const bool per = P.periodic;
  T* Lr = reinterpret_cast<T*>(smem);
  T* Lu = Lr + EX * EY;
  T* Lv = Lu + EX * EY;
  T* Lp = Lv + EX * EY;
  T* Le = Lp + EX * EY;  // internal energy p/(rho (gamma-1))
  uint8_t* raw = reinterpret_cast<uint8_t*>(Le + EX * EY);
  uint8_t* mid = raw + RX * RY;
  uint8_t* outt = mid + RY * kSTX;  // [kSTY][kSTX + 1]
  for (int t = threadIdx.x; t < EX * EY; t += kBlock) {
    const int r = t / EX, c = t - r * EX;
    int gy = y0 - H + r, gx = x0 - H + c;
    if (per) {
      gy = wrapn(gy, ny);
      gx = wrapn(gx, nx);
    } else {
      gy = clampi(gy, 0, ny - 1);
      gx = clampi(gx, 0, nx - 1);
    }
    const Prim<T> pr = prim_at(P, q, N, gy * nx + gx, enforce_first);
    Lr[t] = pr.rho;
    Lu[t] = pr.u;
    Lv[t] = pr.v;
    Lp[t] = pr.p;
    Le[t] = dv(pr.p, pr.rho * P.gm1);
  }
  __syncthreads();
  const T er = (T)scal[S_EPS_R], ep = (T)scal[S_EPS_P], ee = (T)scal[S_EPS_E];
  const bool comp_on = scal[S_COMP_ON] != 0.0, shear_on = scal[S_SHEAR_ON] != 0.0;
  const double div_thr = scal[S_DIV_THR], shear_thr = scal[S_SHEAR_THR], thr = P.jump_thr;
  for (int t = threadIdx.x; t < RX * RY; t += kBlock) {
    const int r = t / RX, c = t - r * RX;
    const int gy = y0 - w + r, gx = x0 - w + c;
    bool m = false;
    if (per || (gx >= 0 && gx < nx && gy >= 0 && gy < ny)) {
      const int ic = (r + 1) * EX + (c + 1);
      const int ixl = ic - 1, ixr = ic + 1, iyl = ic - EX, iyr = ic + EX;
      const bool xin = per || (gx > 0 && gx < nx - 1), yin = per || (gy > 0 && gy < ny - 1);
      const T ux = dv(Lu[ixr] - Lu[ixl], xin ? P.two_dx : P.dx);
      const T vx = dv(Lv[ixr] - Lv[ixl], xin ? P.two_dx : P.dx);
      const T uy = dv(Lu[iyr] - Lu[iyl], yin ? P.two_dy : P.dy);
      const T vy = dv(Lv[iyr] - Lv[iyl], yin ? P.two_dy : P.dy);
      const T div = ux + vy;
      if (comp_on && (double)div < div_thr) m = true;
      if (shear_on && (double)absT(vx - uy) > shear_thr) m = true;
      const bool xlo = !per && gx == 0, xhi = !per && gx == nx - 1;
      const bool ylo = !per && gy == 0, yhi = !per && gy == ny - 1;
      const Jn<T> jr{Lr[ixl], Lr[ic], Lr[ixr], Lr[iyl], Lr[iyr]}, jp{Lp[ixl], Lp[ic], Lp[ixr], Lp[iyl], Lp[iyr]};
      const Jn<T> je{Le[ixl], Le[ic], Le[ixr], Le[iyl], Le[iyr]};
      if (jumps_over(jr, jp, je, er, ep, ee, thr, per, xlo, xhi, ylo, yhi)) m = true;
      if (!per && P.bguard > 0) {
        const int g = P.bguard;
        if (gy < g || gy >= ny - g || gx < g || gx >= nx - g) m = true;
      }
    }
    raw[t] = m ? 1 : 0;
  }
  __syncthreads();
  for (int t = threadIdx.x; t < RY * kSTX; t += kBlock) {
    const int r = t / kSTX, c = t - r * kSTX;
    uint8_t o = 0;
    for (int d = 0; d <= 2 * w; ++d) o |= raw[r * RX + c + d];
    mid[t] = o;
  }
  __syncthreads();
  for (int t = threadIdx.x; t < kSTY * kSTX; t += kBlock) {
    const int r = t / kSTX, c = t - r * kSTX;
    uint8_t o = 0;
    for (int d = 0; d <= 2 * w; ++d) o |= mid[(r + d) * kSTX + c];
    outt[r * (kSTX + 1) + c] = o;
    const int gy = y0 + r, gx = x0 + c;
    if (gy < ny && gx < nx) mask[gy * nx + gx] = o;
  }
  __syncthreads();
  for (int t = threadIdx.x; t < kSTY * kSTX; t += kBlock) {
    const int c = t / kSTY, r = t - c * kSTY;  // consecutive threads -> consecutive y
    const int gy = y0 + r, gx = x0 + c;
    if (gy < ny && gx < nx) maskT[(size_t)gx * ny + gy] = outt[r * (kSTX + 1) + c];
  }
  if (qT == nullptr) return;
  // enforced state, transposed, staged through the (now free) primitive area
  T* st = Lr;  // [4][kSTY][kSTX + 1]
  for (int t = threadIdx.x; t < kSTY * kSTX; t += kBlock) {
    const int r = t / kSTX, c = t - r * kSTX;
    const int gy = y0 + r, gx = x0 + c;
    if (gy < ny && gx < nx) {
      T a, b, cc, d;
      load4(q, N, gy * nx + gx, a, b, cc, d);
      if (enforce_first) enforce(P, a, b, cc, d);
      const int o = r * (kSTX + 1) + c;
      st[o] = a;
      st[kSTY * (kSTX + 1) + o] = b;
      st[2 * kSTY * (kSTX + 1) + o] = cc;
      st[3 * kSTY * (kSTX + 1) + o] = d;
    }
  }
  __syncthreads();
  for (int t = threadIdx.x; t < kSTY * kSTX; t += kBlock) {
    const int c = t / kSTY, r = t - c * kSTY;
    const int gy = y0 + r, gx = x0 + c;
    if (gy < ny && gx < nx) {
      const int o = r * (kSTX + 1) + c;
      const size_t g = (size_t)gx * ny + gy;
#pragma unroll
      for (int k = 0; k < 4; ++k) qT[k * (size_t)N + g] = st[k * kSTY * (kSTX + 1) + o];
    }
  }
}

// Split form of the sensor for large batches (the fused tile recomputes
// primitives over a 2.1x halo and the undilated mask over a 1.9x halo).
// 1. Column walk (halo 1, P.rb rows per wave): the undilated mask once per
//    node into raw[ny][nx] (same per-node arithmetic as the fused tile).
template <class T>
__device__ __forceinline__ void rawmask_body(const Params<T>& P, const T* __restrict__ q, int enforce_first,
                                             const double* __restrict__ scal, uint8_t* __restrict__ raw, int bid) {
  const int nx = P.nx, ny = P.ny, N = nx * ny;
  const Walk W = walk_setup(nx, ny, P.periodic, P.rb, 1, bid);
  if (!W.live) return;
  const int lane = threadIdx.x & 63;
  const bool per = P.periodic;
  const T er = (T)scal[S_EPS_R], ep = (T)scal[S_EPS_P], ee = (T)scal[S_EPS_E];
  const bool comp_on = scal[S_COMP_ON] != 0.0, shear_on = scal[S_SHEAR_ON] != 0.0;
  const double div_thr = scal[S_DIV_THR], shear_thr = scal[S_SHEAR_THR], thr = P.jump_thr;
  auto row_of = [&](int y) { return per ? (y < 0 ? y + ny : (y >= ny ? y - ny : y)) : clampi(y, 0, ny - 1); };
  struct Node {
    T rho, u, v, p, e;
  };
  auto node = [&](T a, T b, T c, T d) {
    if (enforce_first) enforce(P, a, b, c, d);
    const Prim<T> pr = prim(P, a, b, c, d);
    return Node{pr.rho, pr.u, pr.v, pr.p, dv(pr.p, pr.rho * P.gm1)};
  };
  auto load_node = [&](int y) {
    T a, b, c, d;
    load4(q, N, row_of(y) * nx + W.xl, a, b, c, d);
    return node(a, b, c, d);
  };
  Node pv = load_node(W.ya - 1), cu = load_node(W.ya), nn = load_node(W.ya + 1);
  const int x = W.x;
  const bool xin = per || (x > 0 && x < nx - 1);
  const bool xlo = !per && x == 0, xhi = !per && x == nx - 1;
  for (int y = W.ya; y < W.yb; ++y) {
    T n0 = T(0), n1 = T(0), n2 = T(0), n3 = T(0);
    if (y + 1 < W.yb) load4(q, N, row_of(y + 2) * nx + W.xl, n0, n1, n2, n3);
    const T ul = lane_step(cu.u, -1), ur = lane_step(cu.u, 1);
    const T vl = lane_step(cu.v, -1), vr = lane_step(cu.v, 1);
    const T rl = lane_step(cu.rho, -1), rr = lane_step(cu.rho, 1);
    const T pl = lane_step(cu.p, -1), pr = lane_step(cu.p, 1);
    const T el = lane_step(cu.e, -1), erg = lane_step(cu.e, 1);
    if (W.own) {
      bool m = false;
      const bool yin = per || (y > 0 && y < ny - 1);
      const T ux = dv(ur - ul, xin ? P.two_dx : P.dx);
      const T vx = dv(vr - vl, xin ? P.two_dx : P.dx);
      const T uy = dv(nn.u - pv.u, yin ? P.two_dy : P.dy);
      const T vy = dv(nn.v - pv.v, yin ? P.two_dy : P.dy);
      const T div = ux + vy;
      if (comp_on && (double)div < div_thr) m = true;
      if (shear_on && (double)absT(vx - uy) > shear_thr) m = true;
      const bool ylo = !per && y == 0, yhi = !per && y == ny - 1;
      const Jn<T> jr{rl, cu.rho, rr, pv.rho, nn.rho}, jp{pl, cu.p, pr, pv.p, nn.p}, je{el, cu.e, erg, pv.e, nn.e};
      if (jumps_over(jr, jp, je, er, ep, ee, thr, per, xlo, xhi, ylo, yhi)) m = true;
      if (!per && P.bguard > 0) {
        const int g = P.bguard;
        if (y < g || y >= ny - g || x < g || x >= nx - g) m = true;
      }
      raw[y * nx + x] = m ? 1 : 0;
    }
    pv = cu;
    cu = nn;
    if (y + 1 < W.yb) nn = node(n0, n1, n2, n3);
  }
}

// 1'. Tile form of step 1: a 32x16 tile with a one-node halo, every node's
//     state loaded at once (all of a thread's loads in flight together) and
//     its primitives formed once into LDS, then the undilated mask of the
//     tile's nodes into raw (the same per-node arithmetic as the column walk
//     and the fused tile).  FIN (single plans): the stage finalize folded in:
//     the block's first partial rows are loaded with the tile, and the block
//     reduces the scalar columns of the nblk partial rows (row stride
//     `stride`) in the finalize's order (the same 256-thread reduction, so
//     every block and a separate finalize agree to the bit) into the stage
//     scalars it then uses; block `first` also writes them to scal_out for
//     the later kernels of the stage.
template <class T, bool PER, bool FIN = false>
__device__ __forceinline__ void rawtile_body(const Params<T>& P, const T* __restrict__ q, int enforce_first,
                                             const double* __restrict__ scal, uint8_t* __restrict__ raw,
                                             T* __restrict__ qT, int bx, int by,
                                             const double* __restrict__ part = nullptr, int nblk = 0,
                                             int stride = 0, double* __restrict__ scal_out = nullptr,
                                             bool first = false) {
  constexpr int EX = kSTX + 2, EY = kSTY + 2, NE = EX * EY, NIT = (NE + kBlock - 1) / kBlock;
  constexpr int PF = 2;  // partial rows per thread loaded with the tile
  __shared__ T Lr[EX * EY], Lu[EX * EY], Lv[EX * EY], Lp[EX * EY], Le[EX * EY];
  // the tile's (enforced) state, transposed, when qT is wanted: column cc
  // holds its 16 rows at cc*16 + (r ^ ((cc >> 1) & 15)) (a row of 32 columns
  // and a column of 16 rows each touch 32 distinct banks; no padding, so
  // eight 4-wavefront blocks fit a CU in FP32)
  static_assert(kSTY == 16, "the swizzle spans 16 rows");
  __shared__ T st[4 * kSTX * kSTY];
  auto sw = [](int cc, int r) { return cc * kSTY + (r ^ ((cc >> 1) & 15)); };
  const int nx = P.nx, ny = P.ny, N = nx * ny;
  const int x0 = bx * kSTX, y0 = by * kSTY;
  constexpr bool per = PER;  // periodic grids take no clamps and no end tests
  double pf[PF][R_DRAG];
  if (FIN) {
#pragma unroll
    for (int j = 0; j < PF; ++j) {
      const int b = threadIdx.x + j * kBlock;
      if (b < nblk) {
#pragma unroll
        for (int k = 0; k < R_DRAG; ++k) pf[j][k] = part[(size_t)b * stride + k];
      }
    }
  }
  T qa[NIT], qb[NIT], qc[NIT], qd[NIT];
#pragma unroll
  for (int s = 0; s < NIT; ++s) {
    const int t = threadIdx.x + s * kBlock;
    if (t < NE) {
      const int r = t / EX, c = t - r * EX;
      int gy = y0 - 1 + r, gx = x0 - 1 + c;
      if (per) {
        gy = gy < 0 ? gy + ny : (gy >= ny ? gy - ny : gy);
        gx = gx < 0 ? gx + nx : (gx >= nx ? gx - nx : gx);
      } else {
        gy = clampi(gy, 0, ny - 1);
        gx = clampi(gx, 0, nx - 1);
      }
      load4(q, N, gy * nx + gx, qa[s], qb[s], qc[s], qd[s]);
    }
  }
#pragma unroll
  for (int s = 0; s < NIT; ++s) {
    const int t = threadIdx.x + s * kBlock;
    if (t < NE) {
      T a = qa[s], b = qb[s], c = qc[s], d = qd[s];
      if (enforce_first) enforce(P, a, b, c, d);
      const Prim<T> pr = prim(P, a, b, c, d);
      Lr[t] = pr.rho;
      Lu[t] = pr.u;
      Lv[t] = pr.v;
      Lp[t] = pr.p;
      Le[t] = dv(pr.p, pr.rho * P.gm1);
      const int r = t / EX - 1, cc = t - (r + 1) * EX - 1;
      if (qT != nullptr && r >= 0 && r < kSTY && cc >= 0 && cc < kSTX) {
        const int o = sw(cc, r);
        st[o] = a;
        st[kSTX * kSTY + o] = b;
        st[2 * kSTX * kSTY + o] = c;
        st[3 * kSTX * kSTY + o] = d;
      }
    }
  }
  if (FIN) {
    __shared__ double rr[R_DRAG];
    __shared__ double sc[S_DRAG];
    constexpr uint64_t MM = 0x1Full;  // R_MAXR .. R_AY are maxima
    double v[R_DRAG];
#pragma unroll
    for (int k = 0; k < R_DRAG; ++k) v[k] = 0.0;
#pragma unroll
    for (int j = 0; j < PF; ++j) {
      if (threadIdx.x + j * kBlock < nblk) {
#pragma unroll
        for (int k = 0; k < R_DRAG; ++k) v[k] = ((MM >> k) & 1) ? fmax(v[k], pf[j][k]) : v[k] + pf[j][k];
      }
    }
    for (int b = threadIdx.x + PF * kBlock; b < nblk; b += kBlock) {
#pragma unroll
      for (int k = 0; k < R_DRAG; ++k) {
        const double a = part[(size_t)b * stride + k];
        v[k] = ((MM >> k) & 1) ? fmax(v[k], a) : v[k] + a;
      }
    }
    block_reduce_write<R_DRAG, MM>(v, rr);
    __syncthreads();
    stage_scal_write<T, false>(P, rr, sc);
    if (first) stage_scal_write<T, false>(P, rr, scal_out);
    scal = sc;
  }
  __syncthreads();
  if (qT != nullptr) {  // enforced state, transposed for the y-lines: per column 16 consecutive y
    for (int t = threadIdx.x; t < kSTX * kSTY; t += kBlock) {
      const int c = t / kSTY, r = t - c * kSTY;
      const int gy = y0 + r, gx = x0 + c;
      if (gy < ny && gx < nx) {
        const int o = sw(c, r);
        const size_t g = (size_t)gx * ny + gy;
#pragma unroll
        for (int k = 0; k < 4; ++k) qT[k * (size_t)N + g] = st[k * kSTX * kSTY + o];
      }
    }
  }
  const T er = (T)scal[S_EPS_R], ep = (T)scal[S_EPS_P], ee = (T)scal[S_EPS_E];
  const bool comp_on = scal[S_COMP_ON] != 0.0, shear_on = scal[S_SHEAR_ON] != 0.0;
  const Divisor<T> d2x = divisor(P.two_dx), d2y = divisor(P.two_dy);
  const double div_thr = scal[S_DIV_THR], shear_thr = scal[S_SHEAR_THR], thr = P.jump_thr;
  for (int t = threadIdx.x; t < kSTX * kSTY; t += kBlock) {
    const int r = t / kSTX, c = t - r * kSTX;
    const int gy = y0 + r, gx = x0 + c;
    if (gx >= nx || gy >= ny) continue;
    bool m = false;
    const int ic = (r + 1) * EX + (c + 1);
    const int ixl = ic - 1, ixr = ic + 1, iyl = ic - EX, iyr = ic + EX;
    const bool xin = per || (gx > 0 && gx < nx - 1), yin = per || (gy > 0 && gy < ny - 1);
    T ux, vx, uy, vy;
    if (per) {
      ux = dv(Lu[ixr] - Lu[ixl], d2x);
      vx = dv(Lv[ixr] - Lv[ixl], d2x);
      uy = dv(Lu[iyr] - Lu[iyl], d2y);
      vy = dv(Lv[iyr] - Lv[iyl], d2y);
    } else {
      ux = dv(Lu[ixr] - Lu[ixl], xin ? P.two_dx : P.dx);
      vx = dv(Lv[ixr] - Lv[ixl], xin ? P.two_dx : P.dx);
      uy = dv(Lu[iyr] - Lu[iyl], yin ? P.two_dy : P.dy);
      vy = dv(Lv[iyr] - Lv[iyl], yin ? P.two_dy : P.dy);
    }
    const T div = ux + vy;
    if (comp_on && (double)div < div_thr) m = true;
    if (shear_on && (double)absT(vx - uy) > shear_thr) m = true;
    const bool xlo = !per && gx == 0, xhi = !per && gx == nx - 1;
    const bool ylo = !per && gy == 0, yhi = !per && gy == ny - 1;
    const Jn<T> jr{Lr[ixl], Lr[ic], Lr[ixr], Lr[iyl], Lr[iyr]}, jp{Lp[ixl], Lp[ic], Lp[ixr], Lp[iyl], Lp[iyr]};
    const Jn<T> je{Le[ixl], Le[ic], Le[ixr], Le[iyl], Le[iyr]};
    if (jumps_over(jr, jp, je, er, ep, ee, thr, per, xlo, xhi, ylo, yhi)) m = true;
    if (!per && P.bguard > 0) {
      const int g = P.bguard;
      if (gy < g || gy >= ny - g || gx < g || gx >= nx - g) m = true;
    }
    raw[gy * nx + gx] = m ? 1 : 0;
  }
}

// Tile form of the stage reduction without the drag projection (single
// plans): a 32x16 tile with a one-node halo loaded in one round, the same
// per-node quantities as the column walk (maxima of |rho|, |p|, |e|,
// |u|+c, |v|+c; sums of div^2, vort^2 and, for cooling, the pressure of the
// state as given), one partial row per tile.
// Drag projection of one 32x16 tile (OOP/hit2d.py:391-394, q as given):
// per column the sums of rho u and rho v times 1, cos/sin(ky y) (ky = 1, 2)
// over the tile's rows (threads of a column: lanes c and c+32 of each wave,
// then the four waves through LDS), then lane j < 2*ndrag of wave 0 forms
// mode j/2, field j%2 over the 32 columns with the columns' twiddles of
// |kx| = 1, 2 staged in LDS (cos(-a) = cos(a), sin(-a) = -sin(a)), as the
// column walk's end phase does.
template <class T>
__device__ __forceinline__ void drag_tile_project(const Params<T>& P, const T* __restrict__ q,
                                                  double* __restrict__ out, int x0, int y0) {
  __shared__ double cw[kWaves][10][32];
  __shared__ double cs[10][32];
  __shared__ double tw[4][32];
  const int nx = P.nx, ny = P.ny, N = nx * ny;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  double bs[10];
#pragma unroll
  for (int j = 0; j < 10; ++j) bs[j] = 0.0;
#pragma unroll
  for (int s = 0; s < kSTX * kSTY / kBlock; ++s) {
    const int t = threadIdx.x + s * kBlock, r = t / kSTX, c = t - r * kSTX;
    const int gy = y0 + r, gx = x0 + c;
    if (gx < nx && gy < ny) {
      const double fb = (double)at32(q + (size_t)N, gy * nx + gx), fc = (double)at32(q + 2 * (size_t)N, gy * nx + gx);
      bs[0] += fb;
      bs[5] += fc;
#pragma unroll
      for (int k = 1; k <= 2; ++k) {
        if (k <= P.dky) {
          const double* ty = P.twy + ((size_t)(k + P.kmy) * ny + gy) * 2;
          const double cy = ty[0], sy = ty[1];
          bs[2 * k - 1] += fb * cy;
          bs[2 * k] += fb * sy;
          bs[5 + 2 * k - 1] += fc * cy;
          bs[5 + 2 * k] += fc * sy;
        }
      }
    }
  }
  // lanes c and c + 32 hold the same column
#pragma unroll
  for (int j = 0; j < 10; ++j) {
    const double o = __shfl_xor(bs[j], 32, 64);
    if (lane < 32) cw[wid][j][lane] = bs[j] + o;
  }
  if (threadIdx.x < 4 * 32) {  // the columns' twiddles of |kx| = 1, 2 (any column past the grid)
    const int k = threadIdx.x >> 6, e = (threadIdx.x >> 5) & 1, c = threadIdx.x & 31;
    const int xc = min(x0 + c, nx - 1);
    if (k + 1 <= P.dkx) tw[2 * k + e][c] = P.twx[((size_t)(k + 1 + P.kmx) * nx + xc) * 2 + e];
  }
  __syncthreads();
  for (int j = threadIdx.x; j < 10 * 32; j += kBlock) {
    const int f = j >> 5, c = j & 31;
    double a = cw[0][f][c];
#pragma unroll
    for (int w = 1; w < kWaves; ++w) a += cw[w][f][c];
    cs[f][c] = a;
  }
  __syncthreads();
  if (threadIdx.x < 2 * P.ndrag) {
    const int m = threadIdx.x >> 1, f = threadIdx.x & 1;
    const int kx = P.drag_modes[2 * m], ky = P.drag_modes[2 * m + 1];
    const int ka = ky < 0 ? -ky : ky, kxa = kx < 0 ? -kx : kx;
    const int sc = f * 5 + (ka == 0 ? 0 : 2 * ka - 1);
    double re = 0.0, im = 0.0;
#pragma unroll 8
    for (int c = 0; c < kSTX; ++c) {  // columns past the grid have zero sums
      const double cx = kxa == 0 ? 1.0 : tw[2 * kxa - 2][c];
      const double sa = kxa == 0 ? 0.0 : tw[2 * kxa - 1][c];
      const double sx = kx < 0 ? -sa : sa;
      const double br = cs[sc][c];
      const double bi = ka == 0 ? 0.0 : (ky > 0 ? -cs[sc + 1][c] : cs[sc + 1][c]);
      re += cx * br + sx * bi;
      im += cx * bi - sx * br;
    }
    out[4 * m + 2 * f] = re;
    out[4 * m + 2 * f + 1] = im;
  }
}

template <class T, bool DRAG>
__device__ __forceinline__ void stage_tile_body(const Params<T>& P, const T* __restrict__ q, int enforce_first_in,
                                                double* __restrict__ part, int bx, int by, int bid) {
  const int enforce_first = reduce_enforce_first<T>(enforce_first_in);
  constexpr int NR = DRAG ? R_DRAG + kDragCols : R_DRAG;  // partial row: the scalars, then the drag projection
  constexpr int EX = kSTX + 2, EY = kSTY + 2;
  __shared__ T Lu[EX * EY], Lv[EX * EY];
  double v[R_DRAG];
#pragma unroll
  for (int k = 0; k < R_DRAG; ++k) v[k] = 0.0;
  const int nx = P.nx, ny = P.ny, N = nx * ny;
  const int x0 = bx * kSTX, y0 = by * kSTY;
  const bool per = P.periodic;
  for (int t = threadIdx.x; t < EX * EY; t += kBlock) {
    const int r = t / EX, c = t - r * EX;
    const int uy = y0 - 1 + r, ux = x0 - 1 + c;  // unwrapped
    int gy = uy, gx = ux;
    if (per) {
      gy = gy < 0 ? gy + ny : (gy >= ny ? gy - ny : gy);
      gx = gx < 0 ? gx + nx : (gx >= nx ? gx - nx : gx);
    } else {
      gy = clampi(gy, 0, ny - 1);
      gx = clampi(gx, 0, nx - 1);
    }
    T a, b, cc, d;
    load4(q, N, gy * nx + gx, a, b, cc, d);
    const bool own = r >= 1 && r <= kSTY && c >= 1 && c <= kSTX && uy < ny && ux < nx;
    if (own && P.cool_on && enforce_first) v[R_PSUM] += (double)prim(P, a, b, cc, d).p;
    if (enforce_first) enforce(P, a, b, cc, d);
    const Prim<T> pr = prim(P, a, b, cc, d);
    if (own && P.cool_on && !enforce_first) v[R_PSUM] += (double)pr.p;
    Lu[t] = pr.u;
    Lv[t] = pr.v;
    if (own) {
      const T cs = sqrt_alg(dv(P.gamma * pr.p, pr.rho));
      v[R_MAXR] = fmax(v[R_MAXR], (double)absT(pr.rho));
      v[R_MAXP] = fmax(v[R_MAXP], (double)absT(pr.p));
      const T e = dv(pr.p, pr.rho * P.gm1);
      v[R_MAXE] = fmax(v[R_MAXE], (double)absT(e));
      v[R_AX] = fmax(v[R_AX], (double)(absT(pr.u) + cs));
      v[R_AY] = fmax(v[R_AY], (double)(absT(pr.v) + cs));
    }
  }
  __syncthreads();
  for (int t = threadIdx.x; t < kSTX * kSTY; t += kBlock) {
    const int r = t / kSTX, c = t - r * kSTX;
    const int gy = y0 + r, gx = x0 + c;
    if (gx >= nx || gy >= ny) continue;
    const int ic = (r + 1) * EX + (c + 1);
    const bool xin = per || (gx > 0 && gx < nx - 1), yin = per || (gy > 0 && gy < ny - 1);
    const T ux = dv(Lu[ic + 1] - Lu[ic - 1], xin ? P.two_dx : P.dx);
    const T vx = dv(Lv[ic + 1] - Lv[ic - 1], xin ? P.two_dx : P.dx);
    const T uy = dv(Lu[ic + EX] - Lu[ic - EX], yin ? P.two_dy : P.dy);
    const T vy = dv(Lv[ic + EX] - Lv[ic - EX], yin ? P.two_dy : P.dy);
    const T div = ux + vy;
    v[R_DIV2] += (double)div * (double)div;
    if (P.shear_on) {
      const T w = vx - uy;
      v[R_VORT2] += (double)w * (double)w;
    }
  }
  block_reduce_write<R_DRAG, 0x1Full>(v, part + (size_t)bid * NR);
  if (DRAG && P.drag_on) drag_tile_project(P, q, part + (size_t)bid * NR + R_DRAG, x0, y0);
}

// 2. 32x16 tile: the (2w+1)^2 OR dilation of raw (out-of-domain nodes of the
//    non-periodic grid false), mask and maskT stores, and (qT not null) the
//    enforced state written transposed for the y sweep.
template <class T>
__device__ __forceinline__ void dilate_body(const Params<T>& P, const T* __restrict__ q, int enforce_first,
                                            const uint8_t* __restrict__ rawg, uint8_t* __restrict__ mask,
                                            uint8_t* __restrict__ maskT, T* __restrict__ qT, int bx, int by) {
  __shared__ uint8_t raw[(kSTX + 16) * (kSTY + 16)];
  __shared__ uint8_t mid[(kSTY + 16) * kSTX];
  __shared__ uint8_t outt[kSTY * (kSTX + 1)];
  __shared__ T st[4 * kSTY * (kSTX + 1)];
  const int nx = P.nx, ny = P.ny, N = nx * ny, w = P.width;
  const int RX = kSTX + 2 * w, RY = kSTY + 2 * w;
  const int x0 = bx * kSTX, y0 = by * kSTY;
  const bool per = P.periodic;
  for (int t = threadIdx.x; t < RX * RY; t += kBlock) {
    const int r = t / RX, c = t - r * RX;
    int gy = y0 - w + r, gx = x0 - w + c;
    uint8_t m = 0;
    if (per) {
      gy = wrapn(gy, ny);
      gx = wrapn(gx, nx);
      m = rawg[gy * nx + gx];
    } else if (gx >= 0 && gx < nx && gy >= 0 && gy < ny) {
      m = rawg[gy * nx + gx];
    }
    raw[t] = m;
  }
  __syncthreads();
  for (int t = threadIdx.x; t < RY * kSTX; t += kBlock) {
    const int r = t / kSTX, c = t - r * kSTX;
    uint8_t o = 0;
    for (int d = 0; d <= 2 * w; ++d) o |= raw[r * RX + c + d];
    mid[t] = o;
  }
  __syncthreads();
  for (int t = threadIdx.x; t < kSTY * kSTX; t += kBlock) {
    const int r = t / kSTX, c = t - r * kSTX;
    uint8_t o = 0;
    for (int d = 0; d <= 2 * w; ++d) o |= mid[(r + d) * kSTX + c];
    outt[r * (kSTX + 1) + c] = o;
    const int gy = y0 + r, gx = x0 + c;
    if (gy < ny && gx < nx) mask[gy * nx + gx] = o;
    if (qT != nullptr && gy < ny && gx < nx) {
      T a, b, cc, d;
      load4(q, N, gy * nx + gx, a, b, cc, d);
      if (enforce_first) enforce(P, a, b, cc, d);
      const int oo = r * (kSTX + 1) + c;
      st[oo] = a;
      st[kSTY * (kSTX + 1) + oo] = b;
      st[2 * kSTY * (kSTX + 1) + oo] = cc;
      st[3 * kSTY * (kSTX + 1) + oo] = d;
    }
  }
  __syncthreads();
  for (int t = threadIdx.x; t < kSTY * kSTX; t += kBlock) {
    const int c = t / kSTY, r = t - c * kSTY;  // consecutive threads -> consecutive y
    const int gy = y0 + r, gx = x0 + c;
    if (gy < ny && gx < nx) {
      const int oo = r * (kSTX + 1) + c;
      maskT[(size_t)gx * ny + gy] = outt[oo];
      if (qT != nullptr) {
        const size_t g = (size_t)gx * ny + gy;
#pragma unroll
        for (int k = 0; k < 4; ++k) qT[k * (size_t)N + g] = st[k * kSTY * (kSTX + 1) + oo];
      }
    }
  }
}

template <class T>
__global__ void __launch_bounds__(kBlock) k_rawmask(Params<T> P, const T* __restrict__ q, int enforce_first,
                                                    const double* __restrict__ scal, uint8_t* __restrict__ raw) {
  rawmask_body<T>(P, q, enforce_first, scal, raw, blockIdx.x);
}

template <class T, bool DRAG>
__global__ void __launch_bounds__(kBlock) k_stage_tile(Params<T> P, const T* __restrict__ q, int enforce_first,
                                                       double* __restrict__ part) {
  const Tile2 t_ = tile_xcd();
  stage_tile_body<T, DRAG>(P, q, enforce_first, part, t_.bx, t_.by, t_.by * (int)gridDim.x + t_.bx);
}

template <class T, bool PER>
__global__ void __launch_bounds__(kBlock) k_rawtile_fin(Params<T> P, const T* __restrict__ q, int enforce_first,
                                                        const double* __restrict__ part, int nblk, int stride,
                                                        double* __restrict__ scal, uint8_t* __restrict__ raw,
                                                        T* __restrict__ qT) {
  const Tile2 t_ = tile_xcd();
  rawtile_body<T, PER, true>(P, q, enforce_first, nullptr, raw, qT, t_.bx, t_.by, part, nblk, stride, scal,
                             blockIdx.x == 0 && blockIdx.y == 0);
}

// Raw-mask tiles that read the stage scalars written by the finalize kernel
// (single plans with many partial rows, where re-reducing them in every tile
// would cost more than the separate finalize).
template <class T, bool PER>
__global__ void __launch_bounds__(kBlock) k_rawtile(Params<T> P, const T* __restrict__ q, int enforce_first,
                                                    const double* __restrict__ scal, uint8_t* __restrict__ raw,
                                                    T* __restrict__ qT) {
  const Tile2 t_ = tile_xcd();
  rawtile_body<T, PER>(P, q, enforce_first, scal, raw, qT, t_.bx, t_.by);
}

// ---------------------------------------------------------------- bit dilation
// The (2w+1)^2 OR dilation of the undilated mask on bit words: one wavefront
// per 64-row x 32-column tile, lane l on row y0 + l.  A row's 32 columns and
// 8 more on each side become the bits of one word (interior rows read as
// aligned 8- or 4-byte loads), the horizontal OR is 2w+1 shifts, the
// vertical OR reads the 2w+1 rows' words from the wave's LDS slots, the mask
// row is stored as bytes and each maskT column comes from one ballot over the
// tile's rows.  Out-of-domain nodes wrap (periodic) or are false
// (non-periodic), as in OOP/spatial_operator.py:169-176 and
// OOP/problems/riemann_config3.py:149-162.  Undilated mask bytes are 0 or 1.
constexpr int kDBRows = 64, kDBCols = 32, kDBWaves = 4;
__device__ __forceinline__ uint32_t gather4(uint32_t d) {  // 4 bytes of 0/1 -> 4 bits
  return (d & 1u) | ((d >> 7) & 2u) | ((d >> 14) & 4u) | ((d >> 21) & 8u);
}
__device__ __forceinline__ uint32_t spread4(uint32_t t) {  // 4 bits -> 4 bytes of 0/1
  return (t & 1u) | ((t & 2u) << 7) | ((t & 4u) << 14) | ((t & 8u) << 21);
}

// Horizontally dilated bits of row gy, columns x0 .. x0+31 (bit c = column x0+c).
template <class T>
__device__ __forceinline__ uint32_t dilate_row_bits(const Params<T>& P, const uint8_t* __restrict__ raw, int gy,
                                                    int x0) {
  const int nx = P.nx, ny = P.ny, w = P.width;
  if (P.periodic)
    gy = wrapn(gy, ny);
  else if (gy < 0 || gy >= ny)
    return 0u;
  const uint8_t* r = raw + (size_t)gy * nx;
  uint64_t bits = 0;  // bit k: column x0 - 8 + k, k < 48
  // aligned groups of 8 (or 4) columns never straddle the grid's end when nx
  // is a multiple of 8 (or 4): each group is one load from its wrapped start,
  // or zero outside a non-periodic grid
  auto group = [&](int g0, bool& in) {
    in = true;
    if (P.periodic)
      g0 = wrapn(g0, nx);
    else
      in = g0 >= 0 && g0 < nx;
    return in ? g0 : 0;
  };
  if ((nx & 7) == 0) {
    uint2 d[6];
    bool in[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) d[k] = *reinterpret_cast<const uint2*>(r + group(x0 - 8 + 8 * k, in[k]));
#pragma unroll
    for (int k = 0; k < 6; ++k)
      if (in[k]) bits |= (uint64_t)(gather4(d[k].x) | (gather4(d[k].y) << 4)) << (8 * k);
  } else if ((nx & 3) == 0) {
    uint32_t d[12];
    bool in[12];
#pragma unroll
    for (int k = 0; k < 12; ++k) d[k] = *reinterpret_cast<const uint32_t*>(r + group(x0 - 8 + 4 * k, in[k]));
#pragma unroll
    for (int k = 0; k < 12; ++k)
      if (in[k]) bits |= (uint64_t)gather4(d[k]) << (4 * k);
  } else {
    for (int k = 0; k < 48; ++k) {
      int gx = x0 - 8 + k;
      bool in = true;
      if (P.periodic)
        gx = wrapn(gx, nx);
      else
        in = gx >= 0 && gx < nx;
      if (in && r[gx] != 0) bits |= (uint64_t)1 << k;
    }
  }
  uint32_t h = 0;
  for (int s = -w; s <= w; ++s) h |= (uint32_t)(bits >> (8 + s));
  return h;
}

// One tile (tx, ty) by the calling wavefront; hs: its kDBRows + 16 LDS words.
template <class T>
__device__ __forceinline__ void dilate_bits_tile(const Params<T>& P, const uint8_t* __restrict__ raw,
                                                 uint8_t* __restrict__ mask, uint8_t* __restrict__ maskT, int tx,
                                                 int ty, uint32_t* __restrict__ hs) {
  const int nx = P.nx, ny = P.ny, w = P.width;
  const int lane = threadIdx.x & 63;
  const int x0 = tx * kDBCols, y0 = ty * kDBRows;
  // slot i holds row y0 - w + i: own rows at w .. w+63, halo rows above and below
  hs[w + lane] = dilate_row_bits(P, raw, y0 + lane, x0);
  if (lane < 2 * w) {
    const int slot = lane < w ? lane : kDBRows + lane;
    const int gy = lane < w ? y0 - w + lane : y0 + kDBRows + (lane - w);
    hs[slot] = dilate_row_bits(P, raw, gy, x0);
  }
  wave_sync();
  uint32_t v = 0;
  for (int d = 0; d <= 2 * w; ++d) v |= hs[lane + d];
  const int ncol = nx - x0 < kDBCols ? nx - x0 : kDBCols, nrow = ny - y0 < kDBRows ? ny - y0 : kDBRows;
  if (ncol < 32) v &= (1u << ncol) - 1u;
  if (lane >= nrow) v = 0u;
  if (lane < nrow) {
    uint8_t* o = mask + (size_t)(y0 + lane) * nx + x0;
    if (ncol == kDBCols && (nx & 3) == 0) {
      uint32_t* o4 = reinterpret_cast<uint32_t*>(o);
#pragma unroll
      for (int k = 0; k < 8; ++k) o4[k] = spread4((v >> (4 * k)) & 15u);
    } else {
      for (int k = 0; k < ncol; ++k) o[k] = (uint8_t)((v >> k) & 1u);
    }
  }
  uint64_t col = 0;  // lane c < 32: column x0 + c over the tile's rows
#pragma unroll
  for (int c = 0; c < kDBCols; ++c) {
    const uint64_t b = __ballot((v >> c) & 1u);
    col = lane == c ? b : col;
  }
  if (lane < ncol) {
    uint8_t* o = maskT + (size_t)(x0 + lane) * ny + y0;
    if (nrow == kDBRows && (ny & 3) == 0) {
      uint32_t* o4 = reinterpret_cast<uint32_t*>(o);
#pragma unroll
      for (int k = 0; k < 16; ++k) o4[k] = spread4((uint32_t)(col >> (4 * k)) & 15u);
    } else {
      for (int k = 0; k < nrow; ++k) o[k] = (uint8_t)((col >> k) & 1u);
    }
  }
}

template <class T>
__global__ void __launch_bounds__(64 * kDBWaves) k_dilate_bits(Params<T> P, const uint8_t* __restrict__ raw,
                                                               uint8_t* __restrict__ mask,
                                                               uint8_t* __restrict__ maskT) {
  __shared__ uint32_t hs[kDBWaves][kDBRows + 16];
  const int ntx = (P.nx + kDBCols - 1) / kDBCols, nty = (P.ny + kDBRows - 1) / kDBRows;
  const int wv = threadIdx.x >> 6, t = (int)blockIdx.x * kDBWaves + wv;
  if (t >= ntx * nty) return;
  dilate_bits_tile<T>(P, raw, mask, maskT, t % ntx, t / ntx, hs[wv]);
}

template <class T>
__global__ void __launch_bounds__(kBlock) k_sensor(Params<T> P, const T* __restrict__ q, int enforce_first,
                                                   const double* __restrict__ scal, uint8_t* __restrict__ mask,
                                                   uint8_t* __restrict__ maskT, T* __restrict__ qT) {
  const Tile2 t_ = tile_xcd();
  sensor_body<T>(P, q, enforce_first, scal, mask, maskT, qT, t_.bx, t_.by);
}

// ================================================================ line sweep
// Power of rho from the LDS table (zero beyond the table: |rho|^66 < 1e-22).
template <class T>
__device__ __forceinline__ T pwk(const T* PW, int k) { return k <= 65 ? PW[k] : T(0); }

// The one-chunk shape of the compact solve (chunks of L <= 4, nch = n/L <=
// 128 per line, 4 lines): component c takes threads 128c .. 128c+127, two
// wavefronts, thread 128c + j holding chunk j (none for j >= nch).  Each
// recurrence sweeps the chunks locally, joins them with the wavefront affine
// scan (decay a = rho^L) and one exchange of the two wavefronts' end values
// through LDS, and closes the periodic wrap in closed form: carry into chunk
// 0 = S_{nch-1} / (1 - rho^n).  Powers rho^k come from the table (zero beyond
// k = 65, as the truncated carries of the general shape).  Same result as
// the general shape to rounding.
template <class T, bool PER>
__device__ __forceinline__ void solve_store_scan(const T* __restrict__ F, int S, T h, const LineK<T>& K,
                                                 const T* __restrict__ PW, T* __restrict__ scratch,
                                                 T* __restrict__ out, size_t NL, size_t g0) {
  const int n = K.n, L = K.L, nch = K.nch;
  const T rho = K.rho;
  const int t = threadIdx.x, line = t >> 7, ch = t & 127, wv = ch >> 6, lane = t & 63;
  const bool act = ch < nch;
  const int s0 = ch * L;
  T* xf = scratch;        // [4][2] wavefront end values of the causal scan
  T* xb = scratch + 8;    // [4][2] wavefront start values of the anti-causal scan
  T* bnd = scratch + 16;  // [4][3]: z_n, p_0, p_{n-1}
  const T ih = T(1) / h;  // divided differences times the reciprocal (within 2 ulp of the quotient)
  auto raw = [&](int c, int i) -> T {
    const T* f = F + c * S;
    return PER ? (f[i] - f[wrap(i - 1, n)]) * ih : (f[i + 1] - f[i]) * ih;
  };
  // scan powers a^1, a^2, a^4, a^8, a^16 and the lane's a^(r+1), a^(16-r), r = lane % 16
  T apw[5];
#pragma unroll
  for (int k = 0; k < 5; ++k) apw[k] = pwk(PW, L << k);
  const T a_f = pwk(PW, L * ((lane & 15) + 1)), a_b = pwk(PW, L * (16 - (lane & 15)));
  const int last = nch - 1;  // last chunk of the line
  T y[4];
  T r0 = T(0), rn1 = T(0);
  if (!PER) {
    r0 = raw(line, 0);
    rn1 = raw(line, n - 1);
  }
  // causal pass: local sweep, then the scan of the chunk end values
  T acc = T(0);
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    if (k < L && act) {
      const int i = s0 + k;
      T xv = raw(line, i);
      if (!PER) {
        if (i == 0 || i == n - 1) {
          xv = T(0);
        } else {
          if (i == 1) xv = madd(-K.alpha_c, r0, xv);
          if (i == n - 2) xv = madd(-K.alpha_c, rn1, xv);
        }
      }
      acc = madd(rho, acc, xv);
      y[k] = acc;
    }
  }
  T sc = scan_fwd(act ? acc : T(0), apw, a_f, lane);  // inclusive, within the wavefront
  {
    const int lw = wv == 0 ? (last < 63 ? last : 63) : last - 64;  // last live lane of this wavefront
    if (lw >= 0) {
      const T e = rdlane(sc, lw);
      if (lane == 0) xf[line * 2 + wv] = e;
    }
  }
  __syncthreads();
  const T e0 = xf[line * 2];
  if (wv == 1) sc = madd(pwk(PW, L * (lane + 1)), e0, sc);  // chunks 0..63 into 64 + lane
  T carry = dpp<0x138>(sc);                                  // exclusive
  if (wv == 1 && lane == 0) carry = e0;
  if (PER) {
    // S_{nch-1} over the line, then the wrap: C_0 * a^ch
    const T tot = nch > 64 ? madd(pwk(PW, L * (nch - 64)), e0, xf[line * 2 + 1]) : e0;
    carry = madd(pwk(PW, L * ch), tot * K.inv1mrn, carry);
  }
#pragma unroll
  for (int k = 0; k < 4; ++k)
    if (k < L) y[k] = madd(pwk(PW, k + 1), carry, y[k]);
  if (!PER && ch == last) bnd[line * 3 + 0] = (rho * y[L - 1]) / K.one_m_rho2;
  // anti-causal pass: local sweep, then the reversed scan of the chunk start values
  acc = T(0);
#pragma unroll
  for (int k = 3; k >= 0; --k) {
    if (k < L && act) {
      acc = madd(rho, acc, y[k]);
      y[k] = acc;
    }
  }
  sc = scan_bwd(act ? acc : T(0), apw, a_b, lane);  // inclusive, within the wavefront
  if (wv * 64 <= last && lane == 0) xb[line * 2 + wv] = sc;
  __syncthreads();
  const T b1 = nch > 64 ? xb[line * 2 + 1] : T(0);  // chunk 64's inclusive value
  if (wv == 0 && nch > 64) sc = madd(pwk(PW, L * (64 - lane)), b1, sc);
  carry = dpp<0x130>(sc);  // exclusive
  if (wv == 0 && lane == 63) carry = b1;
  if (PER) {
    const T tot = wv == 0 ? rdlane(sc, 0) : madd(pwk(PW, L * 64), b1, xb[line * 2]);  // T_0 over the line
    carry = madd(pwk(PW, L * (last - ch)), tot * K.inv1mrn, carry);
  } else {
    carry = madd(pwk(PW, L * (last - ch)), bnd[line * 3 + 0], carry);
  }
#pragma unroll
  for (int k = 0; k < 4; ++k)
    if (k < L) y[k] = K.c1p * madd(pwk(PW, L - k), carry, y[k]);
  if (!PER) {
    if (ch == 0) bnd[line * 3 + 1] = y[0];
    if (ch == last) bnd[line * 3 + 2] = y[L - 1];
    __syncthreads();
  }
  if (!act) return;
  T* o = out + line * NL + g0;
  if (PER) {
#pragma unroll
    for (int k = 0; k < 4; ++k)
      if (k < L) o[s0 + k] = y[k];
  } else {
    const T p0 = bnd[line * 3 + 1], pn = bnd[line * 3 + 2];
    const T det = T(1) - K.rnm1 * K.rnm1;
    const T idet = T(1) / det, a = (-p0 + pn * K.rnm1) * idet, bb = (-pn + p0 * K.rnm1) * idet;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      if (k < L) {
        const int i = s0 + k;
        T v;
        if (i == 0)
          v = raw(line, 0);
        else if (i == n - 1)
          v = raw(line, n - 1);
        else
          v = (y[k] + a * pwk(PW, i)) + bb * pwk(PW, n - 1 - i);
        o[i] = v;
      }
    }
  }
}

// Divided difference of the interface flux F[c*S + .] and compact solve of
// the four (alpha, 1, alpha) systems, stored to out[c*NL + g0 + i].
// Periodic: cyclic matrix (OOP/spatial_operator.py:227-249 solves it by FFT).
// Non-periodic: identity end rows (OOP/problems/riemann_config3.py:237-265,
// Thomas).  Method: A = K (1 - rho S^-1)(1 - rho S), K = 1/(1+rho^2), so
// A^-1 r is a causal then an anti-causal first-order recurrence with
// coefficient rho = (-1 + sqrt(1 - 4 alpha^2)) / (2 alpha) = -0.4514.  Each
// recurrence runs as a chunked scan: every thread holds one chunk of L <= 8
// points of one line in registers, sweeps it with zero carry-in, then adds
// rho^k times the exact carry, a short geometric sum over the preceding
// chunks' local results read from LDS (|rho|^52 < 1e-18 truncates it; the
// periodic wrap is summed in closed form).  The non-periodic system is the
// zero-Dirichlet interior problem plus the homogeneous solutions
// a rho^i + b rho^(n-1-i) that restore the end rows.
// Requires 4*n/L <= 2*blockDim.x (at most two chunks per thread); SMALL
// (lines the plan cuts into chunks of L <= 4, at most one per thread) keeps
// one chunk of four in registers, which holds the kernel to 64 VGPRs.
template <class T, bool PER, bool SMALL>
__device__ __forceinline__ void solve_store(const T* __restrict__ F, int S, T h, const LineK<T>& K,
                                            const T* __restrict__ PW, T* __restrict__ scratch, T* __restrict__ out,
                                            size_t NL, size_t g0) {
  const int n = K.n, L = K.L, nch = K.nch, M = K.M, ntask = 4 * nch, B = blockDim.x;
  const T rho = K.rho;
  T* e1 = scratch;             // [4*nch] local causal results
  T* e2 = scratch + 4 * nch;   // [4*nch] local anti-causal results
  T* bnd = scratch + 8 * nch;  // [4][3]: z_n, p_0, p_{n-1}
  const T ih = T(1) / h;  // divided differences times the reciprocal (within 2 ulp of the quotient)
  auto raw = [&](int c, int i) -> T {
    const T* f = F + c * S;
    return PER ? (f[i] - f[wrap(i - 1, n)]) * ih : (f[i + 1] - f[i]) * ih;
  };
  if constexpr (SMALL) {
    solve_store_scan<T, PER>(F, S, h, K, PW, scratch, out, NL, g0);
    return;
  }
  constexpr int NT = SMALL ? 1 : 2, LM = SMALL ? 4 : 8;
  T y[NT][LM];
  // causal pass, local
#pragma unroll
  for (int u = 0; u < NT; ++u) {
    const int t = threadIdx.x + u * B;
    if (t < ntask) {
      const int line = t / nch, s0 = (t - line * nch) * L;
      T r0 = T(0), rn1 = T(0);
      if (!PER) {
        r0 = raw(line, 0);
        rn1 = raw(line, n - 1);
      }
      T acc = T(0);
#pragma unroll
      for (int k = 0; k < LM; ++k) {
        if (k < L) {
          const int i = s0 + k;
          T xv = raw(line, i);
          if (!PER) {
            if (i == 0 || i == n - 1) {
              xv = T(0);
            } else {
              if (i == 1) xv = madd(-K.alpha_c, r0, xv);
              if (i == n - 2) xv = madd(-K.alpha_c, rn1, xv);
            }
          }
          acc = madd(rho, acc, xv);
          y[u][k] = acc;
        }
      }
      e1[t] = acc;
    }
  }
  __syncthreads();
  // causal carries, then the anti-causal local pass
#pragma unroll
  for (int u = 0; u < NT; ++u) {
    const int t = threadIdx.x + u * B;
    if (t < ntask) {
      const int line = t / nch, ch = t - line * nch;
      const T* el = e1 + line * nch;
      T carry = T(0);
      if (PER) {
        const int mm = nch < M ? nch : M;
        int j = ch;
        for (int m = 1; m <= mm; ++m) {
          j = (j == 0) ? nch - 1 : j - 1;
          carry = madd(pwk(PW, (m - 1) * L), el[j], carry);
        }
        carry = carry * K.inv1mrn;
      } else {
        const int mm = ch < M ? ch : M;
        for (int m = 1; m <= mm; ++m) carry = madd(pwk(PW, (m - 1) * L), el[ch - m], carry);
      }
#pragma unroll
      for (int k = 0; k < LM; ++k)
        if (k < L) y[u][k] = madd(pwk(PW, k + 1), carry, y[u][k]);
      if (!PER && ch == nch - 1) bnd[line * 3 + 0] = (rho * y[u][L - 1]) / K.one_m_rho2;
      T acc = T(0);
#pragma unroll
      for (int k = LM - 1; k >= 0; --k) {
        if (k < L) {
          acc = madd(rho, acc, y[u][k]);
          y[u][k] = acc;
        }
      }
      e2[t] = acc;
    }
  }
  __syncthreads();
  // anti-causal carries, scale by 1 + rho^2
#pragma unroll
  for (int u = 0; u < NT; ++u) {
    const int t = threadIdx.x + u * B;
    if (t < ntask) {
      const int line = t / nch, ch = t - line * nch;
      const T* el = e2 + line * nch;
      T carry = T(0);
      if (PER) {
        const int mm = nch < M ? nch : M;
        int j = ch;
        for (int m = 1; m <= mm; ++m) {
          j = (j == nch - 1) ? 0 : j + 1;
          carry = madd(pwk(PW, (m - 1) * L), el[j], carry);
        }
        carry = carry * K.inv1mrn;
      } else {
        const int rem = nch - 1 - ch;
        const int mm = rem < M ? rem : M;
        for (int m = 1; m <= mm; ++m) carry = madd(pwk(PW, (m - 1) * L), el[ch + m], carry);
        if (rem < M) carry = madd(pwk(PW, rem * L), bnd[line * 3 + 0], carry);
      }
#pragma unroll
      for (int k = 0; k < LM; ++k)
        if (k < L) y[u][k] = K.c1p * madd(pwk(PW, L - k), carry, y[u][k]);
      if (!PER) {
        if (ch == 0) bnd[line * 3 + 1] = y[u][0];
        if (ch == nch - 1) bnd[line * 3 + 2] = y[u][L - 1];
      }
    }
  }
  if (!PER) __syncthreads();
  // (non-periodic: end-row corrections) and store
#pragma unroll
  for (int u = 0; u < NT; ++u) {
    const int t = threadIdx.x + u * B;
    if (t < ntask) {
      const int line = t / nch, s0 = (t - line * nch) * L;
      T* o = out + line * NL + g0;
      if (PER) {
#pragma unroll
        for (int k = 0; k < LM; ++k)
          if (k < L) o[s0 + k] = y[u][k];
      } else {
        const T p0 = bnd[line * 3 + 1], pn = bnd[line * 3 + 2];
        const T det = T(1) - K.rnm1 * K.rnm1;
        const T idet = T(1) / det, a = (-p0 + pn * K.rnm1) * idet, bb = (-pn + p0 * K.rnm1) * idet;
#pragma unroll
        for (int k = 0; k < LM; ++k) {
          if (k < L) {
            const int i = s0 + k;
            T v;
            if (i == 0)
              v = raw(line, 0);
            else if (i == n - 1)
              v = raw(line, n - 1);
            else
              v = (y[u][k] + a * pwk(PW, i)) + bb * pwk(PW, n - 1 - i);
            o[i] = v;
          }
        }
      }
    }
  }
}

// One block per line: point fluxes, hybrid interface flux (smooth compact
// flux, or compactified WENO7 of the Lax-Friedrichs split where the
// interface is flagged), divided difference, compact solve.
// axis 0: x-lines of q [4][ny][nx], flux_x, alpha_x, dx.
// axis 1: y-lines of qT [4][nx][ny], flux_y, alpha_y, dy.
// OOP/spatial_operator.py:179-249, :321-342; non-periodic
// OOP/problems/riemann_config3.py:165-265, :344-357.
// LDS holds two 4 x (n+1) planes, the state and the point flux; the WENO7
// values of one or two components at a time are staged in the compact
// solve's scratch and the hybrid fluxes overwrite those components' state
// planes (a 512-node line: 42 KiB for double, three blocks per CU; 21 KiB for
// float).  The WENO work list is built in interface order (wave ballots), so
// neighbouring lanes read neighbouring LDS words.
constexpr int kLineBlock = 512;
template <class T, bool PER, bool SMALL>
__device__ __forceinline__ void line_body(const Params<T>& P, const LineK<T>& K, int axis, int enforced_input,
                                                     const T* __restrict__ ql, const uint8_t* __restrict__ ml,
                                                     const double* __restrict__ scal, T* __restrict__ out, int line, int nl) {
  extern __shared__ __align__(16) unsigned char smem[];
  constexpr int BL = kLineBlock, NWV = kLineBlock / 64;
  const int n = K.n, S = n + 1;
  T* A = reinterpret_cast<T*>(smem);
  T* B = A + 4 * S;
  T* PW = B + 4 * S;                // 66
  T* scratch = PW + 66;             // one-chunk shape: max(CPP*S, 28); else 8*nch + 20
  constexpr int CPP = SMALL ? 2 : 1;
  const int nscr = SMALL ? (CPP * S > 28 ? CPP * S : 28) : 8 * K.nch + 20;
  uint16_t* wl = reinterpret_cast<uint16_t*>(scratch + nscr);  // WENO work list, n+1
  uint8_t* Mk = reinterpret_cast<uint8_t*>(wl + S);
  __shared__ int wcnt[NWV];
  const size_t NL = (size_t)nl * n;
  const T alpha = (T)scal[axis == 0 ? S_ALPHA_X : S_ALPHA_Y];
  const T h = axis == 0 ? P.dx : P.dy;
  const int wv_id = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int k = threadIdx.x; k < 66; k += BL) PW[k] = K.pw[k];
  // 1. (enforce), primitives, point flux
  for (int i = threadIdx.x; i < n; i += BL) {
    const size_t g = (size_t)line * n + i;
    T a = ql[g], b = ql[NL + g], c = ql[2 * NL + g], d = ql[3 * NL + g];
    if (!enforced_input) enforce(P, a, b, c, d);
    const Prim<T> pr = prim(P, a, b, c, d);
    A[i] = a;
    A[S + i] = b;
    A[2 * S + i] = c;
    A[3 * S + i] = d;
    if (axis == 0) {  // OOP/equations.py:183-190
      B[i] = b;
      B[S + i] = pr.rho * (pr.u * pr.u) + pr.p;
      B[2 * S + i] = (pr.rho * pr.u) * pr.v;
      B[3 * S + i] = (d + pr.p) * pr.u;
    } else {  // OOP/equations.py:192-199
      B[i] = c;
      B[S + i] = (pr.rho * pr.u) * pr.v;
      B[2 * S + i] = pr.rho * (pr.v * pr.v) + pr.p;
      B[3 * S + i] = (d + pr.p) * pr.v;
    }
    Mk[i] = ml[g];
  }
  __syncthreads();
  const int nI = PER ? n : n + 1;
  auto iflag = [&](int k) -> bool {
    if (PER) {
      k = wrap(k, n);
      return (Mk[k] | Mk[wrap(k + 1, n)]) != 0;
    }
    k = clampi(k, 0, n);
    return (Mk[clampi(k - 1, 0, n - 1)] | Mk[clampi(k, 0, n - 1)]) != 0;
  };
  auto nidx = [&](int j) -> int { return PER ? wrap(j, n) : clampi(j, 0, n - 1); };
  // 2a. the interfaces whose WENO7 flux a flagged interface's compactification
  //     reads (i-1, i, i+1 of every flagged i), in increasing order; each
  //     thread takes two consecutive interfaces (one round up to 1024)
  int nw = 0;
#pragma unroll 1
  for (int k0 = 0; k0 < nI; k0 += 2 * BL) {
    const int ka = k0 + 2 * threadIdx.x, kb = ka + 1;
    bool fa = false, fb = false;
    if (ka < nI) {
      if (P.weno_only) {
        fa = true;
        fb = kb < nI;
      } else {
        const bool g0 = iflag(ka - 1), g1 = iflag(ka), g2 = iflag(kb);
        fa = g0 || g1 || g2;
        fb = kb < nI && (g1 || g2 || iflag(kb + 1));
      }
    }
    const uint64_t ba = __ballot(fa), bb = __ballot(fb);
    const int pre = (int)(__builtin_amdgcn_mbcnt_hi((unsigned)(ba >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)ba, 0u)) +
                          __builtin_amdgcn_mbcnt_hi((unsigned)(bb >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)bb, 0u)));
    if (lane == 0) wcnt[wv_id] = __popcll(ba) + __popcll(bb);
    __syncthreads();
    int off = nw, tot = 0;
#pragma unroll
    for (int j = 0; j < NWV; ++j) {
      off += j < wv_id ? wcnt[j] : 0;
      tot += wcnt[j];
    }
    if (fa) wl[off + pre] = (uint16_t)ka;
    if (fb) wl[off + pre + (fa ? 1 : 0)] = (uint16_t)kb;
    nw += tot;
    __syncthreads();
  }
  // 2b/3. CPP components per pass: WENO7 of the Lax-Friedrichs split at the
  //       listed interfaces into W (the solve's scratch, free until step 4;
  //       two planes in the one-chunk shape, one otherwise; one plane for a
  //       512-node double line would fit four blocks per CU, but the two
  //       extra passes cost more than the occupancy gains), then the hybrid
  //       interface flux -- compactified WENO where flagged, smooth compact
  //       flux elsewhere -- over those components' state planes.  Each WENO
  //       item runs on a lane pair: the even lane the f+ stencil, the odd
  //       lane f-, summed (f+ first) through a lane exchange.
  T* W = scratch;  // CPP * (n + 1) values
  const T c3 = T(-1.0 / 480.0), c2 = T(1.0 / 20.0 + (-1.0 / 480.0)),
          c1 = T((25.0 / 32.0 + 1.0 / 20.0) + (-1.0 / 480.0));
  // unrolled for float, rolled for double: each keeps the kernel near
  // 64 VGPRs (eight wavefronts per SIMD)
  constexpr int CU = sizeof(T) == 4 ? 4 / CPP : 1;
#pragma unroll CU
  for (int c0 = 0; c0 < 4; c0 += CPP) {
    const int nt = 2 * CPP * nw;  // even: both lanes of a pair are live together
#pragma unroll 1
    for (int t = threadIdx.x; t < nt; t += BL) {
      const int cc = t / (2 * nw), r = t - cc * (2 * nw);
      const int k = wl[r >> 1], side = r & 1;
      const int il = PER ? k : k - 1;
      const T* qa = A + (c0 + cc) * S;
      const T* fa = B + (c0 + cc) * S;
      T vs[7];
#pragma unroll
      for (int j = 0; j < 7; ++j) {
        // f+ stencil i-3 .. i+3: 0.5 (f + alpha q); f- stencil i+4 .. i-2: 0.5 (f - alpha q)
        const int ix = side ? nidx(il + 4 - j) : nidx(il - 3 + j);
        vs[j] = T(0.5) * madd(side ? -alpha : alpha, qa[ix], fa[ix]);
      }
      const T wv = weno7(vs[0], vs[1], vs[2], vs[3], vs[4], vs[5], vs[6]);
      const T wo = __shfl_xor(wv, 1, 64);
      if (!side) W[cc * S + k] = wv + wo;
    }
    __syncthreads();
#pragma unroll 1
    for (int t = threadIdx.x; t < CPP * nI; t += BL) {
      const int cc = t / nI, k = t - cc * nI;
      const T* Wc = W + cc * S;
      const T* fa = B + (c0 + cc) * S;
      const bool fl = P.weno_only || iflag(k);
      const int il = PER ? k : k - 1;
      T F;
      if (fl) {
        const int kr = PER ? wrap(k + 1, n) : clampi(k + 1, 0, n);
        const int kl = PER ? wrap(k - 1, n) : clampi(k - 1, 0, n);
        F = madd(K.alpha_c, Wc[kl], madd(K.alpha_c, Wc[kr], Wc[k]));
      } else {
        F = madd(c1, fa[nidx(il + 1)] + fa[nidx(il)],
                 madd(c3, fa[nidx(il + 3)] + fa[nidx(il - 2)], c2 * (fa[nidx(il + 2)] + fa[nidx(il - 1)])));
      }
      A[(c0 + cc) * S + k] = F;  // the state planes of this pass are no longer read
    }
    __syncthreads();
  }
  // 4. divided difference, compact solve, store
  solve_store<T, PER, SMALL>(A, S, h, K, PW, scratch, out, NL, (size_t)line * n);
}

// Line-sweep launch.  AX = 2: both sweeps of a stage in one launch, blocks
// [0, nlx) take the x-lines of q (enforced on load), blocks [nlx, nlx + nly)
// the y-lines of the enforced transposed state qT (either count may be 0);
// AX = 0 or 1: one sweep (one inlined body: the general line shape needs
// more registers than two bodies in one kernel leave it).
template <class T, bool PER, bool SMALL, int AX>
__global__ void __launch_bounds__(kLineBlock, 4)
    k_line(Params<T> P, LineK<T> KX, LineK<T> KY, const T* __restrict__ qx, const uint8_t* __restrict__ mx,
           const T* __restrict__ qy, const uint8_t* __restrict__ my, const double* __restrict__ scal,
           T* __restrict__ ox, T* __restrict__ oy, int nlx, int nly) {
  const int b = blockIdx.x;
  if (AX == 0 || (AX == 2 && b < nlx))
    line_body<T, PER, SMALL>(P, KX, 0, 0, qx, mx, scal, ox, b, P.ny);
  else
    line_body<T, PER, SMALL>(P, KY, 1, 1, qy, my, scal, oy, AX == 2 ? b - nlx : b, P.nx);
}

// ================================================================ stage update
// Periodic Navier-Stokes terms (OOP/equations.py:232-267) in the reference's
// flux form: the stresses tau_xx, tau_yy, tau_xy and the energy fluxes
// u tau_xx + v tau_xy + k dT/dx, u tau_xy + v tau_yy + k dT/dy are central
// differences of the primitives at each node, and the viscous terms their
// central differences.  A tile forms the five flux values once per node of
// its one-node halo ring (from the primitives u, v, p/(rho R) of its
// two-node halo in LDS), then each node differences its neighbours' values:
// the same expressions, evaluated once instead of at every neighbour.
// Update tiles are kUTX x kUTY nodes, one thread each: a row of the tile is
// one 128-byte line of an FP32 plane (two of an FP64 plane).
constexpr int kUTX = 32, kUTY = 16, kUBlock = kUTX * kUTY;
// LDS of an update tile of TX x TY nodes (TX * TY = kUBlock threads): the
// primitives with their 2-node halo, the viscous fluxes with their 1-node
// halo, the tile's y-derivative staged transposed, the drag tables of the
// tile's rows and columns.
template <class T, int TX, int TY>
struct UpdLds {
  static constexpr int VW = TX + 4, VH = TY + 4, VS = VW + 1;  // primitives tile
  static constexpr int FW = TX + 2, FH = TY + 2, FS = FW + 1;  // flux tile
  T U[VH * VS], V[VH * VS], Tt[VH * VS];
  T xx[FH * FS], yy[FH * FS], xy[FH * FS], ex[FH * FS], ey[FH * FS];
  T D[4][TX][TY + 1];
  // the tile rows' drag table rows ([TY][10]), then its columns' twiddles of
  // |kx| = 1, 2 ([TX][4])
  double drag[TY * 10 + TX * 4];
};
// The fluxes at the FW x FH halo-ring nodes (the primitives in LDS); the
// caller synchronises before and after.
template <class T, int TX, int TY>
__device__ __forceinline__ void viscous_fluxes(const Params<T>& P, UpdLds<T, TX, TY>& L) {
  using Ls = UpdLds<T, TX, TY>;
  for (int i = threadIdx.x; i < Ls::FW * Ls::FH; i += kUBlock) {
    const int fr = i / Ls::FW, fc = i - fr * Ls::FW;
    const int b = (fr + 1) * Ls::VS + (fc + 1);  // the node in the primitives tile
    const T ux = dv(L.U[b + 1] - L.U[b - 1], P.two_dx);
    const T uy = dv(L.U[b + Ls::VS] - L.U[b - Ls::VS], P.two_dy);
    const T vx = dv(L.V[b + 1] - L.V[b - 1], P.two_dx);
    const T vy = dv(L.V[b + Ls::VS] - L.V[b - Ls::VS], P.two_dy);
    const T txx = madd(-P.tt_mu, ux + vy, P.two_mu * ux);
    const T tyy = madd(-P.tt_mu, ux + vy, P.two_mu * vy);
    const T txy = P.mu * (uy + vx);
    const T tx = dv(L.Tt[b + 1] - L.Tt[b - 1], P.two_dx);
    const T ty = dv(L.Tt[b + Ls::VS] - L.Tt[b - Ls::VS], P.two_dy);
    const T u = L.U[b], v = L.V[b];
    const int f = fr * Ls::FS + fc;
    L.xx[f] = txx;
    L.yy[f] = tyy;
    L.xy[f] = txy;
    L.ex[f] = madd(P.kcond, tx, madd(v, txy, u * txx));
    L.ey[f] = madd(P.kcond, ty, madd(v, tyy, u * txy));
  }
}
// Viscous terms at the tile node (ty, tx) from the fluxes in LDS.
template <class T, int TX, int TY>
__device__ __forceinline__ void viscous_div(const Params<T>& P, const UpdLds<T, TX, TY>& L, int ty, int tx,
                                            T out[4]) {
  constexpr int FS = UpdLds<T, TX, TY>::FS;
  const int c = (ty + 1) * FS + (tx + 1);
  out[0] = T(0);
  out[1] = dv(L.xx[c + 1] - L.xx[c - 1], P.two_dx) + dv(L.xy[c + FS] - L.xy[c - FS], P.two_dy);
  out[2] = dv(L.xy[c + 1] - L.xy[c - 1], P.two_dx) + dv(L.yy[c + FS] - L.yy[c - FS], P.two_dy);
  out[3] = dv(L.ex[c + 1] - L.ex[c - 1], P.two_dx) + dv(L.ey[c + FS] - L.ey[c - FS], P.two_dy);
}

// Primitives u, v, p/(rho R) of the (TX+4) x (TY+4) tile around a TX x TY
// block (periodic wrap) into LDS.
// Split in two so that a caller can issue the loads, do other work, then
// form the primitives: each thread holds at most two of the tile nodes.
template <class T>
struct VTileLoad {
  T a[2][4];
};
template <class T, int TX, int TY>
__device__ __forceinline__ VTileLoad<T> viscous_tile_load(const Params<T>& P, const T* __restrict__ q, int x0,
                                                           int y0) {
  constexpr int VW = TX + 4, VH = TY + 4;
  static_assert(VW * VH <= 2 * kUBlock, "two tile nodes per thread");
  const int nx = P.nx, ny = P.ny, N = nx * ny;
  VTileLoad<T> v;
#pragma unroll
  for (int s = 0; s < 2; ++s) {
    const int i = threadIdx.x + s * kUBlock;
    if (i < VW * VH) {
      const int ly = i / VW, lx = i - ly * VW;
      const int gy = wrap(y0 - 2 + ly, ny), gx = wrap(x0 - 2 + lx, nx);
      load4(q, N, gy * nx + gx, v.a[s][0], v.a[s][1], v.a[s][2], v.a[s][3]);
    }
  }
  return v;
}
template <class T, int TX, int TY>
__device__ __forceinline__ void viscous_tile_store(const Params<T>& P, const VTileLoad<T>& v,
                                                   UpdLds<T, TX, TY>& L) {
  using Ls = UpdLds<T, TX, TY>;
#pragma unroll
  for (int s = 0; s < 2; ++s) {
    const int i = threadIdx.x + s * kUBlock;
    if (i < Ls::VW * Ls::VH) {
      const int ly = i / Ls::VW, lx = i - ly * Ls::VW;
      const Prim<T> pr = prim(P, v.a[s][0], v.a[s][1], v.a[s][2], v.a[s][3]);
      L.U[ly * Ls::VS + lx] = pr.u;
      L.V[ly * Ls::VS + lx] = pr.v;
      L.Tt[ly * Ls::VS + lx] = dv(pr.p, pr.rho * P.rgas);
    }
  }
}

enum { MODE_RHS = 0, MODE_STAGE = 1, MODE_VISC = 2 };

// rhs = -dF/dx - dG/dy (OOP/spatial_operator.py:367); in MODE_STAGE also the
// HIT closure (OOP/hit2d.py:847-867: + viscous + forcing + drag + cooling),
// the Shu-Osher combination (OOP/time_operator.py:24-26) and clean =
// enforce (+ outflow guard: every node takes the value of its clamped
// interior source, OOP/problems/riemann_config3.py:419-429).
// One TX x TY tile per block (kUBlock threads, thread t at (t % TX, t / TX)):
// the y-derivative (stored transposed, [c][x][y]) is staged through LDS so
// both derivative reads are coalesced.  The node arithmetic does not depend
// on the tile shape.  Every thread passes the tile's barriers.
template <class T, bool PER_ONLY, int TX, int TY>  // PER_ONLY: a periodic grid (no outflow guard), known at compile time
__device__ __forceinline__ void update_tile(const Params<T>& P, int mode, int stage, const T* __restrict__ q0,
                                            const T* __restrict__ qs, const T* __restrict__ dfx,
                                            const T* __restrict__ dfyT, const T* __restrict__ fx,
                                            const T* __restrict__ fy, const double* __restrict__ scal,
                                            const double* __restrict__ drows, T dt, T* __restrict__ out, int x0,
                                            int y0, UpdLds<T, TX, TY>& L, T fsc, T fal) {
  static_assert(TX * TY == kUBlock, "one thread per tile node");
  const int nx = P.nx, ny = P.ny, N = nx * ny;
  const int tx = threadIdx.x % TX, ty = threadIdx.x / TX;
  const int x = x0 + tx, y = y0 + ty;
  const bool in = x < nx && y < ny;
  int cx = x, cy = y;
  if (!PER_ONLY && P.guard > 0) {
    cx = clampi(x, P.guard, nx - 1 - P.guard);
    cy = clampi(y, P.guard, ny - 1 - P.guard);
  }
  const int idx = cy * nx + cx, oidx = y * nx + x;
  const bool visc = mode == MODE_VISC || (mode == MODE_STAGE && P.visc_on);
  // the node's own operands are loaded before the tile phase so their
  // latency overlaps it
  T gx[4] = {T(0), T(0), T(0), T(0)}, a0[4] = {T(0), T(0), T(0), T(0)};
  T s0 = T(0), s1 = T(0), s2 = T(0), s3 = T(0), f1r = T(0), f2r = T(0);
  if (in && mode != MODE_VISC) {
#pragma unroll
    for (int c = 0; c < 4; ++c) gx[c] = at32(dfx + (size_t)c * N, idx);
    if (mode == MODE_STAGE) {
      load4(qs, N, idx, s0, s1, s2, s3);
#pragma unroll
      for (int c = 0; c < 4; ++c) a0[c] = at32(q0 + (size_t)c * N, idx);
      if (P.force_on) {
        f1r = at32(fx, idx);
        f2r = at32(fy, idx);
      }
    }
  }
  VTileLoad<T> vt;
  if (visc) vt = viscous_tile_load<T, TX, TY>(P, qs, x0, y0);
  T dsh[4];
  if (mode != MODE_VISC) {
    // dfyT tile: thread t reads x-row x0 + t / TY, y = y0 + t % TY (unit stride in y)
    const int rx = x0 + (int)threadIdx.x / TY, ry = y0 + (int)threadIdx.x % TY;
#pragma unroll
    for (int c = 0; c < 4; ++c) dsh[c] = (rx < nx && ry < ny) ? at32(dfyT + (size_t)c * N, rx * ny + ry) : T(0);
  }
  // drag low-pass tables (|kx| <= 2 for drag modes), one entry per thread
  // for LDS: the per-row table's rows of the tile, then the column twiddles
  // of |kx| = 1, 2 (cos, sin of |kx| = 1, then of 2)
  constexpr int kDragTab = TY * 10 + TX * 4;
  static_assert(kDragTab <= kUBlock, "one drag table entry per thread");
  const bool drag = mode == MODE_STAGE && P.drag_on;
  double dtab = 0.0;
  if (drag && threadIdx.x < kDragTab) {
    const bool clamp = !PER_ONLY && P.guard > 0;
    const int i = threadIdx.x;
    if (i < TY * 10) {
      const int r = i / 10, j = i - r * 10;
      int yr = min(y0 + r, ny - 1);  // rows past the grid: any in-range row (unused)
      if (clamp) yr = clampi(yr, P.guard, ny - 1 - P.guard);
      dtab = drows[(size_t)yr * 10 + j];
    } else {
      const int e = i - TY * 10, c = e >> 2, k = ((e >> 1) & 1) + 1;
      int xc = min(x0 + c, nx - 1);
      if (clamp) xc = clampi(xc, P.guard, nx - 1 - P.guard);
      const int kk = k <= P.dkx ? k : 0;  // an in-range row when |kx| = k is absent
      dtab = P.twx[((size_t)(kk + P.kmx) * nx + xc) * 2 + (e & 1)];
    }
  }
  // node-local source terms while the tile loads are in flight
  T fsrc[4] = {T(0), T(0), T(0), T(0)};
  if (in && mode == MODE_STAGE && P.force_on) {  // OOP/hit2d.py:264-276; f = (f_raw*scale)*alpha (OOP/forcing.py:91-112)
    const Prim<T> st = prim(P, s0, s1, s2, s3);
    const T f1 = (f1r * fsc) * fal, f2 = (f2r * fsc) * fal;
    fsrc[1] = st.rho * f1;
    fsrc[2] = st.rho * f2;
    fsrc[3] = st.rho * madd(st.v, f2, st.u * f1);
  }
  if (visc) viscous_tile_store<T, TX, TY>(P, vt, L);
  if (mode != MODE_VISC) {
#pragma unroll
    for (int c = 0; c < 4; ++c) L.D[c][threadIdx.x / TY][threadIdx.x % TY] = dsh[c];
  }
  if (drag && threadIdx.x < kDragTab) L.drag[threadIdx.x] = dtab;
  __syncthreads();
  if (visc) {
    viscous_fluxes<T, TX, TY>(P, L);
    __syncthreads();
  }
  if (!in) return;
  if (mode == MODE_VISC) {  // standalone viscous_rhs of qs
    T vv[4];
    viscous_div<T, TX, TY>(P, L, ty, tx, vv);
#pragma unroll
    for (int c = 0; c < 4; ++c) at32(out + (size_t)c * N, oidx) = vv[c];
    return;
  }
  const bool own = PER_ONLY || (cx == x && cy == y);  // guard nodes take their interior source's values
  T Lv[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const T gy = own ? L.D[c][tx][ty] : ld_global(dfyT + c * N + (size_t)cx * ny + cy);
    Lv[c] = (-gx[c]) - gy;
  }
  if (mode == MODE_RHS) {
#pragma unroll
    for (int c = 0; c < 4; ++c) at32(out + (size_t)c * N, oidx) = Lv[c];
    return;
  }
  if (P.visc_on) {
    T vv[4];
    viscous_div<T, TX, TY>(P, L, ty, tx, vv);
#pragma unroll
    for (int c = 0; c < 4; ++c) Lv[c] = Lv[c] + vv[c];
  }
  // the reference's order: rhs + viscous, + forcing, + drag, + cooling
  if (P.force_on) {
#pragma unroll
    for (int c = 1; c < 4; ++c) Lv[c] = Lv[c] + fsrc[c];
  }
  if (drag) {  // OOP/hit2d.py:370-399, low-pass field from the per-row table
    const Prim<T> st = prim(P, s0, s1, s2, s3);
    const double* H = L.drag + ty * 10;
    const double* tw = L.drag + TY * 10 + tx * 4;
    double lx = H[0], ly = H[5];
#pragma unroll
    for (int k = 1; k <= 2; ++k) {
      if (k > P.dkx) break;
      const double c = tw[2 * k - 2], sn = tw[2 * k - 1];
      lx += H[2 * k - 1] * c - H[2 * k] * sn;
      ly += H[5 + 2 * k - 1] * c - H[5 + 2 * k] * sn;
    }
    const T d1 = P.neg_drag * (T)lx, d2 = P.neg_drag * (T)ly;
    Lv[1] = Lv[1] + d1;
    Lv[2] = Lv[2] + d2;
    Lv[3] = Lv[3] + madd(st.v, d2, st.u * d1);
  }
  if (P.cool_on) Lv[3] = Lv[3] + (T)scal[S_COOL];  // OOP/hit2d.py:402-426
  T o[4];
  const T qsv[4] = {s0, s1, s2, s3};
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    if (stage == 0)
      o[c] = madd(dt, Lv[c], a0[c]);
    else if (stage == 1)
      o[c] = madd(T(0.75), a0[c], T(0.25) * madd(dt, Lv[c], qsv[c]));
    else
      o[c] = madd(T(1.0 / 3.0), a0[c], T(2.0 / 3.0) * madd(dt, Lv[c], qsv[c]));
  }
  enforce(P, o[0], o[1], o[2], o[3]);
#pragma unroll
  for (int c = 0; c < 4; ++c) at32(out + (size_t)c * N, oidx) = o[c];
}

// The kUTX x kUTY tile (bx, by) with its LDS in static shared memory.
template <class T, bool PER_ONLY = false>
__device__ __forceinline__ void update_body(const Params<T>& P, int mode, int stage, const T* __restrict__ q0,
                                            const T* __restrict__ qs, const T* __restrict__ dfx,
                                            const T* __restrict__ dfyT, const T* __restrict__ fx,
                                            const T* __restrict__ fy, const double* __restrict__ scal,
                                            const double* __restrict__ drows, T dt, T* __restrict__ out, int bx,
                                            int by, T fsc = T(1), T fal = T(1)) {
  __shared__ UpdLds<T, kUTX, kUTY> L;
  update_tile<T, PER_ONLY, kUTX, kUTY>(P, mode, stage, q0, qs, dfx, dfyT, fx, fy, scal, drows, dt, out, bx * kUTX,
                                       by * kUTY, L, fsc, fal);
}

template <class T>
__global__ void __launch_bounds__(kUBlock) k_update(Params<T> P, int mode, int stage, const T* __restrict__ q0,
                                                   const T* __restrict__ qs, const T* __restrict__ dfx,
                                                   const T* __restrict__ dfyT, const T* __restrict__ fx,
                                                   const T* __restrict__ fy, const double* __restrict__ scal,
                                                   const double* __restrict__ drows, T dt, T* __restrict__ out) {
  const Tile2 t_ = tile_xcd();
  update_body<T>(P, mode, stage, q0, qs, dfx, dfyT, fx, fy, scal, drows, dt, out, t_.bx, t_.by);
}

// ================================================================ hyperviscosity
// q - mn*w*h^4*active*lap(lap(q)), then enforce (+ guard).  Periodic
// Laplacian with wrap (OOP/spatial_operator.py:394-435) or edge-padded
// (OOP/problems/riemann_config3.py:390-416).
template <class T>
__device__ __forceinline__ T lap_at(const Params<T>& P, const T* __restrict__ a, int y, int x) {
  const int nx = P.nx, ny = P.ny;
  int xl, xr, yl, yr;
  if (P.periodic) {
    xl = wrap(x - 1, nx); xr = wrap(x + 1, nx); yl = wrap(y - 1, ny); yr = wrap(y + 1, ny);
  } else {
    xl = clampi(x - 1, 0, nx - 1); xr = clampi(x + 1, 0, nx - 1);
    yl = clampi(y - 1, 0, ny - 1); yr = clampi(y + 1, 0, ny - 1);
  }
  const T c = a[y * nx + x];
  const T lx = dv((a[y * nx + xr] - T(2.0) * c) + a[y * nx + xl], P.dx2);
  const T ly = dv((a[yr * nx + x] - T(2.0) * c) + a[yl * nx + x], P.dy2);
  return lx + ly;
}

// 16x16 output tile per block; per component the state on the tile plus a
// 2-node halo and its Laplacian on the tile plus a 1-node halo go through
// LDS.  Edge padding (non-periodic) means the Laplacian at an out-of-grid
// neighbour is the Laplacian at the clamped node, so every halo position
// evaluates at its wrapped or clamped node.  Guard nodes whose source lies
// outside the tile take the direct path.
template <class T>
__device__ __forceinline__ void hv_body(const Params<T>& P, const T* __restrict__ q, const uint8_t* __restrict__ active,
                                        int guard, int invert, T* __restrict__ out, int bx, int by) {
  __shared__ T sq[4][20][21];
  __shared__ T sl[18][19];
  const int nx = P.nx, ny = P.ny, N = nx * ny;
  const int x0 = bx * 16, y0 = by * 16;
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  const int x = x0 + tx, y = y0 + ty;
  const bool per = P.periodic;
  auto gmap = [&](int g, int n) {
    return per ? wrapn(g, n) : clampi(g, 0, n - 1);
  };
  int cx = x, cy = y;
  if (guard && P.guard > 0) {
    cx = clampi(x, P.guard, nx - 1 - P.guard);
    cy = clampi(y, P.guard, ny - 1 - P.guard);
  }
  const bool in = x < nx && y < ny;
  const bool tiled = cx >= x0 && cx < x0 + 16 && cy >= y0 && cy < y0 + 16;
  const bool act = in && ((active[gmap(cy, ny) * nx + gmap(cx, nx)] != 0) != (invert != 0));
  // tile coordinate of the node a tile position evaluates at (wrapped: the
  // position itself; clamped: the clamped node)
  auto tcx = [&](int g) { return per ? g - (x0 - 2) : clampi(g, 0, nx - 1) - (x0 - 2); };
  auto tcy = [&](int g) { return per ? g - (y0 - 2) : clampi(g, 0, ny - 1) - (y0 - 2); };
  // the four components' tiles (state plus a 2-node halo), every load in
  // flight together
  constexpr int NT = 4 * 400, NL = (NT + kBlock - 1) / kBlock;
  T v[NL];
#pragma unroll
  for (int s = 0; s < NL; ++s) {
    const int i = threadIdx.x + s * kBlock;
    if (i < NT) {
      const int c = i / 400, r = i - c * 400, ly = r / 20, lx = r - ly * 20;
      v[s] = q[(size_t)c * N + gmap(y0 - 2 + ly, ny) * nx + gmap(x0 - 2 + lx, nx)];
    }
  }
#pragma unroll
  for (int s = 0; s < NL; ++s) {
    const int i = threadIdx.x + s * kBlock;
    if (i < NT) {
      const int c = i / 400, r = i - c * 400, ly = r / 20, lx = r - ly * 20;
      sq[c][ly][lx] = v[s];
    }
  }
  T o[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const T* a = q + (size_t)c * N;
    T(*sqc)[21] = sq[c];
    __syncthreads();  // tiles stored; sl free (the previous component's biharmonic has read it)
    for (int i = threadIdx.x; i < 324; i += kBlock) {
      const int ly = i / 18, lx = i - ly * 18;
      const int gy = y0 - 1 + ly, gx = x0 - 1 + lx;
      const int ey = tcy(gy), ex = tcx(gx);  // node evaluated, tile coordinates
      const int exl = tcx(per ? gx - 1 : clampi(gx, 0, nx - 1) - 1), exr = tcx(per ? gx + 1 : clampi(gx, 0, nx - 1) + 1);
      const int eyl = tcy(per ? gy - 1 : clampi(gy, 0, ny - 1) - 1), eyr = tcy(per ? gy + 1 : clampi(gy, 0, ny - 1) + 1);
      const T cc = sqc[ey][ex];
      const T lx_ = dv((sqc[ey][exr] - T(2.0) * cc) + sqc[ey][exl], P.dx2);
      const T ly_ = dv((sqc[eyr][ex] - T(2.0) * cc) + sqc[eyl][ex], P.dy2);
      sl[ly][lx] = lx_ + ly_;
    }
    __syncthreads();
    if (in) {
      T bih, ac;
      if (tiled) {
        const int lcx = cx - x0 + 1, lcy = cy - y0 + 1;  // Laplacian tile coordinates
        // neighbours of the (clamped) node: positions whose evaluation node is the neighbour
        int lxl, lxr, lyl, lyr;
        if (per) {
          lxl = lcx - 1; lxr = lcx + 1; lyl = lcy - 1; lyr = lcy + 1;
        } else {
          lxl = clampi(cx - 1, 0, nx - 1) - x0 + 1; lxr = clampi(cx + 1, 0, nx - 1) - x0 + 1;
          lyl = clampi(cy - 1, 0, ny - 1) - y0 + 1; lyr = clampi(cy + 1, 0, ny - 1) - y0 + 1;
        }
        const T Lc = sl[lcy][lcx];
        const T lx_ = dv((sl[lcy][lxr] - T(2.0) * Lc) + sl[lcy][lxl], P.dx2);
        const T ly_ = dv((sl[lyr][lcx] - T(2.0) * Lc) + sl[lyl][lcx], P.dy2);
        bih = lx_ + ly_;
        ac = sqc[cy - y0 + 2][cx - x0 + 2];
      } else {
        int xl, xr, yl, yr;
        if (per) {
          xl = wrap(cx - 1, nx); xr = wrap(cx + 1, nx); yl = wrap(cy - 1, ny); yr = wrap(cy + 1, ny);
        } else {
          xl = clampi(cx - 1, 0, nx - 1); xr = clampi(cx + 1, 0, nx - 1);
          yl = clampi(cy - 1, 0, ny - 1); yr = clampi(cy + 1, 0, ny - 1);
        }
        const T Lc = lap_at(P, a, cy, cx);
        const T lx_ = dv((lap_at(P, a, cy, xr) - T(2.0) * Lc) + lap_at(P, a, cy, xl), P.dx2);
        const T ly_ = dv((lap_at(P, a, yr, cx) - T(2.0) * Lc) + lap_at(P, a, yl, cx), P.dy2);
        bih = lx_ + ly_;
        ac = ld_global(a + cy * nx + cx);
      }
      const T coef = act ? P.hv_coef[c] : T(0);
      o[c] = ac - (coef * bih);
    }
  }
  if (!in) return;
  enforce(P, o[0], o[1], o[2], o[3]);
#pragma unroll
  for (int c = 0; c < 4; ++c) out[(size_t)c * N + y * nx + x] = o[c];
}

template <class T>
__global__ void __launch_bounds__(kBlock) k_hv(Params<T> P, const T* __restrict__ q, const uint8_t* __restrict__ active,
                                               int guard, int invert, T* __restrict__ out) {
  const Tile2 t_ = tile_xcd();
  hv_body<T>(P, q, active, guard, invert, out, t_.bx, t_.by);
}

// clean(): enforce_physical_state, then the outflow guard if requested.
template <class T>
__device__ __forceinline__ void enforce_body(const Params<T>& P, const T* __restrict__ q, int guard,
                                                    T* __restrict__ out, int bx, int by) {
  const int nx = P.nx, ny = P.ny, N = nx * ny;
  const int x = bx * 16 + (threadIdx.x & 15), y = by * 16 + (threadIdx.x >> 4);
  if (x >= nx || y >= ny) return;
  int cx = x, cy = y;
  if (guard && P.guard > 0) {
    cx = clampi(x, P.guard, nx - 1 - P.guard);
    cy = clampi(y, P.guard, ny - 1 - P.guard);
  }
  T a, b, c, d;
  load4(q, N, cy * nx + cx, a, b, c, d);
  enforce(P, a, b, c, d);
  const int o = y * nx + x;
  out[o] = a;
  out[N + o] = b;
  out[2 * N + o] = c;
  out[3 * N + o] = d;
}

template <class T>
__global__ void __launch_bounds__(kBlock) k_enforce(Params<T> P, const T* __restrict__ q, int guard,
                                                    T* __restrict__ out) {
  enforce_body<T>(P, q, guard, out, blockIdx.x, blockIdx.y);
}

// ================================================================ forcing
constexpr int kFK = 12;  // largest |kx|, |ky| index of a forcing mode
constexpr int kProjChunk = 16;  // rows of row twiddles a projection wave stages in LDS at a time

// fft2(noise) at the shell modes (OOP/forcing.py:78-79): per block and mode
// the partial sum of noise*exp(-i(kx x + ky y)), re and im.  Column walk
// (no halo, P.rbp rows per wave): each lane sums noise*cos/sin(ky y) for
// ky = 0..fky over its rows; the lanes' sums go to LDS (dynamic, 2*fky+1
// rows of 64 per wave) and lane j then forms mode j over the wave's columns
// with exp(-i kx x).
template <class T>
__device__ __forceinline__ void fproject_body(const Params<T>& P, const double* __restrict__ noise,
                                              double* __restrict__ part, int bid) {
  extern __shared__ double dsh[];
  __shared__ double sh[kWaves][2 * HITX_MAX_MODES];
  const int nx = P.nx, ny = P.ny;
  const Walk W = walk_setup(nx, ny, 1, P.rbp, 0, bid);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int NS = 2 * P.fky + 1;
  double ac[kFK + 1], as[kFK + 1];
#pragma unroll
  for (int k = 0; k <= kFK; ++k) {
    ac[k] = 0.0;
    as[k] = 0.0;
  }
  if (W.live) {
    // rows in groups of 8 with the group's noise loads in flight together;
    // every k up to kFK without a branch: the sums past fky read the table
    // row of fky and are never stored.  The rows' twiddles cos/sin(k y),
    // k = 1..fky, are copied 16 rows at a time into the wave's LDS slice (the
    // end phase's area, free during the walk: 2 fky x 16 <= 64 (2 fky + 1)
    // doubles) and read as broadcasts (no scalar loads).
    const int fk = P.fky, fs = fk > 0 ? fk : 1;
    double* tw = dsh + (size_t)(threadIdx.x >> 6) * (2 * fk + 1) * 64;
    for (int c0 = W.ya; c0 < W.yb; c0 += kProjChunk) {
      const int rows = min(kProjChunk, W.yb - c0);
      __builtin_amdgcn_wave_barrier();  // the previous chunk's reads are done (in order within the wave)
      for (int i = lane; i < rows * fk; i += 64) {
        const int r = i / fk, k = i - r * fk + 1;
        const double* ty = P.twy + ((size_t)(k + P.kmy) * ny + c0 + r) * 2;
        tw[(r * fs + k - 1) * 2] = ty[0];
        tw[(r * fs + k - 1) * 2 + 1] = ty[1];
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      for (int y0 = c0; y0 < c0 + rows; y0 += 8) {
        double nv[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) nv[j] = (W.own && y0 + j < W.yb) ? noise[(size_t)(y0 + j) * nx + W.x] : 0.0;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          const int y = y0 + j;
          if (y < c0 + rows) {
            const double n = nv[j];
            const double* t = tw + (y - c0) * fs * 2;
            ac[0] += n;
#pragma unroll
            for (int k = 1; k <= kFK; ++k) {
              const int kk = min(k, fs);
              ac[k] += n * t[2 * (kk - 1)];
              as[k] += n * t[2 * (kk - 1) + 1];
            }
          }
        }
      }
    }
  }
  double* L = dsh + (size_t)wid * NS * 64;
  L[lane] = ac[0];
#pragma unroll
  for (int k = 1; k <= kFK; ++k) {
    if (k <= P.fky) {
      L[(2 * k - 1) * 64 + lane] = ac[k];
      L[2 * k * 64 + lane] = as[k];
    }
  }
  __syncthreads();
  const int x0 = W.x - lane;
  for (int m = lane; m < P.nforce; m += 64) {
    const int kx = P.force_modes[2 * m], ky = P.force_modes[2 * m + 1];
    const int ka = ky < 0 ? -ky : ky;
    double re = 0.0, im = 0.0;
    if (W.live) {
      // every column in turn (lanes past the grid's edge hold zero sums, and
      // read a clamped twiddle), so the unrolled loads are independent
#pragma unroll 8
      for (int l = 0; l < 64; ++l) {
        const int x = min(x0 + l, nx - 1);
        const double* tx = P.twx + ((size_t)(kx + P.kmx) * nx + x) * 2;
        const double cx = tx[0], sx = tx[1];
        const double br = ka == 0 ? L[l] : L[(2 * ka - 1) * 64 + l];
        const double bi = ka == 0 ? 0.0 : (ky > 0 ? -L[2 * ka * 64 + l] : L[2 * ka * 64 + l]);
        re += cx * br + sx * bi;
        im += cx * bi - sx * br;
      }
    }
    sh[wid][2 * m] = re;
    sh[wid][2 * m + 1] = im;
  }
  __syncthreads();
  for (int j = threadIdx.x; j < 2 * P.nforce; j += blockDim.x) {
    double a = sh[0][j];
    for (int w = 1; w < kWaves; ++w) a += sh[w][j];
    part[(size_t)bid * 2 * P.nforce + j] = a;
  }
}

template <class T>
__global__ void __launch_bounds__(kBlock) k_fproject(Params<T> P, const double* __restrict__ noise,
                                                     double* __restrict__ part) {
  fproject_body<T>(P, noise, part, blockIdx.x);
}

// Per-row table of the forcing synthesis (one block): for row y and field f
// (fx, fy), G(kx) = sum over the modes with that kx of fhat_f(m)*exp(i ky y),
// frows[y][f] = {Re G(0), then Re(G(k)+G(-k)), Im(G(k)-G(-k)) for k = 1..fkx},
// so that the real inverse DFT at (x, y) is H0 + sum_k (Hc_k cos(k x) -
// Hs_k sin(k x)) (OOP/forcing.py:84-85, before the 1/N).
template <class T>
__device__ __forceinline__ void frows_body(const Params<T>& P, const double* __restrict__ fhat,
                                           double* __restrict__ frows, int t) {
  const int kx1 = P.fkx + 1, FR = 2 * P.fkx + 1, nf = P.nforce;
  if (t >= P.ny * kx1) return;
  const int y = t / kx1, k = t - y * kx1;
  double hr[2] = {0.0, 0.0}, hs[2] = {0.0, 0.0};
  const int* grp = P.force_bykx;  // the modes with |kx| = k, in mode order
  const int q0 = grp[nf + k], q1 = grp[nf + k + 1];
#pragma unroll 4
  for (int qi = q0; qi < q1; ++qi) {
    const int m = grp[qi];
    const int kx = P.force_modes[2 * m], ky = P.force_modes[2 * m + 1];
    const double* ty = P.twy + ((size_t)(ky + P.kmy) * P.ny + y) * 2;
    const double cy = ty[0], sy = ty[1];
#pragma unroll
    for (int f = 0; f < 2; ++f) {
      const double fr = fhat[f * 2 * nf + 2 * m], fi = fhat[f * 2 * nf + 2 * m + 1];
      const double gr = fr * cy - fi * sy, gi = fr * sy + fi * cy;
      hr[f] += gr;
      hs[f] += kx > 0 ? gi : -gi;
    }
  }
#pragma unroll
  for (int f = 0; f < 2; ++f) {
    double* o = frows + ((size_t)y * 2 + f) * FR;
    if (k == 0) {
      o[0] = hr[f];
    } else {
      o[2 * k - 1] = hr[f];
      o[2 * k] = hs[f];
    }
  }
}

template <class T>
__global__ void __launch_bounds__(kBlock) k_frows(Params<T> P, const double* __restrict__ fhat,
                                                  double* __restrict__ frows) {
  frows_body<T>(P, fhat, frows, blockIdx.x * kBlock + threadIdx.x);
}

// Real inverse DFT of the shell coefficients at one node from the row table.
template <class T>
__device__ __forceinline__ void fsynth_at(const Params<T>& P, const double* __restrict__ frows, int y, int x,
                                          double& ax, double& ay) {
  const int FR = 2 * P.fkx + 1;
  const double* H = frows + (size_t)y * 2 * FR;
  ax = H[0];
  ay = H[FR];
#pragma unroll 4
  for (int k = 1; k <= P.fkx; ++k) {
    const double* tx = P.twx + ((size_t)(k + P.kmx) * P.nx + x) * 2;
    const double c = tx[0], sn = tx[1];
    ax += H[2 * k - 1] * c - H[2 * k] * sn;
    ay += H[FR + 2 * k - 1] * c - H[FR + 2 * k] * sn;
  }
}

// fx, fy = real inverse DFT of (fx_hat, fy_hat) on the shell modes
// (OOP/forcing.py:82-85), plus sum(fx^2+fy^2) and sum(rho*(fx*u+fy*v)).
template <class T>
__device__ __forceinline__ void fsynth_body(const Params<T>& P, const double* __restrict__ frows, T* __restrict__ fxo,
                                            T* __restrict__ fyo, const T* __restrict__ rho,
                                            const T* __restrict__ u, const T* __restrict__ v,
                                            double* __restrict__ part, int bid, int nblk) {
  const int nx = P.nx, ny = P.ny, N = nx * ny;
  const double invN = 1.0 / ((double)nx * (double)ny);
  double acc[2] = {0.0, 0.0};
  for (int idx = bid * kBlock + threadIdx.x; idx < N; idx += nblk * kBlock) {
    const int y = idx / nx, x = idx - y * nx;
    double ax, ay;
    fsynth_at(P, frows, y, x, ax, ay);
    const T fxv = (T)(ax * invN), fyv = (T)(ay * invN);
    fxo[idx] = fxv;
    fyo[idx] = fyv;
    acc[0] += (double)(fxv * fxv + fyv * fyv);
    acc[1] += (double)(rho[idx] * (fxv * u[idx] + fyv * v[idx]));
  }
  block_reduce_write<2, 0>(acc, part + (size_t)bid * 2);
}

template <class T>
__global__ void __launch_bounds__(kBlock) k_fsynth(Params<T> P, const double* __restrict__ frows, T* __restrict__ fxo,
                                                   T* __restrict__ fyo, const T* __restrict__ rho,
                                                   const T* __restrict__ u, const T* __restrict__ v,
                                                   double* __restrict__ part) {
  fsynth_body<T>(P, frows, fxo, fyo, rho, u, v, part, blockIdx.x, gridDim.x);
}

// fx = (fx*scale)*alpha, fy likewise (OOP/forcing.py:91-112), and the
// statistics of OOP/forcing.py:113-118.
template <class T>
__device__ __forceinline__ void fscale_body(const Params<T>& P, T* __restrict__ fx, T* __restrict__ fy, T scale,
                                                   T alpha, const T* __restrict__ rho, const T* __restrict__ u,
                                                   const T* __restrict__ v, double* __restrict__ part, int bid, int nblk) {
  const int N = P.nx * P.ny;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  for (int idx = bid * kBlock + threadIdx.x; idx < N; idx += nblk * kBlock) {
    const T a = (fx[idx] * scale) * alpha, b = (fy[idx] * scale) * alpha;
    fx[idx] = a;
    fy[idx] = b;
    acc[0] += (double)(a * a);
    acc[1] += (double)(b * b);
    acc[2] += (double)(a * b);
    acc[3] += (double)(rho[idx] * (a * u[idx] + b * v[idx]));
  }
  block_reduce_write<4, 0>(acc, part + (size_t)bid * 4);
}

template <class T>
__global__ void __launch_bounds__(kBlock) k_fscale(Params<T> P, T* __restrict__ fx, T* __restrict__ fy, T scale,
                                                   T alpha, const T* __restrict__ rho, const T* __restrict__ u,
                                                   const T* __restrict__ v, double* __restrict__ part) {
  fscale_body<T>(P, fx, fy, scale, alpha, rho, u, v, part, blockIdx.x, gridDim.x);
}

// OOP/equations.py:136-145 on the whole grid.
template <class T>
__global__ void __launch_bounds__(kBlock) k_primitives(Params<T> P, const T* __restrict__ q, T* __restrict__ rho,
                                                       T* __restrict__ u, T* __restrict__ v, T* __restrict__ p) {
  const int N = P.nx * P.ny;
  for (int idx = blockIdx.x * kBlock + threadIdx.x; idx < N; idx += gridDim.x * kBlock) {
    const Prim<T> pr = prim_at(P, q, N, idx, 0);
    rho[idx] = pr.rho;
    u[idx] = pr.u;
    v[idx] = pr.v;
    p[idx] = pr.p;
  }
}

}  // namespace hitx
