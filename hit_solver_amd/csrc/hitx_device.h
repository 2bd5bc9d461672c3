// Device-side building blocks of the hitx kernels (gfx950, wave64).
//
// Every floating-point expression below reproduces the operation order of the
// reference NumPy code it cites, so that FP64 results agree bit-for-bit where
// the algorithm is the same (pointwise algebra, stencils, the sensor) and to
// rounding where it is not (WENO7 weights with one division, line solves,
// reductions).  The library is built with -ffp-contract=off: no multiply-add
// contraction in FP64; FP32 fuses explicitly (madd).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

namespace hitx {

constexpr int kBlock = 256;  // threads per block (4 wave64s)
constexpr int kWaves = kBlock / 64;

// A pointer to global memory held in a descriptor.  Kernels of the batched
// path read their buffers' addresses from per-instance descriptors in device
// memory; a plain pointer loaded that way is generic and compiles to flat
// memory instructions (which also count against the LDS wait counter).
// Stored in the global address space, accesses through it compile to global
// loads and stores.  Converts to and from T* on both host and device.
template <class T>
struct GPtr {
  __attribute__((address_space(1))) T* p;
  __host__ __device__ operator T*() const { return (T*)p; }
  __host__ __device__ GPtr& operator=(T* q) {
    p = (__attribute__((address_space(1))) T*)q;
    return *this;
  }
};
// A load known to read global memory: where one expression reads either LDS
// or global memory, a plain load would become one flat load of a selected
// address.
template <class T>
__device__ __forceinline__ T ld_global(const T* p) {
  return *(const __attribute__((address_space(1))) T*)p;
}

// ---------------------------------------------------------------- parameters
// Per-plan constants passed by value to every kernel.  T is the element type
// of the state (double or float); thresholds compared against FP64
// reductions stay double.
template <class T>
struct Params {
  int nx, ny, periodic, weno_only;
  T dx, dy, two_dx, two_dy;          // dx, dy and 2.0*dx, 2.0*dy
  T gamma, gm1, rho_floor, p_floor, pf_gm1;  // gm1 = gamma-1.0, pf_gm1 = p_floor/(gamma-1.0)
  // sensor
  int width, bguard, shear_on;
  double comp_thr, jump_thr, shear_thr;
  // viscous terms
  int visc_on;
  T mu, two_mu, tt_mu, kcond, rgas;  // mu, 2.0*mu, (2.0/3.0)*mu, mu*cp/Pr, R
  // hyperviscosity
  T hv_coef[4];
  T dx2, dy2;
  int guard;
  // sources
  int drag_on, cool_on, force_on;
  T neg_drag;                         // -drag_coef
  double cool_pref, cool_tc, gamma_d; // cooling target pressure, time; gamma as double
  int ndrag, nforce;
  int kmx, kmy;                       // twiddle half-ranges (|k index| <= kmx, kmy)
  GPtr<const double> twx;             // [(2*kmx+1)][nx][2]  cos, sin of 2*pi*k*i/nx
  GPtr<const double> twy;             // [(2*kmy+1)][ny][2]
  GPtr<const int> drag_modes;         // [ndrag][2]
  GPtr<const int> force_modes;        // [nforce][2]
  GPtr<const int> force_bykx;         // [nforce] mode indices grouped by |kx|, then [fkx+2] group offsets
  // separable mode sums (column walks and per-row tables)
  int rb, rbp;                        // rows per wave: stage reductions, forcing projection
  int dkx, dky;                       // max |kx|, |ky| of the drag modes (<= 2)
  int fkx, fky;                       // max |kx|, |ky| of the forcing modes (<= 12)
};

// Constants of the line solve along one axis (see solve_lines()).
template <class T>
struct LineK {
  int n, L, nch, M;
  int small;                              // 4 n / L <= 512 with L <= 4: one chunk per thread
  T rho, c1p, inv1mrn, rnm1, one_m_rho2;  // rho, 1+rho^2, 1/(1-rho^n), rho^(n-1), 1-rho^2
  T alpha_c;                              // compact alpha 3/8
  T pw[66];                               // rho^k, k = 0..65 (0 beyond)
};

// Stage scalars written by the finalize kernel (double slots).
enum {
  S_EPS_R = 0, S_EPS_P, S_EPS_E, S_DIV_THR, S_COMP_ON, S_SHEAR_THR, S_SHEAR_ON,
  S_ALPHA_X, S_ALPHA_Y, S_COOL, S_DRAG = 16,  // drag: [ndrag][2 comps][re, im] / N
  S_COUNT = 16 + 4 * 24
};

// Stage reduction slots (block partials).
enum { R_MAXR = 0, R_MAXP, R_MAXE, R_AX, R_AY, R_DIV2, R_VORT2, R_PSUM, R_DRAG };
constexpr int kDragCols = 4 * 12;  // drag projection columns of a partial row (after R_DRAG)

// ---------------------------------------------------------------- launch mapping
// Workgroups of a 1-D grid are dispatched round-robin over the 8 XCDs (each
// with its own L2).  Tile kernels remap so that XCD j processes one contiguous
// range of tile indices: neighbouring tiles then share halo and half-used
// cache lines in one L2 instead of fetching them again from HBM.  A bijection
// of [0, total).
constexpr int kXCDs = 8;
__device__ __forceinline__ int xcd_remap(int b, int total) {
  const int q = total / kXCDs, r = total - q * kXCDs;
  const int j = b % kXCDs, i = b / kXCDs;
  return j * q + (j < r ? j : r) + i;
}

// ---------------------------------------------------------------- algebra
// Division in the pointwise algebra: IEEE for double (the FP64 mode
// reproduces the reference's quotients bit for bit); for float, the hardware
// reciprocal times the numerator (<= 2 ulp; the FP32 mode is held to its
// error budget, not to bit parity).
template <class T>
__device__ __forceinline__ T dv(T a, T b) {
  return a / b;
}
template <>
__device__ __forceinline__ float dv<float>(float a, float b) {
  return a * __builtin_amdgcn_rcpf(b);
}

// A divisor used many times: dv(a, d) with the FP32 reciprocal formed once
// (the same bits as dv, which forms it per call).
template <class T>
struct Divisor {
  T d, r;
};
template <class T>
__device__ __forceinline__ Divisor<T> divisor(T d) {
  return {d, d};
}
template <>
__device__ __forceinline__ Divisor<float> divisor<float>(float d) {
  return {d, __builtin_amdgcn_rcpf(d)};
}
template <class T>
__device__ __forceinline__ T dv(T a, const Divisor<T>& d) {
  return a / d.d;
}
template <>
__device__ __forceinline__ float dv<float>(float a, const Divisor<float>& d) {
  return a * d.r;
}

// Square root in the pointwise algebra: IEEE for double; for float the
// hardware square root (<= 1 ulp, no correction sequence), as the FP32 mode
// is held to its error budget.
__device__ __forceinline__ double sqrt_alg(double x) { return sqrt(x); }
__device__ __forceinline__ float sqrt_alg(float x) { return __builtin_amdgcn_sqrtf(x); }

// a*b + c: fused for float (one rounding; the FP32 mode is held to its error
// budget), two IEEE operations for double (the library is built without
// contraction, so the FP64 mode keeps the reference's rounding).  Call sites
// only reorder the operands of additions, which IEEE addition leaves exact.
__device__ __forceinline__ float madd(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double madd(double a, double b, double c) { return a * b + c; }

// np.maximum: NaN-propagating maximum (a if a is NaN, else b if b is NaN,
// else the larger).  Written as selects of precomputed values so that it
// compiles to compares and v_cndmask, not to branches that stop the
// scheduler from moving loads across it.
template <class T>
__device__ __forceinline__ T npmax(T a, T b) {
  const T m = a > b ? a : b;
  const T nb = (b != b) ? b : m;
  return (a != a) ? a : nb;
}

template <class T>
struct Prim {
  T rho, u, v, p;
};

// OOP/equations.py:136-145 primitive_from_conservative
template <class T>
__device__ __forceinline__ Prim<T> prim(const Params<T>& P, T q0, T q1, T q2, T q3) {
  Prim<T> r;
  r.rho = npmax(q0, P.rho_floor);
  r.u = dv(q1, r.rho);
  r.v = dv(q2, r.rho);
  T ke = (T(0.5) * r.rho) * madd(r.u, r.u, r.v * r.v);
  T p = P.gm1 * (q3 - ke);
  r.p = npmax(p, P.p_floor);
  return r;
}

// OOP/equations.py:165-181 enforce_physical_state (in registers)
template <class T>
__device__ __forceinline__ void enforce(const Params<T>& P, T& q0, T& q1, T& q2, T& q3) {
  T rho = npmax(q0, P.rho_floor);
  T u = dv(q1, rho);
  T v = dv(q2, rho);
  T ke = (T(0.5) * rho) * madd(u, u, v * v);
  T p = P.gm1 * (q3 - ke);
  q0 = rho;
  q1 = rho * u;
  q2 = rho * v;
  q3 = (p > P.p_floor) ? q3 : (P.pf_gm1 + ke);
}

// Element i of a plane as a base pointer plus a 32-bit byte offset: global
// memory instructions then take the (uniform) base in SGPRs and the offset
// in one VGPR, with no 64-bit address arithmetic per lane.  Every plane here
// has fewer than 2^29 elements (nx*ny*4 <= INT32_MAX), so i * sizeof(T) <
// 2^32.
template <class T>
__device__ __forceinline__ const T& at32(const T* p, uint32_t i) {
  return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(p) + (size_t)(i * (uint32_t)sizeof(T)));
}
template <class T>
__device__ __forceinline__ T& at32(T* p, uint32_t i) {
  return *reinterpret_cast<T*>(reinterpret_cast<char*>(p) + (size_t)(i * (uint32_t)sizeof(T)));
}

template <class T>
__device__ __forceinline__ void load4(const T* q, int N, int idx, T& a, T& b, T& c, T& d) {
  a = at32(q, idx);
  b = at32(q + (size_t)N, idx);
  c = at32(q + 2 * (size_t)N, idx);
  d = at32(q + 3 * (size_t)N, idx);
}

// Primitives of q (enforce_first: of enforce(q)).
template <class T>
__device__ __forceinline__ Prim<T> prim_at(const Params<T>& P, const T* q, int N, int idx,
                                           int enforce_first) {
  T a, b, c, d;
  load4(q, N, idx, a, b, c, d);
  if (enforce_first) enforce(P, a, b, c, d);
  return prim(P, a, b, c, d);
}

// ---------------------------------------------------------------- WENO7
// OOP/spatial_operator.py:87-124 (IS0 carries the reference's 544).
template <class T>
__device__ __forceinline__ T weno7(T v1, T v2, T v3, T v4, T v5, T v6, T v7) {
  const T eps = T(1e-10);
  // (each madd(a, b, c) is a*b + c; a difference x - k*v is madd(-k, v, x))
  T q0 = madd(T(25.0 / 12.0), v4, madd(T(-23.0 / 12.0), v3, madd(T(13.0 / 12.0), v2, T(-(1.0 / 4.0)) * v1)));
  T q1 = madd(T(1.0 / 4.0), v5, madd(T(13.0 / 12.0), v4, madd(T(-5.0 / 12.0), v3, T(1.0 / 12.0) * v2)));
  T q2 = madd(T(-1.0 / 12.0), v6, madd(T(7.0 / 12.0), v5, madd(T(7.0 / 12.0), v4, T(-(1.0 / 12.0)) * v3)));
  T q3 = madd(T(1.0 / 12.0), v7, madd(T(-5.0 / 12.0), v6, madd(T(13.0 / 12.0), v5, T(1.0 / 4.0) * v4)));
  // sum of a row (v_a * (c1 v_a + c2 v_b + ...)) in the reference's order
  T b0 = madd(T(2107), v4 * v4,
              madd(v3, madd(T(-9402), v4, T(11003) * v3),
                   madd(v2, madd(T(7042), v4, madd(T(-17246), v3, T(7043) * v2)),
                        v1 * madd(T(-1854), v4, madd(T(4642), v3, madd(T(-3882), v2, T(544) * v1))))));
  T b1 = madd(T(547), v5 * v5,
              madd(v4, madd(T(-2522), v5, T(3443) * v4),
                   madd(v3, madd(T(1922), v5, madd(T(-5966), v4, T(2843) * v3)),
                        v2 * madd(T(-494), v5, madd(T(1602), v4, madd(T(-1642), v3, T(267) * v2))))));
  T b2 = madd(T(267), v6 * v6,
              madd(v5, madd(T(-1642), v6, T(2843) * v5),
                   madd(v4, madd(T(1602), v6, madd(T(-5966), v5, T(3443) * v4)),
                        v3 * madd(T(-494), v6, madd(T(1922), v5, madd(T(-2522), v4, T(547) * v3))))));
  T b3 = madd(T(547), v7 * v7,
              madd(v6, madd(T(-3882), v7, T(7043) * v6),
                   madd(v5, madd(T(4642), v7, madd(T(-17246), v6, T(11003) * v5)),
                        v4 * madd(T(-1854), v7, madd(T(7042), v6, madd(T(-9402), v5, T(2107) * v4))))));
  T e0 = eps + b0, e1 = eps + b1, e2 = eps + b2, e3 = eps + b3;
  if constexpr (sizeof(T) == 8) {
    // one division: the weights d_k / E_k (E_k = e_k^2) times E0 E1 E2 E3 are
    // d_k * prod_{j != k} E_j, and the normalised sum is unchanged (within a
    // few ulp of the reference's quotients; E_k >= 1e-20 and products of
    // three stay in range for any flux below ~1e20)
    const T E0 = e0 * e0, E1 = e1 * e1, E2 = e2 * e2, E3 = e3 * e3;
    const T E01 = E0 * E1, E23 = E2 * E3;
    const T a0 = T(1.0 / 35.0) * (E1 * E23), a1 = T(12.0 / 35.0) * (E0 * E23);
    const T a2 = T(18.0 / 35.0) * (E3 * E01), a3 = T(4.0 / 35.0) * (E2 * E01);
    const T asum = ((a0 + a1) + a2) + a3;
    return madd(a3, q3, madd(a2, q2, madd(a1, q1, a0 * q0))) / asum;
  } else {
    T a0 = dv(T(1.0 / 35.0), e0 * e0);
    T a1 = dv(T(12.0 / 35.0), e1 * e1);
    T a2 = dv(T(18.0 / 35.0), e2 * e2);
    T a3 = dv(T(4.0 / 35.0), e3 * e3);
    T asum = ((a0 + a1) + a2) + a3;
    return dv(madd(a3, q3, madd(a2, q2, madd(a1, q1, a0 * q0))), asum);
  }
}

// ---------------------------------------------------------------- reductions
// DPP move of a double within rows of 16 lanes (GFX9 control CTRL, e.g.
// row_shr:d = 0x110 + d); lanes without a source receive 0.
template <int CTRL>
__device__ __forceinline__ double dpp_row_d(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffll), CTRL, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xf, 0xf, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ double readlane_d(double v, int l) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), l);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
// Sum (MAX false) or maximum (MAX true, of values >= 0: the lanes a row shift
// leaves without a source read 0) over the 64 lanes of a wavefront, every
// lane active; wave-uniform result in a fixed order (the same bits run to
// run): inclusive scans within the rows of 16 lanes by DPP row shifts (one
// VALU move each, no LDS round trip), then the four row totals in row order.
template <bool MAX>
__device__ __forceinline__ double wave_reduce(double v) {
  auto op = [](double a, double b) { return MAX ? fmax(a, b) : a + b; };
  v = op(v, dpp_row_d<0x111>(v));
  v = op(v, dpp_row_d<0x112>(v));
  v = op(v, dpp_row_d<0x114>(v));
  v = op(v, dpp_row_d<0x118>(v));
  return op(op(op(readlane_d(v, 15), readlane_d(v, 31)), readlane_d(v, 47)), readlane_d(v, 63));
}

// Block-wide reduction of NR doubles (every thread of a kBlock block calls
// it); slot k is a max (of values >= 0) when bit k of MAXMASK is set, else a
// sum.  Thread 0..NR-1 write the block's partials.
template <int NR, uint64_t MAXMASK>
__device__ __forceinline__ void block_reduce_write(double (&v)[NR], double* out) {
  __shared__ double sh[kWaves][NR];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < NR; ++k) {
    const double a = ((MAXMASK >> k) & 1) ? wave_reduce<true>(v[k]) : wave_reduce<false>(v[k]);
    if (lane == 0) sh[w][k] = a;
  }
  __syncthreads();
  for (int k = threadIdx.x; k < NR; k += blockDim.x) {
    double a = sh[0][k];
    for (int j = 1; j < kWaves; ++j) a = ((MAXMASK >> k) & 1) ? fmax(a, sh[j][k]) : a + sh[j][k];
    out[k] = a;
  }
}

// Reduce columns col0 .. col0+NC-1 of nblk partial rows (row stride
// `stride`) with one block; result in out[0..NC).  Each column is reduced on
// its own in a fixed order, so any subset of a row's columns reduces to the
// same bits as the whole row, run to run and block to block.
template <int NC, uint64_t MAXMASK>
__device__ __forceinline__ void reduce_partial_cols(const double* part, int nblk, int stride, int col0, double* out) {
  double v[NC];
#pragma unroll
  for (int k = 0; k < NC; ++k) v[k] = 0.0;
  for (int b = threadIdx.x; b < nblk; b += blockDim.x) {
#pragma unroll
    for (int k = 0; k < NC; ++k) {
      double a = part[(size_t)b * stride + col0 + k];
      v[k] = ((MAXMASK >> k) & 1) ? fmax(v[k], a) : v[k] + a;
    }
  }
  block_reduce_write<NC, MAXMASK>(v, out);
}
// All NR values of the rows.
template <int NR, uint64_t MAXMASK>
__device__ __forceinline__ void reduce_partials(const double* part, int nblk, double* out) {
  reduce_partial_cols<NR, MAXMASK>(part, nblk, NR, 0, out);
}

}  // namespace hitx
