// Line sweep with one wavefront per grid line (lines of n = 64*L nodes,
// L = 1..8): the x- or y-derivative of the hybrid compact/WENO7 flux
// (OOP/spatial_operator.py:321-342 periodic, OOP/problems/riemann_config3.py:329-360
// non-periodic) without block barriers.
//
// Lane l holds nodes l*L .. l*L+L-1 of its line in registers.  Per
// component, the conservative variable and the flux are staged in the wave's
// LDS slice so interface stencils read neighbours directly; the WENO7 flux of
// the interfaces around flagged ones is computed from a wave-compacted work
// list (ballot + prefix count), so a shock cluster is shared by all 64 lanes.
// The compact solve A w = raw, A = (1+rho^2)^-1 (1 - rho S^-1)(1 - rho S),
// runs the causal and anti-causal first-order recurrences locally over each
// lane's L nodes and joins the lanes with a 6-step affine scan
// (C_{l+1} = e_l + rho^L C_l) across the wavefront; the periodic wrap is
// closed in closed form, C_0 = (e_63 + rho^L P_63) / (1 - rho^n).
// Non-periodic lines use identity end rows, with the end contributions moved
// to the right-hand side and the homogeneous corrections a*rho^i +
// b*rho^(n-1-i) fixed by the two end values, as the block form does.
#pragma once

#include "hitx_kernels.h"

namespace hitx {

constexpr int kLineWaveMaxL = 8;

// LDS bytes of one wave's slice for a line of n nodes: q, f and the WENO
// fluxes with kLineHalo halo entries on each side (wrapped or clamped copies,
// so stencils index without wrap or clamp), the work list, the mask (halo 2).
// q and f are stored padded: node i at i + pad*((i + kLineHalo) / L), so the
// lanes' chunks start an odd number of words apart (no LDS bank conflicts
// for 4-byte elements when all lanes read the same chunk offset).
constexpr int kLineHalo = 4;
__host__ __device__ constexpr int line_pad(int L) { return L == 1 ? 0 : (L % 2 == 0 ? 1 : 2); }
__host__ __device__ inline int line_row(int n) {  // padded row length of q and f
  const int L = n / 64;
  return n + 1 + 2 * kLineHalo + line_pad(L) * ((n + 1 + 2 * kLineHalo) / (L > 0 ? L : 1) + 1);
}
__host__ __device__ inline size_t line_wave_lds(int n, int esize) {
  return (size_t)(2 * line_row(n) + (n + 1 + 2 * kLineHalo)) * esize + (size_t)(n + 1) * 2 + (size_t)(n + 4) + 16;
}

// Column slab (SLAB): a block of waves sweeps adjacent y-lines (grid
// columns) of the row-major state.  The block first copies the columns from
// HBM in row segments (coalesced) into an LDS slab, component c of column w
// at slab[(c * waves + w) * cstride + slab_pos(i)]; each wave then reads its
// column from the slab instead of HBM.  The wave slices overlay the slab, so
// the waves meet at one block barrier after reading it (every wave of the
// block, also one without a column, must reach that barrier once).
template <int L>
__host__ __device__ constexpr int slab_pos(int i) {
  return i + line_pad(L) * (i / (L > 0 ? L : 1));
}
__host__ __device__ inline int slab_cstride(int n) {  // odd, so the block's copy-in spreads over the banks
  const int L = n / 64;
  const int last = n - 1 + line_pad(L) * ((n - 1) / (L > 0 ? L : 1));
  return (last + 1) | 1;
}

template <class T, bool PER, int L, bool SLAB = false>
__device__ __forceinline__ void line_wave(const Params<T>& P, const LineK<T>& K, int axis, int enforced_input,
                                          const T* __restrict__ ql, const uint8_t* __restrict__ ml,
                                          const double* __restrict__ scal, T* __restrict__ out, int line, int nl,
                                          unsigned char* slice, int slab_cs = 0) {
  const int n = 64 * L, S = n + 1 + 2 * kLineHalo, R = line_row(n);
  constexpr int PAD = line_pad(L);
  const int lane = threadIdx.x & 63;
  // halo-padded rows: node i of the line at [i + kLineHalo] (+ bank padding for q and f)
  T* const qb = reinterpret_cast<T*>(slice) + kLineHalo;
  T* const fb = qb + R;
  T* ws = fb + R;
  uint16_t* wl = reinterpret_cast<uint16_t*>(ws - kLineHalo + S);
  auto A = [&](int i) { return i + PAD * ((i + kLineHalo) / L); };
  uint8_t* Mk = reinterpret_cast<uint8_t*>(wl + n + 1) + 2;  // mask, halo 2
  const size_t NL = (size_t)nl * n;
  const T alpha = (T)scal[axis == 0 ? S_ALPHA_X : S_ALPHA_Y];
  const T h = axis == 0 ? P.dx : P.dy;
  const T ih = T(1) / h;  // divided differences times the reciprocal (within 2 ulp of the quotient)
  const T rho = K.rho;
  // 1. (enforce), primitives, point flux of the lane's L nodes
  T q[4][L], f[4][L];
  const size_t g0 = (size_t)line * n + lane * L;
#pragma unroll
  for (int j = 0; j < L; ++j) {
    T a, b, c, d;
    if (SLAB) {  // the wave's column in the slab (ql), components slab_cs apart
      const int o = slab_pos<L>(lane * L + j);
      a = ql[o];
      b = ql[slab_cs + o];
      c = ql[2 * slab_cs + o];
      d = ql[3 * slab_cs + o];
    } else {
      a = ql[g0 + j];
      b = ql[NL + g0 + j];
      c = ql[2 * NL + g0 + j];
      d = ql[3 * NL + g0 + j];
    }
    if (!enforced_input) enforce(P, a, b, c, d);
    const Prim<T> pr = prim(P, a, b, c, d);
    q[0][j] = a;
    q[1][j] = b;
    q[2][j] = c;
    q[3][j] = d;
    if (axis == 0) {  // OOP/equations.py:183-190
      f[0][j] = b;
      f[1][j] = pr.rho * (pr.u * pr.u) + pr.p;
      f[2][j] = (pr.rho * pr.u) * pr.v;
      f[3][j] = (d + pr.p) * pr.u;
    } else {  // OOP/equations.py:192-199
      f[0][j] = c;
      f[1][j] = (pr.rho * pr.u) * pr.v;
      f[2][j] = pr.rho * (pr.v * pr.v) + pr.p;
      f[3][j] = (d + pr.p) * pr.v;
    }
    if (!SLAB) Mk[lane * L + j] = ml[g0 + j];
  }
  if (SLAB) {
    __syncthreads();  // every wave has read the slab: the slices may be written
#pragma unroll
    for (int j = 0; j < L; ++j) Mk[lane * L + j] = ml[g0 + j];
  }
  wave_sync();
  if (lane < 2) {  // mask halo: wrapped (periodic) or clamped copies
    Mk[-1 - lane] = PER ? Mk[n - 1 - lane] : Mk[0];
    Mk[n + lane] = PER ? Mk[lane] : Mk[n - 1];
  }
  wave_sync();
  const int nI = PER ? n : n + 1;
  // interface flag: periodic k = x_{k+1/2} (nodes k, k+1); non-periodic
  // k = x_{k-1/2} (nodes k-1, k); k in [-1, nI] reads the mask halo
  auto iflag = [&](int k) -> bool { return PER ? (Mk[k] | Mk[k + 1]) != 0 : (Mk[k - 1] | Mk[k]) != 0; };
  // 2. work list: interfaces whose WENO7 flux a flagged interface's
  //    compactification reads (i-1, i, i+1 of every flagged i)
  int nw = 0;
  for (int k0 = 0; k0 < nI; k0 += 64) {
    const int k = k0 + lane;
    bool need = false;
    if (k < nI) {
      // non-periodic ends: the neighbours of interface 0 and n are clamped
      const int km = PER ? k - 1 : (k > 0 ? k - 1 : 0), kp = PER ? k + 1 : (k < n ? k + 1 : n);
      need = P.weno_only || iflag(km) || iflag(k) || iflag(kp);
    }
    const unsigned long long bal = __ballot(need);
    if (need) wl[nw + __popcll(bal & ((1ull << lane) - 1ull))] = (uint16_t)k;
    nw += __popcll(bal);
  }
  // flags of the lane's interfaces kb .. kb+L (weno-only: all)
  const int kb = PER ? lane * L - 1 : lane * L;
  const int il0 = PER ? kb : kb - 1;  // first node of interface kb's smooth stencil, + 2
  const bool wo = P.weno_only != 0;
  uint32_t fl = 0u;
#pragma unroll
  for (int j = 0; j <= L; ++j) fl |= (uint32_t)(wo | iflag(kb + j)) << j;
  // 3. per component: interface fluxes, divided difference, compact solve, store
  const T c3 = T(-1.0 / 480.0), c2 = T(1.0 / 20.0 + (-1.0 / 480.0)),
          c1 = T((25.0 / 32.0 + 1.0 / 20.0) + (-1.0 / 480.0));
  const T aL = ipow(rho, L);
  // per-lane powers of the scans (fixed for the line), as products of the
  // uniform squares aL^(2^k) selected by the exponent's bits
  T apw[6];
  apw[0] = aL;
#pragma unroll
  for (int k = 1; k < 6; ++k) apw[k] = apw[k - 1] * apw[k - 1];
  auto lpow = [&](int e) {  // aL^e, 0 <= e < 64
    T r = T(1);
#pragma unroll
    for (int b = 0; b < 6; ++b) r = ((e >> b) & 1) ? r * apw[b] : r;
    return r;
  };
  const T a_f = lpow((lane & 15) + 1), a_b = lpow(16 - (lane & 15));
  const T a_l = PER ? lpow(lane) : T(0), a_r = lpow(63 - lane);
#pragma unroll 1
  for (int c = 0; c < 4; ++c) {
    wave_sync();  // previous component's LDS reads are done
    // component c of the lane's state and point flux into the LDS rows:
    // one branch on the uniform c, each arm storing its own registers (the
    // empty asm with a distinct text per arm keeps the compiler from merging
    // the arms into per-value selects)
    auto stage = [&](const T(&qq)[L], const T(&ff)[L]) {
#pragma unroll
      for (int j = 0; j < L; ++j) {
        qb[A(lane * L + j)] = qq[j];
        fb[A(lane * L + j)] = ff[j];
      }
    };
    if (c == 0) {
      stage(q[0], f[0]);
      asm volatile("; stage component 0" ::: "memory");
    } else if (c == 1) {
      stage(q[1], f[1]);
      asm volatile("; stage component 1" ::: "memory");
    } else if (c == 2) {
      stage(q[2], f[2]);
      asm volatile("; stage component 2" ::: "memory");
    } else {
      stage(q[3], f[3]);
      asm volatile("; stage component 3" ::: "memory");
    }
    wave_sync();
    if (lane < kLineHalo) {  // halos: wrapped or clamped copies of the line ends
      const int e = lane + 1;
      qb[A(-e)] = PER ? qb[A(n - e)] : qb[A(0)];
      fb[A(-e)] = PER ? fb[A(n - e)] : fb[A(0)];
      qb[A(n - 1 + e)] = PER ? qb[A(e - 1)] : qb[A(n - 1)];
      fb[A(n - 1 + e)] = PER ? fb[A(e - 1)] : fb[A(n - 1)];
    }
    wave_sync();
    // WENO7 of the Lax-Friedrichs split on the listed interfaces
    for (int t = lane; t < nw; t += 64) {
      const int k = wl[t];
      const int il = PER ? k : k - 1;
      T fp[7], fm[7];
#pragma unroll
      for (int j = 0; j < 7; ++j) {
        const int ip = il - 3 + j;  // f+ stencil i-3 .. i+3
        const int im = il + 4 - j;  // f- stencil i+4 .. i-2
        fp[j] = T(0.5) * madd(alpha, qb[A(ip)], fb[A(ip)]);
        fm[j] = T(0.5) * madd(-alpha, qb[A(im)], fb[A(im)]);
      }
      ws[k] = weno7(fp[0], fp[1], fp[2], fp[3], fp[4], fp[5], fp[6]) +
              weno7(fm[0], fm[1], fm[2], fm[3], fm[4], fm[5], fm[6]);
    }
    wave_sync();
    if (lane < 2) {  // WENO flux halo: wrapped (periodic, 2 each side) or clamped (interfaces 0..n)
      const int e = lane + 1;
      if (PER) {
        ws[-e] = ws[n - e];
        ws[n - 1 + e] = ws[e - 1];
      } else if (lane == 0) {
        ws[-1] = ws[0];
        ws[n + 1] = ws[n];
      }
    }
    wave_sync();
    // hybrid interface flux of interfaces kb .. kb+L (one beyond the lane's
    // own; periodic lane 0 starts at interface -1, the halo copy of n-1):
    // the 8th-order smooth flux of all L+1 from one window of L+6 point
    // fluxes (interface kb+j reads window entries j .. j+5), then the
    // compactified WENO flux where flagged
    T Fi[L + 1];
    {
      T Fw[L + 6];
#pragma unroll
      for (int m = 0; m < L + 6; ++m) Fw[m] = fb[A(il0 - 2 + m)];
#pragma unroll
      for (int j = 0; j <= L; ++j)
        Fi[j] = madd(c1, Fw[j + 3] + Fw[j + 2], madd(c3, Fw[j + 5] + Fw[j], c2 * (Fw[j + 4] + Fw[j + 1])));
    }
    if (fl != 0u) {  // window of WENO fluxes kb-1 .. kb+L+1 (entries of unlisted interfaces are not used)
      T Ww[L + 3];
#pragma unroll
      for (int m = 0; m < L + 3; ++m) Ww[m] = ws[kb - 1 + m];
#pragma unroll
      for (int j = 0; j <= L; ++j) {
        const T Fc = madd(K.alpha_c, Ww[j], madd(K.alpha_c, Ww[j + 2], Ww[j + 1]));
        Fi[j] = ((fl >> j) & 1u) ? Fc : Fi[j];
      }
    }
    T x[L];
#pragma unroll
    for (int j = 0; j < L; ++j)
      x[j] = (Fi[j + 1] - Fi[j]) * ih;  // periodic (F_i - F_{i-1})/dx; non-periodic (F_{j+1} - F_j)/dx
    // non-periodic: identity end rows, end contributions on the right-hand side
    T r0 = T(0), rn1 = T(0);
    if (!PER) {
      r0 = rdlane(x[0], 0);
      rn1 = rdlane(x[L - 1], 63);
    }
    T y[L];
    // causal pass, local
    T acc = T(0);
#pragma unroll
    for (int j = 0; j < L; ++j) {
      T xv = x[j];
      if (!PER) {
        const int i = lane * L + j;
        if (i == 0 || i == n - 1) {
          xv = T(0);
        } else {
          if (i == 1) xv = madd(-K.alpha_c, r0, xv);
          if (i == n - 2) xv = madd(-K.alpha_c, rn1, xv);
        }
      }
      acc = madd(rho, acc, xv);
      y[j] = acc;
    }
    // carries: inclusive scan of C_{l+1} = e_l + a C_l, a = rho^L
    T sc = scan_fwd(acc, apw, a_f, lane);
    T carry = dpp<0x138>(sc);  // exclusive: P_l (lane 0: 0)
    if (PER) {
      // C_0 = (e_63 + a P_63) / (1 - rho^n) = S_63 / (1 - rho^n); C_l = P_l + a^l C_0
      const T c0 = rdlane(sc, 63) * K.inv1mrn;
      carry = madd(a_l, c0, carry);
    }
    {
      T pw = rho;
#pragma unroll
      for (int j = 0; j < L; ++j) {
        y[j] = madd(pw, carry, y[j]);
        pw = pw * rho;
      }
    }
    T zn = T(0);
    if (!PER) zn = rdlane((rho * y[L - 1]) / K.one_m_rho2, 63);
    // anti-causal pass, local
    acc = T(0);
#pragma unroll
    for (int j = L - 1; j >= 0; --j) {
      acc = madd(rho, acc, y[j]);
      y[j] = acc;
    }
    // carries from the right: D_{l-1} = g_l + a D_l
    sc = scan_bwd(acc, apw, a_b, lane);
    carry = dpp<0x130>(sc);  // exclusive: Q_l (lane 63: 0)
    // wrap (periodic) or the value beyond the end (non-periodic): a^(63-l) times
    if (PER)
      carry = madd(a_r, rdlane(sc, 0) * K.inv1mrn, carry);
    else
      carry = madd(a_r, zn, carry);
    {
      T pw = rho;
#pragma unroll
      for (int j = L - 1; j >= 0; --j) {
        y[j] = K.c1p * madd(pw, carry, y[j]);
        pw = pw * rho;
      }
    }
    T* o = out + (size_t)c * NL + g0;
    if (PER) {
#pragma unroll
      for (int j = 0; j < L; ++j) o[j] = y[j];
    } else {
      const T p0 = rdlane(y[0], 0), pn = rdlane(y[L - 1], 63);
      const T det = T(1) - K.rnm1 * K.rnm1;
      const T idet = T(1) / det, a = (-p0 + pn * K.rnm1) * idet, bb = (-pn + p0 * K.rnm1) * idet;
      T pl = ipow(rho, lane * L), pr = ipow(rho, n - 1 - lane * L);  // rho^i, rho^(n-1-i)
      const T irho = T(1) / rho;
#pragma unroll
      for (int j = 0; j < L; ++j) {
        const int i = lane * L + j;
        T v;
        if (i == 0)
          v = r0;
        else if (i == n - 1)
          v = rn1;
        else
          v = (y[j] + a * pl) + bb * pr;
        o[j] = v;
        pl = pl * rho;
        pr = pr * irho;
      }
    }
  }
}

// Dispatch on L = n / 64 (the caller checks n % 64 == 0 and n <= 512).
template <class T, bool PER>
__device__ __forceinline__ void line_wave_any(const Params<T>& P, const LineK<T>& K, int axis, int enforced_input,
                                              const T* __restrict__ ql, const uint8_t* __restrict__ ml,
                                              const double* __restrict__ scal, T* __restrict__ out, int line,
                                              int nl, unsigned char* slice) {
  switch (K.n >> 6) {
    case 1: line_wave<T, PER, 1>(P, K, axis, enforced_input, ql, ml, scal, out, line, nl, slice); break;
    case 2: line_wave<T, PER, 2>(P, K, axis, enforced_input, ql, ml, scal, out, line, nl, slice); break;
    case 3: line_wave<T, PER, 3>(P, K, axis, enforced_input, ql, ml, scal, out, line, nl, slice); break;
    case 4: line_wave<T, PER, 4>(P, K, axis, enforced_input, ql, ml, scal, out, line, nl, slice); break;
    case 5: line_wave<T, PER, 5>(P, K, axis, enforced_input, ql, ml, scal, out, line, nl, slice); break;
    case 6: line_wave<T, PER, 6>(P, K, axis, enforced_input, ql, ml, scal, out, line, nl, slice); break;
    case 7: line_wave<T, PER, 7>(P, K, axis, enforced_input, ql, ml, scal, out, line, nl, slice); break;
    default: line_wave<T, PER, 8>(P, K, axis, enforced_input, ql, ml, scal, out, line, nl, slice); break;
  }
}

// Slab form of the y-sweep (one column per wave, kSlabWaves columns per
// block); the caller has filled the slab and passed its wave's column.
template <class T, bool PER>
__device__ __forceinline__ void line_wave_slab_any(const Params<T>& P, const LineK<T>& K, const T* __restrict__ col,
                                                   int cs, const uint8_t* __restrict__ ml,
                                                   const double* __restrict__ scal, T* __restrict__ out, int line,
                                                   int nl, unsigned char* slice) {
  switch (K.n >> 6) {
    case 1: line_wave<T, PER, 1, true>(P, K, 1, 0, col, ml, scal, out, line, nl, slice, cs); break;
    case 2: line_wave<T, PER, 2, true>(P, K, 1, 0, col, ml, scal, out, line, nl, slice, cs); break;
    case 3: line_wave<T, PER, 3, true>(P, K, 1, 0, col, ml, scal, out, line, nl, slice, cs); break;
    case 4: line_wave<T, PER, 4, true>(P, K, 1, 0, col, ml, scal, out, line, nl, slice, cs); break;
    case 5: line_wave<T, PER, 5, true>(P, K, 1, 0, col, ml, scal, out, line, nl, slice, cs); break;
    case 6: line_wave<T, PER, 6, true>(P, K, 1, 0, col, ml, scal, out, line, nl, slice, cs); break;
    case 7: line_wave<T, PER, 7, true>(P, K, 1, 0, col, ml, scal, out, line, nl, slice, cs); break;
    default: line_wave<T, PER, 8, true>(P, K, 1, 0, col, ml, scal, out, line, nl, slice, cs); break;
  }
}

constexpr int kSlabWaves = 8;  // columns (wavefronts) per block of the slab y-sweep

// Copy columns x0 .. x0 + kSlabWaves - 1 of the row-major state (all rows,
// four components) into the slab; columns beyond nx are left unwritten.
template <class T>
__device__ __forceinline__ void slab_fill(const T* __restrict__ q, int nx, int ny, int x0, T* __restrict__ slab,
                                          int cs) {
  const size_t N = (size_t)nx * ny;
  const int L = ny >> 6, pad = line_pad(L);
  const int total = ny * kSlabWaves;
  constexpr int U = 4;  // rows in flight per thread
  for (int base = (int)threadIdx.x; base < total; base += U * (int)blockDim.x) {
    T v[U][4];
    int pos[U], w[U];
    bool ok[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int idx = base + u * (int)blockDim.x;
      const int y = idx / kSlabWaves;
      w[u] = idx - y * kSlabWaves;
      ok[u] = idx < total && x0 + w[u] < nx;
      pos[u] = y + pad * (y / L);
      const size_t g = (size_t)y * nx + x0 + w[u];
#pragma unroll
      for (int c = 0; c < 4; ++c) v[u][c] = ok[u] ? q[c * N + g] : T(0);
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (ok[u]) {
#pragma unroll
        for (int c = 0; c < 4; ++c) slab[(c * kSlabWaves + w[u]) * cs + pos[u]] = v[u][c];
      }
  }
}

constexpr int kLineWaves = 2;  // lines (wavefronts) per block

// Periodic FP32 lines fit 4 wavefronts per SIMD (<= 128 VGPRs) without
// spilling; the others keep their registers (a forced bound would spill).
template <class T, bool PER>
struct LineWaveOcc {
  static constexpr int waves = 1;
};
template <>
struct LineWaveOcc<float, true> {
  static constexpr int waves = 4;
};

#define HITX_LINE_WAVE_ATTR(T, PER) __attribute__((amdgpu_waves_per_eu(LineWaveOcc<T, PER>::waves)))

template <class T, bool PER>
__global__ void __launch_bounds__(64 * kLineWaves) HITX_LINE_WAVE_ATTR(T, PER)
    k_line_wave(Params<T> P, LineK<T> K, int axis, int enforced_input, const T* __restrict__ ql,
                const uint8_t* __restrict__ ml, const double* __restrict__ scal, T* __restrict__ out, int nl,
                int slice_bytes) {
  extern __shared__ __align__(16) unsigned char wsm[];
  const int w = threadIdx.x >> 6;
  const int line = blockIdx.x * kLineWaves + w;
  if (line >= nl) return;
  // one body per axis (a constant axis folds the flux selection)
  if (axis == 0)
    line_wave_any<T, PER>(P, K, 0, enforced_input, ql, ml, scal, out, line, nl, wsm + (size_t)w * slice_bytes);
  else
    line_wave_any<T, PER>(P, K, 1, enforced_input, ql, ml, scal, out, line, nl, wsm + (size_t)w * slice_bytes);
}

// The block form handles lines that are not a multiple of 64 nodes or longer
// than 512.
inline bool line_wave_ok(int n) { return n % 64 == 0 && n >= 64 && n <= 64 * kLineWaveMaxL; }

}  // namespace hitx
