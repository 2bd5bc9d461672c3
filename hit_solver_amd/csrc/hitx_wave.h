// Wavefront-level helpers (wave64): wave-local LDS ordering, DPP lane
// moves, readlane, and the affine scans of the line solves.
#pragma once

#include "hitx_device.h"

namespace hitx {

// Orders a wavefront's LDS accesses (no block barrier).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// b^e for 0 <= e < 2^10 by squaring (uniform or per-lane e).
template <class T>
__device__ __forceinline__ T ipow(T b, int e) {
  T r = T(1);
#pragma unroll
  for (int bit = 0; bit < 10; ++bit) {
    if ((e >> bit) & 1) r = r * b;
    b = b * b;
  }
  return r;
}

// Cross-lane moves.  DPP (GFX9 controls): row_shr:d 0x110+d (lane l takes
// l-d within its row of 16), row_shl:d 0x100+d (l+d), wave_shr:1 0x138 (l-1),
// wave_shl:1 0x130 (l+1); lanes without a source receive 0.  readlane for a
// uniform lane's value.
template <int CTRL>
__device__ __forceinline__ float dpp(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xf, 0xf, false));
}
template <int CTRL>
__device__ __forceinline__ double dpp(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffll), CTRL, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xf, 0xf, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ float rdlane(float v, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ __forceinline__ double rdlane(double v, int l) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), l);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Affine scans over the wavefront with a fixed decay a (pow: a^1, a^2, a^4,
// a^8, a^16; pf = a^((l&15)+1), pb = a^(16-(l&15)) for this lane):
// forward  S_l = sum_{m=0..l}    a^m v_{l-m},
// backward T_l = sum_{m=0..63-l} a^m v_{l+m}.
// Kogge-Stone within rows of 16 by DPP row shifts, then the rows joined
// through their end values (readlane).
template <class T>
__device__ __forceinline__ T scan_fwd(T v, const T* pw, T pf, int lane) {
  v = madd(pw[0], dpp<0x111>(v), v);
  v = madd(pw[1], dpp<0x112>(v), v);
  v = madd(pw[2], dpp<0x114>(v), v);
  v = madd(pw[3], dpp<0x118>(v), v);
  const T s0 = rdlane(v, 15);
  const T s1 = madd(pw[4], s0, rdlane(v, 31));
  const T s2 = madd(pw[4], s1, rdlane(v, 47));
  const int r = lane >> 4;
  const T prev = r == 0 ? T(0) : (r == 1 ? s0 : (r == 2 ? s1 : s2));
  return madd(pf, prev, v);
}
template <class T>
__device__ __forceinline__ T scan_bwd(T v, const T* pw, T pb, int lane) {
  v = madd(pw[0], dpp<0x101>(v), v);
  v = madd(pw[1], dpp<0x102>(v), v);
  v = madd(pw[2], dpp<0x104>(v), v);
  v = madd(pw[3], dpp<0x108>(v), v);
  const T t3 = rdlane(v, 48);
  const T t2 = madd(pw[4], t3, rdlane(v, 32));
  const T t1 = madd(pw[4], t2, rdlane(v, 16));
  const int r = lane >> 4;
  const T next = r == 3 ? T(0) : (r == 2 ? t3 : (r == 1 ? t2 : t1));
  return madd(pb, next, v);
}

}  // namespace hitx
