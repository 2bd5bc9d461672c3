// Device reproduction of NumPy's Generator(PCG64).standard_normal stream, the
// forcing noise of the reference (OOP/forcing.py:66 default_rng(seed), :77
// standard_normal((ny, nx))).
//
// Generator: PCG64 (O'Neill 2014), 128-bit LCG state s <- s*M + inc, output
// XSL-RR of the new state; jump-ahead by the arbitrary-stride LCG recurrence
// (F. Brown, "Random number generation with arbitrary strides", 1994).
// Normals: 256-layer ziggurat (Marsaglia & Tsang 2000) in NumPy's form: one
// 64-bit word per attempt gives the layer (low 8 bits), the sign (bit 8) and
// a 52-bit magnitude; the wedge test takes one more word as a double, the
// tail two per try.
//
// The stream is sequential (a rejected attempt shifts every later normal), so
// the parallel form works on attempts, not normals:
//   1. every word position p of the stream is evaluated as if an attempt
//      started there: its length in words and whether it yields a value;
//   2. positions are cut in chunks of kRngChunk words; for each chunk and each
//      entry offset e < kRngEntries (how far the previous chunk's last attempt
//      spilled into it) the attempt chain is walked to the chunk end, giving
//      the spill into the next chunk and the number of values;
//   3. these small maps are composed by a block-wide scan, which fixes the
//      true entry offset and the index of the first value of every chunk;
//   4. each chunk regenerates its words from its first word's LCG state
//      (stored by step 1) and writes the values of the attempts on its chain
//      at their indices; the thread that writes value n-1 records how many
//      words the n values consumed.
// A spill of kRngEntries words or more, or too few values in the words
// evaluated, sets a flag, and one thread then regenerates the whole request
// sequentially (tens of milliseconds for 2^18 values).  Both forms give the
// same numbers.  A spill of s words needs a tail attempt (1 in 256) of at
// least s + 1 words ending past the chunk, i.e. (s - 1)/2 rejected tail tries
// (about 6% each): with 8 entry offsets that happened about once per 7e8
// normals, with 16 it is about 1e-5 times as rare.
#pragma once

#include "hitx_device.h"
#include "hitx_normal_tables.h"

namespace hitx {

typedef unsigned __int128 u128;

constexpr int kRngChunk = 32;    // words per chunk (one thread)
constexpr int kRngEntries = 16;  // entry offsets tracked per chunk
constexpr int kRngScan = 256;    // threads of the scan block (four wavefronts: a request holds few CU slots)
constexpr int kRngAttemptThreads = 256;  // threads of an attempt block (one chunk each)
constexpr uint64_t kMask52 = (1ull << 52) - 1;

__device__ __forceinline__ u128 pcg_mult() {
  return ((u128)0x2360ED051FC65DA4ull << 64) | (u128)0x4385DF649FCCF645ull;
}

struct Pcg {
  u128 s, inc;
  __device__ __forceinline__ uint64_t next() {
    s = s * pcg_mult() + inc;
    const uint64_t x = (uint64_t)(s >> 64) ^ (uint64_t)s;
    const unsigned rot = (unsigned)(s >> 122);
    return (x >> rot) | (x << ((64u - rot) & 63u));
  }
};

// s after `delta` steps of s <- s*M + inc.
__device__ __forceinline__ u128 pcg_advance(u128 s, u128 inc, unsigned long long delta) {
  u128 cm = pcg_mult(), cp = inc, am = 1, ap = 0;
  while (delta) {
    if (delta & 1ull) {
      am = am * cm;
      ap = ap * cm + cp;
    }
    cp = (cm + 1) * cp;
    cm = cm * cm;
    delta >>= 1;
  }
  return am * s + ap;
}

// Jumps of the LCG by 32*t steps, t < kRngAttemptThreads: s -> a s + inc m
// with a = M^(32t), m = sum_{i < 32t} M^i (mod 2^128).  Jumps compose as
// (a1, m1) then (a2, m2) = (a1 a2, a2 m1 + m2).
constexpr u128 kPcgM = ((u128)0x2360ED051FC65DA4ull << 64) | (u128)0x4385DF649FCCF645ull;
struct PcgJumps {
  u128 a[256], m[256];
};
constexpr PcgJumps make_pcg_jumps() {
  u128 a32 = 1, m32 = 0;
  for (int i = 0; i < 32; ++i) {
    m32 = m32 + a32;
    a32 = a32 * kPcgM;
  }
  PcgJumps t{};
  u128 a = 1, m = 0;
  for (int i = 0; i < 256; ++i) {
    t.a[i] = a;
    t.m[i] = m;
    m = a32 * m + m32;
    a = a * a32;
  }
  return t;
}
__device__ constexpr PcgJumps kPcgJumps = make_pcg_jumps();

__device__ __forceinline__ Pcg pcg_load(const uint64_t* st) {
  Pcg g;
  g.s = ((u128)st[0] << 64) | (u128)st[1];
  g.inc = ((u128)st[2] << 64) | (u128)st[3];
  return g;
}

__device__ __forceinline__ double pcg_double(uint64_t w) { return (double)(w >> 11) * (1.0 / 9007199254740992.0); }

struct ZigAttempt {
  int len;     // words consumed by the attempt
  bool out;    // the attempt yields a value
  double v;
};

// Attempt starting with word w; g is positioned after w and is advanced only
// on the local copy.  CAP bounds the tail loop so that the length fits the
// 7-bit code of the parallel form (a tail run that long is reported as a
// spill and the request is regenerated serially without the bound).
// Ki, Wi: the layer tables (the device tables, or copies in LDS).
template <bool CAP>
__device__ __forceinline__ ZigAttempt zig_attempt(uint64_t w, Pcg g, const uint64_t* Ki = kZigKi,
                                                  const double* Wi = kZigWi) {
  const int idx = (int)(w & 0xff);
  const uint64_t r = w >> 8;
  const uint64_t rabs = (r >> 1) & kMask52;
  double x = (double)rabs * Wi[idx];
  if (r & 1ull) x = -x;
  ZigAttempt a;
  a.v = x;
  if (rabs < Ki[idx]) {
    a.len = 1;
    a.out = true;
    return a;
  }
  if (idx == 0) {
    int len = 1;
    for (;;) {
      const double xx = -kZigInvR * log1p(-pcg_double(g.next()));
      const double yy = -log1p(-pcg_double(g.next()));
      len += 2;
      if (yy + yy > xx * xx) {
        a.v = ((rabs >> 8) & 1ull) ? -(kZigR + xx) : kZigR + xx;
        a.len = len;
        a.out = true;
        return a;
      }
      if (CAP && len > 125) {
        a.len = 127;
        a.out = false;
        return a;
      }
    }
  }
  const double u = pcg_double(g.next());
  a.len = 2;
  a.out = ((kZigFi[idx - 1] - kZigFi[idx]) * u + kZigFi[idx]) < exp(-0.5 * x * x);
  return a;
}

// The value of an attempt that yields one (word w, length len; g positioned
// after w): the layer's x for the first test and the wedge, the tail's value
// (replayed) for a base-layer attempt that went to the tail.
__device__ __forceinline__ double zig_value(uint64_t w, int len, const Pcg& g, const double* Wi = kZigWi) {
  const int idx = (int)(w & 0xff);
  if (idx == 0 && len > 1) return zig_attempt<true>(w, g).v;
  const uint64_t r = w >> 8;
  const double x = (double)((r >> 1) & kMask52) * Wi[idx];
  return (r & 1ull) ? -x : x;
}

// Buffers of one request of n normals (all in device memory).
struct RngJob {
  GPtr<uint64_t> state;  // s hi, s lo, inc hi, inc lo; advanced past the n values
  GPtr<double> out;      // n values
  GPtr<uint64_t> cst;    // [nch][2] LCG state (hi, lo) before each chunk's first word
  GPtr<uint8_t> att;     // [nch][kRngChunk] attempt length | yields << 7
  GPtr<uint16_t> map;    // [nch][kRngEntries] spill | count << 8 (spill 0xff: overflow)
  GPtr<uint8_t> entry;   // [nch] resolved entry offset
  GPtr<int> base;        // [nch] index of the chunk's first value
  GPtr<int> stat;        // [0] flag: 0 parallel result, 1 overflow / too few words, 2 forced serial;
                    // [1] words consumed by the n values
  long long n;      // values requested
  int W, nch;       // words evaluated, chunks
};

__host__ __device__ inline int rng_words(long long n) { return (int)(n + n / 16 + 64); }
__host__ __device__ inline int rng_chunks(long long n) { return (rng_words(n) + kRngChunk - 1) / kRngChunk; }

// Steps 1 and 2 for chunk c = c0 + threadIdx.x of a block whose first chunk
// is c0 (every thread of the block calls it).  One thread advances the state
// to the block's first word; each thread then jumps by 32 threadIdx.x words
// with the table.  lds: kRngChunk bytes per thread.
// lds: the thread's 3 * kRngChunk bytes (codes, chain spills, chain counts).
__device__ __forceinline__ void rng_attempt_body(const RngJob& J, int c0, uint8_t* lds) {
  __shared__ uint64_t sb[2];
  // the layer tables in LDS: every word is tested against them
  __shared__ uint64_t sKi[256];
  __shared__ double sWi[256];
  static_assert(kRngAttemptThreads == 256, "one table entry per thread");
  sKi[threadIdx.x] = kZigKi[threadIdx.x];
  sWi[threadIdx.x] = kZigWi[threadIdx.x];
  const int c = c0 + threadIdx.x;
  Pcg g = pcg_load(J.state);
  if (threadIdx.x == 0) {
    const u128 s0 = pcg_advance(g.s, g.inc, (unsigned long long)c0 * kRngChunk);
    sb[0] = (uint64_t)(s0 >> 64);
    sb[1] = (uint64_t)s0;
  }
  __syncthreads();
  if (c >= J.nch) return;
  g.s = kPcgJumps.a[threadIdx.x] * (((u128)sb[0] << 64) | (u128)sb[1]) + g.inc * kPcgJumps.m[threadIdx.x];
  J.cst[2 * (size_t)c] = (uint64_t)(g.s >> 64);
  J.cst[2 * (size_t)c + 1] = (uint64_t)g.s;
  for (int j = 0; j < kRngChunk; ++j) {
    const uint64_t w = g.next();
    const ZigAttempt a = zig_attempt<true>(w, g, sKi, sWi);
    lds[j] = (uint8_t)(a.len | (a.out ? 0x80 : 0));
  }
  {  // the chunk's codes, contiguous: two 16-byte stores
    uint4* dst = reinterpret_cast<uint4*>(static_cast<uint8_t*>(J.att) + (size_t)c * kRngChunk);
    const uint4* src = reinterpret_cast<const uint4*>(lds);
    dst[0] = src[0];
    dst[1] = src[1];
  }
  // Chain ends from every position by one backward pass (the chain from p
  // continues at p + len(p)): sp[p] = words past the chunk where it leaves,
  // ct[p] = values it yields on the way; the map's entries are those of
  // positions 0..7.  (Walking each entry's chain on its own costs about
  // eight times the LDS round trips.)
  uint8_t* sp = lds + kRngChunk;
  uint8_t* ct = lds + 2 * kRngChunk;
  for (int p = kRngChunk - 1; p >= 0; --p) {
    const uint8_t code = lds[p];
    const int nx = p + (code & 0x7f), y = code >> 7;
    if (nx >= kRngChunk) {
      sp[p] = (uint8_t)min(nx - kRngChunk, 0xff);
      ct[p] = (uint8_t)y;
    } else {
      sp[p] = sp[nx];
      ct[p] = (uint8_t)(ct[nx] + y);
    }
  }
  for (int e = 0; e < kRngEntries; ++e) {
    const int spill = sp[e];
    J.map[(size_t)c * kRngEntries + e] = (uint16_t)((spill < kRngEntries ? spill : 0xff) | (ct[e] << 8));
  }
}

// A chunk's map (16 uint16) held as two 16-byte loads; entry e of it.
static_assert(kRngEntries == 16, "a chunk's map is two 16-byte loads");
struct ChunkMap {
  uint4 a, b;
};
__device__ __forceinline__ ChunkMap load_map(const RngJob& J, int c) {
  const uint4* m = reinterpret_cast<const uint4*>(static_cast<const uint16_t*>(J.map)) + 2 * (size_t)c;
  return {m[0], m[1]};
}
// (selects between registers: a data-dependent index into the map would
// put the maps in scratch memory)
__device__ __forceinline__ uint32_t map_entry(const ChunkMap& m, int e) {
  const bool hi = e >= 8;
  const uint32_t x = hi ? m.b.x : m.a.x, y = hi ? m.b.y : m.a.y, z = hi ? m.b.z : m.a.z, w = hi ? m.b.w : m.a.w;
  const int h = (e >> 1) & 3;
  const uint32_t v = h == 0 ? x : (h == 1 ? y : (h == 2 ? z : w));
  return (e & 1) ? (v >> 16) : (v & 0xffffu);
}
constexpr int kRngScanBatch = 8;  // chunk maps loaded together (one latency per batch)

// Step 3 (one block of kRngScan threads per request).  A chunk's map sends
// the entry offset e to (spill, count); spill 0xff marks an overflow.  Each
// thread composes the maps of its run of chunks, the block scans the run maps
// (Hillis-Steele in LDS), and every run then knows its entry offset and the
// index of its first value.  force_serial makes the last step regenerate
// serially (a check of that path).

__device__ __forceinline__ void rng_scan_body(const RngJob& J, int force_serial) {
  // entry-major ([e][t]): lanes touch consecutive words, also where a lane
  // reads its own map at a data-dependent entry
  __shared__ uint8_t sx[kRngEntries * kRngScan];
  __shared__ int sn[kRngEntries * kRngScan];
  const int t = threadIdx.x;
  const int per = (J.nch + kRngScan - 1) / kRngScan;
  const int c0 = min(t * per, J.nch), c1 = min(c0 + per, J.nch);
  uint8_t mx[kRngEntries];
  int mn[kRngEntries];
#pragma unroll
  for (int e = 0; e < kRngEntries; ++e) {
    mx[e] = (uint8_t)e;
    mn[e] = 0;
  }
  for (int cb = c0; cb < c1; cb += kRngScanBatch) {
    ChunkMap mv[kRngScanBatch];
#pragma unroll
    for (int j = 0; j < kRngScanBatch; ++j)
      if (cb + j < c1) mv[j] = load_map(J, cb + j);
#pragma unroll
    for (int j = 0; j < kRngScanBatch; ++j) {
      if (cb + j < c1) {
#pragma unroll
        for (int e = 0; e < kRngEntries; ++e) {
          if (mx[e] < kRngEntries) {
            const uint32_t v = map_entry(mv[j], mx[e]);
            mx[e] = (uint8_t)(v & 0xff);
            mn[e] += (int)(v >> 8);
          }
        }
      }
    }
  }
#pragma unroll
  for (int e = 0; e < kRngEntries; ++e) {
    sx[e * kRngScan + t] = mx[e];
    sn[e * kRngScan + t] = mn[e];
  }
  __syncthreads();
  for (int d = 1; d < kRngScan; d <<= 1) {
    const bool has = t >= d;
    if (has) {
      // (left run then own run): own map read from LDS at the left's spill
#pragma unroll
      for (int e = 0; e < kRngEntries; ++e) {
        const int x1 = sx[e * kRngScan + (t - d)];
        if (x1 < kRngEntries) {
          mx[e] = sx[x1 * kRngScan + t];
          mn[e] = sn[e * kRngScan + (t - d)] + sn[x1 * kRngScan + t];
        } else {
          mx[e] = 0xff;
          mn[e] = 0;
        }
      }
    }
    __syncthreads();
    if (has) {
#pragma unroll
      for (int e = 0; e < kRngEntries; ++e) {
        sx[e * kRngScan + t] = mx[e];
        sn[e * kRngScan + t] = mn[e];
      }
    }
    __syncthreads();
  }
  // exclusive prefix: the stream enters chunk 0 at offset 0
  int e = 0, cnt = 0;
  if (t > 0) {
    e = sx[t - 1];
    cnt = sn[t - 1];
  }
  for (int cb = c0; cb < c1 && e < kRngEntries; cb += kRngScanBatch) {
    ChunkMap mv[kRngScanBatch];
#pragma unroll
    for (int j = 0; j < kRngScanBatch; ++j)
      if (cb + j < c1) mv[j] = load_map(J, cb + j);
#pragma unroll
    for (int j = 0; j < kRngScanBatch; ++j) {
      const int c = cb + j;
      if (c < c1 && e < kRngEntries) {
        J.entry[c] = (uint8_t)e;
        J.base[c] = cnt;
        const uint32_t v = map_entry(mv[j], e);
        e = (int)(v & 0xff);
        cnt += (int)(v >> 8);
      }
    }
  }
  if (t == kRngScan - 1) {
    const bool overflow = sx[t] >= kRngEntries;
    const bool short_ = sn[t] < J.n;
    J.stat[0] = (overflow || short_) ? 1 : (force_serial ? 2 : 0);
  }
}

// Step 4, one wavefront per 64 consecutive chunks: their values occupy one
// contiguous index range (at most 64*kRngChunk), staged in the wave's LDS
// slice (lds: kRngChunk*64 doubles) and written out coalesced.  The lane
// that writes value n-1 records how many words the n values consumed.
constexpr int kRngScatterThreads = 128;
// Stages the wave's values [lo, hi) in lds (value i at i - lo); false when
// the wave has nothing to do (serial path, chunks past the request).
__device__ __forceinline__ bool rng_scatter_stage(const RngJob& J, int c, double* lds, long long& lo_out,
                                                  long long& hi_out, const double* Wi = kZigWi) {
  if (J.stat[0] != 0) return false;
  const int lane = threadIdx.x & 63;
  const int c0 = c - lane;  // chunk of lane 0
  if (c0 >= J.nch) return false;
  const long long lo = J.base[c0];
  long long k = lo;
  if (c < J.nch) {
    // the chunk's 32 codes in two 16-byte loads, then the attempt chain from
    // the entry offset walked over them in registers (position q is an
    // attempt start when the chain reaches it)
    const uint4* ap = reinterpret_cast<const uint4*>(static_cast<const uint8_t*>(J.att) + (size_t)c * kRngChunk);
    const uint4 a0 = ap[0], a1 = ap[1];
    const uint32_t cw[8] = {a0.x, a0.y, a0.z, a0.w, a1.x, a1.y, a1.z, a1.w};
    int nxt = J.entry[c];
    k = J.base[c];
    // the chunk's words regenerated from its first word's state
    Pcg g;
    g.s = ((u128)J.cst[2 * (size_t)c] << 64) | (u128)J.cst[2 * (size_t)c + 1];
    g.inc = ((u128)J.state[2] << 64) | (u128)J.state[3];
#pragma unroll
    for (int q = 0; q < kRngChunk; ++q) {
      const uint64_t w = g.next();
      if (q == nxt) {
        const uint32_t code = (cw[q >> 2] >> ((q & 3) * 8)) & 0xffu;
        const int len = (int)(code & 0x7f);
        if (code & 0x80) {
          const double v = zig_value(w, len, g, Wi);
          if (k < J.n) lds[k - lo] = v;
          if (k == J.n - 1) J.stat[1] = c * kRngChunk + q + len;  // words consumed
          ++k;
        }
        nxt = q + len;
      }
    }
  }
  long long hi = k;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const long long o = __shfl_xor(hi, off, 64);
    hi = o > hi ? o : hi;
  }
  if (hi > J.n) hi = J.n;
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  lo_out = lo;
  hi_out = hi;
  return true;
}
__device__ __forceinline__ void rng_scatter_body(const RngJob& J, int c, double* lds) {
  long long lo, hi;
  if (!rng_scatter_stage(J, c, lds, lo, hi)) return;
  const int lane = threadIdx.x & 63;
  for (long long i = lo + lane; i < hi; i += 64) J.out[i] = lds[i - lo];
}

// Last step (one thread): advance the state, or regenerate serially.
// Returns the number of 64-bit words the request consumed.
__device__ __forceinline__ unsigned long long rng_finish_body(const RngJob& J) {
  Pcg g = pcg_load(J.state);
  unsigned long long words = 0;
  if (J.stat[0] == 0) {
    words = (unsigned long long)(unsigned)J.stat[1];
    g.s = pcg_advance(g.s, g.inc, words);
  } else {
    for (long long k = 0; k < J.n;) {
      const uint64_t w = g.next();
      const ZigAttempt a = zig_attempt<false>(w, g);
      for (int j = 1; j < a.len; ++j) g.next();
      words += (unsigned long long)a.len;
      if (a.out) J.out[k++] = a.v;
    }
  }
  J.state[0] = (uint64_t)(g.s >> 64);
  J.state[1] = (uint64_t)g.s;
  J.stat[0] = 0;
  return words;
}

static_assert(kRngAttemptThreads <= 256, "jump table covers 256 threads");
__global__ void __launch_bounds__(kRngAttemptThreads) k_rng_attempts(RngJob J) {
  __shared__ __align__(16) uint8_t lds[kRngAttemptThreads * 3 * kRngChunk];
  rng_attempt_body(J, blockIdx.x * kRngAttemptThreads, lds + threadIdx.x * 3 * kRngChunk);
}
__global__ void __launch_bounds__(kRngScan) k_rng_scan(RngJob J, int force_serial) { rng_scan_body(J, force_serial); }
__global__ void __launch_bounds__(kRngScatterThreads) k_rng_scatter(RngJob J) {
  __shared__ double lds[kRngScatterThreads / 64][64 * kRngChunk];
  rng_scatter_body(J, blockIdx.x * kRngScatterThreads + threadIdx.x, lds[threadIdx.x >> 6]);
}
__global__ void k_rng_finish(RngJob J) {
  if (threadIdx.x == 0) rng_finish_body(J);
}

}  // namespace hitx
