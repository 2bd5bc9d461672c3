// Batched launches: many instances of possibly different grid sizes and kinds
// (forced HIT and Riemann, one element type) advance together, one launch
// per kernel kind per stage for all of them (the line sweeps: one per
// periodicity present).  Each instance has a descriptor (Inst) in device
// memory; a block finds its instance and local block index in a per-kind
// block map (one 32-bit entry per block, built on the host) and runs the same
// per-block body as the single-instance kernels.
// Per-instance scalar control (dt, Mach control, the forcing's OU recurrence
// and power rescale, step counting) runs in small device kernels, so a whole
// step needs no host synchronisation and can be captured in a HIP graph.
#pragma once

#include "hitx_kernels.h"
#include "hitx_rng.h"
#include "hitx_line.h"

namespace hitx {

// per-instance control scalars (double)
enum {
  C_ACTIVE = 0, C_TIME, C_TFINAL, C_DT, C_STEP, C_MAXSTEPS, C_HVNOW, C_HVAPPS,
  C_MAXSR, C_SUV, C_SC, C_POWER, C_ALPHA, C_DECAY, C_INCR, C_SCALE,
  C_SFF, C_SP, C_FXX, C_FYY, C_FXY, C_PIN, C_PBEFORE, C_ATGT, C_MACH, C_PDES, C_NSTEPS,
  C_NPRE,   // 1: the noise and projection partials of the next step are drawn (prefetched)
  C_NPREW,  // PCG64 words that prefetched draw consumed (a copy of the instance rewinds by them)
  C_NFALL,  // 1: the last draw took the serial path (noise in the noise buffer, projected by b_fproject)
  C_COUNT = 32
};
// per-instance run parameters (double, constant); NaN encodes "None"
enum {
  K_CFL = 0, K_VCLIP, K_FIXED_DT, K_HVINT, K_HVON, K_MACHON, K_MTGT, K_MMEM, K_MEXP, K_MMIN, K_MMAX,
  K_PTGT, K_MINP, K_MAXR, K_AMEM, K_FRMS, K_TCORR, K_FORCE, K_TFINAL, K_MAXSTEPS, K_NSTEPS,
  K_COUNT = 24
};

template <class T>
struct Inst {
  Params<T> P;
  LineK<T> Kx, Ky;
  GPtr<T> Q, Q1, Q2, S;  // state, stage 1 and 2 results, stage 3 result (before hyperviscosity)
  GPtr<T> qT, dfx, dfyT, fx, fy;
  GPtr<uint8_t> mask, maskT, rawm;  // dilated mask, its transpose, undilated mask (split sensor)
  GPtr<double> part, scal, ctl, kpar, phi, fhat, noise, fpart, kxy;
  GPtr<uint64_t> rng;  // PCG64 state: state hi, state lo, increment hi, increment lo
  RngJob R;            // forcing noise request: nx*ny normals of the rng stream into noise
  GPtr<double> drows, frows;  // per-row synthesis tables: drag low-pass, forcing
  int nblk_stage, nblk_sr, nblk_proj;
  int nsw;     // wavefronts of the noise scatter (one projection partial row each)
  int ntile;   // 32x16 tiles of the grid
  int use_qT;  // the y-sweep reads the transposed state (block form or wave form without the slab)
};

// offsets kinds (rows of the offsets table, each ninst+1 ints)
enum {
  OFF_RED = 0, OFF_SENS, OFF_LX, OFF_LY, OFF_T16, OFF_PROJ, OFF_RNG, OFF_SR, OFF_DROWS, OFF_FROWS, OFF_LXW, OFF_LYW,
  OFF_RNGS, OFF_DB, OFF_FS, OFF_XU, OFF_KINDS
};

// rows per wavefront of the batch forcing synthesis (column walk)
constexpr int kFsRows = 16;

constexpr int kMapLbBits = 20;  // local blocks per instance and kind < 2^20; instances < 2^12

// (selected as global pointers, converted once)
template <class T>
__device__ __forceinline__ GPtr<T> stage_in(const Inst<T>& I, int st) {
  return st == 0 ? I.Q : st == 1 ? I.Q1 : st == 2 ? I.Q2 : I.S;
}
template <class T>
__device__ __forceinline__ GPtr<T> stage_out(const Inst<T>& I, int st) {
  return st == 0 ? I.Q1 : st == 1 ? I.Q2 : I.S;
}

// The block map of kind KIND (built on the host) gives a block's instance
// and local block in one load: entry = instance << kMapLbBits | local block.
#define HITX_BATCH_INST_AT(KIND, GB)                                                  \
  const int gb_ = (GB);                                                               \
  const uint32_t code_ = (uint32_t)offs[offs[OFF_KINDS * (ninst + 1) + (KIND)] + gb_]; \
  const int ii_ = (int)(code_ >> kMapLbBits);                                         \
  const Inst<T>& I = tab[ii_];                                                        \
  if (I.ctl[C_ACTIVE] == 0.0) return;                                                 \
  const int lb = (int)(code_ & ((1u << kMapLbBits) - 1u));
#define HITX_BATCH_INST(KIND) HITX_BATCH_INST_AT(KIND, (int)blockIdx.x)
// tile kernels: XCD-contiguous block order
#define HITX_BATCH_TILE(KIND) HITX_BATCH_INST_AT(KIND, xcd_remap((int)blockIdx.x, (int)gridDim.x))

// Batch-wide word after the block-map starts of the offsets buffer: nonzero
// when some instance runs hyperviscosity this step (set by b_ctl_begin,
// cleared by b_ctl_end).  The post-step launches read it first, so on the
// steps without hyperviscosity every block exits after one load instead of
// the block map, instance table and control loads.
__device__ __forceinline__ int hv_any(const int* offs, int ninst) {
  return offs[OFF_KINDS * (ninst + 1) + OFF_KINDS];
}
#define HITX_POST_STEP_EXIT \
  if (st == 3 && hv_any(offs, ninst) == 0) return;

// ---------------------------------------------------------------- stage path
// st = 0, 1, 2: RK stages on Q, Q1, Q2 (enforced input, transposed copy);
// st = 3: the post-step sensor on S for hyperviscosity (as given, no copy).
template <class T, bool DRAG>
__global__ void __launch_bounds__(kBlock) b_stage_reduce(const Inst<T>* __restrict__ tab, const int* __restrict__ offs,
                                                         int ninst, int st) {
  HITX_POST_STEP_EXIT
  HITX_BATCH_INST(OFF_SR)
  if (st == 3 && I.ctl[C_HVNOW] == 0.0) return;
  if (DRAG && !I.P.drag_on) return;  // instances without drag reduce on tiles (b_stage_tile)
  stage_reduce_body<T, DRAG>(I.P, stage_in(I, st), st < 3 ? 2 : 0, I.part, lb);
}

// The stage reduction (with the drag projection when DRAG) runs on the
// sensor's 32x16 tiles, as in single plans (the same partial rows, so a
// batch instance and its single plan reduce in the same order).
template <class T, bool DRAG>
__global__ void __launch_bounds__(kBlock) b_stage_tile(const Inst<T>* __restrict__ tab, const int* __restrict__ offs,
                                                       int ninst, int st, int walk) {
  HITX_POST_STEP_EXIT
  HITX_BATCH_TILE(OFF_SENS)
  if (st == 3 && I.ctl[C_HVNOW] == 0.0) return;
  if (walk && I.P.drag_on) return;  // reduced by the column walk (b_stage_reduce)
  const int ntx = (I.P.nx + kSTX - 1) / kSTX;
  const int by = lb / ntx, bx = lb - by * ntx;
  stage_tile_body<T, DRAG>(I.P, stage_in(I, st), st < 3 ? 2 : 0, I.part, bx, by, lb);
}

// Grid (instances, 1 + kFinDragGroups with drag): block y = 0 reduces the
// stage scalar columns and writes the stage scalars, block y = g >= 1 the
// drag columns [(g-1) kFinDragCols, g kFinDragCols).  Each column is reduced
// on its own in the fixed order, so the split gives the one-block bits.
constexpr int kFinDragCols = 12, kFinDragGroups = kDragCols / kFinDragCols;
static_assert(kDragCols % kFinDragCols == 0, "drag columns in whole groups");
template <class T, bool DRAG>
__global__ void __launch_bounds__(kBlock) b_stage_finalize(const Inst<T>* __restrict__ tab, int st, int walk) {
  const Inst<T>& I = tab[blockIdx.x];
  if (I.ctl[C_ACTIVE] == 0.0 || (st == 3 && I.ctl[C_HVNOW] == 0.0)) return;
  const int nblk = (walk && I.P.drag_on) ? I.nblk_sr : I.ntile;  // instances without drag reduce on tiles
  constexpr int NR = DRAG ? (R_DRAG + kDragCols) : R_DRAG;
  __shared__ double r[DRAG ? (R_DRAG > kFinDragCols ? R_DRAG : kFinDragCols) : R_DRAG];
  if (blockIdx.y == 0) {
    reduce_partial_cols<R_DRAG, 0x1Full>(I.part, nblk, NR, 0, r);
    __syncthreads();
    stage_scal_write<T, false>(I.P, r, I.scal);
    return;
  }
  if (DRAG) {
    const int c0 = ((int)blockIdx.y - 1) * kFinDragCols;
    if (!I.P.drag_on || c0 >= 4 * I.P.ndrag) return;
    reduce_partial_cols<kFinDragCols, 0>(I.part, nblk, NR, R_DRAG + c0, r);
    __syncthreads();
    const double N = (double)I.P.nx * (double)I.P.ny;
    for (int k = threadIdx.x; k < kFinDragCols && c0 + k < 4 * I.P.ndrag; k += blockDim.x)
      I.scal[S_DRAG + c0 + k] = r[k] / N;
  }
}

template <class T>
__global__ void __launch_bounds__(kBlock) b_drows(const Inst<T>* __restrict__ tab, const int* __restrict__ offs,
                                                  int ninst, int st) {
  HITX_POST_STEP_EXIT
  HITX_BATCH_INST(OFF_DROWS)
  if (st == 3 && I.ctl[C_HVNOW] == 0.0) return;
  drows_body<T>(I.P, I.scal + S_DRAG, I.drows, lb * kBlock + threadIdx.x);
}

template <class T>
__global__ void __launch_bounds__(kBlock) b_sensor(const Inst<T>* __restrict__ tab, const int* __restrict__ offs,
                                                   int ninst, int st) {
  HITX_POST_STEP_EXIT
  HITX_BATCH_TILE(OFF_SENS)
  if (st == 3 && I.ctl[C_HVNOW] == 0.0) return;
  const int ntx = (I.P.nx + kSTX - 1) / kSTX;
  const int by = lb / ntx, bx = lb - by * ntx;
  sensor_body<T>(I.P, stage_in(I, st), st < 3, I.scal, I.mask, I.maskT, st < 3 && I.use_qT ? I.qT : nullptr, bx, by);
}

// Split sensor: undilated mask by column walk, then dilation + transposes.
template <class T>
__global__ void __launch_bounds__(kBlock) b_rawmask(const Inst<T>* __restrict__ tab, const int* __restrict__ offs,
                                                    int ninst, int st) {
  HITX_POST_STEP_EXIT
  HITX_BATCH_INST(OFF_SR)
  if (st == 3 && I.ctl[C_HVNOW] == 0.0) return;
  rawmask_body<T>(I.P, stage_in(I, st), st < 3, I.scal, I.rawm, lb);
}

template <class T, bool PER>
__global__ void __launch_bounds__(kBlock) b_rawtile(const Inst<T>* __restrict__ tab, const int* __restrict__ offs,
                                                    int ninst, int st, int ndrows) {
  HITX_POST_STEP_EXIT
  // the first ndrows blocks build the drag row table (stages 0-2 of a batch
  // with drag: it reads the stage scalars the finalize wrote, as the raw mask
  // does, so it rides in this launch instead of a launch on another stream)
  if ((int)blockIdx.x < ndrows) {
    HITX_BATCH_INST_AT(OFF_DROWS, (int)blockIdx.x)
    drows_body<T>(I.P, I.scal + S_DRAG, I.drows, lb * kBlock + threadIdx.x);
    return;
  }
  HITX_BATCH_INST_AT(OFF_SENS, xcd_remap((int)blockIdx.x - ndrows, (int)gridDim.x - ndrows))
  if (st == 3 && I.ctl[C_HVNOW] == 0.0) return;
  if (I.P.periodic != (PER ? 1 : 0)) return;  // the other periodicity's launch
  const int ntx = (I.P.nx + kSTX - 1) / kSTX;
  const int by = lb / ntx, bx = lb - by * ntx;
  rawtile_body<T, PER>(I.P, stage_in(I, st), st < 3, I.scal, I.rawm, st < 3 && I.use_qT ? I.qT : nullptr, bx, by);
}

template <class T>
__global__ void __launch_bounds__(kBlock) b_dilate(const Inst<T>* __restrict__ tab, const int* __restrict__ offs,
                                                   int ninst, int st) {
  HITX_POST_STEP_EXIT
  HITX_BATCH_TILE(OFF_SENS)
  if (st == 3 && I.ctl[C_HVNOW] == 0.0) return;
  const int ntx = (I.P.nx + kSTX - 1) / kSTX;
  const int by = lb / ntx, bx = lb - by * ntx;
  dilate_body<T>(I.P, stage_in(I, st), st < 3, I.rawm, I.mask, I.maskT, st < 3 && I.use_qT ? I.qT : nullptr, bx,
                 by);
}

// Bit dilation: kDBWaves tiles of 64 rows x 32 columns per block (the raw
// mask and, at stages 0-2 with the tile raw mask, the transposed state are
// written by the raw-mask tiles).
template <class T>
__global__ void __launch_bounds__(64 * kDBWaves) b_dilate_bits(const Inst<T>* __restrict__ tab,
                                                               const int* __restrict__ offs, int ninst, int st) {
  __shared__ uint32_t hs[kDBWaves][kDBRows + 16];
  HITX_POST_STEP_EXIT
  HITX_BATCH_INST(OFF_DB)
  if (st == 3 && I.ctl[C_HVNOW] == 0.0) return;
  const int ntx = (I.P.nx + kDBCols - 1) / kDBCols, nty = (I.P.ny + kDBRows - 1) / kDBRows;
  const int wv = threadIdx.x >> 6, t = lb * kDBWaves + wv;
  if (t >= ntx * nty) return;
  dilate_bits_tile<T>(I.P, I.rawm, I.mask, I.maskT, t % ntx, t / ntx, hs[wv]);
}

template <class T, bool PER, bool SMALL>
__global__ void __launch_bounds__(kLineBlock, 4) b_line(const Inst<T>* __restrict__ tab, const int* __restrict__ offs,
                                                     int ninst, int axis, int st) {
  HITX_BATCH_INST(axis == 0 ? OFF_LX : OFF_LY)
  if (I.P.periodic != (PER ? 1 : 0)) return;  // the other periodicity's launch
  if (axis == 0)
    line_body<T, PER, SMALL>(I.P, I.Kx, 0, 0, stage_in(I, st), I.mask, I.scal, I.dfx, lb, I.P.ny);
  else
    line_body<T, PER, SMALL>(I.P, I.Ky, 1, 1, I.qT, I.maskT, I.scal, I.dfyT, lb, I.P.nx);
}

// Wave form of the sweep: kLineWaves lines per block, one per wavefront.
template <class T, bool PER>
__global__ void __launch_bounds__(64 * kLineWaves) HITX_LINE_WAVE_ATTR(T, PER)
    b_line_wave(const Inst<T>* __restrict__ tab, const int* __restrict__ offs, int ninst, int axis, int st,
                int slice_bytes) {
  extern __shared__ __align__(16) unsigned char wsm[];
  HITX_BATCH_INST(axis == 0 ? OFF_LXW : OFF_LYW)
  if (I.P.periodic != (PER ? 1 : 0)) return;  // the other periodicity's launch
  const int w = threadIdx.x >> 6;
  const int line = lb * kLineWaves + w;
  const int nl = axis == 0 ? I.P.ny : I.P.nx;
  if (line >= nl) return;
  if (axis == 0)
    line_wave_any<T, PER>(I.P, I.Kx, 0, 0, stage_in(I, st), I.mask, I.scal, I.dfx, line, nl,
                          wsm + (size_t)w * slice_bytes);
  else
    line_wave_any<T, PER>(I.P, I.Ky, 1, 1, I.qT, I.maskT, I.scal, I.dfyT, line, nl, wsm + (size_t)w * slice_bytes);
}

// Slab form of the y-sweep: kSlabWaves adjacent columns of one instance per
// block, read from the stage state (enforced on load) through the LDS slab
// instead of a transposed copy; derivatives go to dfyT as in the other forms.
template <class T, bool PER>
__global__ void __launch_bounds__(64 * kSlabWaves) HITX_LINE_WAVE_ATTR(T, PER)
    b_line_wave_cols(const Inst<T>* __restrict__ tab, const int* __restrict__ offs, int ninst, int st,
                     int slice_bytes) {
  extern __shared__ __align__(16) unsigned char wsm[];
  // XCD-aware: the blocks of neighbouring columns share the 128-byte lines
  // of each row, so they run on one XCD (one L2)
  HITX_BATCH_TILE(OFF_LYW)
  if (I.P.periodic != (PER ? 1 : 0)) return;  // the other periodicity's launch (whole block, before any barrier)
  const int nx = I.P.nx, ny = I.P.ny;
  const int x0 = lb * kSlabWaves, cs = slab_cstride(ny);
  T* slab = reinterpret_cast<T*>(wsm);
  slab_fill<T>(stage_in(I, st), nx, ny, x0, slab, cs);
  __syncthreads();
  const int w = threadIdx.x >> 6;
  const int line = x0 + w;
  if (line >= nx) {  // no column: meet the sweep's one barrier, then leave
    __syncthreads();
    return;
  }
  line_wave_slab_any<T, PER>(I.P, I.Ky, slab + w * cs, kSlabWaves * cs, I.maskT, I.scal, I.dfyT, line, nx,
                             wsm + (size_t)w * slice_bytes);
}

// update tiles are the sensor's 32x16 tiles (the same block offsets)
static_assert(kUTX == kSTX && kUTY == kSTY, "update and sensor tiles differ");
// FP32 update tiles fit 8 wavefronts per SIMD (<= 64 VGPRs, four 512-thread
// blocks and their 30 KB of LDS per CU) without spilling; FP64 keeps its
// registers.
template <class T>
struct UpdateOcc {
  static constexpr int waves = 1;
};
template <>
struct UpdateOcc<float> {
  static constexpr int waves = 8;
};
template <class T, bool PER>
__global__ void __launch_bounds__(kUBlock) __attribute__((amdgpu_waves_per_eu(UpdateOcc<T>::waves))) b_update(const Inst<T>* __restrict__ tab, const int* __restrict__ offs,
                                                    int ninst, int st) {
  HITX_BATCH_TILE(OFF_SENS)
  if (I.P.periodic != (PER ? 1 : 0)) return;  // the other periodicity's launch
  const int ntx = (I.P.nx + kUTX - 1) / kUTX;
  const int by = lb / ntx, bx = lb - by * ntx;
  // the last stage writes Q in place when no hyperviscosity pass follows and
  // there is no outflow guard (each node then reads Q only at itself)
  const GPtr<T> out = (st == 2 && I.ctl[C_HVNOW] == 0.0 && I.P.guard == 0) ? I.Q : stage_out(I, st);
  update_body<T, PER>(I.P, MODE_STAGE, st, I.Q, stage_in(I, st), I.dfx, I.dfyT, I.fx, I.fy, I.scal, I.drows, (T)I.ctl[C_DT],
                 out, bx, by, (T)I.ctl[C_SCALE], (T)I.ctl[C_ALPHA]);
}

// The x-sweep and the stage update of a panel of kXUWaves rows in one block
// (periodic grids, lines in the wave form): each wave sweeps one row into
// dfx, then the block updates the panel in kXTX x kXUWaves tiles, reading
// the x-derivatives its own waves just wrote (L2 hits, not HBM) and the
// panel's state again (its rows were just read by the sweep) instead of in a
// separate update launch.  The same arithmetic as b_line_wave (axis 0)
// followed by b_update, so the results are bit-identical; the line slices
// and the update tiles share the dynamic LDS.
constexpr int kXUWaves = 8, kXTX = kUBlock / kXUWaves;
template <class T>
__global__ void __launch_bounds__(64 * kXUWaves) HITX_LINE_WAVE_ATTR(T, true)
    b_xupdate(const Inst<T>* __restrict__ tab, const int* __restrict__ offs, int ninst, int st, int slice_bytes) {
  extern __shared__ __align__(16) unsigned char wsm[];
  HITX_BATCH_TILE(OFF_XU)
  const int nx = I.P.nx, ny = I.P.ny;
  const int y0 = lb * kXUWaves;
  const int w = threadIdx.x >> 6;
  const int line = y0 + w;
  if (line < ny)
    line_wave_any<T, true>(I.P, I.Kx, 0, 0, stage_in(I, st), I.mask, I.scal, I.dfx, line, ny,
                           wsm + (size_t)w * slice_bytes);
  __syncthreads();  // the panel's x-derivatives are written; the slices are free
  // the last stage writes Q in place when no hyperviscosity pass follows
  // (each node then reads Q only at itself; these grids have no guard)
  const GPtr<T> out = (st == 2 && I.ctl[C_HVNOW] == 0.0) ? I.Q : stage_out(I, st);
  UpdLds<T, kXTX, kXUWaves>& L = *reinterpret_cast<UpdLds<T, kXTX, kXUWaves>*>(wsm);
  const T dt = (T)I.ctl[C_DT], fsc = (T)I.ctl[C_SCALE], fal = (T)I.ctl[C_ALPHA];
  for (int x0 = 0; x0 < nx; x0 += kXTX) {
    if (x0 > 0) __syncthreads();  // the previous tile's LDS reads are done
    update_tile<T, true, kXTX, kXUWaves>(I.P, MODE_STAGE, st, I.Q, stage_in(I, st), I.dfx, I.dfyT, I.fx, I.fy,
                                         I.scal, I.drows, dt, out, x0, y0, L, fsc, fal);
  }
}

// After the stages: hyperviscosity S -> Q with the fresh mask, or the S -> Q
// copy when the last stage could not write Q in place.
template <class T>
__global__ void __launch_bounds__(kBlock) b_finish(const Inst<T>* __restrict__ tab, const int* __restrict__ offs,
                                                   int ninst, int guard_any) {
  if (!guard_any && hv_any(offs, ninst) == 0) return;  // nothing to do in any instance
  HITX_BATCH_TILE(OFF_T16)
  const int ntx = (I.P.nx + 15) / 16;
  const int by = lb / ntx, bx = lb - by * ntx;
  if (I.ctl[C_HVNOW] != 0.0) {
    hv_body<T>(I.P, I.S, I.mask, 1, 1, I.Q, bx, by);
    return;
  }
  if (I.P.guard == 0) return;
  const int x = bx * 16 + (threadIdx.x & 15), y = by * 16 + (threadIdx.x >> 4);
  if (x >= I.P.nx || y >= I.P.ny) return;
  const int N = I.P.nx * I.P.ny, o = y * I.P.nx + x;
#pragma unroll
  for (int c = 0; c < 4; ++c) I.Q[c * N + o] = I.S[c * N + o];
}

// ---------------------------------------------------------------- step control
template <class T>
__global__ void __launch_bounds__(kBlock) b_step_reduce(const Inst<T>* __restrict__ tab, const int* __restrict__ offs,
                                                        int ninst) {
  HITX_BATCH_INST(OFF_RED)
  step_reduce_body<T>(I.P, I.Q, I.part, lb, I.nblk_stage);
}

__device__ __forceinline__ bool is_none(double v) { return v != v; }

// One block per instance: finish the step reductions, then (thread 0) the
// reference's scalar control: dt with the clip to the remaining time and the
// viscous limit (OOP/hit2d.py:829-832; fixed dt for the Riemann loop), the
// Mach-controlled target power (OOP/hit2d.py:293-349), the OU decay and
// increment (OOP/forcing.py:75-76), and whether hyperviscosity runs this step.
template <class T>
__global__ void __launch_bounds__(kBlock) b_ctl_begin(const Inst<T>* __restrict__ tab, int* __restrict__ hvany) {
  const Inst<T>& I = tab[blockIdx.x];
  double* c = I.ctl;
  if (c[C_ACTIVE] == 0.0) return;
  __shared__ double r[3];
  reduce_partials<3, 1>(I.part, I.nblk_stage, r);
  __syncthreads();
  if (threadIdx.x != 0) return;
  const double* k = I.kpar;
  c[C_MAXSR] = r[0];
  c[C_SUV] = r[1];
  c[C_SC] = r[2];
  const double remaining = c[C_TFINAL] - c[C_TIME];
  if (k[K_FIXED_DT] == 0.0) {
    double dt = k[K_CFL] / r[0];
    dt = fmin(dt, remaining);
    if (!is_none(k[K_VCLIP])) dt = fmin(fmin(dt, k[K_VCLIP]), remaining);
    c[C_DT] = dt;
  }
  if (k[K_FORCE] != 0.0) {
    const double N = (double)I.P.nx * (double)I.P.ny;
    const double mach = sqrt(r[1] / N) / (r[2] / N);
    c[C_MACH] = mach;
    const double power = c[C_POWER];
    c[C_PDES] = power;
    if (k[K_MACHON] != 0.0 && !is_none(power)) {
      const double ratio = pow(k[K_MTGT] / fmax(mach, 1.0e-8), k[K_MEXP]);
      double desired = power * ratio;
      if (!is_none(k[K_MMIN])) desired = fmax(desired, k[K_MMIN]);
      if (!is_none(k[K_MMAX])) desired = fmin(desired, k[K_MMAX]);
      c[C_PDES] = desired;
      c[C_POWER] = k[K_MMEM] * power + (1.0 - k[K_MMEM]) * desired;
    }
    const double decay = exp(-c[C_DT] / k[K_TCORR]);
    c[C_DECAY] = decay;
    c[C_INCR] = sqrt(fmax(1.0 - decay * decay, 0.0));
  }
  const int hv_int = (int)k[K_HVINT];
  c[C_HVNOW] = (k[K_HVON] != 0.0 && hv_int > 0 && (((long long)c[C_STEP] + 1) % hv_int) == 0) ? 1.0 : 0.0;
  if (c[C_HVNOW] != 0.0) atomicOr(hvany, 1);
}

// Fixed dt of the Riemann loop from the initial state: n = ceil(T/dt0),
// dt = T/n (OOP/problems/riemann_config3.py:371-378).  One block per
// instance; instances that already hold their step count are left alone.
template <class T>
__global__ void __launch_bounds__(kBlock) b_fixed_dt(const Inst<T>* __restrict__ tab) {
  const Inst<T>& I = tab[blockIdx.x];
  if (I.kpar[K_FIXED_DT] == 0.0 || I.ctl[C_NSTEPS] > 0.0 || I.ctl[C_ACTIVE] == 0.0) return;
  __shared__ double r[3];
  reduce_partials<3, 1>(I.part, I.nblk_stage, r);
  __syncthreads();
  if (threadIdx.x != 0) return;
  double* c = I.ctl;
  const double dt0 = I.kpar[K_CFL] / r[0];
  const double nsteps = ceil(c[C_TFINAL] / dt0);
  c[C_NSTEPS] = nsteps;
  c[C_DT] = c[C_TFINAL] / nsteps;
  c[C_ACTIVE] = nsteps > 0.0 ? 1.0 : 0.0;
}

// time, step and activity after the step (OOP/hit2d.py:870-871; the Riemann
// loop's time = (step+1)*dt, OOP/problems/riemann_config3.py:741)
template <class T>
__global__ void b_ctl_end(const Inst<T>* __restrict__ tab, int ninst, int* __restrict__ hvany) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) *hvany = 0;
  if (i >= ninst) return;
  double* c = tab[i].ctl;
  const double* k = tab[i].kpar;
  if (c[C_ACTIVE] == 0.0) return;
  c[C_STEP] = c[C_STEP] + 1.0;
  c[C_HVAPPS] = c[C_HVAPPS] + c[C_HVNOW];
  if (k[K_FIXED_DT] != 0.0) {
    c[C_TIME] = c[C_STEP] * c[C_DT];
    c[C_ACTIVE] = (c[C_STEP] < c[C_NSTEPS] && c[C_STEP] < k[K_MAXSTEPS]) ? 1.0 : 0.0;
  } else {
    c[C_TIME] = c[C_TIME] + c[C_DT];
    c[C_ACTIVE] = (c[C_TIME] < c[C_TFINAL] && c[C_STEP] < k[K_MAXSTEPS]) ? 1.0 : 0.0;
  }
}

// ---------------------------------------------------------------- forcing
// The projection of a draw that took the serial path (its noise is in the
// noise buffer); the parallel draw projects in its scatter.  (After
// b_rng_finish: with pre = 1 the mark is set and this projection is part of
// the prefetch, so only pre = 0 reads it.)
template <class T>
__global__ void __launch_bounds__(kBlock) b_fproject(const Inst<T>* __restrict__ tab, const int* __restrict__ offs,
                                                     int ninst, int pre) {
  HITX_BATCH_INST(OFF_PROJ)
  if (rng_skip(I.ctl, I.kpar, pre) || I.ctl[C_NFALL] == 0.0) return;
  fproject_body<T>(I.P, I.noise, I.fpart, lb);
}

// The instance's next nx*ny standard normals (OOP/forcing.py:77), in place
// of the host draw: four launches, see hitx_rng.h; then their projection.
// The draw for step n+1 runs during step n (pre = 1, after stage 0's update,
// beside stages 1-2) and marks the instance (C_NPRE); at the start of a step
// the same launches with pre = 0 draw only for instances not marked (the
// first step, a copied instance), and the OU update clears the mark.
// The noise depends only on the PCG64 state, so the values are those of a
// draw at the start of the step.
__device__ __forceinline__ bool rng_skip(const double* ctl, const double* kpar, int pre) {
  return kpar[K_FORCE] == 0.0 || (pre == 0 && ctl[C_NPRE] != 0.0);
}
template <class T>
__global__ void __launch_bounds__(kRngAttemptThreads) b_rng_attempts(const Inst<T>* __restrict__ tab,
                                                                     const int* __restrict__ offs, int ninst,
                                                                     int pre) {
  HITX_BATCH_INST(OFF_RNG)
  if (rng_skip(I.ctl, I.kpar, pre)) return;
  __shared__ __align__(16) uint8_t lds[kRngAttemptThreads * 3 * kRngChunk];
  rng_attempt_body(I.R, lb * kRngAttemptThreads, lds + threadIdx.x * 3 * kRngChunk);
}

template <class T>
__global__ void __launch_bounds__(kRngScan) b_rng_scan(const Inst<T>* __restrict__ tab, int pre, int force_serial) {
  const Inst<T>& I = tab[blockIdx.x];
  if (I.ctl[C_ACTIVE] == 0.0 || rng_skip(I.ctl, I.kpar, pre)) return;
  rng_scan_body(I.R, force_serial);
}

// The noise draw's scatter with the projection formed in place of the
// noise field (OOP/forcing.py:77-79: fft2 of the normals at the shell
// modes).  Each wavefront's values are a contiguous index range [lo, hi)
// staged in its LDS slice (rng_scatter_body); row by row the wave sums
// v exp(-i k x) over its part of the row for k = 0..fkx (twiddles of k from
// the k = 1 entry by the angle-addition recurrence; lane partial sums added
// over the wave in a fixed order, wave_reduce), then lane m adds exp(-i ky y) times the
// row sum at its mode's kx (a negative kx takes the conjugate), and writes
// the wave's partial row of fft2 at the modes.  The noise never goes to
// memory; a draw that falls back to the serial form writes the noise and
// b_fproject projects it.
template <class T>
__device__ __forceinline__ void rng_scatter_project_body(const Params<T>& P, const RngJob& J, int c, double* lds,
                                                          double* __restrict__ part, const double* Wi) {
  long long lo, hi;
  if (!rng_scatter_stage(J, c, lds, lo, hi, Wi)) return;
  const int lane = threadIdx.x & 63;
  const int nx = P.nx, nf = P.nforce, fk = P.fkx;
  int mkx = 0, mky = 0;
  if (lane < nf) {
    mkx = P.force_modes[2 * lane];
    mky = P.force_modes[2 * lane + 1];
  }
  const int mka = mkx < 0 ? -mkx : mkx, mkya = mky < 0 ? -mky : mky;
  double mre = 0.0, mim = 0.0;
  for (long long i0 = lo; i0 < hi;) {
    const int y = (int)(i0 / nx);
    const long long row0 = (long long)y * nx, rend = min(row0 + nx, hi);
    double rc[kFK + 1], rs[kFK + 1];
#pragma unroll
    for (int k = 0; k <= kFK; ++k) {
      rc[k] = 0.0;
      rs[k] = 0.0;
    }
    for (long long i = i0 + lane; i < rend; i += 64) {
      const double v = lds[i - lo];
      const int x = (int)(i - row0);
      double c1 = 1.0, s1 = 0.0;
      if (fk >= 1) {  // (the table holds |k| <= kmx, and kmx >= fkx)
        const double* t1 = P.twx + ((size_t)(1 + P.kmx) * nx + x) * 2;
        c1 = t1[0];
        s1 = t1[1];
      }
      rc[0] += v;
      double ck = c1, sk = s1;
#pragma unroll
      for (int k = 1; k <= kFK; ++k) {
        if (k <= fk) {
          rc[k] = fma(v, ck, rc[k]);
          rs[k] = fma(v, sk, rs[k]);
          const double cn = ck * c1 - sk * s1, sn = sk * c1 + ck * s1;
          ck = cn;
          sk = sn;
        }
      }
    }
#pragma unroll
    for (int k = 0; k <= kFK; ++k) {
      if (k <= fk) {
        rc[k] = wave_reduce<false>(rc[k]);
        rs[k] = wave_reduce<false>(rs[k]);
      }
    }
    if (lane < nf) {
      double Rc = 0.0, Rs = 0.0;
#pragma unroll
      for (int k = 0; k <= kFK; ++k)
        if (k == mka) {
          Rc = rc[k];
          Rs = rs[k];
        }
      if (mkx < 0) Rs = -Rs;  // sin(kx x) is odd in kx
      const double* ty = P.twy + ((size_t)(mkya + P.kmy) * P.ny + y) * 2;
      const double cy = ty[0], sy = mky < 0 ? -ty[1] : ty[1];
      // (cy - i sy)(Rc - i Rs)
      mre += cy * Rc - sy * Rs;
      mim -= cy * Rs + sy * Rc;
    }
    i0 = rend;
  }
  if (lane < nf) {
    part[2 * lane] = mre;
    part[2 * lane + 1] = mim;
  }
}

template <class T>
__global__ void __launch_bounds__(kRngScatterThreads) b_rng_scatter(const Inst<T>* __restrict__ tab,
                                                                    const int* __restrict__ offs, int ninst, int pre) {
  __shared__ double lds[kRngScatterThreads / 64][64 * kRngChunk];
  __shared__ double sWi[256];  // the layer table of the values, in LDS
  HITX_BATCH_INST(OFF_RNGS)
  if (rng_skip(I.ctl, I.kpar, pre)) return;  // (uniform over the block)
  for (int k = threadIdx.x; k < 256; k += kRngScatterThreads) sWi[k] = kZigWi[k];
  __syncthreads();
  const int c = lb * kRngScatterThreads + threadIdx.x;
  rng_scatter_project_body<T>(I.P, I.R, c, lds[threadIdx.x >> 6],
                              I.fpart + (size_t)((c >> 6)) * 2 * I.P.nforce, sWi);
}

template <class T>
__global__ void b_rng_finish(const Inst<T>* __restrict__ tab, int pre) {
  const Inst<T>& I = tab[blockIdx.x];
  if (threadIdx.x != 0 || I.ctl[C_ACTIVE] == 0.0 || rng_skip(I.ctl, I.kpar, pre)) return;
  const bool serial = I.R.stat[0] != 0;
  const unsigned long long words = rng_finish_body(I.R);
  I.ctl[C_NFALL] = serial ? 1.0 : 0.0;
  if (pre) {
    I.ctl[C_NPREW] = (double)words;
    I.ctl[C_NPRE] = 1.0;
  }
}

// One block per instance: sum the projection partials, OU recurrence of the
// shell coefficients (OOP/forcing.py:80) and the curl (:82-83).
template <class T>
__global__ void __launch_bounds__(kBlock) b_fou(const Inst<T>* __restrict__ tab) {
  const Inst<T>& I = tab[blockIdx.x];
  if (I.ctl[C_ACTIVE] == 0.0 || I.kpar[K_FORCE] == 0.0) return;
  const int nm = I.P.nforce;
  const double decay = I.ctl[C_DECAY], incr = I.ctl[C_INCR];
  // partial rows: the scatter's wavefronts, or the serial path's projection blocks
  const int nb = I.ctl[C_NFALL] != 0.0 ? I.nblk_proj : I.nsw;
  for (int m = threadIdx.x; m < nm; m += blockDim.x) {
    // the partials in row order, eight rows' loads in flight at a time
    double re = 0.0, im = 0.0;
    const double* fp = I.fpart + (size_t)m * 2;
    const size_t stride = (size_t)nm * 2;
    int b = 0;
    for (; b + 8 <= nb; b += 8) {
      double r8[8], i8[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        r8[j] = fp[(size_t)(b + j) * stride];
        i8[j] = fp[(size_t)(b + j) * stride + 1];
      }
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        re += r8[j];
        im += i8[j];
      }
    }
    for (; b < nb; ++b) {
      re += fp[(size_t)b * stride];
      im += fp[(size_t)b * stride + 1];
    }
    const double pr = decay * I.phi[2 * m] + incr * re;
    const double pi = decay * I.phi[2 * m + 1] + incr * im;
    I.phi[2 * m] = pr;
    I.phi[2 * m + 1] = pi;
    const double kx = I.kxy[2 * m], ky = I.kxy[2 * m + 1];
    I.fhat[2 * m] = 0.0 * pr - ky * pi;  // (1j*ky) * phi
    I.fhat[2 * m + 1] = 0.0 * pi + ky * pr;
    I.fhat[2 * nm + 2 * m] = 0.0 * pr - (-kx) * pi;  // (-1j*kx) * phi
    I.fhat[2 * nm + 2 * m + 1] = 0.0 * pi + (-kx) * pr;
  }
  if (threadIdx.x == 0) I.ctl[C_NPRE] = 0.0;  // the drawn noise is used
}

template <class T>
__global__ void __launch_bounds__(kBlock) b_frows(const Inst<T>* __restrict__ tab, const int* __restrict__ offs,
                                                  int ninst) {
  HITX_BATCH_INST(OFF_FROWS)
  if (I.kpar[K_FORCE] == 0.0) return;
  frows_body<T>(I.P, I.fhat, I.frows, lb * kBlock + threadIdx.x);
}

// Raw fx, fy from the shell coefficients (the rescale by scale*alpha is
// applied where the stage update reads them), with the sums the rescale and
// the statistics need: sum(fx^2+fy^2), sum(rho*(fx*u+fy*v)) at Q, sum(fx^2),
// sum(fy^2), sum(fx*fy).
// Column walk: a wavefront owns 64 columns and kFsRows rows.  Each lane keeps
// its column's twiddles cos/sin(k x), k = 1..fkx, in registers; the wave's
// rows of the synthesis table (2 fields x (2 fkx + 1) doubles per row) are
// copied to its LDS slice once and read as broadcasts, so a node costs the
// two inverse-DFT sums as fused multiply-adds (the summation order of
// fsynth_at, fused: within a few ulp of the single-plan kernel).
template <class T>
__global__ void __launch_bounds__(kBlock) b_fsynth(const Inst<T>* __restrict__ tab, const int* __restrict__ offs,
                                                   int ninst) {
  constexpr int kRow = 2 * (2 * kFK + 1);  // table entries of a row at most
  __shared__ double hs[kWaves][kFsRows * kRow];
  HITX_BATCH_INST(OFF_FS)
  if (I.kpar[K_FORCE] == 0.0) return;
  const Params<T>& P = I.P;
  const int nx = P.nx, ny = P.ny, N = nx * ny;
  const double invN = 1.0 / ((double)nx * (double)ny);
  const Walk W = walk_setup(nx, ny, 1, kFsRows, 0, lb);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int fk = P.fkx, FR = 2 * fk + 1;
  double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  if (W.live) {
    double* H0 = hs[wid];
    {
      const double* src = I.frows + (size_t)W.ya * 2 * FR;
      const int cnt = (W.yb - W.ya) * 2 * FR;
      for (int i = lane; i < cnt; i += 64) H0[i] = src[i];
    }
    double tc[kFK], ts[kFK];
    {
      const int xc = W.own ? W.x : 0;  // any column for lanes past the grid
#pragma unroll
      for (int k = 1; k <= kFK; ++k) {  // past fkx: the row of fkx, never used
        const double* tx = P.twx + ((size_t)(min(k, fk) + P.kmx) * nx + xc) * 2;
        tc[k - 1] = tx[0];
        ts[k - 1] = tx[1];
      }
    }
    // the slice is the wave's own: its stores complete before any lane reads
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    // the state of the current row for the power sums, the next row's loads
    // in flight while this row is synthesised
    T a0 = T(0), a1 = T(0), a2 = T(0), a3 = T(0);
    if (W.own) load4<T>(I.Q, N, W.ya * nx + W.x, a0, a1, a2, a3);
    for (int y = W.ya; y < W.yb; ++y) {
      T b0 = T(0), b1 = T(0), b2 = T(0), b3 = T(0);
      if (W.own && y + 1 < W.yb) load4<T>(I.Q, N, (y + 1) * nx + W.x, b0, b1, b2, b3);
      const double* H = H0 + (y - W.ya) * 2 * FR;
      double ax = H[0], ay = H[FR];
#pragma unroll
      for (int k = 1; k <= kFK; ++k) {
        if (k <= fk) {
          const double c = tc[k - 1], sn = ts[k - 1];
          ax = fma(H[2 * k - 1], c, ax);
          ax = fma(-H[2 * k], sn, ax);
          ay = fma(H[FR + 2 * k - 1], c, ay);
          ay = fma(-H[FR + 2 * k], sn, ay);
        }
      }
      if (W.own) {
        const int idx = y * nx + W.x;
        const T fxv = (T)(ax * invN), fyv = (T)(ay * invN);
        I.fx[idx] = fxv;
        I.fy[idx] = fyv;
        const Prim<T> pr = prim(P, a0, a1, a2, a3);
        acc[0] += (double)(fxv * fxv + fyv * fyv);
        acc[1] += (double)(pr.rho * (fxv * pr.u + fyv * pr.v));
        acc[2] += (double)(fxv * fxv);
        acc[3] += (double)(fyv * fyv);
        acc[4] += (double)(fxv * fyv);
      }
      a0 = b0;
      a1 = b1;
      a2 = b2;
      a3 = b3;
    }
  }
  // partials in fpart (free after the OU update): the stage reduction may
  // use part meanwhile
  block_reduce_write<5, 0>(acc, I.fpart + (size_t)lb * 5);
}

// One block per instance: rms normalisation and the power rescale with the
// alpha memory (OOP/forcing.py:87-112), and the statistics of the rescaled
// forcing (OOP/forcing.py:113-118) from the raw sums: the rescaled field is
// scale*alpha times the raw one.
template <class T>
__global__ void __launch_bounds__(kBlock) b_falpha(const Inst<T>* __restrict__ tab) {
  const Inst<T>& I = tab[blockIdx.x];
  if (I.ctl[C_ACTIVE] == 0.0 || I.kpar[K_FORCE] == 0.0) return;
  __shared__ double r[5];
  reduce_partials<5, 0>(I.fpart, walk_blocks(I.P.nx, I.P.ny, kFsRows, 0), r);
  __syncthreads();
  if (threadIdx.x != 0) return;
  double* c = I.ctl;
  const double* k = I.kpar;
  const double N = (double)I.P.nx * (double)I.P.ny;
  c[C_SFF] = r[0];
  c[C_SP] = r[1];
  const double vrms = sqrt(r[0] / N);
  double scale;
  if (k[K_FRMS] <= 0.0)
    scale = 0.0;
  else if (vrms > 2.220446049250313e-16)
    scale = k[K_FRMS] / vrms;
  else
    scale = 1.0;
  c[C_SCALE] = scale;
  const double pb = scale * r[1] / N;
  c[C_PBEFORE] = pb;
  const double target = c[C_POWER];
  double a = c[C_ALPHA];
  double atgt = a;
  if (!is_none(target) && fabs(pb) > k[K_MINP]) {
    atgt = target / pb;
    if (!is_none(k[K_MAXR])) atgt = fmin(fmax(atgt, -k[K_MAXR]), k[K_MAXR]);
    a = k[K_AMEM] * a + (1.0 - k[K_AMEM]) * atgt;
  } else if (is_none(target)) {
    a = 1.0;
  }
  c[C_ALPHA] = a;
  c[C_ATGT] = atgt;
  const double g = scale * a;
  c[C_FXX] = g * g * r[2] / N;
  c[C_FYY] = g * g * r[3] / N;
  c[C_FXY] = g * g * r[4] / N;
  c[C_PIN] = g * r[1] / N;
}

}  // namespace hitx
