// hitx C ABI: plan creation, workspace layout and launch sequences.
// Declarations and the reference interfaces they replace: include/hitx.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/hitx.h"
#include "hitx_batch.h"
#include "hitx_diag.h"
#include "hitx_init.h"
#include "hitx_line.h"
#include "hitx_post.h"

using namespace hitx;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIPCHK(expr)                                                                    \
  do {                                                                                  \
    hipError_t e_ = (expr);                                                             \
    if (e_ != hipSuccess) return fail(HITX_EDEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define LAUNCHCHK()                                                                     \
  do {                                                                                  \
    hipError_t e_ = hipGetLastError();                                                  \
    if (e_ != hipSuccess) return fail(HITX_EDEVICE, std::string("kernel launch: ") + hipGetErrorString(e_)); \
  } while (0)

size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

// bounds of the column-walk tile heights a batch may be tuned to
constexpr int kTuneRbMin = 2, kTuneRbMax = 64, kTuneRbpMin = 8, kTuneRbpMax = 256;
// a plan's own tile heights: enough waves for the grid alone
int default_rb(int nx, int ny) { return std::max(2, std::min(32, ((nx + 61) / 62) * ny / 2048)); }
int default_rbp(int nx, int ny) { return std::max(8, std::min(64, ((nx + 63) / 64) * ny / 1024)); }

// Workspace of one normal-variate request (see hitx_rng.h).
struct RngLayout {
  size_t cst, att, map, entry, base, stat, total;
};
RngLayout rng_layout(long long n) {
  const size_t nch = (size_t)rng_chunks(n), words = nch * kRngChunk;
  RngLayout L;
  L.cst = 0;
  L.att = align256(L.cst + nch * 2 * sizeof(uint64_t));
  L.map = align256(L.att + words);
  L.entry = align256(L.map + nch * kRngEntries * sizeof(uint16_t));
  L.base = align256(L.entry + nch);
  L.stat = align256(L.base + nch * sizeof(int));
  L.total = align256(L.stat + 4 * sizeof(int));
  return L;
}
RngJob make_rngjob(uint8_t* ws, const RngLayout& L, uint64_t* state, double* out, long long n) {
  RngJob J;
  J.state = state;
  J.out = out;
  J.cst = reinterpret_cast<uint64_t*>(ws + L.cst);
  J.att = ws + L.att;
  J.map = reinterpret_cast<uint16_t*>(ws + L.map);
  J.entry = ws + L.entry;
  J.base = reinterpret_cast<int*>(ws + L.base);
  J.stat = reinterpret_cast<int*>(ws + L.stat);
  J.n = n;
  J.W = rng_words(n);
  J.nch = rng_chunks(n);
  return J;
}
constexpr long long kRngMaxN = 1ll << 30;

constexpr double kAlphaC = 3.0 / 8.0;  // OOP/spatial_operator.py:240

template <class T>
LineK<T> make_linek(int n) {
  LineK<T> k{};
  const double a = kAlphaC;
  const double rho = (-1.0 + std::sqrt(1.0 - 4.0 * a * a)) / (2.0 * a);
  k.n = n;
  // one chunk of L <= 4 per thread of the line block when the line allows it
  // (the smallest such L: most threads busy); otherwise the largest chunk
  // length that still gives every thread a task (at most two)
  k.L = 0;
  for (int L : {1, 2, 4})
    if (n % L == 0 && 4 * (n / L) <= kLineBlock) {
      k.L = L;
      break;
    }
  k.small = k.L != 0;
  if (!k.small) {
    k.L = 1;
    for (int L : {8, 4, 2, 1})
      if (n % L == 0 && 4 * (n / L) >= kLineBlock) {
        k.L = L;
        break;
      }
  }
  k.nch = n / k.L;
  k.M = (52 + k.L - 1) / k.L;
  k.rho = (T)rho;
  k.c1p = (T)(1.0 + rho * rho);
  k.inv1mrn = (T)(1.0 / (1.0 - std::pow(rho, (double)n)));
  k.rnm1 = (T)std::pow(rho, (double)(n - 1));
  k.one_m_rho2 = (T)(1.0 - rho * rho);
  k.alpha_c = (T)a;
  double p = 1.0;
  for (int i = 0; i < 66; ++i) {
    k.pw[i] = (T)p;
    p *= rho;
  }
  return k;
}

}  // namespace

struct hitx_plan {
  hitx_desc d{};
  int device = 0;
  int esize = 8;
  int N = 0;
  int kmx = 0, kmy = 0;
  std::vector<int32_t> drag_modes, force_modes;
  double* d_tw = nullptr;   // twx then twy
  int* d_modes = nullptr;   // drag modes then force modes
  double* d_dk = nullptr;   // spectral derivative tables, x then y (periodic plans)
  LineK<double> lx_d{}, ly_d{};
  LineK<float> lx_f{}, ly_f{};
  // workspace layout
  size_t off_raw = 0, off_mask = 0, off_maskT = 0, off_qT = 0, off_dfx = 0, off_dfyT = 0, off_part = 0,
         off_scal = 0, off_misc = 0, ws_bytes = 0;
  size_t sensor_lds = 0;
  int nblk_red = 0, nblk_proj = 0, nblk_stage = 0, nblk_sr = 0;
  int rb = 2, rbp = 4;                // rows per wave of the column walks
  int line_form = 0;                  // 0: auto, 1: block form, 2: wave form where the line allows it
  int sensor_form = 0;                // 0: auto, 1: fused tile, 2: split (raw-mask column walk + dilation tile)
  int dkx = 0, dky = 0, fkx = 0, fky = 0;
  size_t off_drows = 0, off_frows = 0;
  size_t line_lds_x = 0, line_lds_y = 0;
  // optional per-launch timing of the line-sweep kernel (HIP events on the launch stream)
  int timing = 0;
  std::vector<hipEvent_t> ev;
  size_t ev_used = 0;
};

namespace {

template <class T>
Params<T> make_params(const hitx_plan* pl) {
  const hitx_desc& d = pl->d;
  Params<T> P{};
  P.nx = d.nx;
  P.ny = d.ny;
  P.periodic = d.periodic;
  P.weno_only = d.scheme == HITX_WENO;
  P.dx = (T)d.dx;
  P.dy = (T)d.dy;
  P.two_dx = (T)(2.0 * d.dx);
  P.two_dy = (T)(2.0 * d.dy);
  P.gamma = (T)d.gamma;
  P.gm1 = (T)(d.gamma - 1.0);
  P.rho_floor = (T)d.rho_floor;
  P.p_floor = (T)d.p_floor;
  P.pf_gm1 = (T)(d.p_floor / (d.gamma - 1.0));
  P.width = d.sensor_width;
  P.bguard = d.periodic ? 0 : d.boundary_guard;
  P.shear_on = d.shear_enabled;
  P.comp_thr = d.compression_threshold;
  P.jump_thr = d.jump_threshold;
  P.shear_thr = d.shear_threshold;
  P.visc_on = d.periodic && d.viscosity > 0.0;
  const double mu = d.viscosity;
  const double cp = d.gamma * d.gas_constant / (d.gamma - 1.0);
  P.mu = (T)mu;
  P.two_mu = (T)(2.0 * mu);
  P.tt_mu = (T)((2.0 / 3.0) * mu);
  P.kcond = (T)(mu * cp / d.prandtl);
  P.rgas = (T)d.gas_constant;
  for (int c = 0; c < 4; ++c) P.hv_coef[c] = (T)d.hv_coef[c];
  P.dx2 = (T)d.dx2;
  P.dy2 = (T)d.dy2;
  P.guard = d.periodic ? 0 : d.guard_cells;
  P.drag_on = d.periodic && d.drag_enabled && d.n_drag_modes > 0;
  P.cool_on = d.periodic && d.cooling_enabled;
  P.force_on = d.periodic && d.forcing_enabled;
  P.neg_drag = (T)(-d.drag_coef);
  P.cool_pref = d.cooling_p_ref;
  P.cool_tc = d.cooling_time;
  P.gamma_d = d.gamma;
  P.ndrag = d.n_drag_modes;
  P.nforce = d.n_force_modes;
  P.kmx = pl->kmx;
  P.kmy = pl->kmy;
  P.twx = pl->d_tw;
  P.twy = pl->d_tw ? pl->d_tw + (size_t)(2 * pl->kmx + 1) * d.nx * 2 : nullptr;
  P.drag_modes = pl->d_modes;
  P.force_modes = pl->d_modes ? pl->d_modes + 2 * d.n_drag_modes : nullptr;
  P.force_bykx = pl->d_modes ? pl->d_modes + 2 * d.n_drag_modes + 2 * d.n_force_modes : nullptr;
  P.rb = pl->rb;
  P.rbp = pl->rbp;
  P.dkx = pl->dkx;
  P.dky = pl->dky;
  P.fkx = pl->fkx;
  P.fky = pl->fky;
  return P;
}

// dynamic LDS of the forcing projection: the lanes' column sums, 2*fky+1 per lane
size_t proj_lds_bytes(const hitx_plan* pl) { return (size_t)kWaves * 64 * (2 * pl->fky + 1) * sizeof(double); }

// LDS of a line block (line_body): state and flux planes 8 (n+1), the power
// table 66, the scratch (one-chunk shape: the WENO staging of CPP planes and
// the solve's exchange slots; otherwise the chunk results 8 nch + 20, which
// also hold one staging plane), the work list (uint16, n+1), the mask.
size_t line_lds_scratch(int n, int nch, bool small, int esize) {
  const size_t S = (size_t)n + 1, cpp = small ? 2 : 1;
  return small ? std::max(cpp * S, (size_t)28) : 8 * (size_t)nch + 20;
}
size_t line_lds_bytes(int n, int nch, bool small, int esize) {
  return (size_t)(8 * (n + 1) + 66 + line_lds_scratch(n, nch, small, esize)) * esize + 2 * (size_t)(n + 1) + n + 64;
}

// Lines in one-chunk shape (every launched line small) go in one launch;
// otherwise one launch per sweep.
template <class T, bool PER>
void launch_k_line(bool small, size_t lds, hipStream_t s, const Params<T>& P, const LineK<T>& KX, const LineK<T>& KY,
                   const T* qx, const uint8_t* mx, const T* qy, const uint8_t* my, const double* scal, T* ox, T* oy,
                   int nlx, int nly) {
  if (small) {
    k_line<T, PER, true, 2><<<nlx + nly, kLineBlock, lds, s>>>(P, KX, KY, qx, mx, qy, my, scal, ox, oy, nlx, nly);
    return;
  }
  if (nlx > 0)
    k_line<T, PER, false, 0><<<nlx, kLineBlock, lds, s>>>(P, KX, KY, qx, mx, qy, my, scal, ox, oy, nlx, nly);
  if (nly > 0)
    k_line<T, PER, false, 1><<<nly, kLineBlock, lds, s>>>(P, KX, KY, qx, mx, qy, my, scal, ox, oy, nlx, nly);
}

template <class T>
hipError_t allow_line_lds(size_t bytes) {
  hipError_t e = hipSuccess;
  const int b = (int)bytes;
  const hipFuncAttribute a = hipFuncAttributeMaxDynamicSharedMemorySize;
  const void* fs[] = {
      reinterpret_cast<const void*>(&k_line<T, true, true, 2>), reinterpret_cast<const void*>(&k_line<T, true, false, 0>),
      reinterpret_cast<const void*>(&k_line<T, true, false, 1>), reinterpret_cast<const void*>(&k_line<T, false, true, 2>),
      reinterpret_cast<const void*>(&k_line<T, false, false, 0>),
      reinterpret_cast<const void*>(&k_line<T, false, false, 1>),
      reinterpret_cast<const void*>(&b_line<T, true, true>), reinterpret_cast<const void*>(&b_line<T, true, false>),
      reinterpret_cast<const void*>(&b_line<T, false, true>), reinterpret_cast<const void*>(&b_line<T, false, false>)};
  for (const void* f : fs)
    if (e == hipSuccess) e = hipFuncSetAttribute(f, a, b);
  return e;
}

template <class T>
hipError_t allow_slab_lds(size_t bytes) {
  hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&b_line_wave_cols<T, true>),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
  if (e == hipSuccess)
    e = hipFuncSetAttribute(reinterpret_cast<const void*>(&b_line_wave_cols<T, false>),
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
  return e;
}

size_t sensor_lds_bytes(int w, int esize) {
  const int H = w + 1, EX = kSTX + 2 * H, EY = kSTY + 2 * H;
  return (size_t)5 * EX * EY * esize + (size_t)(kSTX + 2 * w) * (kSTY + 2 * w) + (size_t)(kSTY + 2 * w) * kSTX +
         (size_t)kSTY * (kSTX + 1) + 64;
}

// Spectral derivative kernel of a periodic line of n nodes spaced h apart,
// extended for the tiled convolution: t[i] = d((i - kDiagPad) mod n),
// i < 2 n + 2 kDiagPad, with d(0) = 0 and, for m != 0,
// d(m) = (2 pi / (n h)) (-1)^m / 2 * (cot(pi m / n) for even n, 1/sin(pi m / n) for odd n),
// evaluated in long double.
std::vector<double> spectral_derivative_table(int n, double h) {
  const long double pi = 3.14159265358979323846264338327950288L;
  std::vector<double> d(n, 0.0);
  const long double scale = 2.0L * pi / ((long double)n * (long double)h);
  for (int m = 1; m < n; ++m) {
    const long double th = pi * (long double)m / (long double)n;
    const long double sg = (m & 1) ? -0.5L : 0.5L;
    const long double v = (n % 2 == 0) ? std::cos(th) / std::sin(th) : 1.0L / std::sin(th);
    d[m] = (double)(scale * sg * v);
  }
  std::vector<double> t((size_t)2 * n + 2 * kDiagPad);
  for (size_t i = 0; i < t.size(); ++i) t[i] = d[(((long)i - kDiagPad) % n + n) % n];
  return t;
}

template <class T>
const LineK<T>& linek(const hitx_plan* pl, int axis);
template <>
const LineK<double>& linek<double>(const hitx_plan* pl, int axis) { return axis == 0 ? pl->lx_d : pl->ly_d; }
template <>
const LineK<float>& linek<float>(const hitx_plan* pl, int axis) { return axis == 0 ? pl->lx_f : pl->ly_f; }

struct WS {
  uint8_t *raw, *mask, *maskT;
  void *qT, *dfx, *dfyT;
  double *part, *scal, *misc, *drows, *frows;
};

WS carve(const hitx_plan* pl, void* ws) {
  uint8_t* b = static_cast<uint8_t*>(ws);
  WS w;
  w.raw = b + pl->off_raw;
  w.mask = b + pl->off_mask;
  w.maskT = b + pl->off_maskT;
  w.qT = b + pl->off_qT;
  w.dfx = b + pl->off_dfx;
  w.dfyT = b + pl->off_dfyT;
  w.part = reinterpret_cast<double*>(b + pl->off_part);
  w.scal = reinterpret_cast<double*>(b + pl->off_scal);
  w.misc = reinterpret_cast<double*>(b + pl->off_misc);
  w.drows = reinterpret_cast<double*>(b + pl->off_drows);
  w.frows = reinterpret_cast<double*>(b + pl->off_frows);
  return w;
}

int check_ws(const hitx_plan* pl, void* ws, size_t ws_bytes) {
  if (!pl) return fail(HITX_EINVAL, "null plan");
  if (!ws || ws_bytes < pl->ws_bytes)
    return fail(HITX_ENOMEM, "workspace too small: need " + std::to_string(pl->ws_bytes) + " bytes");
  return HITX_OK;
}

dim3 grid16(const hitx_plan* pl) { return dim3((pl->d.nx + 15) / 16, (pl->d.ny + 15) / 16); }
dim3 grid_update(const hitx_plan* pl) { return dim3((pl->d.nx + kUTX - 1) / kUTX, (pl->d.ny + kUTY - 1) / kUTY); }

// stage reductions (one kernel, last block finalizes) + fused sensor tile:
// mask and maskT of q, and the enforced transposed state when qT != nullptr
// The fused tile recomputes primitives and jump indicators over its halo
// (2.1x); the split form forms the undilated mask on tiles with a one-node
// halo (1.2x), then dilates.
constexpr long long kSplitSensorMinNodes = 0;
// The drag stage reduction runs on the 32x16 sensor tiles while a launch has
// few of them (a tile's drag end phase is a serial loop, cheap when tiles are
// spread thin over the CUs); launches of many tiles use the column walk,
// whose end phase covers 62 columns of rb rows at once.
constexpr int kStageTileMaxTiles = 4096;
bool stage_walk_form(long long tiles) { return tiles > kStageTileMaxTiles; }
bool use_split_sensor(const hitx_plan* pl) {
  if (pl->sensor_form != 0) return pl->sensor_form == 2;
  return (long long)pl->N >= kSplitSensorMinNodes;
}

// Every raw-mask tile re-reducing the stage partials costs tiles x rows:
// above this many partial rows the finalize runs once in its own launch.
constexpr int kFoldMaxRows = 1024;

template <class T>
int run_sensor(hitx_plan* pl, const T* q, int enforce_first, uint8_t* mask_out, T* qT, const WS& w, hipStream_t s) {
  const Params<T> P = make_params<T>(pl);
  const dim3 g((pl->d.nx + kSTX - 1) / kSTX, (pl->d.ny + kSTY - 1) / kSTY);
  const bool walk = P.drag_on && stage_walk_form((long long)g.x * g.y);
  const int nblk = walk ? pl->nblk_sr : (int)(g.x * g.y), stride = P.drag_on ? R_DRAG + kDragCols : R_DRAG;
  // the split sensor's raw-mask tiles (and the drag rows) finalize the
  // reduction themselves while the partial rows are few; otherwise, and for
  // the fused sensor, the finalize has its own launch
  const bool split = use_split_sensor(pl);
  const bool fold = split && nblk <= kFoldMaxRows;
  if (P.drag_on) {
    if (walk)
      k_stage_reduce<T, true><<<pl->nblk_sr, kBlock, 0, s>>>(P, q, enforce_first, w.part);
    else
      k_stage_tile<T, true><<<g, kBlock, 0, s>>>(P, q, enforce_first, w.part);
    if (fold) {
      k_drows_fin<T><<<(2 * pl->d.ny + kBlock - 1) / kBlock, kBlock, 0, s>>>(P, w.part, nblk, w.scal, w.drows);
    } else {
      k_stage_finalize<T, true><<<1, kBlock, 0, s>>>(P, w.part, nblk, w.scal);
      k_drows<T><<<(2 * pl->d.ny + kBlock - 1) / kBlock, kBlock, 0, s>>>(P, w.scal, w.drows);
    }
  } else {
    k_stage_tile<T, false><<<g, kBlock, 0, s>>>(P, q, enforce_first, w.part);
    if (!fold) k_stage_finalize<T, false><<<1, kBlock, 0, s>>>(P, w.part, nblk, w.scal);
  }
  if (split) {
    const bool per = pl->d.periodic != 0;
    if (fold && per)
      k_rawtile_fin<T, true><<<g, kBlock, 0, s>>>(P, q, enforce_first, w.part, nblk, stride, w.scal, w.raw, qT);
    else if (fold)
      k_rawtile_fin<T, false><<<g, kBlock, 0, s>>>(P, q, enforce_first, w.part, nblk, stride, w.scal, w.raw, qT);
    else if (per)
      k_rawtile<T, true><<<g, kBlock, 0, s>>>(P, q, enforce_first, w.scal, w.raw, qT);
    else
      k_rawtile<T, false><<<g, kBlock, 0, s>>>(P, q, enforce_first, w.scal, w.raw, qT);
    const int ntiles = ((pl->d.nx + kDBCols - 1) / kDBCols) * ((pl->d.ny + kDBRows - 1) / kDBRows);
    k_dilate_bits<T><<<(ntiles + kDBWaves - 1) / kDBWaves, 64 * kDBWaves, 0, s>>>(P, w.raw, mask_out, w.maskT);
  } else {
    k_sensor<T><<<g, kBlock, pl->sensor_lds, s>>>(P, q, enforce_first, w.scal, mask_out, w.maskT, qT);
  }
  LAUNCHCHK();
  return HITX_OK;
}

// Event pair around one sweep launch when timing is on.
int tmark(hitx_plan* pl, hipStream_t s) {
  if (!pl->timing) return HITX_OK;
  if (pl->ev_used == pl->ev.size()) {
    hipEvent_t e;
    HIPCHK(hipEventCreate(&e));
    pl->ev.push_back(e);
  }
  HIPCHK(hipEventRecord(pl->ev[pl->ev_used++], s));
  return HITX_OK;
}

// both sweeps of the operator on state qs: dfx and dfyT in the workspace
size_t wave_slice_bytes(int n, int esize) { return (line_wave_lds(n, esize) + 15) & ~size_t(15); }
// The wave form needs many lines to fill the device (one wavefront each);
// with few lines the block form's 8 wavefronts per line finish sooner.
constexpr int kWaveFormMinLines = 2048;
bool use_wave_form(int form, int n, long long lines) {
  if (!line_wave_ok(n) || form == 1) return false;
  return form == 2 || lines >= kWaveFormMinLines;
}
bool use_wave(const hitx_plan* pl, int n) {
  const int nl = n == pl->d.nx ? pl->d.ny : pl->d.nx;
  return use_wave_form(pl->line_form, n, nl);
}

// The two line sweeps of a stage: x-lines of qs (enforced on load) and
// y-lines of the enforced transposed state.  Sweeps in the block form go in
// one launch (the x- and y-line blocks fill the device together); a sweep in
// the wave form has its own launch.
template <class T>
void launch_sweeps(const hitx_plan* pl, const Params<T>& P, const T* qs, const WS& w, hipStream_t s) {
  const int nx = pl->d.nx, ny = pl->d.ny;
  const bool wx = use_wave(pl, nx), wy = use_wave(pl, ny);
  const int sb = (int)wave_slice_bytes(std::max(nx, ny), pl->esize);
  for (int axis = 0; axis < 2; ++axis) {
    if (!(axis == 0 ? wx : wy)) continue;
    const int nl = axis == 0 ? ny : nx;
    const int nb = (nl + kLineWaves - 1) / kLineWaves;
    const T* in = axis == 0 ? qs : static_cast<const T*>(w.qT);
    const uint8_t* mk = axis == 0 ? w.mask : w.maskT;
    T* out = static_cast<T*>(axis == 0 ? w.dfx : w.dfyT);
    if (P.periodic)
      k_line_wave<T, true><<<nb, 64 * kLineWaves, kLineWaves * sb, s>>>(P, linek<T>(pl, axis), axis, axis, in, mk,
                                                                          w.scal, out, nl, sb);
    else
      k_line_wave<T, false><<<nb, 64 * kLineWaves, kLineWaves * sb, s>>>(P, linek<T>(pl, axis), axis, axis, in, mk,
                                                                           w.scal, out, nl, sb);
  }
  const int nlx = wx ? 0 : ny, nly = wy ? 0 : nx;
  if (nlx + nly == 0) return;
  const size_t lds = std::max(wx ? 0 : pl->line_lds_x, wy ? 0 : pl->line_lds_y);
  const T* qT = static_cast<const T*>(w.qT);
  T* fx = static_cast<T*>(w.dfx);
  T* fyT = static_cast<T*>(w.dfyT);
  const bool small = (wx || linek<T>(pl, 0).small) && (wy || linek<T>(pl, 1).small);
  if (P.periodic)
    launch_k_line<T, true>(small, lds, s, P, linek<T>(pl, 0), linek<T>(pl, 1), qs, w.mask, qT,
                           w.maskT, w.scal, fx, fyT, nlx, nly);
  else
    launch_k_line<T, false>(small, lds, s, P, linek<T>(pl, 0), linek<T>(pl, 1), qs, w.mask, qT,
                            w.maskT, w.scal, fx, fyT, nlx, nly);
}

template <class T>
int run_sweeps(hitx_plan* pl, const T* qs, const WS& w, hipStream_t s) {
  int rc = run_sensor<T>(pl, qs, 1, w.mask, static_cast<T*>(w.qT), w, s);
  if (rc) return rc;
  const Params<T> P = make_params<T>(pl);
  if ((rc = tmark(pl, s))) return rc;
  launch_sweeps<T>(pl, P, qs, w, s);
  if ((rc = tmark(pl, s))) return rc;
  LAUNCHCHK();
  return HITX_OK;
}

template <class T>
int set_line_lds(hitx_plan* pl) {
  const size_t mx = pl->line_lds_x > pl->line_lds_y ? pl->line_lds_x : pl->line_lds_y;
  int maxsh = 0;
  HIPCHK(hipDeviceGetAttribute(&maxsh, hipDeviceAttributeMaxSharedMemoryPerBlock, pl->device));
  if (mx > (size_t)maxsh)
    return fail(HITX_EINVAL, "grid line too long for LDS: need " + std::to_string(mx) + " bytes, device allows " +
                                 std::to_string(maxsh));
  if (mx > 65536) HIPCHK(allow_line_lds<T>(mx));
  return HITX_OK;
}

int validate_desc(const hitx_desc* d) {
  if (!d) return fail(HITX_EINVAL, "null descriptor");
  if (d->abi_version != HITX_ABI_VERSION) return fail(HITX_EINVAL, "descriptor ABI version mismatch");
  if (d->nx < 8 || d->ny < 8) return fail(HITX_EINVAL, "nx and ny must be >= 8");
  if ((int64_t)d->nx * d->ny * 4 > (int64_t)INT32_MAX) return fail(HITX_EINVAL, "grid too large");
  if (d->dtype != HITX_F64 && d->dtype != HITX_F32) return fail(HITX_EINVAL, "dtype must be HITX_F64 or HITX_F32");
  if (d->scheme != HITX_HYBRID && d->scheme != HITX_WENO) return fail(HITX_EINVAL, "scheme must be hybrid or weno");
  if (d->periodic && d->scheme == HITX_WENO) return fail(HITX_EINVAL, "weno-only scheme is non-periodic only");
  if (d->sensor_width < 0 || d->sensor_width > 8) return fail(HITX_EINVAL, "sensor_width must be in [0, 8]");
  if (!(d->dx > 0.0) || !(d->dy > 0.0)) return fail(HITX_EINVAL, "dx and dy must be positive");
  if (!d->periodic && (d->guard_cells < 0 || 2 * d->guard_cells >= d->nx || 2 * d->guard_cells >= d->ny))
    return fail(HITX_EINVAL, "guard_cells out of range");
  if (d->n_drag_modes < 0 || d->n_force_modes < 0 || d->n_drag_modes > 12 ||
      d->n_drag_modes + d->n_force_modes > HITX_MAX_MODES)
    return fail(HITX_EINVAL, "mode list too long (drag <= 12, total <= HITX_MAX_MODES)");
  if ((d->n_drag_modes && !d->drag_modes_host) || (d->n_force_modes && !d->force_modes_host))
    return fail(HITX_EINVAL, "mode list pointer is null");
  for (int m = 0; m < 2 * d->n_drag_modes; ++m)
    if (std::abs(d->drag_modes_host[m]) > 2) return fail(HITX_EINVAL, "drag mode index beyond 2");
  for (int m = 0; m < 2 * d->n_force_modes; ++m)  // projection LDS holds 2*|ky|+1 sums per lane
    if (std::abs(d->force_modes_host[m]) > 12) return fail(HITX_EINVAL, "forcing mode index beyond 12");
  return HITX_OK;
}

template <class T>
int rhs_or_stage(hitx_plan* pl, int mode, int stage, const void* q0, const void* qs, void* out, const void* fx,
                        const void* fy, double dt, void* ws, hipStream_t s) {
  const WS w = carve(pl, ws);
  int rc = run_sweeps<T>(pl, static_cast<const T*>(qs), w, s);
  if (rc) return rc;
  const Params<T> P = make_params<T>(pl);
  k_update<T><<<grid_update(pl), kUBlock, 0, s>>>(P, mode, stage, static_cast<const T*>(q0), static_cast<const T*>(qs),
                                            static_cast<const T*>(w.dfx), static_cast<const T*>(w.dfyT),
                                            static_cast<const T*>(fx), static_cast<const T*>(fy), w.scal, w.drows,
                                            (T)dt, static_cast<T*>(out));
  LAUNCHCHK();
  return HITX_OK;
}

}  // namespace

extern "C" {

const char* hitx_last_error(void) { return g_err.c_str(); }
int hitx_abi_version(void) { return HITX_ABI_VERSION; }

int hitx_plan_create(const hitx_desc* desc, hitx_plan** out) {
  if (!out) return fail(HITX_EINVAL, "null output pointer");
  *out = nullptr;
  int rc = validate_desc(desc);
  if (rc) return rc;
  hitx_plan* pl = new hitx_plan();
  pl->d = *desc;
  pl->d.drag_modes_host = nullptr;
  pl->d.force_modes_host = nullptr;
  pl->esize = desc->dtype == HITX_F64 ? 8 : 4;
  pl->N = desc->nx * desc->ny;
  if (hipGetDevice(&pl->device) != hipSuccess) {
    delete pl;
    return fail(HITX_EDEVICE, "hipGetDevice failed");
  }
  // mode lists and twiddle tables
  int km_x = 0, km_y = 0;
  for (int m = 0; m < desc->n_drag_modes; ++m) {
    pl->drag_modes.push_back(desc->drag_modes_host[2 * m]);
    pl->drag_modes.push_back(desc->drag_modes_host[2 * m + 1]);
  }
  for (int m = 0; m < desc->n_force_modes; ++m) {
    pl->force_modes.push_back(desc->force_modes_host[2 * m]);
    pl->force_modes.push_back(desc->force_modes_host[2 * m + 1]);
  }
  std::vector<int32_t> all(pl->drag_modes);
  all.insert(all.end(), pl->force_modes.begin(), pl->force_modes.end());
  for (size_t i = 0; i < all.size(); i += 2) {
    km_x = std::max(km_x, std::abs(all[i]));
    km_y = std::max(km_y, std::abs(all[i + 1]));
  }
  for (size_t i = 0; i < pl->drag_modes.size(); i += 2) {
    pl->dkx = std::max(pl->dkx, std::abs(pl->drag_modes[i]));
    pl->dky = std::max(pl->dky, std::abs(pl->drag_modes[i + 1]));
  }
  for (size_t i = 0; i < pl->force_modes.size(); i += 2) {
    pl->fkx = std::max(pl->fkx, std::abs(pl->force_modes[i]));
    pl->fky = std::max(pl->fky, std::abs(pl->force_modes[i + 1]));
  }
  if (km_x > 16 || km_y > 16) {
    delete pl;
    return fail(HITX_EINVAL, "mode index beyond 16");
  }
  pl->kmx = km_x;
  pl->kmy = km_y;
  // forcing modes grouped by |kx| (ascending mode index within a group), then
  // the group offsets: the per-row synthesis table visits only its own modes
  {
    const int nf = desc->n_force_modes;
    std::vector<int32_t> order, ofs(1, 0);
    for (int k = 0; k <= pl->fkx; ++k) {
      for (int m = 0; m < nf; ++m)
        if (std::abs(pl->force_modes[2 * m]) == k) order.push_back(m);
      ofs.push_back((int32_t)order.size());
    }
    all.insert(all.end(), order.begin(), order.end());
    all.insert(all.end(), ofs.begin(), ofs.end());
  }
  if (desc->n_drag_modes + desc->n_force_modes > 0) {
    const int nx = desc->nx, ny = desc->ny;
    std::vector<double> tw((size_t)(2 * km_x + 1) * nx * 2 + (size_t)(2 * km_y + 1) * ny * 2);
    const double two_pi = 2.0 * M_PI;
    size_t o = 0;
    for (int k = -km_x; k <= km_x; ++k)
      for (int i = 0; i < nx; ++i) {
        const long r = (((long)k * i) % nx + nx) % nx;  // exact argument reduction
        tw[o++] = std::cos(two_pi * (double)r / nx);
        tw[o++] = std::sin(two_pi * (double)r / nx);
      }
    for (int k = -km_y; k <= km_y; ++k)
      for (int i = 0; i < ny; ++i) {
        const long r = (((long)k * i) % ny + ny) % ny;
        tw[o++] = std::cos(two_pi * (double)r / ny);
        tw[o++] = std::sin(two_pi * (double)r / ny);
      }
    if (hipMalloc(&pl->d_tw, tw.size() * sizeof(double)) != hipSuccess ||
        hipMemcpy(pl->d_tw, tw.data(), tw.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
        hipMalloc(&pl->d_modes, all.size() * sizeof(int)) != hipSuccess ||
        hipMemcpy(pl->d_modes, all.data(), all.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess) {
      hitx_plan_destroy(pl);
      return fail(HITX_EDEVICE, "allocating plan tables failed");
    }
  }
  if (desc->periodic) {
    std::vector<double> dk = spectral_derivative_table(desc->nx, desc->dx);
    const std::vector<double> dky = spectral_derivative_table(desc->ny, desc->dy);
    dk.insert(dk.end(), dky.begin(), dky.end());
    if (hipMalloc(&pl->d_dk, dk.size() * sizeof(double)) != hipSuccess ||
        hipMemcpy(pl->d_dk, dk.data(), dk.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) {
      hitx_plan_destroy(pl);
      return fail(HITX_EDEVICE, "allocating plan tables failed");
    }
  }
  pl->lx_d = make_linek<double>(desc->nx);
  pl->ly_d = make_linek<double>(desc->ny);
  pl->lx_f = make_linek<float>(desc->nx);
  pl->ly_f = make_linek<float>(desc->ny);
  pl->line_lds_x = line_lds_bytes(desc->nx, pl->lx_d.nch, pl->lx_d.small != 0, pl->esize);
  pl->line_lds_y = line_lds_bytes(desc->ny, pl->ly_d.nch, pl->ly_d.small != 0, pl->esize);
  pl->sensor_lds = sensor_lds_bytes(desc->sensor_width, pl->esize);
  rc = pl->esize == 8 ? set_line_lds<double>(pl) : set_line_lds<float>(pl);
  if (rc) {
    std::string keep = g_err;
    hitx_plan_destroy(pl);
    g_err = keep;
    return rc;
  }
  // workspace layout
  const size_t N = pl->N, E = pl->esize;
  pl->nblk_red = (int)std::min<size_t>((N + kBlock - 1) / kBlock, 1024);
  pl->nblk_stage = (int)std::min<size_t>((N + kBlock - 1) / kBlock, 512);  // grid-stride reductions
  // column walks: rows per wave so that a plan alone still gives ~2048 waves
  pl->rb = default_rb(desc->nx, desc->ny);
  pl->rbp = default_rbp(desc->nx, desc->ny);
  pl->nblk_sr = walk_blocks(desc->nx, desc->ny, pl->rb, 1);
  pl->nblk_proj = walk_blocks(desc->nx, desc->ny, pl->rbp, 0);
  size_t o = 0;
  pl->off_raw = o; o = align256(o + N);
  pl->off_mask = o; o = align256(o + N);
  pl->off_maskT = o; o = align256(o + N);
  pl->off_qT = o; o = align256(o + 4 * N * E);
  pl->off_dfx = o; o = align256(o + 4 * N * E);
  pl->off_dfyT = o; o = align256(o + 4 * N * E);
  const size_t ntiles = (size_t)((desc->nx + kSTX - 1) / kSTX) * ((desc->ny + kSTY - 1) / kSTY);
  const size_t part = std::max({(size_t)pl->nblk_red * 64, (size_t)pl->nblk_sr * 64,
                                (size_t)pl->nblk_proj * 2 * std::max(1, desc->n_force_modes), ntiles * (R_DRAG + kDragCols)});
  pl->off_part = o; o = align256(o + part * sizeof(double));
  pl->off_scal = o; o = align256(o + S_COUNT * sizeof(double));
  pl->off_misc = o; o = align256(o + 64 * sizeof(double));
  // per-row drag table: 10 doubles a row whenever there are drag modes (also
  // when the only mode is k = 0)
  pl->off_drows = o; o = align256(o + (desc->n_drag_modes > 0 ? (size_t)desc->ny * 10 : 1) * sizeof(double));
  pl->off_frows = o; o = align256(o + (desc->n_force_modes > 0 ? (size_t)desc->ny * 2 * (2 * pl->fkx + 1) : 1) *
                                          sizeof(double));
  pl->ws_bytes = o;
  *out = pl;
  return HITX_OK;
}

int hitx_plan_destroy(hitx_plan* pl) {
  if (!pl) return HITX_OK;
  for (hipEvent_t e : pl->ev) (void)hipEventDestroy(e);
  if (pl->d_tw) (void)hipFree(pl->d_tw);
  if (pl->d_modes) (void)hipFree(pl->d_modes);
  if (pl->d_dk) (void)hipFree(pl->d_dk);
  delete pl;
  return HITX_OK;
}

size_t hitx_workspace_bytes(const hitx_plan* pl) { return pl ? pl->ws_bytes : 0; }

int hitx_sensor(hitx_plan* pl, const void* q, int32_t enforce_first, uint8_t* mask_out, void* ws, size_t ws_bytes,
                void* stream) {
  int rc = check_ws(pl, ws, ws_bytes);
  if (rc) return rc;
  if (!q || !mask_out) return fail(HITX_EINVAL, "null array");
  const WS w = carve(pl, ws);
  hipStream_t s = static_cast<hipStream_t>(stream);
  return pl->esize == 8
             ? run_sensor<double>(pl, static_cast<const double*>(q), enforce_first, mask_out, nullptr, w, s)
             : run_sensor<float>(pl, static_cast<const float*>(q), enforce_first, mask_out, nullptr, w, s);
}

int hitx_rhs(hitx_plan* pl, const void* q, void* dq, void* ws, size_t ws_bytes, void* stream) {
  int rc = check_ws(pl, ws, ws_bytes);
  if (rc) return rc;
  if (!q || !dq) return fail(HITX_EINVAL, "null array");
  hipStream_t s = static_cast<hipStream_t>(stream);
  return pl->esize == 8 ? rhs_or_stage<double>(pl, MODE_RHS, 0, q, q, dq, nullptr, nullptr, 0.0, ws, s)
                        : rhs_or_stage<float>(pl, MODE_RHS, 0, q, q, dq, nullptr, nullptr, 0.0, ws, s);
}

int hitx_rk3_stage(hitx_plan* pl, int32_t stage, const void* q0, const void* qs, void* out, const void* fx,
                   const void* fy, double dt, void* ws, size_t ws_bytes, void* stream) {
  int rc = check_ws(pl, ws, ws_bytes);
  if (rc) return rc;
  if (stage < 0 || stage > 2) return fail(HITX_EINVAL, "stage must be 0, 1 or 2");
  if (!q0 || !qs || !out) return fail(HITX_EINVAL, "null array");
  if (out == q0 || out == qs) return fail(HITX_EINVAL, "out must not alias q0 or qs");
  if (pl->d.periodic && pl->d.forcing_enabled && (!fx || !fy)) return fail(HITX_EINVAL, "forcing enabled: fx, fy required");
  hipStream_t s = static_cast<hipStream_t>(stream);
  return pl->esize == 8 ? rhs_or_stage<double>(pl, MODE_STAGE, stage, q0, qs, out, fx, fy, dt, ws, s)
                        : rhs_or_stage<float>(pl, MODE_STAGE, stage, q0, qs, out, fx, fy, dt, ws, s);
}

}  // extern "C"

// ------------------------------------------------- stage without allocation
// hitx_rk3_stage_static runs hitx_rk3_stage with a file-scope device array in
// place of the caller's workspace.  In host code `scratch` names a host-side
// stand-in, not device memory: the array's device address comes from
// hipGetSymbolAddress, taken once per device on the entry point's first call
// there (with that device current) and kept per device.  Every call on a
// device shares that device's array; the stage writes every workspace byte it
// reads within the call (stage partials, mask, transposed state, sweep
// results), so nothing must start at zero.
constexpr size_t kScratchBytes = size_t(32) << 20;
static __device__ __align__(256) uint8_t scratch[kScratchBytes];

namespace {

constexpr int kScratchMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kScratchMaxDevices] = {};

int scratch_address(int dev, void** out) {
  if (dev < 0 || dev >= kScratchMaxDevices) return fail(HITX_EDEVICE, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(HITX_EDEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return HITX_OK;
}

}  // namespace

extern "C" {

int hitx_rk3_stage_static(hitx_plan* pl, int32_t stage, const void* q0, const void* qs, void* out, const void* fx,
                          const void* fy, double dt, void* stream) {
  if (!pl) return fail(HITX_EINVAL, "null plan");
  if (pl->ws_bytes > kScratchBytes)
    return fail(HITX_EINVAL, "grid needs " + std::to_string(pl->ws_bytes) + " bytes of workspace; the static scratch "
                             "holds " + std::to_string(kScratchBytes) + " (use hitx_rk3_stage with a workspace)");
  if (pl->timing) return fail(HITX_EINVAL, "sweep timing creates events: turn it off for hitx_rk3_stage_static");
  int dev = -1;
  HIPCHK(hipGetDevice(&dev));
  if (dev != pl->device)
    return fail(HITX_EINVAL, "plan belongs to device " + std::to_string(pl->device) + ", current device is " +
                             std::to_string(dev));
  void* ws = nullptr;
  const int rc = scratch_address(dev, &ws);
  if (rc) return rc;
  return hitx_rk3_stage(pl, stage, q0, qs, out, fx, fy, dt, ws, kScratchBytes, stream);
}

int hitx_hyperviscosity(hitx_plan* pl, const void* q, const uint8_t* active, int32_t invert, void* out,
                        int32_t apply_guard, void* ws, size_t ws_bytes, void* stream) {
  int rc = check_ws(pl, ws, ws_bytes);
  if (rc) return rc;
  if (!q || !active || !out) return fail(HITX_EINVAL, "null array");
  if (out == q) return fail(HITX_EINVAL, "out must not alias q");
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (pl->esize == 8)
    k_hv<double><<<grid16(pl), kBlock, 0, s>>>(make_params<double>(pl), static_cast<const double*>(q), active,
                                              apply_guard, invert, static_cast<double*>(out));
  else
    k_hv<float><<<grid16(pl), kBlock, 0, s>>>(make_params<float>(pl), static_cast<const float*>(q), active,
                                             apply_guard, invert, static_cast<float*>(out));
  LAUNCHCHK();
  return HITX_OK;
}

int hitx_enforce(hitx_plan* pl, const void* q, void* out, int32_t apply_guard, void* stream) {
  if (!pl) return fail(HITX_EINVAL, "null plan");
  if (!q || !out) return fail(HITX_EINVAL, "null array");
  if (out == q && apply_guard && pl->d.guard_cells > 0 && !pl->d.periodic)
    return fail(HITX_EINVAL, "guarded enforce must not run in place");
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (pl->esize == 8)
    k_enforce<double><<<grid16(pl), kBlock, 0, s>>>(make_params<double>(pl), static_cast<const double*>(q),
                                                   apply_guard, static_cast<double*>(out));
  else
    k_enforce<float><<<grid16(pl), kBlock, 0, s>>>(make_params<float>(pl), static_cast<const float*>(q), apply_guard,
                                                  static_cast<float*>(out));
  LAUNCHCHK();
  return HITX_OK;
}

int hitx_step_reduce(hitx_plan* pl, const void* q, double* out_dev, void* ws, size_t ws_bytes, void* stream) {
  int rc = check_ws(pl, ws, ws_bytes);
  if (rc) return rc;
  if (!q || !out_dev) return fail(HITX_EINVAL, "null array");
  const WS w = carve(pl, ws);
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (pl->esize == 8)
    k_step_reduce<double><<<pl->nblk_red, kBlock, 0, s>>>(make_params<double>(pl), static_cast<const double*>(q),
                                                          w.part);
  else
    k_step_reduce<float><<<pl->nblk_red, kBlock, 0, s>>>(make_params<float>(pl), static_cast<const float*>(q),
                                                         w.part);
  k_finalize_max0<<<3, kBlock, 0, s>>>(w.part, pl->nblk_red, 3, out_dev);
  LAUNCHCHK();
  return HITX_OK;
}

int hitx_forcing_project(hitx_plan* pl, const double* noise, double* modes_out, void* ws, size_t ws_bytes,
                         void* stream) {
  int rc = check_ws(pl, ws, ws_bytes);
  if (rc) return rc;
  if (!noise || !modes_out) return fail(HITX_EINVAL, "null array");
  if (pl->d.n_force_modes <= 0) return fail(HITX_EINVAL, "plan has no forcing modes");
  const WS w = carve(pl, ws);
  hipStream_t s = static_cast<hipStream_t>(stream);
  const size_t lds = proj_lds_bytes(pl);
  if (pl->esize == 8)
    k_fproject<double><<<pl->nblk_proj, kBlock, lds, s>>>(make_params<double>(pl), noise, w.part);
  else
    k_fproject<float><<<pl->nblk_proj, kBlock, lds, s>>>(make_params<float>(pl), noise, w.part);
  k_finalize_sum<<<2 * pl->d.n_force_modes, kBlock, 0, s>>>(w.part, pl->nblk_proj, 2 * pl->d.n_force_modes,
                                                            modes_out);
  LAUNCHCHK();
  return HITX_OK;
}

int hitx_forcing_synthesize(hitx_plan* pl, const double* fhat, void* fx, void* fy, const void* rho, const void* u,
                            const void* v, double* stats_out, void* ws, size_t ws_bytes, void* stream) {
  int rc = check_ws(pl, ws, ws_bytes);
  if (rc) return rc;
  if (!fhat || !fx || !fy || !rho || !u || !v || !stats_out) return fail(HITX_EINVAL, "null array");
  if (pl->d.n_force_modes <= 0) return fail(HITX_EINVAL, "plan has no forcing modes");
  const WS w = carve(pl, ws);
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int nbr = (pl->d.ny * (pl->fkx + 1) + kBlock - 1) / kBlock;
  if (pl->esize == 8)
    k_frows<double><<<nbr, kBlock, 0, s>>>(make_params<double>(pl), fhat, w.frows);
  else
    k_frows<float><<<nbr, kBlock, 0, s>>>(make_params<float>(pl), fhat, w.frows);
  if (pl->esize == 8)
    k_fsynth<double><<<pl->nblk_red, kBlock, 0, s>>>(make_params<double>(pl), w.frows, static_cast<double*>(fx),
                                                     static_cast<double*>(fy), static_cast<const double*>(rho),
                                                     static_cast<const double*>(u), static_cast<const double*>(v),
                                                     w.part);
  else
    k_fsynth<float><<<pl->nblk_red, kBlock, 0, s>>>(make_params<float>(pl), w.frows, static_cast<float*>(fx),
                                                    static_cast<float*>(fy), static_cast<const float*>(rho),
                                                    static_cast<const float*>(u), static_cast<const float*>(v), w.part);
  k_finalize_sum<<<2, kBlock, 0, s>>>(w.part, pl->nblk_red, 2, stats_out);
  LAUNCHCHK();
  return HITX_OK;
}

int hitx_forcing_scale(hitx_plan* pl, void* fx, void* fy, double scale, double alpha, const void* rho, const void* u,
                       const void* v, double* stats_out, void* ws, size_t ws_bytes, void* stream) {
  int rc = check_ws(pl, ws, ws_bytes);
  if (rc) return rc;
  if (!fx || !fy || !rho || !u || !v || !stats_out) return fail(HITX_EINVAL, "null array");
  const WS w = carve(pl, ws);
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (pl->esize == 8)
    k_fscale<double><<<pl->nblk_red, kBlock, 0, s>>>(make_params<double>(pl), static_cast<double*>(fx),
                                                     static_cast<double*>(fy), scale, alpha,
                                                     static_cast<const double*>(rho), static_cast<const double*>(u),
                                                     static_cast<const double*>(v), w.part);
  else
    k_fscale<float><<<pl->nblk_red, kBlock, 0, s>>>(make_params<float>(pl), static_cast<float*>(fx),
                                                    static_cast<float*>(fy), (float)scale, (float)alpha,
                                                    static_cast<const float*>(rho), static_cast<const float*>(u),
                                                    static_cast<const float*>(v), w.part);
  k_finalize_sum<<<4, kBlock, 0, s>>>(w.part, pl->nblk_red, 4, stats_out);
  LAUNCHCHK();
  return HITX_OK;
}

int hitx_primitives(hitx_plan* pl, const void* q, void* rho, void* u, void* v, void* p, void* stream) {
  if (!pl) return fail(HITX_EINVAL, "null plan");
  if (!q || !rho || !u || !v || !p) return fail(HITX_EINVAL, "null array");
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (pl->esize == 8)
    k_primitives<double><<<pl->nblk_red, kBlock, 0, s>>>(make_params<double>(pl), static_cast<const double*>(q),
                                                         static_cast<double*>(rho), static_cast<double*>(u),
                                                         static_cast<double*>(v), static_cast<double*>(p));
  else
    k_primitives<float><<<pl->nblk_red, kBlock, 0, s>>>(make_params<float>(pl), static_cast<const float*>(q),
                                                        static_cast<float*>(rho), static_cast<float*>(u),
                                                        static_cast<float*>(v), static_cast<float*>(p));
  LAUNCHCHK();
  return HITX_OK;
}

size_t hitx_diag_workspace_bytes(const hitx_plan* pl, int32_t with_fields) {
  if (!pl) return 0;
  const size_t N = pl->N, ny = pl->d.ny;
  size_t o = with_fields ? 0 : align256(6 * N * sizeof(double));
  o += align256(ny * kDiagRowW * sizeof(double));
  o += align256(ny * kDiagMomW * sizeof(double));
  o += align256((size_t)kDiagF * kDiagMaxModes * 2 * sizeof(double));
  return o;
}

int hitx_diagnostics(hitx_plan* pl, const void* q, const uint8_t* mask, const int32_t* modes_host,
                     int32_t n_modes, double drag_coef, double* fields, double* out, void* ws, size_t ws_bytes,
                     void* stream) {
  if (!pl) return fail(HITX_EINVAL, "null plan");
  if (!q || !out) return fail(HITX_EINVAL, "null array");
  if (!pl->d.periodic || !pl->d_dk) return fail(HITX_EINVAL, "diagnostics are defined on the periodic grid only");
  if (n_modes < 0 || n_modes > kDiagMaxModes)
    return fail(HITX_EINVAL, "n_lowpass must be in 0.." + std::to_string(kDiagMaxModes));
  if (n_modes > 0 && !modes_host) return fail(HITX_EINVAL, "null low-pass mode list");
  const size_t need = hitx_diag_workspace_bytes(pl, fields != nullptr);
  if (!ws || ws_bytes < need) return fail(HITX_ENOMEM, "workspace too small: need " + std::to_string(need) + " bytes");
  DiagModes M{};
  M.n = n_modes;
  for (int m = 0; m < n_modes; ++m) {
    const int kx = modes_host[2 * m], ky = modes_host[2 * m + 1];
    if (std::abs(kx) > kDiagMaxK || std::abs(ky) > kDiagMaxK)
      return fail(HITX_EINVAL, "low-pass mode index beyond " + std::to_string(kDiagMaxK));
    if (kx == 0 && ky == 0) return fail(HITX_EINVAL, "the low-pass band excludes k = 0");
    M.kx[m] = (int8_t)kx;
    M.ky[m] = (int8_t)ky;
    M.kxmax = std::max(M.kxmax, std::abs(kx));
    M.kymax = std::max(M.kymax, std::abs(ky));
  }
  const int nx = pl->d.nx, ny = pl->d.ny;
  const size_t N = pl->N;
  uint8_t* b = static_cast<uint8_t*>(ws);
  double* F = fields;
  if (!F) {
    F = reinterpret_cast<double*>(b);
    b += align256(6 * N * sizeof(double));
  }
  double* rowp = reinterpret_cast<double*>(b);
  b += align256((size_t)ny * kDiagRowW * sizeof(double));
  double* rowm = reinterpret_cast<double*>(b);
  b += align256((size_t)ny * kDiagMomW * sizeof(double));
  double* modes = reinterpret_cast<double*>(b);
  double *rho = F, *u = F + N, *v = F + 2 * N, *p = F + 3 * N, *dvg = F + 4 * N, *vort = F + 5 * N;
  hipStream_t s = static_cast<hipStream_t>(stream);
  const Params<double> P = make_params<double>(pl);
  if (pl->esize == 8)
    k_diag_fields<double><<<ny, kBlock, 0, s>>>(P, static_cast<const double*>(q), rho, u, v, p, rowp, M.kxmax);
  else
    k_diag_fields<float><<<ny, kBlock, 0, s>>>(P, static_cast<const float*>(q), rho, u, v, p, rowp, M.kxmax);
  k_finalize_sum<<<kDiagS, kBlock, 0, s>>>(rowp, ny, kDiagRowW, out);
  if (M.n > 0) k_diag_modes<<<kDiagF * M.n, kBlock, 0, s>>>(rowp, ny, M, modes);
  const double* dkx = pl->d_dk;
  const double* dky = pl->d_dk + 2 * (size_t)nx + 2 * kDiagPad;
  k_diag_deriv<0, false><<<dim3((ny + kDerLines - 1) / kDerLines, (nx + kDerPos - 1) / kDerPos), kBlock, 0, s>>>(
      nx, ny, dkx, u, v, dvg, vort);
  k_diag_deriv<1, true><<<dim3((nx + kDerLines - 1) / kDerLines, (ny + kDerPos - 1) / kDerPos), kBlock, 0, s>>>(
      nx, ny, dky, v, u, dvg, vort);
  k_diag_moments<<<ny, kBlock, 0, s>>>(nx, ny, rho, u, v, dvg, vort, mask, out, modes, M, -drag_coef, rowm);
  k_finalize_sum<<<kDiagMomW, kBlock, 0, s>>>(rowm, ny, kDiagMomW, out + kDiagS);
  LAUNCHCHK();
  return HITX_OK;
}

int hitx_hit_initial_state(hitx_plan* pl, const double* modes_host, int32_t n_modes, double target_mach,
                           void* q_out, double* out_dev, void* ws, size_t ws_bytes, void* stream) {
  if (!pl) return fail(HITX_EINVAL, "null plan");
  if (!q_out || !out_dev) return fail(HITX_EINVAL, "null array");
  if (!pl->d.periodic || !pl->d_dk) return fail(HITX_EINVAL, "the HIT initial state needs a periodic plan");
  if (n_modes < 1 || !modes_host) return fail(HITX_EINVAL, "no streamfunction modes");
  const size_t N = pl->N;
  if ((size_t)n_modes * kInitModeW > N)
    return fail(HITX_EINVAL, "more streamfunction modes than nx*ny/4");
  const size_t need = hitx_diag_workspace_bytes(pl, 0);
  if (!ws || ws_bytes < need) return fail(HITX_ENOMEM, "workspace too small: need " + std::to_string(need) + " bytes");
  const int nx = pl->d.nx, ny = pl->d.ny;
  double* F = static_cast<double*>(ws);
  // planes: psi, d/dx, d/dy, discard, block partials + sums (3 nblk + 4 <= N
  // doubles since nblk <= N/256 + 1 and N >= 64), mode table
  double *psi = F, *dpx = F + N, *dpy = F + 2 * N, *dmy = F + 3 * N, *modes = F + 5 * N;
  double* part = F + 4 * N;
  double* sums = part + (size_t)3 * pl->nblk_red;
  hipStream_t s = static_cast<hipStream_t>(stream);
  HIPCHK(hipMemcpyAsync(modes, modes_host, (size_t)n_modes * kInitModeW * sizeof(double), hipMemcpyHostToDevice, s));
  const hitx_desc& d = pl->d;
  const double x0 = 0.0, y0 = 0.0;  // HIT domain origin (OOP/hit2d.py:114-119)
  k_init_psi<<<pl->nblk_red, kBlock, 0, s>>>(nx, ny, d.dx, d.dy, x0, y0, modes, n_modes, psi);
  const double* dkx = pl->d_dk;
  const double* dky = pl->d_dk + 2 * (size_t)nx + 2 * kDiagPad;
  k_diag_deriv<0, false><<<dim3((ny + kDerLines - 1) / kDerLines, (nx + kDerPos - 1) / kDerPos), kBlock, 0, s>>>(
      nx, ny, dkx, psi, psi, dpx, dmy);
  k_diag_deriv<1, false><<<dim3((nx + kDerLines - 1) / kDerLines, (ny + kDerPos - 1) / kDerPos), kBlock, 0, s>>>(
      nx, ny, dky, psi, psi, dpy, dmy);
  const double p0 = 1.0 / d.gamma;
  const double c0 = std::sqrt(d.gamma * p0 / 1.0);
  k_init_sums<<<pl->nblk_red, kBlock, 0, s>>>((int)N, dpx, dpy, c0, part);
  k_finalize_sum<<<3, kBlock, 0, s>>>(part, pl->nblk_red, 3, sums);
  k_init_rms<<<pl->nblk_red, kBlock, 0, s>>>((int)N, dpx, dpy, sums, part);
  k_finalize_sum<<<1, kBlock, 0, s>>>(part, pl->nblk_red, 1, sums + 3);
  k_init_scale<<<1, 64, 0, s>>>((int)N, sums, target_mach, out_dev);
  const double rho0 = std::max(1.0, d.rho_floor), pf = std::max(p0, d.p_floor);
  if (pl->esize == 8)
    k_init_hit_write<double><<<pl->nblk_red, kBlock, 0, s>>>((int)N, dpx, dpy, sums, out_dev, d.gamma, rho0, pf,
                                                             static_cast<double*>(q_out));
  else
    k_init_hit_write<float><<<pl->nblk_red, kBlock, 0, s>>>((int)N, dpx, dpy, sums, out_dev, d.gamma, rho0, pf,
                                                            static_cast<float*>(q_out));
  LAUNCHCHK();
  return HITX_OK;
}

int hitx_riemann_initial_state(hitx_plan* pl, const double* quadrants_host, double x_min, double y_min,
                               double x_split, double y_split, void* q_out, void* stream) {
  if (!pl) return fail(HITX_EINVAL, "null plan");
  if (!q_out || !quadrants_host) return fail(HITX_EINVAL, "null array");
  Quadrants Qd;
  for (int k = 0; k < 4; ++k)
    for (int c = 0; c < 4; ++c) Qd.s[k][c] = quadrants_host[4 * k + c];
  const hitx_desc& d = pl->d;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (pl->esize == 8)
    k_init_riemann<double><<<pl->nblk_red, kBlock, 0, s>>>(d.nx, d.ny, d.dx, d.dy, x_min, y_min, x_split, y_split, Qd,
                                                           d.gamma, d.rho_floor, d.p_floor,
                                                           static_cast<double*>(q_out));
  else
    k_init_riemann<float><<<pl->nblk_red, kBlock, 0, s>>>(d.nx, d.ny, d.dx, d.dy, x_min, y_min, x_split, y_split, Qd,
                                                          d.gamma, d.rho_floor, d.p_floor,
                                                          static_cast<float*>(q_out));
  LAUNCHCHK();
  return HITX_OK;
}

int hitx_riemann_diagnostics(hitx_plan* pl, const void* q, const uint8_t* mask, double* fields, double* out,
                             void* ws, size_t ws_bytes, void* stream) {
  int rc = check_ws(pl, ws, ws_bytes);
  if (rc) return rc;
  if (!q || !out) return fail(HITX_EINVAL, "null array");
  const WS w = carve(pl, ws);
  hipStream_t s = static_cast<hipStream_t>(stream);
  const Params<double> P = make_params<double>(pl);
  if (pl->esize == 8)
    k_riemann_diag<double><<<pl->nblk_red, kBlock, 0, s>>>(P, static_cast<const double*>(q), mask, fields, w.part);
  else
    k_riemann_diag<float><<<pl->nblk_red, kBlock, 0, s>>>(P, static_cast<const float*>(q), mask, fields, w.part);
  k_riemann_diag_finalize<<<1, kBlock, 0, s>>>(w.part, pl->nblk_red, out);
  LAUNCHCHK();
  return HITX_OK;
}

int hitx_viscous_rhs(hitx_plan* pl, const void* q, void* out, void* stream) {
  if (!pl) return fail(HITX_EINVAL, "null plan");
  if (!q || !out) return fail(HITX_EINVAL, "null array");
  if (!pl->d.periodic) return fail(HITX_EINVAL, "viscous terms are defined on the periodic grid only");
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (pl->esize == 8) {
    Params<double> P = make_params<double>(pl);
    P.visc_on = 1;
    k_update<double><<<grid_update(pl), kUBlock, 0, s>>>(P, MODE_VISC, 0, nullptr, static_cast<const double*>(q), nullptr,
                                                   nullptr, nullptr, nullptr, nullptr, nullptr, 0.0,
                                                   static_cast<double*>(out));
  } else {
    Params<float> P = make_params<float>(pl);
    P.visc_on = 1;
    k_update<float><<<grid_update(pl), kUBlock, 0, s>>>(P, MODE_VISC, 0, nullptr, static_cast<const float*>(q), nullptr,
                                                  nullptr, nullptr, nullptr, nullptr, nullptr, 0.0f,
                                                  static_cast<float*>(out));
  }
  LAUNCHCHK();
  return HITX_OK;
}

// ---------------------------------------------------------------- post-processing
}  // extern "C"

namespace {

constexpr int kPostMaxN = 2048;  // line length of the LDS-staged transforms
constexpr int kPostRedBlocks = 256;

struct PostWS {
  double *twx, *twy, *mid, *part, *means, *hist;
};
// Workspace of the post-processing calls on an nx x ny grid with `planes`
// complex intermediate planes: the two twiddle tables, the planes, the mean
// partials, the shell row partials.
size_t post_layout(int nx, int ny, int planes, void* base, PostWS* w) {
  size_t o = 0;
  auto take = [&](size_t bytes) {
    const size_t r = o;
    o = align256(o + bytes);
    return r;
  };
  const size_t tx = take(2 * (size_t)nx * 8), ty = take(2 * (size_t)ny * 8);
  const size_t mid = take((size_t)planes * nx * ny * 16);
  const size_t part = take((size_t)kPostRedBlocks * PM_COUNT * 8), means = take(PM_COUNT * 8);
  const size_t hist = take((size_t)ny * 2 * (std::min(nx, ny) / 2 + 1) * 8);
  if (w && base) {
    uint8_t* b = static_cast<uint8_t*>(base);
    w->twx = reinterpret_cast<double*>(b + tx);
    w->twy = reinterpret_cast<double*>(b + ty);
    w->mid = reinterpret_cast<double*>(b + mid);
    w->part = reinterpret_cast<double*>(b + part);
    w->means = reinterpret_cast<double*>(b + means);
    w->hist = reinterpret_cast<double*>(b + hist);
  }
  return o;
}

int post_check(int nx, int ny, int planes, void* ws, size_t ws_bytes) {
  if (nx < 2 || ny < 2 || nx > kPostMaxN || ny > kPostMaxN)
    return fail(HITX_EINVAL, "post-processing grids are 2..2048 nodes a side");
  if (planes < 1) return fail(HITX_EINVAL, "at least one plane");
  if (!ws || ws_bytes < post_layout(nx, ny, planes, nullptr, nullptr))
    return fail(HITX_ENOMEM, "post-processing workspace too small: need " +
                                 std::to_string(post_layout(nx, ny, planes, nullptr, nullptr)) + " bytes");
  return HITX_OK;
}

// 2-D transform of count planes (in: real or complex, out: complex) through
// the intermediate planes w.mid (count of them).
int post_dft2(int nx, int ny, int count, const double* in, bool in_complex, bool inverse, double* out,
              const PostWS& w, hipStream_t s) {
  k_dft_table<<<(nx + kPostBlock - 1) / kPostBlock, kPostBlock, 0, s>>>(nx, w.twx);
  k_dft_table<<<(ny + kPostBlock - 1) / kPostBlock, kPostBlock, 0, s>>>(ny, w.twy);
  const int sign = inverse ? 1 : -1;
  const double scale = inverse ? 1.0 / ((double)nx * (double)ny) : 1.0;
  const size_t N = (size_t)nx * ny;
  // rows (lines along x) of every plane: line r = plane * ny + y
  if (in_complex)
    k_dft_lines<false><<<count * ny, kDftThreads, 32 * (size_t)nx, s>>>(nx, count * ny, in, 2 * (long long)nx, 1,
                                                                        w.mid, 2 * (long long)nx, 1, w.twx, sign, 1.0);
  else
    k_dft_lines<true><<<count * ny, kDftThreads, 32 * (size_t)nx, s>>>(nx, count * ny, in, nx, 1, w.mid,
                                                                       2 * (long long)nx, 1, w.twx, sign, 1.0);
  // columns (lines along y), plane by plane
  for (int f = 0; f < count; ++f)
    k_dft_lines<false><<<nx, kDftThreads, 32 * (size_t)ny, s>>>(ny, nx, w.mid + f * N * 2, 2, nx, out + f * N * 2, 2,
                                                                nx, w.twy, sign, scale);
  LAUNCHCHK();
  return HITX_OK;
}

}  // namespace

extern "C" {

size_t hitx_post_workspace_bytes(int32_t nx, int32_t ny, int32_t planes) {
  if (nx < 2 || ny < 2 || planes < 1) return 0;
  return post_layout(nx, ny, planes, nullptr, nullptr);
}

int hitx_post_fields(int32_t nx, int32_t ny, double dx, double dy, const double* rho, const double* u,
                     const double* v, const double* p, const double* omega, const double* theta, int32_t favre,
                     double* out, void* ws, size_t ws_bytes, void* stream) {
  int rc = post_check(nx, ny, 1, ws, ws_bytes);
  if (rc) return rc;
  if (!rho || !u || !v || !out) return fail(HITX_EINVAL, "null array");
  if ((omega == nullptr) != (theta == nullptr)) return fail(HITX_EINVAL, "give both omega and theta, or neither");
  if (!(dx > 0.0) || !(dy > 0.0)) return fail(HITX_EINVAL, "dx and dy must be positive");
  PostWS w;
  post_layout(nx, ny, 1, ws, &w);
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int N = nx * ny;
  if (!omega) {  // central differences into the omega' and theta' planes, fluctuations taken in place below
    double* vo = out + 6 * (size_t)N;
    double* th = out + 7 * (size_t)N;
    k_post_divvort<<<std::min(1024, (N + kPostBlock - 1) / kPostBlock), kPostBlock, 0, s>>>(nx, ny, dx, dy, u, v, th,
                                                                                            vo);
    omega = vo;
    theta = th;
  }
  const int nb = std::min(kPostRedBlocks, (N + kPostBlock - 1) / kPostBlock);
  k_post_partials<<<nb, kPostBlock, 0, s>>>(N, rho, u, v, p, omega, theta, w.part);
  k_post_means<<<1, kPostBlock, 0, s>>>(nb, N, w.part, w.means);
  k_post_fields<<<std::min(1024, (N + kPostBlock - 1) / kPostBlock), kPostBlock, 0, s>>>(N, favre, rho, u, v, p, omega,
                                                                                         theta, w.means, out);
  LAUNCHCHK();
  return HITX_OK;
}

int hitx_dft2(int32_t nx, int32_t ny, int32_t count, const double* in, int32_t in_complex, int32_t inverse,
              double* out, void* ws, size_t ws_bytes, void* stream) {
  int rc = post_check(nx, ny, count, ws, ws_bytes);
  if (rc) return rc;
  if (!in || !out) return fail(HITX_EINVAL, "null array");
  PostWS w;
  post_layout(nx, ny, count, ws, &w);
  return post_dft2(nx, ny, count, in, in_complex != 0, inverse != 0, out, w, static_cast<hipStream_t>(stream));
}

int hitx_shell_spectra(int32_t nx, int32_t ny, double dx, double dy, const double* hat, int32_t nshell,
                       double* energy, double* enstrophy, void* ws, size_t ws_bytes, void* stream) {
  int rc = post_check(nx, ny, 1, ws, ws_bytes);
  if (rc) return rc;
  if (!hat || !energy || !enstrophy) return fail(HITX_EINVAL, "null array");
  if (nshell < 1 || nshell > std::min(nx, ny) / 2 + 1) return fail(HITX_EINVAL, "nshell must be 1..min(nx, ny)/2 + 1");
  if (!(dx > 0.0) || !(dy > 0.0)) return fail(HITX_EINVAL, "dx and dy must be positive");
  PostWS w;
  post_layout(nx, ny, 1, ws, &w);
  hipStream_t s = static_cast<hipStream_t>(stream);
  k_shell_rows<<<ny, kPostBlock, 20 * (size_t)nx, s>>>(nx, ny, dx, dy, nshell, hat, w.hist);
  k_shell_sum<<<(nshell + kPostBlock - 1) / kPostBlock, kPostBlock, 0, s>>>(ny, nshell, w.hist, energy, enstrophy);
  LAUNCHCHK();
  return HITX_OK;
}

int hitx_correlations(int32_t nx, int32_t ny, int32_t nmaps, const double* hat, int32_t nplanes,
                      const int32_t* pairs_host, const int32_t* norm_host, double* out, void* ws, size_t ws_bytes,
                      void* stream) {
  if (nmaps < 1 || nmaps > kCorrMaxMaps) return fail(HITX_EINVAL, "1..16 correlation maps per call");
  int rc = post_check(nx, ny, 3 * nmaps, ws, ws_bytes);
  if (rc) return rc;
  if (!hat || !out || !pairs_host || !norm_host) return fail(HITX_EINVAL, "null array");
  CorrMaps M{};
  M.n = nmaps;
  for (int m = 0; m < nmaps; ++m) {
    M.pair[m][0] = pairs_host[2 * m];
    M.pair[m][1] = pairs_host[2 * m + 1];
    M.mode[m] = norm_host[3 * m];
    M.ref[m][0] = norm_host[3 * m + 1];
    M.ref[m][1] = norm_host[3 * m + 2];
    if (M.pair[m][0] < 0 || M.pair[m][0] >= nplanes || M.pair[m][1] < 0 || M.pair[m][1] >= nplanes)
      return fail(HITX_EINVAL, "correlation pair names a plane beyond nplanes");
    if (M.mode[m] < 0 || M.mode[m] > 2) return fail(HITX_EINVAL, "normalisation must be 0, 1 or 2");
    if (M.mode[m] == 2 && (M.ref[m][0] < 0 || M.ref[m][0] >= nmaps || M.ref[m][1] < 0 || M.ref[m][1] >= nmaps))
      return fail(HITX_EINVAL, "normalisation reference beyond the call's maps");
  }
  PostWS w;
  post_layout(nx, ny, 3 * nmaps, ws, &w);
  hipStream_t s = static_cast<hipStream_t>(stream);
  const size_t N = (size_t)nx * ny;
  double* prod = w.mid + 2 * (size_t)nmaps * N * 2;  // the last nmaps planes; the first 2 nmaps take the transform
  double* inv = w.mid + (size_t)nmaps * N * 2;
  const int nb = (int)std::min<size_t>(4096, (N * nmaps + kPostBlock - 1) / kPostBlock);
  k_cross_spectra<<<nb, kPostBlock, 0, s>>>((int)N, M, hat, prod);
  PostWS wt = w;  // intermediate planes of the inverse transform: the first nmaps
  rc = post_dft2(nx, ny, nmaps, prod, true, true, inv, wt, s);
  if (rc) return rc;
  k_corr_finish<<<nb, kPostBlock, 0, s>>>((int)N, M, inv, out);
  LAUNCHCHK();
  return HITX_OK;
}

size_t hitx_normals_workspace_bytes(int64_t n) { return n > 0 && n <= kRngMaxN ? rng_layout(n).total : 0; }

int hitx_normals(uint64_t* state, double* out, int64_t n, int32_t flags, void* ws, size_t ws_bytes, void* stream) {
  if (!state || !out || !ws) return fail(HITX_EINVAL, "null argument");
  if (n <= 0 || n > kRngMaxN) return fail(HITX_EINVAL, "normal count out of range (1 .. 2^30)");
  const RngLayout L = rng_layout(n);
  if (ws_bytes < L.total) return fail(HITX_ENOMEM, "workspace too small for the normal variates");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const RngJob J = make_rngjob(static_cast<uint8_t*>(ws), L, state, out, n);
  const int nb = (J.nch + kRngAttemptThreads - 1) / kRngAttemptThreads;
  k_rng_attempts<<<nb, kRngAttemptThreads, 0, s>>>(J);
  k_rng_scan<<<1, kRngScan, 0, s>>>(J, (flags & HITX_NORMALS_SERIAL) ? 1 : 0);
  k_rng_scatter<<<(J.nch + kRngScatterThreads - 1) / kRngScatterThreads, kRngScatterThreads, 0, s>>>(J);
  k_rng_finish<<<1, 64, 0, s>>>(J);
  LAUNCHCHK();
  return HITX_OK;
}

int hitx_set_sensor_form(hitx_plan* pl, int32_t form) {
  if (!pl) return fail(HITX_EINVAL, "null plan");
  if (form < 0 || form > 2) return fail(HITX_EINVAL, "sensor form must be 0 (auto), 1 (fused) or 2 (split)");
  pl->sensor_form = form;
  return HITX_OK;
}

int hitx_set_line_form(hitx_plan* pl, int32_t form) {
  if (!pl) return fail(HITX_EINVAL, "null plan");
  if (form < 0 || form > 2) return fail(HITX_EINVAL, "line form must be 0 (auto), 1 (block) or 2 (wave)");
  pl->line_form = form;
  return HITX_OK;
}

int hitx_timing_enable(hitx_plan* pl, int32_t on) {
  if (!pl) return fail(HITX_EINVAL, "null plan");
  pl->timing = on ? 1 : 0;
  pl->ev_used = 0;
  return HITX_OK;
}

int hitx_timing_read(hitx_plan* pl, double* total_ms, int32_t* launches) {
  if (!pl || !total_ms || !launches) return fail(HITX_EINVAL, "null argument");
  double t = 0.0;
  for (size_t i = 0; i + 1 < pl->ev_used; i += 2) {
    HIPCHK(hipEventSynchronize(pl->ev[i + 1]));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, pl->ev[i], pl->ev[i + 1]));
    t += ms;
  }
  *total_ms = t;
  *launches = (int32_t)(pl->ev_used / 2);
  pl->ev_used = 0;
  return HITX_OK;
}

}  // extern "C"

// ============================================================== batches
struct hitx_batch {
  int n = 0, esize = 8, drag = 0, device = 0;
  bool any_per = false, any_nonper = false;  // periodic (HIT) and non-periodic (Riemann) instances present
  bool any_nodrag = false;                   // instances without the drag source (with drag: b->drag)
  bool stage_walk = false;  // drag stage reductions by column walk (many tiles) instead of on tiles
  int stage_form = 0;       // 0 auto (column walk above kStageTileMaxTiles tiles), 1 tiles, 2 column walk
  std::vector<hitx_plan*> plans;
  std::vector<double> kpar;  // n * K_COUNT
  struct Buf {
    size_t Q, Q1, Q2, S, qT, dfx, dfyT, fx, fy, mask, maskT, part, scal, ctl, kpar, phi, fhat, noise, fpart, kxy, rng,
        rngws, drows, frows, rawm;
  };
  std::vector<Buf> buf;
  size_t off_tab = 0, off_offs = 0, ws_bytes = 0, offs_cap = 0;
  int total[OFF_KINDS] = {0};
  std::vector<int> offs;  // OFF_KINDS * (n + 1)
  size_t line_lds = 0, sensor_lds = 0, proj_lds = 0, wave_slice = 16;
  bool line_small[2] = {true, true};  // every block-form line of the axis in one-chunk shape
  int line_form = 0;   // 0 auto (by the batch's total line count), 1 block, 2 wave, 3 wave with the slab y-sweep,
                       // 4 wave with the x-sweep fused into the stage update (where the batch allows it)
  bool xfused = false;  // stages run the y-sweep, then b_xupdate (x-sweep + update per row panel)
  int serial_draws = 0;  // every noise draw takes the serial path (a check of that path)
  size_t xu_lds = 0;    // dynamic LDS of b_xupdate
  bool wave_x = false, wave_y = false, slab_y = false;
  size_t slab_lds = 0;  // dynamic LDS of the slab y-sweep (slab or slices, whichever is larger)
  int sensor_form = 0;  // 0 auto (split for >= 1M nodes), 1 fused tile, 2 split (column walk), 3 split (tiles)
  bool split_sensor = false, raw_tile = false;
  void* ws = nullptr;
  // event pairs around the stage-update launches while timing is on
  bool timing = false;
  std::vector<hipEvent_t> ev;
  size_t ev_used = 0;
  // the forcing chain's own stream (non-blocking, default priority) and its
  // events: the chain runs beside stage 0's sensor and sweeps, which do not
  // read the forcing field, the next step's noise draw beside stages 1-2;
  // created when the batch is bound
  hipStream_t side = nullptr;
  hipEvent_t fork_ev = nullptr, join_ev = nullptr, pre_ev = nullptr, prejoin_ev = nullptr;
  // per stage: the drag row table and the x-sweep on their own stream, beside
  // the raw mask, dilation and y-sweep (fork after the finalize, the x-sweep
  // after the dilation, join before the update)
  hipStream_t side2 = nullptr;
  hipEvent_t sfork_ev = nullptr, smask_ev = nullptr, sjoin_ev = nullptr;
};

namespace {

// Grid-stride reduction blocks of a batch instance (b_step_reduce,
// b_fsynth): 1024 nodes per block, at most 128 (a batch runs tens of
// thousands of tile blocks anyway, and each block first resolves its
// instance through the block map, so fewer and fuller blocks pay that once
// per four nodes a thread).
int batch_red_blocks(const hitx_plan* pl) {
  return (int)std::min<size_t>((pl->N + 4 * kBlock - 1) / (4 * kBlock), 128);
}

int batch_blocks(const hitx_batch* b, const hitx_plan* pl, int kind) {
  const int nx = pl->d.nx, ny = pl->d.ny;
  const bool wx = b->wave_x && line_wave_ok(nx), wy = b->wave_y && line_wave_ok(ny);
  switch (kind) {
    case OFF_RED: return batch_red_blocks(pl);
    case OFF_SR: return pl->nblk_sr;
    case OFF_DROWS: return pl->d.drag_enabled && pl->d.n_drag_modes > 0 ? (2 * ny + kBlock - 1) / kBlock : 0;
    case OFF_FROWS: return pl->d.n_force_modes > 0 ? (ny * (pl->fkx + 1) + kBlock - 1) / kBlock : 0;
    case OFF_SENS: return ((nx + kSTX - 1) / kSTX) * ((ny + kSTY - 1) / kSTY);
    case OFF_DB: return (((nx + kDBCols - 1) / kDBCols) * ((ny + kDBRows - 1) / kDBRows) + kDBWaves - 1) / kDBWaves;
    case OFF_LX: return wx ? 0 : ny;
    case OFF_LY: return wy ? 0 : nx;
    case OFF_LXW: return (wx && !b->xfused) ? (ny + kLineWaves - 1) / kLineWaves : 0;
    case OFF_XU: return b->xfused ? (ny + kXUWaves - 1) / kXUWaves : 0;
    case OFF_LYW: return wy ? (b->slab_y ? (nx + kSlabWaves - 1) / kSlabWaves : (nx + kLineWaves - 1) / kLineWaves) : 0;
    case OFF_T16: return ((nx + 15) / 16) * ((ny + 15) / 16);
    case OFF_PROJ: return pl->d.n_force_modes > 0 ? pl->nblk_proj : 0;
    case OFF_FS: return pl->d.n_force_modes > 0 ? walk_blocks(nx, ny, kFsRows, 0) : 0;
    case OFF_RNG:
      return pl->d.n_force_modes > 0 ? (rng_chunks(pl->N) + kRngAttemptThreads - 1) / kRngAttemptThreads : 0;
    case OFF_RNGS:
      return pl->d.n_force_modes > 0 ? (rng_chunks(pl->N) + kRngScatterThreads - 1) / kRngScatterThreads : 0;
    default: return 0;
  }
}

void batch_offsets(hitx_batch* b) {
  const int n = b->n;
  long long lx = 0, ly = 0;  // wave-capable lines per axis
  for (const hitx_plan* pl : b->plans) {
    if (line_wave_ok(pl->d.nx)) lx += pl->d.ny;
    if (line_wave_ok(pl->d.ny)) ly += pl->d.nx;
  }
  long long cells = 0;
  for (const hitx_plan* pl : b->plans) cells += pl->N;
  b->split_sensor = b->sensor_form >= 2 || (b->sensor_form == 0 && cells >= (1ll << 20));
  b->raw_tile = b->sensor_form == 3;
  b->wave_x = b->line_form >= 2 || (b->line_form == 0 && lx >= kWaveFormMinLines);
  b->wave_y = b->line_form >= 2 || (b->line_form == 0 && ly >= kWaveFormMinLines);
  if (b->line_form == 1) b->wave_x = b->wave_y = false;
  b->slab_y = b->wave_y && b->line_form == 3;
  // the fused x-sweep + update (line form 4 only: measured slower than the
  // separate launches on the 64-instance workload, so never chosen by auto):
  // FP32, every grid periodic with both line lengths in the wave form, the
  // y-sweep on the transposed state
  bool fusable = b->esize == 4 && b->xu_lds > 0 && b->wave_x && b->wave_y && !b->slab_y && b->line_form == 4;
  for (const hitx_plan* pl : b->plans)
    fusable = fusable && pl->d.periodic && line_wave_ok(pl->d.nx) && line_wave_ok(pl->d.ny);
  b->xfused = fusable;
  b->offs.assign((size_t)OFF_KINDS * (n + 1), 0);
  for (int k = 0; k < OFF_KINDS; ++k) {
    int acc = 0;
    for (int i = 0; i < n; ++i) {
      b->offs[(size_t)k * (n + 1) + i] = acc;
      acc += batch_blocks(b, b->plans[i], k);
    }
    b->offs[(size_t)k * (n + 1) + n] = acc;
    b->total[k] = acc;
  }
  b->stage_walk = b->stage_form == 2 || (b->stage_form == 0 && stage_walk_form(b->total[OFF_SENS]));
  // block maps: for kind k, entry b = instance << kMapLbBits | local block
  // (the map of kind k starts at the index held in slot OFF_KINDS*(n+1) + k);
  // slot OFF_KINDS*(n+1) + OFF_KINDS is the batch-wide hyperviscosity word
  // (0 between steps)
  const size_t hdr = (size_t)OFF_KINDS * (n + 1);
  b->offs.resize(hdr + OFF_KINDS + 1, 0);
  for (int k = 0; k < OFF_KINDS; ++k) {
    b->offs[hdr + k] = (int)b->offs.size();
    for (int i = 0; i < n; ++i) {
      const int nb = batch_blocks(b, b->plans[i], k);
      for (int l = 0; l < nb; ++l) b->offs.push_back((int)(((uint32_t)i << kMapLbBits) | (uint32_t)l));
    }
  }
}

// Upper bound of the offsets buffer over every launch-shape choice (tile
// heights at their minimum, both line forms counted).
size_t batch_offs_capacity(const hitx_batch* b) {
  const int n = b->n;
  size_t cap = (size_t)OFF_KINDS * (n + 1) + OFF_KINDS + 1;
  for (const hitx_plan* pl : b->plans) {
    const int nx = pl->d.nx, ny = pl->d.ny;
    cap += (size_t)pl->nblk_stage + walk_blocks(nx, ny, kTuneRbMin, 1) + ((nx + kSTX - 1) / kSTX) * ((ny + kSTY - 1) / kSTY);
    cap += (size_t)ny + nx + (ny + kLineWaves - 1) / kLineWaves + (nx + kLineWaves - 1) / kLineWaves;
    cap += (size_t)((nx + 15) / 16) * ((ny + 15) / 16);
    cap += (size_t)(((nx + kDBCols - 1) / kDBCols) * ((ny + kDBRows - 1) / kDBRows) + kDBWaves - 1) / kDBWaves;
    if (pl->d.n_force_modes > 0)
      cap += (size_t)walk_blocks(nx, ny, kTuneRbpMin, 0) + (ny * (pl->fkx + 1) + kBlock - 1) / kBlock +
             2 * ((rng_chunks(pl->N) + kRngScatterThreads - 1) / kRngScatterThreads) +
             (rng_chunks(pl->N) + kRngAttemptThreads - 1) / kRngAttemptThreads + walk_blocks(nx, ny, kFsRows, 0);
    cap += (size_t)(2 * ny + kBlock - 1) / kBlock;
    cap += (size_t)(ny + kXUWaves - 1) / kXUWaves;
  }
  return cap;
}

template <class T>
void fill_inst(const hitx_batch* b, int i, uint8_t* base, Inst<T>& I) {
  const hitx_plan* pl = b->plans[i];
  const hitx_batch::Buf& o = b->buf[i];
  I.P = make_params<T>(pl);
  I.Kx = linek<T>(pl, 0);
  I.Ky = linek<T>(pl, 1);
  auto t = [&](size_t off) { return reinterpret_cast<T*>(base + off); };
  auto d = [&](size_t off) { return reinterpret_cast<double*>(base + off); };
  I.Q = t(o.Q); I.Q1 = t(o.Q1); I.Q2 = t(o.Q2); I.S = t(o.S);
  I.qT = t(o.qT); I.dfx = t(o.dfx); I.dfyT = t(o.dfyT); I.fx = t(o.fx); I.fy = t(o.fy);
  I.mask = base + o.mask; I.maskT = base + o.maskT; I.rawm = base + o.rawm;
  I.part = d(o.part); I.scal = d(o.scal); I.ctl = d(o.ctl); I.kpar = d(o.kpar); I.phi = d(o.phi);
  I.fhat = d(o.fhat); I.noise = d(o.noise); I.fpart = d(o.fpart); I.kxy = d(o.kxy);
  I.rng = reinterpret_cast<uint64_t*>(base + o.rng);
  I.R = make_rngjob(base + o.rngws, rng_layout(pl->N), I.rng, I.noise, pl->N);
  I.drows = d(o.drows);
  I.frows = d(o.frows);
  I.nblk_stage = batch_red_blocks(pl);
  I.nblk_sr = pl->nblk_sr;
  I.nblk_proj = pl->nblk_proj;
  I.nsw = (rng_chunks(pl->N) + 63) / 64;
  I.ntile = ((pl->d.nx + kSTX - 1) / kSTX) * ((pl->d.ny + kSTY - 1) / kSTY);
  I.use_qT = !(b->wave_y && b->slab_y && line_wave_ok(pl->d.ny));
}

template <class T>
int batch_bind_t(hitx_batch* b, void* ws, bool constants) {
  uint8_t* base = static_cast<uint8_t*>(ws);
  std::vector<Inst<T>> tab(b->n);
  for (int i = 0; i < b->n; ++i) fill_inst<T>(b, i, base, tab[i]);
  HIPCHK(hipMemcpy(base + b->off_tab, tab.data(), tab.size() * sizeof(Inst<T>), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(base + b->off_offs, b->offs.data(), b->offs.size() * sizeof(int), hipMemcpyHostToDevice));
  if (!constants) return HITX_OK;
  for (int i = 0; i < b->n; ++i) {
    const hitx_plan* pl = b->plans[i];
    HIPCHK(hipMemcpy(base + b->buf[i].kpar, &b->kpar[(size_t)i * K_COUNT], K_COUNT * sizeof(double),
                     hipMemcpyHostToDevice));
    // physical wavenumbers of the forcing modes: 2*pi * fftfreq(n, d) as NumPy forms it
    const int nf = pl->d.n_force_modes;
    if (nf > 0) {
      std::vector<double> kxy(2 * nf);
      const double vx = 1.0 / ((double)pl->d.nx * pl->d.dx), vy = 1.0 / ((double)pl->d.ny * pl->d.dy);
      for (int m = 0; m < nf; ++m) {
        kxy[2 * m] = (2.0 * M_PI) * ((double)pl->force_modes[2 * m] * vx);
        kxy[2 * m + 1] = (2.0 * M_PI) * ((double)pl->force_modes[2 * m + 1] * vy);
      }
      HIPCHK(hipMemcpy(base + b->buf[i].kxy, kxy.data(), kxy.size() * sizeof(double), hipMemcpyHostToDevice));
    }
  }
  return HITX_OK;
}

template <class T>
const Inst<T>* btab(const hitx_batch* b) {
  return reinterpret_cast<const Inst<T>*>(static_cast<uint8_t*>(b->ws) + b->off_tab);
}
const int* boffs(const hitx_batch* b) {
  return reinterpret_cast<const int*>(static_cast<uint8_t*>(b->ws) + b->off_offs);
}

// Event on the batch's stream around each stage-update launch when timing is
// on (eager steps only: events recorded inside a captured graph cannot be
// read back).
int batch_mark(hitx_batch* b, hipStream_t s) {
  if (!b->timing) return HITX_OK;
  if (b->ev_used == b->ev.size()) {
    hipEvent_t e;
    HIPCHK(hipEventCreate(&e));
    b->ev.push_back(e);
  }
  HIPCHK(hipEventRecord(b->ev[b->ev_used++], s));
  return HITX_OK;
}

template <class T, bool DRAG>
int batch_stage_sequence(hitx_batch* b, int st, hipStream_t s, hipEvent_t before_update = nullptr) {
  const Inst<T>* tab = btab<T>(b);
  const int* offs = boffs(b);
  const int n = b->n;
  // (batches keep the one-block-per-instance finalize: with tens of tile
  // blocks per CU, a reduction in every tile block costs more than the launch)
  // drag instances by column walk or on tiles (b->stage_walk); instances
  // without drag always on tiles, so a batch with and without drag gives each
  // instance the reduction it has in a batch of its own kind
  const bool walk = DRAG && b->stage_walk;
  if (walk) b_stage_reduce<T, true><<<b->total[OFF_SR], kBlock, 0, s>>>(tab, offs, n, st);
  if (!walk || b->any_nodrag)
    b_stage_tile<T, DRAG><<<b->total[OFF_SENS], kBlock, 0, s>>>(tab, offs, n, st, walk ? 1 : 0);
  // the post-step sensor (st 3) needs only the sensor scalars: no drag
  // coefficients and no drag row table (the next stage update is a step away
  // and reduces its own)
  const bool drag_now = DRAG && st < 3;
  b_stage_finalize<T, DRAG><<<dim3(n, drag_now ? 1 + kFinDragGroups : 1), kBlock, 0, s>>>(tab, st,
                                                                                     DRAG && b->stage_walk ? 1 : 0);
  // the drag row table rides in the tile raw mask's launch (extra blocks);
  // otherwise (other sensor forms) it runs on the second stream, forked here
  const bool drows_in_raw = drag_now && b->split_sensor && b->raw_tile && b->any_per;
  const int ndrows = drows_in_raw ? b->total[OFF_DROWS] : 0;
  // stages 0-2: the second stream takes the drag row table (when not in the
  // raw-mask launch) and the x-sweep (unless it is fused into the update)
  hipStream_t s2 = b->side2;
  const bool fork = st < 3 && s2 != nullptr && ((drag_now && !drows_in_raw) || !b->xfused);
  const bool early_fork = fork && drag_now && !drows_in_raw;
  if (early_fork) {
    HIPCHK(hipEventRecord(b->sfork_ev, s));
    HIPCHK(hipStreamWaitEvent(s2, b->sfork_ev, 0));
  }
  if (drag_now && !drows_in_raw) b_drows<T><<<b->total[OFF_DROWS], kBlock, 0, fork ? s2 : s>>>(tab, offs, n, st);
  if (b->split_sensor) {
    // the tile raw mask writes the transposed state itself, so the dilation
    // is bit work only; after the column-walk raw mask the byte dilation
    // writes it (stages 0-2)
    if (b->raw_tile) {
      if (b->any_per) b_rawtile<T, true><<<b->total[OFF_SENS] + ndrows, kBlock, 0, s>>>(tab, offs, n, st, ndrows);
      if (b->any_nonper) b_rawtile<T, false><<<b->total[OFF_SENS], kBlock, 0, s>>>(tab, offs, n, st, 0);
      b_dilate_bits<T><<<b->total[OFF_DB], 64 * kDBWaves, 0, s>>>(tab, offs, n, st);
    } else {
      b_rawmask<T><<<b->total[OFF_SR], kBlock, 0, s>>>(tab, offs, n, st);
      if (st == 3)
        b_dilate_bits<T><<<b->total[OFF_DB], 64 * kDBWaves, 0, s>>>(tab, offs, n, st);
      else
        b_dilate<T><<<b->total[OFF_SENS], kBlock, 0, s>>>(tab, offs, n, st);
    }
  } else {
    b_sensor<T><<<b->total[OFF_SENS], kBlock, b->sensor_lds, s>>>(tab, offs, n, st);
  }
  if (st == 3) {
    LAUNCHCHK();
    return HITX_OK;
  }
  const int sb = (int)b->wave_slice;
  if constexpr (sizeof(T) == 4) {
    if (b->xfused) {  // the y-sweep, then the x-sweep and the update per row panel
      b_line_wave<T, true><<<b->total[OFF_LYW], 64 * kLineWaves, kLineWaves * sb, s>>>(tab, offs, n, 1, st, sb);
      if (fork) {  // the drag row table is ready
        HIPCHK(hipEventRecord(b->sjoin_ev, s2));
        HIPCHK(hipStreamWaitEvent(s, b->sjoin_ev, 0));
      }
      if (before_update) HIPCHK(hipStreamWaitEvent(s, before_update, 0));  // the forcing field is ready
      int rc = batch_mark(b, s);
      if (rc) return rc;
      b_xupdate<T><<<b->total[OFF_XU], 64 * kXUWaves, b->xu_lds, s>>>(tab, offs, n, st, sb);
      if ((rc = batch_mark(b, s))) return rc;
      LAUNCHCHK();
      return HITX_OK;
    }
  }
  // the x-sweep on the second stream once the mask is ready, the y-sweep here
  if (fork) {
    HIPCHK(hipEventRecord(b->smask_ev, s));
    HIPCHK(hipStreamWaitEvent(s2, b->smask_ev, 0));
  }
  // one launch per line form and periodicity present in the batch; each
  // skips the instances of the other periodicity
  for (int axis = 0; axis < 2; ++axis) {
    const int kb = axis == 0 ? OFF_LX : OFF_LY, kw = axis == 0 ? OFF_LXW : OFF_LYW;
    const hipStream_t ls = (fork && axis == 0) ? s2 : s;
    for (int per = 1; per >= 0; --per) {
      if (!(per ? b->any_per : b->any_nonper)) continue;
      if (b->total[kw] > 0 && axis == 1 && b->slab_y) {
        if (per)
          b_line_wave_cols<T, true><<<b->total[kw], 64 * kSlabWaves, b->slab_lds, ls>>>(tab, offs, n, st, sb);
        else
          b_line_wave_cols<T, false><<<b->total[kw], 64 * kSlabWaves, b->slab_lds, ls>>>(tab, offs, n, st, sb);
      } else if (b->total[kw] > 0) {
        if (per)
          b_line_wave<T, true><<<b->total[kw], 64 * kLineWaves, kLineWaves * sb, ls>>>(tab, offs, n, axis, st, sb);
        else
          b_line_wave<T, false><<<b->total[kw], 64 * kLineWaves, kLineWaves * sb, ls>>>(tab, offs, n, axis, st, sb);
      }
      if (b->total[kb] > 0) {
        const bool sm = b->line_small[axis];
        if (per) {
          if (sm)
            b_line<T, true, true><<<b->total[kb], kLineBlock, b->line_lds, ls>>>(tab, offs, n, axis, st);
          else
            b_line<T, true, false><<<b->total[kb], kLineBlock, b->line_lds, ls>>>(tab, offs, n, axis, st);
        } else {
          if (sm)
            b_line<T, false, true><<<b->total[kb], kLineBlock, b->line_lds, ls>>>(tab, offs, n, axis, st);
          else
            b_line<T, false, false><<<b->total[kb], kLineBlock, b->line_lds, ls>>>(tab, offs, n, axis, st);
        }
      }
    }
  }
  if (fork) {  // the x-derivatives and the drag row table are ready
    HIPCHK(hipEventRecord(b->sjoin_ev, s2));
    HIPCHK(hipStreamWaitEvent(s, b->sjoin_ev, 0));
  }
  if (before_update) HIPCHK(hipStreamWaitEvent(s, before_update, 0));  // the forcing field is ready
  int rc = batch_mark(b, s);
  if (rc) return rc;
  if (b->any_per) b_update<T, true><<<b->total[OFF_SENS], kUBlock, 0, s>>>(tab, offs, n, st);
  if (b->any_nonper) b_update<T, false><<<b->total[OFF_SENS], kUBlock, 0, s>>>(tab, offs, n, st);
  if ((rc = batch_mark(b, s))) return rc;
  LAUNCHCHK();
  return HITX_OK;
}

// Noise draw and its projection for every forced instance on stream f:
// pre = 1 draws the next step's (and marks the instances), pre = 0 only for
// instances not marked.
template <class T>
void batch_draw(hitx_batch* b, hipStream_t f, int pre) {
  const Inst<T>* tab = btab<T>(b);
  const int* offs = boffs(b);
  const int n = b->n;
  b_rng_attempts<T><<<b->total[OFF_RNG], kRngAttemptThreads, 0, f>>>(tab, offs, n, pre);
  b_rng_scan<T><<<n, kRngScan, 0, f>>>(tab, pre, b->serial_draws);
  b_rng_scatter<T><<<b->total[OFF_RNGS], kRngScatterThreads, 0, f>>>(tab, offs, n, pre);
  b_rng_finish<T><<<n, 64, 0, f>>>(tab, pre);
  b_fproject<T><<<b->total[OFF_PROJ], kBlock, b->proj_lds, f>>>(tab, offs, n, pre);
}

template <class T>
int batch_step_t(hitx_batch* b, hipStream_t s) {
  const Inst<T>* tab = btab<T>(b);
  const int* offs = boffs(b);
  const int n = b->n;
  b_step_reduce<T><<<b->total[OFF_RED], kBlock, 0, s>>>(tab, offs, n);
  // the batch-wide hyperviscosity word (slot after the block-map starts)
  int* hvany = const_cast<int*>(offs) + (size_t)OFF_KINDS * (n + 1) + OFF_KINDS;
  b_ctl_begin<T><<<n, kBlock, 0, s>>>(tab, hvany);
  LAUNCHCHK();
  // The forcing chain (OU update, synthesis, rescale; the noise draw and its
  // projection only for instances whose noise the previous step did not
  // draw) reads the control scalars b_ctl_begin wrote and writes only forcing
  // buffers, its partials included (fpart); stage 0's sensor and sweeps read
  // neither.  So the chain runs on the side stream beside them, and stage 0's
  // update waits for it.  Captured into a graph, the fork and joins become the
  // graph's edges.
  hipEvent_t joined = nullptr;
  if (b->total[OFF_PROJ] > 0) {  // some instance is forced
    if (!b->side) return fail(HITX_EINVAL, "batch not bound to a workspace");
    hipStream_t f = b->side;
    HIPCHK(hipEventRecord(b->fork_ev, s));
    HIPCHK(hipStreamWaitEvent(f, b->fork_ev, 0));
    batch_draw<T>(b, f, 0);  // instances whose noise was not drawn during the previous step
    b_fou<T><<<n, kBlock, 0, f>>>(tab);
    b_frows<T><<<b->total[OFF_FROWS], kBlock, 0, f>>>(tab, offs, n);
    b_fsynth<T><<<b->total[OFF_FS], kBlock, 0, f>>>(tab, offs, n);
    b_falpha<T><<<n, kBlock, 0, f>>>(tab);
    LAUNCHCHK();
    HIPCHK(hipEventRecord(b->join_ev, f));
    joined = b->join_ev;
  }
  for (int st = 0; st < 4; ++st) {
    const hipEvent_t w = st == 0 ? joined : nullptr;
    int rc = b->drag ? batch_stage_sequence<T, true>(b, st, s, w) : batch_stage_sequence<T, false>(b, st, s, w);
    if (rc) return rc;
    if (st == 0 && joined) {
      // the next step's noise and projection on the side stream beside
      // stages 1-2 (the noise and the partials are free once the rescale has
      // read the synthesis sums; the stages read only fx, fy); joined before
      // the step's control end
      HIPCHK(hipEventRecord(b->pre_ev, s));
      HIPCHK(hipStreamWaitEvent(b->side, b->pre_ev, 0));
      batch_draw<T>(b, b->side, 1);
      LAUNCHCHK();
      HIPCHK(hipEventRecord(b->prejoin_ev, b->side));
    }
  }
  if (joined) HIPCHK(hipStreamWaitEvent(s, b->prejoin_ev, 0));
  int guard_any = 0;  // outflow guards copy S -> Q after every step
  for (const hitx_plan* pl : b->plans) guard_any |= (!pl->d.periodic && pl->d.guard_cells > 0) ? 1 : 0;
  b_finish<T><<<b->total[OFF_T16], kBlock, 0, s>>>(tab, offs, n, guard_any);
  b_ctl_end<T><<<(n + 63) / 64, 64, 0, s>>>(tab, n, hvany);
  LAUNCHCHK();
  return HITX_OK;
}

template <class T>
int batch_time_updates_t(hitx_batch* b, int reps, hipStream_t s, double* total_ms) {
  const Inst<T>* tab = btab<T>(b);
  const int* offs = boffs(b);
  const int n = b->n;
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  HIPCHK(hipEventRecord(e0, s));
  for (int r = 0; r < reps; ++r)
    for (int st = 0; st < 3; ++st) {
      if (b->any_per) b_update<T, true><<<b->total[OFF_SENS], kUBlock, 0, s>>>(tab, offs, n, st);
      if (b->any_nonper) b_update<T, false><<<b->total[OFF_SENS], kUBlock, 0, s>>>(tab, offs, n, st);
    }
  LAUNCHCHK();
  HIPCHK(hipEventRecord(e1, s));
  HIPCHK(hipEventSynchronize(e1));
  float ms = 0.f;
  HIPCHK(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  *total_ms = ms;
  return HITX_OK;
}

template <class T>
int batch_fixed_dt_t(hitx_batch* b, hipStream_t s) {
  const Inst<T>* tab = btab<T>(b);
  b_step_reduce<T><<<b->total[OFF_RED], kBlock, 0, s>>>(tab, boffs(b), b->n);
  b_fixed_dt<T><<<b->n, kBlock, 0, s>>>(tab);
  LAUNCHCHK();
  return HITX_OK;
}

template <class T>
int batch_set_lds(hitx_batch* b) {
  int maxsh = 0;
  HIPCHK(hipDeviceGetAttribute(&maxsh, hipDeviceAttributeMaxSharedMemoryPerBlock, b->device));
  if (b->line_lds > (size_t)maxsh) return fail(HITX_EINVAL, "grid line too long for LDS");
  if (b->line_lds > 65536) HIPCHK(allow_line_lds<T>(b->line_lds));
  if (b->slab_lds > (size_t)maxsh) return fail(HITX_EINVAL, "grid column too long for the slab y-sweep");
  if (b->slab_lds > 65536) HIPCHK(allow_slab_lds<T>(b->slab_lds));
  if constexpr (sizeof(T) == 4) {
    if (b->xu_lds > (size_t)maxsh) b->xu_lds = 0;  // no fused form for this batch
    if (b->xu_lds > 65536)
      HIPCHK(hipFuncSetAttribute(reinterpret_cast<const void*>(&b_xupdate<float>),
                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)b->xu_lds));
  }
  return HITX_OK;
}

}  // namespace

extern "C" {

int hitx_batch_create(const hitx_desc* descs, const double* kpar, int32_t n, hitx_batch** out) {
  if (!out || !descs || !kpar || n <= 0) return fail(HITX_EINVAL, "bad batch arguments");
  *out = nullptr;
  if (n >= (1 << (32 - kMapLbBits))) return fail(HITX_EINVAL, "too many batch instances (at most 4095)");
  for (int i = 0; i < n; ++i)  // the block map holds local block indices below 2^kMapLbBits
    if ((long long)((descs[i].nx + 15) / 16) * ((descs[i].ny + 15) / 16) >= (1ll << kMapLbBits))
      return fail(HITX_EINVAL, "batch instance " + std::to_string(i) + " too large (16x16 tiles must be < 2^20)");
  for (int i = 1; i < n; ++i)  // one kernel instantiation per element type; everything else is per instance
    if (descs[i].dtype != descs[0].dtype) return fail(HITX_EINVAL, "batch instances must share the dtype");
  hitx_batch* b = new hitx_batch();
  b->n = n;
  b->esize = descs[0].dtype == HITX_F64 ? 8 : 4;
  b->kpar.assign(kpar, kpar + (size_t)n * K_COUNT);
  (void)hipGetDevice(&b->device);
  for (int i = 0; i < n; ++i) {
    hitx_plan* pl = nullptr;
    int rc = hitx_plan_create(&descs[i], &pl);
    if (rc) {
      std::string keep = g_err;
      hitx_batch_destroy(b);
      g_err = "instance " + std::to_string(i) + ": " + keep;
      return rc;
    }
    b->plans.push_back(pl);
    b->line_lds = std::max({b->line_lds, pl->line_lds_x, pl->line_lds_y});
    b->line_small[0] = b->line_small[0] && pl->lx_d.small;
    b->line_small[1] = b->line_small[1] && pl->ly_d.small;
    if (line_wave_ok(pl->d.nx)) b->wave_slice = std::max(b->wave_slice, wave_slice_bytes(pl->d.nx, pl->esize));
    if (line_wave_ok(pl->d.ny)) b->wave_slice = std::max(b->wave_slice, wave_slice_bytes(pl->d.ny, pl->esize));
    b->sensor_lds = std::max(b->sensor_lds, pl->sensor_lds);
    b->proj_lds = std::max(b->proj_lds, proj_lds_bytes(pl));
  }
  for (const hitx_plan* pl : b->plans) {
    const bool dr = make_params<double>(pl).drag_on;
    b->drag |= dr ? 1 : 0;
    b->any_nodrag |= !dr;
    b->any_per |= pl->d.periodic != 0;
    b->any_nonper |= pl->d.periodic == 0;
  }
  b->slab_lds = (size_t)kSlabWaves * b->wave_slice;
  for (const hitx_plan* pl : b->plans)
    if (line_wave_ok(pl->d.ny))
      b->slab_lds = std::max(b->slab_lds, (size_t)4 * kSlabWaves * slab_cstride(pl->d.ny) * b->esize);
  if (b->esize == 4)
    b->xu_lds = std::max((size_t)kXUWaves * b->wave_slice, sizeof(UpdLds<float, kXTX, kXUWaves>));
  int rc = b->esize == 8 ? batch_set_lds<double>(b) : batch_set_lds<float>(b);
  if (rc) {
    std::string keep = g_err;
    hitx_batch_destroy(b);
    g_err = keep;
    return rc;
  }
  // layout
  size_t o = 0;
  const size_t E = b->esize;
  for (int i = 0; i < n; ++i) {
    const hitx_plan* pl = b->plans[i];
    const size_t N = pl->N;
    const int nf = std::max(1, pl->d.n_force_modes);
    hitx_batch::Buf u;
    auto take = [&](size_t bytes) { size_t r = o; o = align256(o + bytes); return r; };
    u.Q = take(4 * N * E); u.Q1 = take(4 * N * E); u.Q2 = take(4 * N * E); u.S = take(4 * N * E);
    u.qT = take(4 * N * E); u.dfx = take(4 * N * E); u.dfyT = take(4 * N * E);
    u.fx = take(N * E); u.fy = take(N * E);
    u.mask = take(N); u.maskT = take(N); u.rawm = take(N);
    // partials sized for the smallest tile heights hitx_batch_tune accepts
    const int sr_max = walk_blocks(pl->d.nx, pl->d.ny, kTuneRbMin, 1);
    const int pj_max = walk_blocks(pl->d.nx, pl->d.ny, kTuneRbpMin, 0);
    const size_t ntile = (size_t)((pl->d.nx + kSTX - 1) / kSTX) * ((pl->d.ny + kSTY - 1) / kSTY);
    u.part = take(std::max((size_t)std::max({pl->nblk_stage, pl->nblk_sr, sr_max}) * 64, ntile * (R_DRAG + kDragCols)) *
                  sizeof(double));
    u.scal = take(S_COUNT * sizeof(double));
    u.ctl = take(C_COUNT * sizeof(double));
    u.kpar = take(K_COUNT * sizeof(double));
    u.phi = take(2 * nf * sizeof(double));
    u.fhat = take(4 * nf * sizeof(double));
    u.noise = take(pl->d.n_force_modes > 0 ? N * sizeof(double) : 8);
    // forcing partials: the projection's (2 per mode and projection block or
    // scatter wavefront), then the synthesis's (5 per block)
    const size_t prow = (size_t)std::max({pl->nblk_proj, pj_max, (rng_chunks(pl->N) + 63) / 64});
    u.fpart = take(std::max(prow * 2 * nf, (size_t)walk_blocks(pl->d.nx, pl->d.ny, kFsRows, 0) * 5) *
                   sizeof(double));
    u.kxy = take(2 * nf * sizeof(double));
    u.rng = take(4 * sizeof(uint64_t));
    u.rngws = take(pl->d.n_force_modes > 0 ? rng_layout(N).total : 8);
    u.drows = take((pl->d.n_drag_modes > 0 ? (size_t)pl->d.ny * 10 : 1) * sizeof(double));
    u.frows = take((pl->d.n_force_modes > 0 ? (size_t)pl->d.ny * 2 * (2 * pl->fkx + 1) : 1) * sizeof(double));
    b->buf.push_back(u);
  }
  batch_offsets(b);
  b->offs_cap = batch_offs_capacity(b);
  if (b->offs.size() > b->offs_cap) return fail(HITX_EINVAL, "block map larger than its bound");
  const size_t inst_bytes = b->esize == 8 ? sizeof(Inst<double>) : sizeof(Inst<float>);
  b->off_tab = o; o = align256(o + inst_bytes * n);
  b->off_offs = o; o = align256(o + b->offs_cap * sizeof(int));
  b->ws_bytes = o;
  *out = b;
  return HITX_OK;
}

int hitx_batch_destroy(hitx_batch* b) {
  if (!b) return HITX_OK;
  for (hipEvent_t e : b->ev) (void)hipEventDestroy(e);
  for (hipEvent_t e : {b->fork_ev, b->join_ev, b->pre_ev, b->prejoin_ev})
    if (e) (void)hipEventDestroy(e);
  if (b->side) (void)hipStreamDestroy(b->side);
  for (hipEvent_t e : {b->sfork_ev, b->smask_ev, b->sjoin_ev})
    if (e) (void)hipEventDestroy(e);
  if (b->side2) (void)hipStreamDestroy(b->side2);
  for (hitx_plan* p : b->plans) hitx_plan_destroy(p);
  delete b;
  return HITX_OK;
}

size_t hitx_batch_workspace_bytes(const hitx_batch* b) { return b ? b->ws_bytes : 0; }

int hitx_batch_bind(hitx_batch* b, void* ws, size_t ws_bytes) {
  if (!b || !ws) return fail(HITX_EINVAL, "null argument");
  if (ws_bytes < b->ws_bytes) return fail(HITX_ENOMEM, "batch workspace too small");
  b->ws = ws;
  if (!b->side) {  // the side streams and their fork/join events (outside any graph capture)
    HIPCHK(hipStreamCreateWithFlags(&b->side, hipStreamNonBlocking));
    HIPCHK(hipEventCreateWithFlags(&b->fork_ev, hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&b->join_ev, hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&b->pre_ev, hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&b->prejoin_ev, hipEventDisableTiming));
    HIPCHK(hipStreamCreateWithFlags(&b->side2, hipStreamNonBlocking));
    HIPCHK(hipEventCreateWithFlags(&b->sfork_ev, hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&b->smask_ev, hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&b->sjoin_ev, hipEventDisableTiming));
  }
  return b->esize == 8 ? batch_bind_t<double>(b, ws, true) : batch_bind_t<float>(b, ws, true);
}

int hitx_batch_buffer(const hitx_batch* b, int32_t i, int32_t which, void** ptr) {
  if (!b || !ptr || i < 0 || i >= b->n) return fail(HITX_EINVAL, "bad instance index");
  if (!b->ws) return fail(HITX_EINVAL, "batch not bound to a workspace");
  const hitx_batch::Buf& u = b->buf[i];
  size_t off;
  switch (which) {
    case HITX_BUF_STATE: off = u.Q; break;
    case HITX_BUF_NOISE: off = u.noise; break;
    case HITX_BUF_CTL: off = u.ctl; break;
    case HITX_BUF_PHI: off = u.phi; break;
    case HITX_BUF_RNG: off = u.rng; break;
    case HITX_BUF_FX: off = u.fx; break;
    case HITX_BUF_FY: off = u.fy; break;
    default: return fail(HITX_EINVAL, "unknown buffer");
  }
  *ptr = static_cast<uint8_t*>(b->ws) + off;
  return HITX_OK;
}

int hitx_batch_tune(hitx_batch* b, int32_t rb, int32_t rbp, int32_t line_form, int32_t sensor_form) {
  if (!b) return fail(HITX_EINVAL, "null batch");
  if ((rb != 0 && (rb < kTuneRbMin || rb > kTuneRbMax)) || (rbp != 0 && (rbp < kTuneRbpMin || rbp > kTuneRbpMax)))
    return fail(HITX_EINVAL, "tile heights out of range (rb 0 or 2..64, rbp 0 or 8..256)");
  if (line_form < 0 || line_form > 4)
    return fail(HITX_EINVAL,
                "line form must be 0 (auto), 1 (block), 2 (wave), 3 (wave, slab y-sweep) or 4 (wave, x-sweep fused "
                "into the update)");
  if (sensor_form < 0 || sensor_form > 3)
    return fail(HITX_EINVAL, "sensor form must be 0 (auto), 1 (fused), 2 (split, column walk) or 3 (split, tiles)");
  b->line_form = line_form;
  b->sensor_form = sensor_form;
  for (hitx_plan* pl : b->plans) {
    pl->rb = rb > 0 ? rb : default_rb(pl->d.nx, pl->d.ny);
    pl->rbp = rbp > 0 ? rbp : default_rbp(pl->d.nx, pl->d.ny);
    pl->nblk_sr = walk_blocks(pl->d.nx, pl->d.ny, pl->rb, 1);
    pl->nblk_proj = walk_blocks(pl->d.nx, pl->d.ny, pl->rbp, 0);
  }
  batch_offsets(b);
  if (b->offs.size() > b->offs_cap) return fail(HITX_EINVAL, "block map larger than its bound");
  if (b->ws) return b->esize == 8 ? batch_bind_t<double>(b, b->ws, false) : batch_bind_t<float>(b, b->ws, false);
  return HITX_OK;
}

int hitx_batch_timing_enable(hitx_batch* b, int32_t on) {
  if (!b) return fail(HITX_EINVAL, "null batch");
  b->timing = on != 0;
  b->ev_used = 0;
  return HITX_OK;
}

int hitx_batch_timing_read(hitx_batch* b, double* total_ms, int32_t* launches) {
  if (!b || !total_ms || !launches) return fail(HITX_EINVAL, "null argument");
  double t = 0.0;
  for (size_t i = 0; i + 1 < b->ev_used; i += 2) {
    HIPCHK(hipEventSynchronize(b->ev[i + 1]));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, b->ev[i], b->ev[i + 1]));
    t += ms;
  }
  *total_ms = t;
  *launches = (int32_t)(b->ev_used / 2);
  b->ev_used = 0;
  return HITX_OK;
}

int hitx_batch_time_updates(hitx_batch* b, int32_t reps, void* stream, double* total_ms) {
  if (!b || !total_ms) return fail(HITX_EINVAL, "null argument");
  if (!b->ws) return fail(HITX_EINVAL, "batch not bound to a workspace");
  if (reps < 1) return fail(HITX_EINVAL, "reps must be positive");
  hipStream_t s = static_cast<hipStream_t>(stream);
  return b->esize == 8 ? batch_time_updates_t<double>(b, reps, s, total_ms)
                       : batch_time_updates_t<float>(b, reps, s, total_ms);
}

int hitx_batch_serial_draws(hitx_batch* b, int32_t on) {
  if (!b) return fail(HITX_EINVAL, "null batch");
  b->serial_draws = on != 0 ? 1 : 0;
  return HITX_OK;
}

int hitx_batch_forms(const hitx_batch* b, int32_t* forms) {
  if (!b || !forms) return fail(HITX_EINVAL, "null argument");
  forms[0] = b->xfused ? 1 : 0;
  forms[1] = b->wave_x ? 1 : 0;
  forms[2] = b->wave_y ? 1 : 0;
  forms[3] = b->slab_y ? 1 : 0;
  forms[4] = b->split_sensor ? 1 : 0;
  forms[5] = b->stage_walk ? 1 : 0;
  return HITX_OK;
}

int hitx_batch_stage_form(hitx_batch* b, int32_t form) {
  if (!b) return fail(HITX_EINVAL, "null batch");
  if (form < 0 || form > 2) return fail(HITX_EINVAL, "stage form must be 0 (auto), 1 (tiles) or 2 (column walk)");
  b->stage_form = form;
  batch_offsets(b);
  if (b->offs.size() > b->offs_cap) return fail(HITX_EINVAL, "block map larger than its bound");
  if (b->ws) return b->esize == 8 ? batch_bind_t<double>(b, b->ws, false) : batch_bind_t<float>(b, b->ws, false);
  return HITX_OK;
}

int hitx_batch_fixed_dt(hitx_batch* b, void* stream) {
  if (!b || !b->ws) return fail(HITX_EINVAL, "batch not bound");
  hipStream_t s = static_cast<hipStream_t>(stream);
  return b->esize == 8 ? batch_fixed_dt_t<double>(b, s) : batch_fixed_dt_t<float>(b, s);
}

int hitx_batch_step(hitx_batch* b, void* stream) {
  if (!b || !b->ws) return fail(HITX_EINVAL, "batch not bound");
  hipStream_t s = static_cast<hipStream_t>(stream);
  return b->esize == 8 ? batch_step_t<double>(b, s) : batch_step_t<float>(b, s);
}

}  // extern "C"
