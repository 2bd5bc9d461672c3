// hitx C ABI: plan creation, workspace layout and launch sequences.
// Declarations and the reference interfaces they replace: include/hitx.h.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/hitx.h"
#include "hitx_kernels.h"

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

constexpr double kAlphaC = 3.0 / 8.0;  // OOP/spatial_operator.py:240

template <class T>
LineK<T> make_linek(int n) {
  LineK<T> k{};
  const double a = kAlphaC;
  const double rho = (-1.0 + std::sqrt(1.0 - 4.0 * a * a)) / (2.0 * a);
  k.n = n;
  // largest chunk length that still gives every thread of the line block a task
  k.L = 1;
  for (int L : {8, 4, 2, 1})
    if (n % L == 0 && 4 * (n / L) >= kLineBlock) {
      k.L = L;
      break;
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
  LineK<double> lx_d{}, ly_d{};
  LineK<float> lx_f{}, ly_f{};
  // workspace layout
  size_t off_raw = 0, off_mask = 0, off_maskT = 0, off_qT = 0, off_dfx = 0, off_dfyT = 0, off_part = 0,
         off_scal = 0, off_misc = 0, off_ticket = 0, ws_bytes = 0;
  size_t sensor_lds = 0;
  int nblk_red = 0, nblk_proj = 0;
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
  return P;
}

size_t line_lds_bytes(int n, int nch, int esize) {
  return (size_t)(12 * (n + 1) + 66 + 8 * nch + 20) * esize + 4 * (size_t)(n + 1) + n + 64;
}

size_t sensor_lds_bytes(int w, int esize) {
  const int H = w + 1, EX = kSTX + 2 * H, EY = kSTY + 2 * H;
  return (size_t)4 * EX * EY * esize + (size_t)(kSTX + 2 * w) * (kSTY + 2 * w) + (size_t)(kSTY + 2 * w) * kSTX +
         (size_t)kSTY * (kSTX + 1) + 64;
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
  double *part, *scal, *misc;
  unsigned int* ticket;
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
  w.ticket = reinterpret_cast<unsigned int*>(b + pl->off_ticket);
  return w;
}

int check_ws(const hitx_plan* pl, void* ws, size_t ws_bytes) {
  if (!pl) return fail(HITX_EINVAL, "null plan");
  if (!ws || ws_bytes < pl->ws_bytes)
    return fail(HITX_ENOMEM, "workspace too small: need " + std::to_string(pl->ws_bytes) + " bytes");
  return HITX_OK;
}

dim3 grid16(const hitx_plan* pl) { return dim3((pl->d.nx + 15) / 16, (pl->d.ny + 15) / 16); }

// stage reductions (one kernel, last block finalizes) + fused sensor tile:
// mask and maskT of q, and the enforced transposed state when qT != nullptr
template <class T>
int run_sensor(hitx_plan* pl, const T* q, int enforce_first, uint8_t* mask_out, T* qT, const WS& w, hipStream_t s) {
  const Params<T> P = make_params<T>(pl);
  if (P.drag_on)
    k_stage_reduce<T, true><<<pl->nblk_red, kBlock, 0, s>>>(P, q, enforce_first, w.part, w.ticket, w.scal);
  else
    k_stage_reduce<T, false><<<pl->nblk_red, kBlock, 0, s>>>(P, q, enforce_first, w.part, w.ticket, w.scal);
  dim3 g((pl->d.nx + kSTX - 1) / kSTX, (pl->d.ny + kSTY - 1) / kSTY);
  k_sensor<T><<<g, kBlock, pl->sensor_lds, s>>>(P, q, enforce_first, w.scal, mask_out, w.maskT, qT);
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
template <class T>
int run_sweeps(hitx_plan* pl, const T* qs, const WS& w, hipStream_t s) {
  int rc = run_sensor<T>(pl, qs, 1, w.mask, static_cast<T*>(w.qT), w, s);
  if (rc) return rc;
  const Params<T> P = make_params<T>(pl);
  if ((rc = tmark(pl, s))) return rc;
  if (P.periodic)
    k_line<T, true><<<pl->d.ny, kLineBlock, pl->line_lds_x, s>>>(P, linek<T>(pl, 0), 0, 0, qs, w.mask, w.scal,
                                                                 static_cast<T*>(w.dfx));
  else
    k_line<T, false><<<pl->d.ny, kLineBlock, pl->line_lds_x, s>>>(P, linek<T>(pl, 0), 0, 0, qs, w.mask, w.scal,
                                                                  static_cast<T*>(w.dfx));
  if ((rc = tmark(pl, s))) return rc;
  if ((rc = tmark(pl, s))) return rc;
  if (P.periodic)
    k_line<T, true><<<pl->d.nx, kLineBlock, pl->line_lds_y, s>>>(P, linek<T>(pl, 1), 1, 1,
                                                                 static_cast<const T*>(w.qT), w.maskT, w.scal,
                                                                 static_cast<T*>(w.dfyT));
  else
    k_line<T, false><<<pl->d.nx, kLineBlock, pl->line_lds_y, s>>>(P, linek<T>(pl, 1), 1, 1,
                                                                  static_cast<const T*>(w.qT), w.maskT, w.scal,
                                                                  static_cast<T*>(w.dfyT));
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
  if (mx > 65536) {
    HIPCHK(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_line<T, true>),
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)mx));
    HIPCHK(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_line<T, false>),
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)mx));
  }
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
  return HITX_OK;
}

template <class T>
int rhs_or_stage(hitx_plan* pl, int mode, int stage, const void* q0, const void* qs, void* out, const void* fx,
                        const void* fy, double dt, void* ws, hipStream_t s) {
  const WS w = carve(pl, ws);
  int rc = run_sweeps<T>(pl, static_cast<const T*>(qs), w, s);
  if (rc) return rc;
  const Params<T> P = make_params<T>(pl);
  k_update<T><<<grid16(pl), kBlock, 0, s>>>(P, mode, stage, static_cast<const T*>(q0), static_cast<const T*>(qs),
                                            static_cast<const T*>(w.dfx), static_cast<const T*>(w.dfyT),
                                            static_cast<const T*>(fx), static_cast<const T*>(fy), w.scal, (T)dt,
                                            static_cast<T*>(out));
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
  if (km_x > 16 || km_y > 16) {
    delete pl;
    return fail(HITX_EINVAL, "mode index beyond 16");
  }
  pl->kmx = km_x;
  pl->kmy = km_y;
  if (!all.empty()) {
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
  pl->lx_d = make_linek<double>(desc->nx);
  pl->ly_d = make_linek<double>(desc->ny);
  pl->lx_f = make_linek<float>(desc->nx);
  pl->ly_f = make_linek<float>(desc->ny);
  pl->line_lds_x = line_lds_bytes(desc->nx, pl->lx_d.nch, pl->esize);
  pl->line_lds_y = line_lds_bytes(desc->ny, pl->ly_d.nch, pl->esize);
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
  pl->nblk_proj = (int)((N + (size_t)kBlock * kPPT - 1) / ((size_t)kBlock * kPPT));
  size_t o = 0;
  pl->off_raw = o; o = align256(o + N);
  pl->off_mask = o; o = align256(o + N);
  pl->off_maskT = o; o = align256(o + N);
  pl->off_qT = o; o = align256(o + 4 * N * E);
  pl->off_dfx = o; o = align256(o + 4 * N * E);
  pl->off_dfyT = o; o = align256(o + 4 * N * E);
  const size_t part = std::max<size_t>((size_t)pl->nblk_red * 64,
                                       (size_t)pl->nblk_proj * 2 * std::max(1, desc->n_force_modes));
  pl->off_part = o; o = align256(o + part * sizeof(double));
  pl->off_scal = o; o = align256(o + S_COUNT * sizeof(double));
  pl->off_misc = o; o = align256(o + 64 * sizeof(double));
  pl->off_ticket = o; o = align256(o + sizeof(unsigned int));
  pl->ws_bytes = o;
  *out = pl;
  return HITX_OK;
}

int hitx_plan_destroy(hitx_plan* pl) {
  if (!pl) return HITX_OK;
  for (hipEvent_t e : pl->ev) (void)hipEventDestroy(e);
  if (pl->d_tw) (void)hipFree(pl->d_tw);
  if (pl->d_modes) (void)hipFree(pl->d_modes);
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
  if (pl->esize == 8)
    k_fproject<double><<<pl->nblk_proj, kBlock, 0, s>>>(make_params<double>(pl), noise, w.part);
  else
    k_fproject<float><<<pl->nblk_proj, kBlock, 0, s>>>(make_params<float>(pl), noise, w.part);
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
  if (pl->esize == 8)
    k_fsynth<double><<<pl->nblk_red, kBlock, 0, s>>>(make_params<double>(pl), fhat, static_cast<double*>(fx),
                                                     static_cast<double*>(fy), static_cast<const double*>(rho),
                                                     static_cast<const double*>(u), static_cast<const double*>(v),
                                                     w.part);
  else
    k_fsynth<float><<<pl->nblk_red, kBlock, 0, s>>>(make_params<float>(pl), fhat, static_cast<float*>(fx),
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

int hitx_viscous_rhs(hitx_plan* pl, const void* q, void* out, void* stream) {
  if (!pl) return fail(HITX_EINVAL, "null plan");
  if (!q || !out) return fail(HITX_EINVAL, "null array");
  if (!pl->d.periodic) return fail(HITX_EINVAL, "viscous terms are defined on the periodic grid only");
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (pl->esize == 8) {
    Params<double> P = make_params<double>(pl);
    P.visc_on = 1;
    k_update<double><<<grid16(pl), kBlock, 0, s>>>(P, MODE_VISC, 0, nullptr, static_cast<const double*>(q), nullptr,
                                                   nullptr, nullptr, nullptr, nullptr, 0.0, static_cast<double*>(out));
  } else {
    Params<float> P = make_params<float>(pl);
    P.visc_on = 1;
    k_update<float><<<grid16(pl), kBlock, 0, s>>>(P, MODE_VISC, 0, nullptr, static_cast<const float*>(q), nullptr,
                                                  nullptr, nullptr, nullptr, nullptr, 0.0f, static_cast<float*>(out));
  }
  LAUNCHCHK();
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
