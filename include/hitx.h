/* hitx — MI355X (gfx950) RHS evaluation and SSP-RK3 step of the 2-D
 * compressible hybrid compact/WENO7 solver of HIT_Solver.
 *
 * C ABI: plain pointers and sizes, no exceptions, no torch types.  Every
 * array pointer is a DEVICE pointer (e.g. torch.Tensor.data_ptr() of a
 * contiguous tensor on the current HIP device) unless named *_host.  Arrays
 * are C-order: state q is [4][ny][nx] (rho, rho*u, rho*v, E; x fastest),
 * masks are uint8 [ny][nx] (1 = WENO node).  The element type of q is
 * double when desc.dtype == HITX_F64 and float when HITX_F32.  Reductions
 * are always accumulated in double.
 *
 * Calls are asynchronous on `stream` (a hipStream_t, 0 = default stream).
 * Return value 0 = success, negative = error; hitx_last_error() returns a
 * thread-local message for the last failure.  A plan is used by one host
 * thread at a time.  The caller owns all memory; a plan owns only constant
 * tables (twiddles, mode lists).
 *
 * Reference interfaces replaced (paths in the HIT_Solver tree):
 *   OOP/parallel/backend.py:17-32        select_array_module -> new backend "hip"
 *   OOP/spatial_operator.py:263-304      PeriodicEulerShockSensor2D.detect   -> hitx_sensor
 *   OOP/problems/riemann_config3.py:288  NonPeriodicEulerShockSensor2D.detect -> hitx_sensor
 *   OOP/spatial_operator.py:344-367      PeriodicHybridEuler2DOperator.rhs    -> hitx_rhs
 *   OOP/problems/riemann_config3.py:359  NonPeriodicHybridEuler2DOperator.rhs -> hitx_rhs
 *   OOP/time_operator.py:21-26           SSPRK3.step (one stage each call)   -> hitx_rk3_stage
 *   OOP/hit2d.py:847-867                 rhs_with_forcing sources (in stage)  -> hitx_rk3_stage
 *   OOP/spatial_operator.py:407-435      PeriodicHyperviscosity2D.apply       -> hitx_hyperviscosity
 *   OOP/problems/riemann_config3.py:398  LocalHyperviscosity2D.apply          -> hitx_hyperviscosity
 *   OOP/equations.py:165-181             enforce_physical_state               -> hitx_enforce
 *   OOP/problems/riemann_config3.py:419  apply_outflow_guard                  -> hitx_enforce (guard flag)
 *   OOP/spatial_operator.py:369-377      stable_time_step (and hit2d.py:279-290 Mt) -> hitx_step_reduce
 *   OOP/problems/riemann_config3.py:371  fixed_time_step                      -> hitx_step_reduce
 *   OOP/forcing.py:77-79                 fft2 + shell mask of the noise       -> hitx_forcing_project
 *   OOP/forcing.py:82-87                 curl + ifft2 + rms                   -> hitx_forcing_synthesize
 *   OOP/forcing.py:91-118                rescale and statistics               -> hitx_forcing_scale
 *   OOP/hit2d.py:132-136                 conservative_to_primitive            -> hitx_primitives
 *   OOP/equations.py:232-267             CompressibleNavierStokes2D.viscous_rhs -> hitx_viscous_rhs
 */
#ifndef HITX_H
#define HITX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HITX_ABI_VERSION 1

enum { HITX_F64 = 0, HITX_F32 = 1 };
enum { HITX_HYBRID = 0, HITX_WENO = 1 };

enum {
  HITX_OK = 0,
  HITX_EINVAL = -1,   /* bad argument or descriptor (Python: ValueError) */
  HITX_EDEVICE = -2,  /* HIP runtime failure (Python: RuntimeError)     */
  HITX_ENOMEM = -3,   /* workspace too small                             */
};

#define HITX_MAX_MODES 128

typedef struct hitx_desc {
  int32_t abi_version;    /* = HITX_ABI_VERSION */
  int32_t nx, ny;         /* grid; both >= 8 */
  int32_t periodic;       /* 1: periodic HIT operator; 0: non-periodic Riemann operator */
  int32_t dtype;          /* HITX_F64 or HITX_F32 */
  int32_t scheme;         /* HITX_HYBRID or HITX_WENO (WENO: non-periodic only) */
  double dx, dy;          /* Domain2D.dx/dy */
  double dx2, dy2;        /* dx**2, dy**2 as the host language rounds them */
  double gamma, rho_floor, p_floor;
  /* shock sensor */
  int32_t sensor_width;   /* dilation half-width, <= 8 */
  int32_t boundary_guard; /* non-periodic: forced band before dilation */
  int32_t shear_enabled;  /* 0: no shear term */
  int32_t pad0;
  double compression_threshold, jump_threshold, shear_threshold;
  /* Navier-Stokes terms (periodic only); viscosity 0 disables them */
  double viscosity, prandtl, gas_constant;
  /* hyperviscosity: per-component coefficient mn*w_k*h**4 */
  double hv_coef[4];
  /* non-periodic outflow guard cells (0 = none) */
  int32_t guard_cells;
  /* HIT sources */
  int32_t drag_enabled, cooling_enabled, forcing_enabled;
  double drag_coef, cooling_time, cooling_p_ref;
  int32_t n_drag_modes, n_force_modes;
  /* (kx index, ky index) pairs in FFT index space, |index| <= 16; host pointers,
   * copied at plan creation */
  const int32_t* drag_modes_host;   /* [n_drag_modes][2] */
  const int32_t* force_modes_host;  /* [n_force_modes][2] */
} hitx_desc;

typedef struct hitx_plan hitx_plan;

/* Plan lifetime.  The plan binds to the HIP device current at creation. */
int hitx_plan_create(const hitx_desc* desc, hitx_plan** out);
int hitx_plan_destroy(hitx_plan* plan);
/* Bytes of device workspace that every compute entry point below needs.
 * A workspace must not be shared by calls that can run concurrently. */
size_t hitx_workspace_bytes(const hitx_plan* plan);
const char* hitx_last_error(void);
int hitx_abi_version(void);

/* Shock sensor: mask_out[ny][nx] = 1 at WENO nodes.  enforce_first = 1
 * applies enforce_physical_state to q before detection, as the operator's
 * rhs does; 0 detects on q as given (the post-step detect calls). */
int hitx_sensor(hitx_plan* plan, const void* q, int32_t enforce_first, uint8_t* mask_out,
                void* ws, size_t ws_bytes, void* stream);

/* Operator RHS only (no viscous or forcing sources): dq = -dF/dx - dG/dy. */
int hitx_rhs(hitx_plan* plan, const void* q, void* dq, void* ws, size_t ws_bytes, void* stream);

/* One Shu-Osher stage (0, 1, 2) of SSP-RK3 with the full HIT source closure
 * (viscous, forcing fx/fy, drag, cooling as enabled in the descriptor):
 *   stage 0: out = clean(q0 + dt L(q0))                     (qs == q0)
 *   stage 1: out = clean(3/4 q0 + 1/4 (qs + dt L(qs)))
 *   stage 2: out = clean(1/3 q0 + 2/3 (qs + dt L(qs)))
 * clean = enforce_physical_state, then the outflow guard when guard_cells > 0.
 * fx, fy: [ny][nx] forcing accelerations (ignored unless forcing_enabled).
 * out must not alias q0 or qs. */
int hitx_rk3_stage(hitx_plan* plan, int32_t stage, const void* q0, const void* qs, void* out,
                   const void* fx, const void* fy, double dt, void* ws, size_t ws_bytes,
                   void* stream);

/* hitx_rk3_stage without a workspace argument and without any allocation.
 * No counterpart in the reference: it is for callers that cannot allocate
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The workspace is a static device array of 32 MiB
 * (33554432 bytes) that each device the library
 * runs on holds until the process ends; grids whose hitx_workspace_bytes
 * exceeds it are rejected with HITX_EINVAL (512x512 FP64 and 768x768 FP32
 * fit, 1024x1024 does not).  Every call on one device shares that device's array, so
 * calls that overlap on one GPU (other streams, other threads) race: issue
 * them in order on one stream.  Sweep timing (hitx_timing_enable) must be
 * off.  Results equal hitx_rk3_stage's bit for bit. */
int hitx_rk3_stage_static(hitx_plan* plan, int32_t stage, const void* q0, const void* qs, void* out,
                          const void* fx, const void* fy, double dt, void* stream);

/* Biharmonic filter on the nodes where (active[ny][nx] != 0) xor invert --
 * invert = 1 takes the sensor's WENO mask and filters the compact nodes --
 * then enforce_physical_state, then the outflow guard if apply_guard.  out
 * must not alias q. */
int hitx_hyperviscosity(hitx_plan* plan, const void* q, const uint8_t* active, int32_t invert, void* out,
                        int32_t apply_guard, void* ws, size_t ws_bytes, void* stream);

/* out = enforce_physical_state(q), then the outflow guard if apply_guard. */
int hitx_enforce(hitx_plan* plan, const void* q, void* out, int32_t apply_guard, void* stream);

/* Step-level reductions of q into out_dev (device, double[3]):
 *   [0] max((|u|+c)/dx + (|v|+c)/dy)   [1] sum(u^2+v^2)   [2] sum(c)   */
int hitx_step_reduce(hitx_plan* plan, const void* q, double* out_dev, void* ws, size_t ws_bytes,
                     void* stream);

/* Forcing.  noise: device double[ny][nx].  modes_out: device double[n_force_modes][2]
 * = sum over the grid of noise * exp(-i k.x) (unnormalised forward DFT at each
 * shell mode). */
int hitx_forcing_project(hitx_plan* plan, const double* noise, double* modes_out, void* ws,
                         size_t ws_bytes, void* stream);
/* fhat: device double[2][n_force_modes][2] (fx_hat, fy_hat, complex), spectral
 * coefficients of the inverse DFT (1/(nx*ny) normalisation applied here).
 * Writes fx, fy (element type of q) and stats_out (device double[2]):
 *   [0] sum(fx^2 + fy^2)   [1] sum(rho*(fx*u + fy*v)).
 * rho, u, v: [ny][nx] arrays of the element type of q. */
int hitx_forcing_synthesize(hitx_plan* plan, const double* fhat, void* fx, void* fy, const void* rho,
                            const void* u, const void* v, double* stats_out, void* ws, size_t ws_bytes,
                            void* stream);
/* In place fx = (fx*scale)*alpha (same for fy), then stats_out (double[4]):
 *   [0] sum(fx^2) [1] sum(fy^2) [2] sum(fx*fy) [3] sum(rho*(fx*u + fy*v)). */
int hitx_forcing_scale(hitx_plan* plan, void* fx, void* fy, double scale, double alpha, const void* rho,
                       const void* u, const void* v, double* stats_out, void* ws, size_t ws_bytes,
                       void* stream);

/* rho, u, v, p = primitive_from_conservative(q), each [ny][nx]. */
int hitx_primitives(hitx_plan* plan, const void* q, void* rho, void* u, void* v, void* p, void* stream);

/* Periodic viscous and heat-conduction terms of q (the same terms the stage
 * adds when desc.viscosity > 0): out [4][ny][nx], out[0] = 0. */
int hitx_viscous_rhs(hitx_plan* plan, const void* q, void* out, void* stream);

/* HIT diagnostics of a periodic state, computed in double whatever the
 * plan's dtype.  Replaces the host reductions of
 *   OOP/hit2d.py:471-548   compute_diagnostics (means, moments, Mt, rms)
 *   OOP/hit2d.py:459-468   divergence_and_vorticity (spectral derivatives)
 *   OOP/hit2d.py:429-453   large_scale_kinetic_fraction
 *   OOP/hit2d.py:370-399, 452-456  source_power(large_scale_drag_source)
 *   OOP/hit2d.py:806, 882  mean(detect(q)) for the WENO fraction (mask count)
 *   OOP/hit2d.py:571-596   the save_snapshot fields.
 * out: device double[HITX_DIAG_COUNT], the sums below (means are sum / N,
 *   N = nx*ny; u' = u - sum_u/N, v' likewise).
 * lowpass_modes_host: n_lowpass (kx, ky) index pairs of the low-pass band
 *   0 < |k| <= kmax (|kx|, |ky| <= 4, n_lowpass <= 64; 0 = no band: the
 *   LOW and DRAG_E sums are 0).
 * drag_coef: c of the drag source s = -c * lowpass(momentum - mean).
 * mask: uint8 [ny][nx] WENO mask to count, or NULL (MASK_COUNT = 0).
 * fields: device double [6][ny][nx] = rho, u, v, p, divergence, vorticity,
 *   or NULL (then the workspace holds them).
 * ws: hitx_diag_workspace_bytes(plan, fields != NULL) bytes. */
enum {
  HITX_DIAG_RHO = 0,      /* sum rho */
  HITX_DIAG_U = 1,        /* sum u */
  HITX_DIAG_V = 2,        /* sum v */
  HITX_DIAG_P = 3,        /* sum p */
  HITX_DIAG_C = 4,        /* sum sqrt(gamma p / rho) */
  HITX_DIAG_U2V2 = 5,     /* sum u^2 + v^2 */
  HITX_DIAG_RUU = 6,      /* sum rho u'^2 */
  HITX_DIAG_RVV = 7,      /* sum rho v'^2 */
  HITX_DIAG_RUV = 8,      /* sum rho u' v' */
  HITX_DIAG_RKE = 9,      /* sum rho (u'^2 + v'^2) */
  HITX_DIAG_DIV2 = 10,    /* sum divergence^2 */
  HITX_DIAG_VORT2 = 11,   /* sum vorticity^2 */
  HITX_DIAG_LOW = 12,     /* sum rho (u_low^2 + v_low^2), low-pass of u', v' */
  HITX_DIAG_DRAG_E = 13,  /* sum u s_x + v s_y (drag energy source) */
  HITX_DIAG_MASK = 14,    /* number of mask nodes */
  HITX_DIAG_COUNT = 15
};
size_t hitx_diag_workspace_bytes(const hitx_plan* plan, int32_t with_fields);
int hitx_diagnostics(hitx_plan* plan, const void* q, const uint8_t* mask, const int32_t* lowpass_modes_host,
                     int32_t n_lowpass, double drag_coef, double* fields, double* out, void* ws, size_t ws_bytes,
                     void* stream);

/* Riemann run checks, diagnostics and output fields of q, in double whatever
 * the plan's dtype.  Replaces
 *   OOP/problems/riemann_config3.py:432-452  validate_raw_physical_state
 *   OOP/problems/riemann_config3.py:629-650  _diagnostics (RiemannDiagnostics)
 *   OOP/problems/riemann_config3.py:459-463  vorticity_z (finite differences of :268-275)
 *   OOP/problems/riemann_config3.py:497-519  the save_riemann_npz planes.
 * Raw density and pressure are q0 and (gamma-1)(q3 - rho/2 (u^2+v^2)) with
 * u = q1/q0, v = q2/q0 (no floors); the vorticity uses the floored
 * primitives.  Minima and maxima propagate NaN.
 * out: device double[HITX_RDIAG_COUNT].  mask: uint8 WENO mask to count, or
 * NULL.  fields: device double [5][ny][nx] = rho, u, v, p (floored
 * primitives), omega_z; or NULL.  ws: hitx_workspace_bytes(plan). */
enum {
  HITX_RDIAG_RHO_MIN = 0,
  HITX_RDIAG_RHO_MAX = 1,
  HITX_RDIAG_P_MIN = 2,
  HITX_RDIAG_P_MAX = 3,
  HITX_RDIAG_VORT_MAX = 4,     /* max |omega_z| */
  HITX_RDIAG_NONFINITE_Q = 5,  /* nodes with a non-finite component of q */
  HITX_RDIAG_NONFINITE_RP = 6, /* nodes with non-finite raw rho or p */
  HITX_RDIAG_RHO_NONPOS = 7,   /* nodes where raw rho > 0 fails */
  HITX_RDIAG_P_NONPOS = 8,     /* nodes where raw p > 0 fails */
  HITX_RDIAG_MASK = 9,         /* mask nodes */
  HITX_RDIAG_COUNT = 10
};
int hitx_riemann_diagnostics(hitx_plan* plan, const void* q, const uint8_t* mask, double* fields, double* out,
                             void* ws, size_t ws_bytes, void* stream);

/* Initial states on the device, formed in double and stored in the plan's
 * dtype (the reference builds them with NumPy on the host).
 * hitx_hit_initial_state replaces OOP/hit2d.py:158-204 initialize_hit_2d
 *   (periodic plan, domain origin 0):
 *   modes_host: n_modes rows (kx, ky, amplitude, phase) in the reference's
 *     loop order, phases drawn on the host from its generator;
 *   psi = sum a cos(kx x + ky y + phase) at the cell centres; u = dpsi/dy,
 *   v = -dpsi/dx (spectral); mean removal; scale = target_mach * mean(c) /
 *   rms(u, v) with rho = 1, p = 1/gamma; conservative variables with the
 *   floors of OOP/equations.py:147-163.
 *   out_dev: device double[2] = rms before scaling (the caller raises when it
 *   is <= machine epsilon, as the reference does), scale.
 *   n_modes * 4 <= nx * ny; ws: hitx_diag_workspace_bytes(plan, 0) bytes.
 * hitx_riemann_initial_state replaces OOP/problems/riemann_config3.py:60-99
 *   (riemann_config3_08.py:9-19, riemann_config6.py:40-87):
 *   quadrants_host: 4 rows (rho, u, v, p) for I top-right, II top-left,
 *   III bottom-left, IV bottom-right; top = y >= y_split, right = x >= x_split
 *   at the cell centres x_min + dx (i + 1/2), y_min + dy (j + 1/2). */
int hitx_hit_initial_state(hitx_plan* plan, const double* modes_host, int32_t n_modes, double target_mach,
                           void* q_out, double* out_dev, void* ws, size_t ws_bytes, void* stream);
int hitx_riemann_initial_state(hitx_plan* plan, const double* quadrants_host, double x_min, double y_min,
                               double x_split, double y_split, void* q_out, void* stream);

/* Forcing noise: the next n standard normals of a NumPy Generator(PCG64)
 * stream, bit for bit as numpy.random.default_rng(seed).standard_normal(n)
 * returns them; replaces the host draw self._rng.standard_normal((ny, nx))
 * of OOP/forcing.py:77 (generator created at OOP/forcing.py:66).
 * state: device uint64[4] = PCG64 state (high, low word) and increment (high,
 * low word), as bit_generator.state holds them; advanced past the n values.
 * out: device double[n].  flags: HITX_NORMALS_SERIAL forces the one-thread
 * form (the same numbers; a check of the parallel form).  n in 1 .. 2^30. */
#define HITX_NORMALS_SERIAL 1
size_t hitx_normals_workspace_bytes(int64_t n);
int hitx_normals(uint64_t* state, double* out, int64_t n, int32_t flags, void* ws, size_t ws_bytes, void* stream);

/* Launch-shape support (no reference counterpart): which form of the line
 * sweep the plan launches.  1: one block of 512 threads per line; 2: one
 * wavefront per line where the line length is a multiple of 64 up to 512
 * (the block form otherwise); 0 (default): the wave form when the sweep has
 * at least 2048 such lines.  Both compute the same derivative to rounding. */
int hitx_set_line_form(hitx_plan* plan, int32_t form);

/* Launch-shape support (no reference counterpart): which form of the shock
 * sensor the plan launches.  1: one fused tile kernel (primitives and jump
 * indicators recomputed over the dilation halo); 2: a tile kernel writing
 * the undilated mask (one-node halo), then a dilation tile; 0 (default): the
 * split form.  Both give the same mask. */
int hitx_set_sensor_form(hitx_plan* plan, int32_t form);

/* Measurement support (no reference counterpart): when enabled, the line
 * sweeps of each stage (x and y in one launch when every line is in the
 * one-chunk shape, n <= 512) are bracketed by HIP events on their stream.
 * hitx_timing_read synchronises those events, returns the summed time and
 * the number of bracketed stages since the last read, and resets. */
int hitx_timing_enable(hitx_plan* plan, int32_t on);
int hitx_timing_read(hitx_plan* plan, double* total_ms, int32_t* launches);

/* ---------------------------------------------------------------- batches
 * Many instances (grids of different sizes; forced HIT and Riemann, schemes,
 * sensor widths and drag settings mixed; one dtype) advance together: one
 * launch per kernel kind per stage covers all of them (the line sweeps: one
 * per periodicity present), each instance runs the reductions and sweeps it
 * runs in a batch of its own kind, and per-instance scalar control (dt, Mach
 * control, the forcing's OU recurrence and power rescale, step counting)
 * runs on the device, so hitx_batch_step needs no host synchronisation and
 * can be captured in a HIP graph.  No reference counterpart: the reference
 * runs one instance per process (OOP/hit2d.py:764, OOP/problems/riemann_config3.py:681).
 *
 * kpar: [n][HITX_KPAR] run parameters per instance, doubles, NaN = None:
 *   0 cfl, 1 viscous dt clip 0.25 h^2/nu (NaN: none), 2 fixed dt (1: Riemann loop),
 *   3 hyperviscosity interval, 4 hyperviscosity on, 5 Mach control on, 6 Mach target,
 *   7 Mach memory, 8 Mach exponent, 9 min power, 10 max power, 11 target power,
 *   12 min forcing power, 13 max rescale, 14 alpha memory, 15 force rms,
 *   16 correlation time, 17 forcing on, 18 tfinal, 19 max steps.
 * Control scalars (HITX_BUF_CTL, doubles): 0 active, 1 time, 2 tfinal, 3 dt, 4 step,
 *   7 hyperviscosity applications, 11 target power, 12 alpha, 18-21 Fxx Fyy Fxy P_in,
 *   22 power before rescale, 23 alpha target, 24 Mach, 25 desired power, 26 n steps (fixed dt),
 *   27 the next step's noise is drawn, 28 PCG64 words that draw consumed,
 *   29 that draw took the serial path.
 * The caller writes each instance's initial state (HITX_BUF_STATE), control
 * scalars and, for forced HIT, the PCG64 state of its noise stream
 * (HITX_BUF_RNG, 4 uint64 words as for hitx_normals); a step draws the
 * instance's normals of the next step on the device while it runs (slot 27
 * set), and a step that finds slot 27 clear draws its own first.  The draw
 * projects the normals onto the forcing modes as it scatters them; only a
 * draw that takes the serial path writes them to HITX_BUF_NOISE.  To continue
 * an instance elsewhere with slot 27 set, move its PCG64 state back by slot
 * 28's words and clear slot 27. */
#define HITX_KPAR 24
#define HITX_CTL 32
enum { HITX_BUF_STATE = 0, HITX_BUF_NOISE, HITX_BUF_CTL, HITX_BUF_PHI, HITX_BUF_RNG, HITX_BUF_FX, HITX_BUF_FY };

/* ---------------------------------------------------------------- post-processing
 * Shell spectra and two-point correlations of snapshot planes (double,
 * [ny][nx], x fastest), on grids of 2..2048 nodes a side.  No plan: sizes
 * are arguments; ws is caller device memory of hitx_post_workspace_bytes(nx,
 * ny, planes) bytes, planes as each call states.  Complex planes are
 * interleaved (re, im) doubles [ny][nx][2].
 *
 * hitx_post_fields (planes 1): out[8][ny][nx] = sqrt(rho) u', sqrt(rho) v',
 *   u', v', rho', p', omega', theta' with u', v' the Reynolds (favre = 0) or
 *   Favre fluctuations; p may be null (its plane is not written); omega and
 *   theta both given, or both null: then the second-order central vorticity
 *   and divergence of u, v on the periodic grid.  Replaces
 *   HIT2DSpectra._fluctuations (OOP/postprocess/spectra.py:203-214),
 *   TwoPointCorrelation2D._velocity_fluctuations, _divergence_and_vorticity
 *   and the mean removals of compute (OOP/postprocess/
 *   two_point_correlation.py:406-422, 472-485, 190-204).
 * hitx_dft2 (planes count): numpy fft2 (inverse = 0) or ifft2 (inverse = 1,
 *   scaled by 1/(nx ny)) of count planes, real (in_complex = 0) or complex
 *   input, complex output (spectra.py:99-105; two_point_correlation.py:111,
 *   131-133).
 * hitx_shell_spectra (planes 1): from the transforms of the first four
 *   hitx_post_fields planes ([4] complex planes), the energy 0.5 (|w_u|^2 +
 *   |w_v|^2) / (nx ny)^2 and enstrophy 0.5 |kx v - ky u|^2 / (nx ny)^2 summed
 *   over shells rint(|k| / dk), nshell = min(nx, ny)/2 + 1 shells
 *   (spectra.py:97-126).
 * hitx_correlations (planes 3 nmaps): for map m < nmaps <= 16, the real part
 *   of ifft2(hat[a_m] conj(hat[b_m])) / (nx ny) with (a_m, b_m) =
 *   pairs_host[2m..2m+1] planes of hat (nplanes of them), normalised by
 *   norm_host[3m]: 0 none, 1 its own zero lag (autocorrelation), 2
 *   sqrt(R(0) R'(0)) of the maps norm_host[3m+1], norm_host[3m+2] of this call
 *   (cross-correlation); out[nmaps][ny][nx] (two_point_correlation.py:107-139).
 */
size_t hitx_post_workspace_bytes(int32_t nx, int32_t ny, int32_t planes);
int hitx_post_fields(int32_t nx, int32_t ny, double dx, double dy, const double* rho, const double* u,
                     const double* v, const double* p, const double* omega, const double* theta, int32_t favre,
                     double* out, void* ws, size_t ws_bytes, void* stream);
int hitx_dft2(int32_t nx, int32_t ny, int32_t count, const double* in, int32_t in_complex, int32_t inverse,
              double* out, void* ws, size_t ws_bytes, void* stream);
int hitx_shell_spectra(int32_t nx, int32_t ny, double dx, double dy, const double* hat, int32_t nshell,
                       double* energy, double* enstrophy, void* ws, size_t ws_bytes, void* stream);
int hitx_correlations(int32_t nx, int32_t ny, int32_t nmaps, const double* hat, int32_t nplanes,
                      const int32_t* pairs_host, const int32_t* norm_host, double* out, void* ws,
                      size_t ws_bytes, void* stream);

typedef struct hitx_batch hitx_batch;
int hitx_batch_create(const hitx_desc* descs, const double* kpar, int32_t n, hitx_batch** out);
int hitx_batch_destroy(hitx_batch* batch);
size_t hitx_batch_workspace_bytes(const hitx_batch* batch);
/* Carve the instances' buffers from ws and upload the instance table (blocking). */
int hitx_batch_bind(hitx_batch* batch, void* ws, size_t ws_bytes);
/* Launch shapes of every instance (no reference counterpart; the batch
 * autotuner's knobs): rb rows per wave of the stage reductions and of the
 * split sensor (2..64), rbp rows per wave of the forcing projection (8..256)
 * (0 for either: each instance's own default, set by its grid alone),
 * the line-sweep form (0 auto, 1 block, 2 one wavefront per line where the
 * line allows it, 3 as 2 with the y-lines read through an LDS slab, 4 as 2
 * with each row panel's x-sweep and stage update in one launch where the
 * batch allows it: FP32, periodic grids, both line lengths in the wave form)
 * and the sensor form (0 auto, 1 fused tile, 2 split: the
 * undilated mask by column walk, then the dilation; 3 split with the
 * undilated mask on tiles with a one-node halo).  Changes the summation
 * order of the reductions and the rounding of the line solve only.  Blocking
 * when bound (re-uploads the instance table); no work may be in flight. */
int hitx_batch_tune(hitx_batch* batch, int32_t rb, int32_t rbp, int32_t line_form, int32_t sensor_form);
/* The launch forms in effect (no reference counterpart; for tests and
 * measurements): forms[0] x-sweep fused into the stage update, [1] x-lines in
 * the wave form, [2] y-lines in the wave form, [3] y-lines through the slab,
 * [4] split sensor, [5] drag stage sums by column walk (each 0 or 1). */
int hitx_batch_forms(const hitx_batch* batch, int32_t* forms);
/* Every noise draw of the batch's steps takes the serial path (no reference
 * counterpart; a check of that path: the values are the same, the projection
 * is formed from the noise buffer in another summation order). */
int hitx_batch_serial_draws(hitx_batch* batch, int32_t on);
/* Form of the drag stage reductions (no reference counterpart): 0 auto (on
 * the sensor tiles while the batch has at most 4096 of them, by column walk
 * above), 1 always on tiles, 2 always by column walk.  The two forms sum the
 * stage scalars in different orders; fixing the form makes an instance's
 * arithmetic independent of what else is in its batch.  Blocking like
 * hitx_batch_tune. */
int hitx_batch_stage_form(hitx_batch* batch, int32_t form);
/* Event timing of the stage-update launches of eager steps (no reference
 * counterpart; measurement support): enable clears the record; read returns
 * the summed milliseconds and the launch count since, and clears them. */
int hitx_batch_timing_enable(hitx_batch* batch, int32_t on);
int hitx_batch_timing_read(hitx_batch* batch, double* total_ms, int32_t* launches);
/* Stage-update kernel time (no reference counterpart; measurement support):
 * launches the three stage updates of every instance reps times back to back
 * on the stream between two events and returns the milliseconds (3*reps
 * launches per periodicity present).  The updates run on whatever the stage
 * buffers hold and the last stage may write the state in place, so the
 * batch's state is not a solution afterwards. */
int hitx_batch_time_updates(hitx_batch* batch, int32_t reps, void* stream, double* total_ms);
/* Device pointer of one of instance i's buffers inside the bound workspace. */
int hitx_batch_buffer(const hitx_batch* batch, int32_t i, int32_t which, void** ptr);
/* Fixed dt of every instance from its current state (Riemann loop). */
int hitx_batch_fixed_dt(hitx_batch* batch, void* stream);
/* One full step of every active instance. */
int hitx_batch_step(hitx_batch* batch, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* HITX_H */
