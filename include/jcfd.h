/* jcfd.h -- C ABI of libjcfd_amd, the MI355X (gfx950) time-step hot path of
 * jax-cfd.
 *
 * The reference (google/jax-cfd) is pure Python on JAX; its "interface" for
 * this path is the step-function contract
 *     step_fn(state) -> state
 * built by jax_cfd/base/equations.py:120-151 (semi_implicit_navier_stokes,
 * FVM state = tuple of two staggered GridVariables) and by
 * jax_cfd/spectral/time_stepping.py:186-197 over
 * jax_cfd/spectral/equations.py:129-220 (crank_nicolson_rk4 of
 * NavierStokes2D, state = complex64 rfft2 of the vorticity), consumed by
 * jax_cfd/base/funcutils.py:82-126 (repeated / trajectory).  Every entry
 * point below replaces one of those Python-level operations; the citation
 * on each names the reference function whose arithmetic it performs.
 *
 * Conventions
 *   - Arrays are passed as jcfd_tensor, laid out exactly like DLPack's
 *     DLTensor (data, device, ndim, dtype, shape, strides, byte_offset), so a
 *     DLPack producer's DLTensor* can be passed directly.  Arrays must be
 *     C-contiguous (strides NULL or row-major), on a ROCm device
 *     (device_type 10).  FVM fields: float32 (n, n), axis 1 contiguous.
 *     Spectral fields: complex64 (n, n/2 + 1).
 *   - All work is enqueued on the given HIP stream (hipStream_t passed as
 *     void*; NULL = default stream) and is asynchronous.
 *   - Functions return JCFD_OK (0) or a negative error code; the message of
 *     the last error on the calling thread is returned by jcfd_last_error().
 *     Nothing throws across the ABI.
 *   - A plan owns its device constants and scratch; one plan must not be
 *     used from two streams at once.  Distinct plans are independent.
 */
#ifndef JCFD_H_
#define JCFD_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------ errors */
#define JCFD_OK 0
#define JCFD_ERR_INVALID_ARGUMENT (-1) /* bad shape / dtype / stride / value */
#define JCFD_ERR_UNSUPPORTED (-2)      /* valid for the reference, not built here */
#define JCFD_ERR_DEVICE (-3)           /* HIP runtime error */
#define JCFD_ERR_OUT_OF_MEMORY (-4)
#define JCFD_ERR_BUDGET (-5)           /* would exceed a caller-set byte budget */

/* ------------------------------------------------ DLPack-layout tensor */
typedef struct {
  int32_t device_type; /* 1 = CPU, 10 = ROCm */
  int32_t device_id;
} jcfd_device;

typedef struct {
  uint8_t code; /* 2 = float, 5 = complex */
  uint8_t bits; /* 32 (float32) or 64 (complex64) */
  uint16_t lanes;
} jcfd_dtype;

typedef struct {
  void* data;
  jcfd_device device;
  int32_t ndim;
  jcfd_dtype dtype;
  int64_t* shape;
  int64_t* strides; /* in elements; NULL = C-contiguous */
  uint64_t byte_offset;
} jcfd_tensor;

/* Message of the last failed call on this thread ("" if none). */
const char* jcfd_last_error(void);
/* Library version string. */
const char* jcfd_version(void);

/* ------------------------------------------------------------ FVM path */
#define JCFD_CONVECT_VAN_LEER 0          /* advection.py:210-342 */
#define JCFD_CONVECT_VAN_LEER_LIMITERS 1 /* advection.py:387-395 */
#define JCFD_CONVECT_LINEAR 2            /* convect_linear, advection.py:119-123, 184-207 */
#define JCFD_CONVECT_UPWIND 3            /* advect_upwind, advection.py:126-130 */
#define JCFD_PRESSURE_FAST_DIAG 0        /* pressure.py:115-157 */
#define JCFD_PRESSURE_CG 1               /* pressure.py:67-112 */
#define JCFD_DIFFUSION_EXPLICIT 0        /* semi_implicit_navier_stokes: diffuse (diffusion.py:35-37) */
#define JCFD_DIFFUSION_IMPLICIT_FAST_DIAG 1 /* implicit_diffusion_navier_stokes (equations.py:154-195)
                                                with diffusion.solve_fast_diag (diffusion.py:170-212) */
#define JCFD_DIFFUSION_IMPLICIT_CG 2     /* ... with diffusion.solve_cg (diffusion.py:131-167):
                                            CG on (1 - viscosity dt L) x = u per component, x0 = u */

typedef struct {
  int32_t n;               /* cells per side; square periodic grid */
  double dx, dy;           /* grid steps (grids.py:551-555) */
  double dt;               /* time step */
  double density;          /* rho */
  double viscosity;        /* nu; ignored unless has_viscosity */
  int32_t has_viscosity;   /* 0 mirrors viscosity=None */
  int32_t convect;         /* JCFD_CONVECT_* */
  int32_t pressure;        /* JCFD_PRESSURE_* */
  double cg_rtol, cg_atol; /* solve_cg tolerances */
  int32_t cg_maxiter;      /* <= 0: 10 * n * n, as jax.scipy.sparse.linalg.cg */
  int32_t has_forcing;     /* 0 mirrors forcing=None */
  double linear_coefficient;     /* linear_forcing coefficient (forcings.py:107-113) */
  const float* forcing_x;  /* host array, n floats: constant x-forcing f_x(y_j) of
                              kolmogorov_forcing (forcings.py:63-104); may be NULL
                              (= zero) */
  int32_t batch;           /* instances the plan holds scratch for (<= 1: one);
                              batched fields are (batch, n, n) */
  int32_t diffusion;       /* JCFD_DIFFUSION_*: explicit nu/rho lap in the explicit
                              terms, or the step of implicit_diffusion_navier_stokes
                              (convection + forcing, Euler, projection, then
                              (1 - viscosity dt L)^-1 per component); needs
                              has_viscosity */
  int32_t rk_stages;       /* Butcher tableau of navier_stokes_rk (time_stepping.py:59-106):
                              0 or 1 with rk_b[0] = 1 is forward Euler (fused
                              kernels); up to 4 stages otherwise */
  double rk_a[6];          /* a[i-1][j], i = 1 .. stages-1, j < i, packed row by row */
  double rk_b[4];
  double diff_cg_rtol, diff_cg_atol; /* diffusion.solve_cg tolerances (JCFD_DIFFUSION_IMPLICIT_CG) */
  int32_t diff_cg_maxiter;           /* <= 0: 10 * n * n */
} jcfd_fvm_config;

typedef struct jcfd_fvm_plan jcfd_fvm_plan;

/* Build a plan: validates the config, computes Lambda^+ of
 * fast_diagonalization.pseudoinverse (fast_diagonalization.py:199-266, cutoff
 * 10 * eps(float32)), FFT twiddles, and allocates scratch on the current
 * device.  Replaces the factory semi_implicit_navier_stokes
 * (equations.py:120-151) with time_stepper=forward_euler. */
int jcfd_fvm_plan_create(const jcfd_fvm_config* config, jcfd_fvm_plan** plan);
int jcfd_fvm_plan_destroy(jcfd_fvm_plan* plan);
/* Device bytes held by the plan (constants + scratch). */
int64_t jcfd_fvm_plan_device_bytes(const jcfd_fvm_plan* plan);
/* Multi-step calls replay steps 2..nsteps from a captured hipGraph of 32
 * steps when enabled (default: off for a single FVM instance, on for batched
 * spectral plans); 0 launches every kernel individually.  Same results. */
int jcfd_fvm_plan_set_graphs(jcfd_fvm_plan* plan, int32_t enabled);
/* CG solves (pressure solve_cg, diffusion solve_cg) as one cooperative
 * launch per solve with grid barriers between the iteration steps (where the
 * device keeps all the solve's workgroups resident), or as two launches per
 * iteration with the convergence flag read every 64 iterations (default:
 * on MI355X a grid barrier, with the cache write-back and invalidation that
 * make other XCDs' writes visible, costs more than a launch: 21.7 against
 * 7.5 us per iteration at 256^2).  No counterpart in the reference; same
 * results. */
int jcfd_fvm_plan_set_cg_persistent(jcfd_fvm_plan* plan, int32_t enabled);

/* Tile shapes and FFT radix schedule of the three step kernels.  No
 * counterpart in the reference (XLA picks its own kernels); set by the
 * project's autotuner (jax_cfd_amd.autotune), which times the menu on the
 * device.  Tile shapes never change results; radix 4 instead of 8 changes
 * the FFT rounding (results stay within fp32 round-off of each other). */
typedef struct {
  int32_t rows_forward; /* rows per workgroup of the stencil kernel: 1, 2, 4, 8; 16, 32
                           march over the rows with a ring of staged rows (n 256, 512) */
  int32_t poisson_cols; /* spectrum columns per workgroup: 1, 2, 4, 8, 16 */
  int32_t rows_project; /* rows per workgroup of the projection kernel: 1, 2, 4, 8 */
  int32_t radix;        /* largest FFT radix: 8 or 4 */
} jcfd_fvm_tiles;
int jcfd_fvm_plan_get_tiles(const jcfd_fvm_plan* plan, jcfd_fvm_tiles* tiles);
/* Rejects (JCFD_ERR_UNSUPPORTED) a shape that is not built for the plan's n. */
int jcfd_fvm_plan_set_tiles(jcfd_fvm_plan* plan, const jcfd_fvm_tiles* tiles);

/* nsteps applications of step_fn (time_stepping.py:88-104 forward Euler +
 * pressure.py:181-198 projection), i.e. funcutils.repeated(step_fn, nsteps)
 * (funcutils.py:82-88).  u_out/v_out may alias u/v. */
int jcfd_fvm_step(jcfd_fvm_plan* plan, const jcfd_tensor* u, const jcfd_tensor* v,
                  jcfd_tensor* u_out, jcfd_tensor* v_out, int64_t nsteps, void* stream);

/* Continuous-batching form of jcfd_fvm_step: advances, in place, only the
 * instances listed in `slots` (int32, 1-D, on the device, distinct values in
 * [0, batch)) of the (batch, n, n) fields u and v, by nsteps steps.  Other
 * slots are neither read nor written.  Every listed instance gets exactly
 * the arithmetic of jcfd_fvm_step on that instance alone. */
int jcfd_fvm_step_slots(jcfd_fvm_plan* plan, jcfd_tensor* u, jcfd_tensor* v,
                        const jcfd_tensor* slots, int64_t nsteps, void* stream);

/* jcfd_fvm_step without any allocation, for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  No counterpart in the reference.  The
 * step's scratch (u*, v*, the packed divergence spectrum: 12 * batch * n^2
 * bytes) lives in a static device array of 16 MiB that each GPU holds until
 * the process ends, instead of the plan's scratch; larger plans are rejected
 * (JCFD_ERR_INVALID_ARGUMENT), as are plans other than forward Euler with
 * solve_fast_diag (JCFD_ERR_UNSUPPORTED).  Steps launch directly (never a
 * graph).  Results equal jcfd_fvm_step's bit for bit.  All calls on one GPU
 * share that GPU's array: calls that overlap in time on one GPU (two
 * streams, two threads) race and must be serialised by the caller.  The
 * first call on a device looks the array up (hipGetSymbolAddress). */
int jcfd_fvm_step_noalloc(jcfd_fvm_plan* plan, const jcfd_tensor* u, const jcfd_tensor* v,
                          jcfd_tensor* u_out, jcfd_tensor* v_out, int64_t nsteps, void* stream);

/* Explicit terms dv/dt = convection + nu/rho * laplacian + forcing / rho
 * (equations.py:77-116, the explicit_terms_with_same_bcs seam). */
int jcfd_fvm_explicit_terms(jcfd_fvm_plan* plan, const jcfd_tensor* u, const jcfd_tensor* v,
                            jcfd_tensor* du, jcfd_tensor* dv, void* stream);

/* Iterations of the plan's last CG solve (pressure == JCFD_PRESSURE_CG).
 * Diagnostic with no counterpart in the reference (jax's cg does not report
 * them).  The CG plan synchronises its stream between chunks of 64
 * iterations to read its convergence flag. */
int32_t jcfd_fvm_cg_iterations(const jcfd_fvm_plan* plan);

/* Iterations of the plan's last CG diffusion solve of velocity component
 * `component` (0: u, 1: v; diffusion == JCFD_DIFFUSION_IMPLICIT_CG).  The same
 * kind of diagnostic; the first convergence check comes after 8 iterations. */
int32_t jcfd_fvm_diffusion_cg_iterations(const jcfd_fvm_plan* plan, int32_t component);

/* pressure.projection(v, solve) (pressure.py:181-198) with the plan's solve
 * (solve_fast_diag or solve_cg). */
int jcfd_fvm_projection(jcfd_fvm_plan* plan, const jcfd_tensor* u, const jcfd_tensor* v,
                        jcfd_tensor* u_out, jcfd_tensor* v_out, void* stream);

/* diffusion.solve_fast_diag(v, viscosity, dt) (diffusion.py:170-212),
 * periodic: each component becomes (1 - viscosity dt L)^-1 u, computed as
 * u + irfft2(D rfft2(u)); or, for JCFD_DIFFUSION_IMPLICIT_CG plans,
 * diffusion.solve_cg (diffusion.py:137-167), CG from x0 = u with the plan's
 * diff_cg_* tolerances.  Implicit-diffusion plans only (the plan's viscosity
 * and dt).  u_out/v_out may alias u/v. */
int jcfd_fvm_diffuse(jcfd_fvm_plan* plan, const jcfd_tensor* u, const jcfd_tensor* v,
                     jcfd_tensor* u_out, jcfd_tensor* v_out, void* stream);

/* fd.divergence of a staggered velocity (finite_differences.py:136-143). */
int jcfd_fvm_divergence(jcfd_fvm_plan* plan, const jcfd_tensor* u, const jcfd_tensor* v,
                        jcfd_tensor* out, void* stream);

/* Reverse (vector-Jacobian) step of the forward-Euler FVM step with
 * advect_van_leer: given the state (u, v) a step started from and the
 * cotangent (lam_u, lam_v) of the state it produced, writes the cotangent
 * of (u, v): lam* = P lam, out = lam* + dt J(u, v)^T lam*.  The quantity
 * jax.grad forms through one step of semi_implicit_navier_stokes
 * (equations.py:120-151, time_stepping.py:88-118); J^T follows jax's
 * differentiation rules for where / sign / abs / division.  Single-instance
 * fast-diagonal plans.  Outputs must not alias inputs. */
int jcfd_fvm_step_vjp(jcfd_fvm_plan* plan, const jcfd_tensor* u, const jcfd_tensor* v,
                      const jcfd_tensor* lam_u, const jcfd_tensor* lam_v, jcfd_tensor* out_u,
                      jcfd_tensor* out_v, void* stream);

/* Cotangent of fd.curl_2d: (u_bar, v_bar) from w_bar. */
int jcfd_fvm_curl_vjp(jcfd_fvm_plan* plan, const jcfd_tensor* w_bar, jcfd_tensor* u_bar,
                      jcfd_tensor* v_bar, void* stream);

/* advection.advect_step_semilagrangian(c, v, dt) (advection.py:345-382) on
 * a periodic grid whose domain starts at zero: c (offset offset_x, offset_y)
 * sampled at the departure points x - dt v(x) by bilinear interpolation with
 * periodic wrap (map_coordinates order 1, mode 'wrap').  Returns the advected
 * field itself, not a time derivative.  u, v: the staggered velocity at
 * offsets (1, 0.5) and (0.5, 1); out must not alias an input. */
int jcfd_fvm_advect_semilagrangian(jcfd_fvm_plan* plan, const jcfd_tensor* c, double offset_x,
                                   double offset_y, const jcfd_tensor* u, const jcfd_tensor* v,
                                   double dt, jcfd_tensor* out, void* stream);

/* fd.curl_2d of a staggered velocity (finite_differences.py:186-193),
 * offset (1, 1). */
int jcfd_fvm_curl(jcfd_fvm_plan* plan, const jcfd_tensor* u, const jcfd_tensor* v,
                  jcfd_tensor* out, void* stream);

/* out = irfft2(multiplier * rfft2(x)), multiplier float32 (n, n/2 + 1) on the
 * device: the spectral filter of filter_utils.filter (filter_utils.py:32-42)
 * for a multiplier that is even in k (then ifftn(fftn(x) m).real equals it). */
int jcfd_fvm_spectral_filter(jcfd_fvm_plan* plan, const jcfd_tensor* x,
                             const jcfd_tensor* multiplier, jcfd_tensor* out, void* stream);

/* Mixed-size groups: one step of many FVM instances of different grid sizes
 * (and different physics), each of them a slot of some plan, as ONE
 * launch per kernel class (one workgroup per work item).  The reference steps one field per
 * call (funcutils.repeated over step_fn, funcutils.py:82-88; a batch of
 * same-size fields would be jax.vmap of that); there is no counterpart for
 * uneven sizes.  Every instance gets exactly the arithmetic of jcfd_fvm_step
 * on that instance alone. */
typedef struct jcfd_fvm_group jcfd_fvm_group;

/* One plan's share of a group step: advance, in place, slots[0..n_slots) of
 * the plan's (batch, n, n) fields u and v.  `slots` is HOST memory (int32,
 * distinct, in [0, batch)).  Fast-diagonal plans with n <= 512. */
typedef struct {
  jcfd_fvm_plan* plan;
  jcfd_tensor* u;
  jcfd_tensor* v;
  const int32_t* slots;
  int32_t n_slots;
} jcfd_fvm_group_member;

/* The group holds the device work list (re-uploaded only when the member
 * set changes) and a captured graph of 32 steps (on by default). */
int jcfd_fvm_group_create(jcfd_fvm_group** group);
int jcfd_fvm_group_destroy(jcfd_fvm_group* group);
int jcfd_fvm_group_set_graphs(jcfd_fvm_group* group, int32_t enabled);
/* nsteps steps of every member instance.  No (plan, slot) and no (u, slot)
 * may appear twice.  Calls on one group must be ordered on one stream. */
int jcfd_fvm_group_step(jcfd_fvm_group* group, const jcfd_fvm_group_member* members,
                        int32_t n_members, int64_t nsteps, void* stream);

/* ------------------------------------------- learned interpolation (ml/) */
/* The learned-interpolation FVM step: modular_navier_stokes_model
 * (ml/equations.py:92-146) with self_advection(modular_self_advection(
 * FusedLearnedInterpolation)) (ml/advections.py:48-85,
 * ml/interpolations.py:31-121) as the convection, forward Euler with
 * explicit diffusion (semi_implicit_navier_stokes) or
 * implicit_diffusion_navier_stokes with diffusion.solve_fast_diag, and the
 * fast-diagonal projection.  Each step runs the CNN tower
 * (ml/towers.py:144-262: periodic 3x3 convolutions with relu between them)
 * on (u, v) on the matrix cores, turns its output into the 8
 * interpolations' stencil coefficients, interpolates u and v to the
 * control-volume faces (4 x 4 stencils, ml/layers.py:586-642) and advects
 * (base/advection.py:34-116). */
#define JCFD_LI_MAX_LAYERS 16

typedef struct {
  jcfd_fvm_config fvm;       /* physics: n >= 32, batch, dt, density, viscosity,
                                forcing, diffusion (explicit or implicit fast-
                                diagonal); pressure must be fast-diagonal, the
                                stepper forward Euler; convect is ignored */
  int32_t n_layers;          /* convolutions of the tower (1 .. 16): 2 -> hidden,
                                hidden -> hidden ..., hidden -> output */
  int32_t hidden_channels;   /* 16, 32 or 64 */
  int32_t n_logits;          /* the tower's own output channels (8 interpolations
                                x subspace size: 120 for 4 x 4 stencils and two
                                tags), <= 128 */
  int32_t interp_key[8];     /* first coefficient channel (16 per interpolation,
                                in the tower's split order) of the interpolation
                                each face flux uses: u's x face (advected u,
                                velocity u), u's y face (advected u, velocity
                                v), v's x face (advected v, velocity u), v's y
                                face (advected v, velocity v) */
} jcfd_li_config;

typedef struct jcfd_li_plan jcfd_li_plan;

/* Build a plan (device scratch for the tower's activations, the 128
 * coefficient planes per instance and the explicit terms).  Replaces the
 * factory modular_navier_stokes_model (ml/equations.py:92-146). */
int jcfd_li_plan_create(const jcfd_li_config* config, jcfd_li_plan** plan);
int jcfd_li_plan_destroy(jcfd_li_plan* plan);
int64_t jcfd_li_plan_device_bytes(const jcfd_li_plan* plan);
/* Weights of convolution `layer` as HOST float32 arrays in haiku's Conv2D
 * layout (3, 3, cin, cout) and its bias (cout) (ml/layers.py:24-116).  For
 * the output layer (layer n_layers - 1) these are the coefficient layer:
 * cout = 128, the tower's output weights times each interpolation's
 * constraint nullspace (and bias . nullspace + the constraint's bias),
 * which the caller forms (PolynomialConstraint, ml/layers.py:451-544). */
int jcfd_li_set_layer(jcfd_li_plan* plan, int32_t layer, const float* w, const float* b);
/* The tower's own output layer (3, 3, cin, n_logits), (n_logits): only
 * jcfd_li_tower's logits output uses it. */
int jcfd_li_set_logits_layer(jcfd_li_plan* plan, const float* w, const float* b);
/* nsteps learned-interpolation steps of (u, v) -> (u_out, v_out), fields
 * (batch, n, n) (or (n, n) for one instance).  Replaces
 * funcutils.repeated(step_fn, nsteps) of the modular model. */
int jcfd_li_step(jcfd_li_plan* plan, const jcfd_tensor* u, const jcfd_tensor* v,
                 jcfd_tensor* u_out, jcfd_tensor* v_out, int64_t nsteps, void* stream);
/* The tower on (u, v) into out (batch * 128, n, n) planes (instance b's
 * channel c at plane 128 b + c): the logits (first n_logits channels; the
 * tower network's output, interpolations.py:93-96) when coefficients == 0,
 * else the 8 x 16 stencil coefficients (bias + logits . nullspace). */
int jcfd_li_tower(jcfd_li_plan* plan, const jcfd_tensor* u, const jcfd_tensor* v,
                  jcfd_tensor* out, int32_t coefficients, void* stream);
/* The explicit terms of the step: convection by the learned interpolations
 * plus diffusion (explicit plans) and forcing, divided by density
 * (equations.py:77-116). */
int jcfd_li_explicit_terms(jcfd_li_plan* plan, const jcfd_tensor* u, const jcfd_tensor* v,
                           jcfd_tensor* du, jcfd_tensor* dv, void* stream);
/* One interpolation (0 .. 7, the tower's split order) of field x with the
 * coefficient planes of jcfd_li_tower(coefficients = 1): the interpolator
 * of FusedLearnedInterpolation.__call__ (interpolations.py:98-121). */
int jcfd_li_interpolate(jcfd_li_plan* plan, const jcfd_tensor* coefficients, const jcfd_tensor* x,
                        int32_t interpolation, jcfd_tensor* out, void* stream);

/* ------------------------------------------------------- spectral path */
#define JCFD_MAX_RK_STAGES 8

typedef struct {
  int32_t n;               /* cells per side; square periodic grid */
  double dx, dy;           /* grid steps */
  double viscosity, drag;  /* NavierStokes2D(viscosity, drag) */
  int32_t smooth;          /* 1: apply the 2/3 filter (the reference applies it
                              whenever smooth is not None) */
  double dt;               /* crank_nicolson_rk4 time step */
  int32_t batch;           /* instances advanced per call (same n), >= 1 */
  int32_t n_forcing;       /* non-zero entries of the constant forcing curl */
  const int32_t* forcing_index; /* host, 2 * n_forcing ints: (p, c), 0 <= c <= n/2 */
  const float* forcing_value;   /* host, 2 * n_forcing floats: (re, im) */
  /* Low-storage Runge-Kutta / Crank-Nicolson tableau
   * (spectral/time_stepping.py:117-170): rk_stages in 1..JCFD_MAX_RK_STAGES
   * with rk_stages + 1 alphas, rk_stages betas and rk_stages gammas; 0
   * selects crank_nicolson_rk4's Carpenter-Kennedy coefficients
   * (spectral/time_stepping.py:186-197) and ignores the arrays. */
  int32_t rk_stages;
  double rk_alpha[JCFD_MAX_RK_STAGES + 1];
  double rk_beta[JCFD_MAX_RK_STAGES];
  double rk_gamma[JCFD_MAX_RK_STAGES];
} jcfd_ns2d_config;

typedef struct jcfd_ns2d_plan jcfd_ns2d_plan;

/* Plan for low_storage_runge_kutta_crank_nicolson(..., NavierStokes2D(...),
 * dt) -- crank_nicolson_rk4 by default, crank_nicolson_rk3 or any other
 * tableau of up to JCFD_MAX_RK_STAGES stages
 * (spectral/time_stepping.py:117-197, spectral/equations.py:129-220): the
 * wavenumber tables, 2/3 filter bounds and the per-stage scalars. */
int jcfd_ns2d_plan_create(const jcfd_ns2d_config* config, jcfd_ns2d_plan** plan);
int jcfd_ns2d_plan_destroy(jcfd_ns2d_plan* plan);
int64_t jcfd_ns2d_plan_device_bytes(const jcfd_ns2d_plan* plan);
/* As jcfd_fvm_plan_set_graphs, for the spectral plan. */
int jcfd_ns2d_plan_set_graphs(jcfd_ns2d_plan* plan, int32_t enabled);

/* Tile shapes and FFT radix schedule of the step kernels (see
 * jcfd_fvm_tiles; no counterpart in the reference).  Tile shapes never
 * change results; radix 4 changes the FFT rounding. */
typedef struct {
  int32_t cols;  /* packed spectrum columns per workgroup of the column kernels: 1..16 */
  int32_t rows;  /* rows per workgroup of the row (nonlinear product) kernel: 1, 2, 4 */
  int32_t radix; /* largest FFT radix: 8 or 4 */
} jcfd_ns2d_tiles;
int jcfd_ns2d_plan_get_tiles(const jcfd_ns2d_plan* plan, jcfd_ns2d_tiles* tiles);
int jcfd_ns2d_plan_set_tiles(jcfd_ns2d_plan* plan, const jcfd_ns2d_tiles* tiles);

/* nsteps steps (of the plan's tableau) of every instance; w_hat/w_out complex64 of shape
 * (batch, n, n/2 + 1), or (n, n/2 + 1) when batch == 1.  w_out may alias
 * w_hat.  funcutils.repeated(step_fn, nsteps) (funcutils.py:82-88). */
int jcfd_ns2d_step(jcfd_ns2d_plan* plan, const jcfd_tensor* w_hat, jcfd_tensor* w_out,
                   int64_t nsteps, void* stream);

/* Continuous-batching form: advances in place only the listed slots of a
 * (batch, n, n/2 + 1) state (see jcfd_fvm_step_slots). */
int jcfd_ns2d_step_slots(jcfd_ns2d_plan* plan, jcfd_tensor* w_hat, const jcfd_tensor* slots,
                         int64_t nsteps, void* stream);

/* NavierStokes2D.explicit_terms (spectral/equations.py:160-184). */
int jcfd_ns2d_explicit_terms(jcfd_ns2d_plan* plan, const jcfd_tensor* w_hat, jcfd_tensor* out,
                             void* stream);

/* jnp.fft.rfftn / irfftn over the last two axes of a batch (the transforms
 * spectral/equations.py:163-180 applies; also how the reference's spectral
 * runs build their initial state, spectral/equations_test.py:216-219).
 * x: float32 (batch, n, n) or (n, n); x_hat: complex64 (batch, n, n/2 + 1).
 * irfft2 ignores the imaginary parts of the DC and Nyquist columns after the
 * axis-0 inverse, as pocketfft's c2r does. */
int jcfd_ns2d_rfft2(jcfd_ns2d_plan* plan, const jcfd_tensor* x, jcfd_tensor* x_hat, void* stream);
int jcfd_ns2d_irfft2(jcfd_ns2d_plan* plan, const jcfd_tensor* x_hat, jcfd_tensor* x, void* stream);

/* Mixed-size spectral groups: one step of many NavierStokes2D instances of
 * different grid sizes (and physics), each a slot of some plan, as ONE
 * launch per kernel class and stage over a work list of (instance, tile)
 * items of equal work, one workgroup per item.  No counterpart in the reference
 * (one field per call, funcutils.repeated, funcutils.py:82-88); every
 * instance gets exactly the arithmetic of jcfd_ns2d_step on it alone. */
typedef struct jcfd_ns2d_group jcfd_ns2d_group;

/* One plan's share: advance, in place, slots[0..n_slots) of the plan's
 * (batch, n, n/2 + 1) complex64 state w_hat.  `slots` is HOST memory
 * (int32, distinct, in [0, batch)).  Radix-8 plans (the default); every
 * member's plan must step the same number of tableau stages. */
typedef struct {
  jcfd_ns2d_plan* plan;
  jcfd_tensor* w_hat;
  const int32_t* slots;
  int32_t n_slots;
} jcfd_ns2d_group_member;

/* The group holds the device work list (re-uploaded only when the member
 * set changes) and a captured graph of 32 steps (on by default). */
int jcfd_ns2d_group_create(jcfd_ns2d_group** group);
int jcfd_ns2d_group_destroy(jcfd_ns2d_group* group);
int jcfd_ns2d_group_set_graphs(jcfd_ns2d_group* group, int32_t enabled);
/* nsteps steps of every member instance.  No (plan, slot) and no
 * (w_hat, slot) may appear twice.  Calls on one group must be ordered on one
 * stream, and a plan must not be stepped on another stream meanwhile (the
 * group steps its instances in the plan's internal buffers). */
int jcfd_ns2d_group_step(jcfd_ns2d_group* group, const jcfd_ns2d_group_member* members,
                         int32_t n_members, int64_t nsteps, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* JCFD_H_ */
