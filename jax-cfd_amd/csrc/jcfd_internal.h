// Internal declarations shared by the HIP translation units of libjcfd_amd.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <string>

#include "jcfd.h"

namespace jcfd {

// Thread-local error message behind jcfd_last_error().
void set_error(const std::string& msg);
int fail(int code, const std::string& msg);
// Map a HIP error to JCFD_ERR_DEVICE with context.
int check_hip(hipError_t e, const char* what);

// Device-side parameters of one FVM instance (all fp32 as used in the
// reference's fp32 arithmetic; host-side fp64 values are rounded once).
struct FvmParams {
  int n;             // cells per side
  float dx, dy;      // grid steps
  float rcp_dx, rcp_dy;  // RN(1/dx), RN(1/dy) (host fp32 division), for over_step
  float inv_dx_sq, inv_dy_sq, lap_sum;  // (1/dx)^2, (1/dy)^2 and their fp32 sum
  float dt;          // time step
  float dt_over_dx, dt_over_dy;         // Courant factors dt/h (rounded from fp64)
  float nu;          // viscosity / density
  float rho;         // density
  float lin;         // linear forcing coefficient
  int has_diffusion;
  int has_forcing;
  int convect;       // JCFD_CONVECT_*
  // Step-kernel tile shapes and FFT radix schedule (host side; see
  // fvm_default_tiles / fvm_tile_supported).
  int rb_fwd, w_cols, rb_proj, radix;
};

// x / h correctly rounded, from r = RN(1/h): q = RN(x r) lies within one ulp
// of x / h, and one remainder step RN(q + RN(x - q h) r) is then the
// correctly rounded quotient (Markstein's theorem).  The theorem needs the
// remainder x - q h to be exact; it is a multiple of about 2^(e_x - 47), so
// it can fall under the subnormal grid once |x| < 2^-102.  Then the
// correction is lost and the result is RN(x r), at most one ulp of the
// quotient from x / h; and x = -0 gives +0.  So the result is bitwise x / h
// for |x| >= 2^-100 (checked against IEEE division on the CPU,
// tests/test_over_step_cpu.py), not unconditionally: flow differences that
// small without being zero are the exception.  Three instructions where the
// IEEE division sequence takes about ten.
__device__ __forceinline__ float over_step(float x, float h, float r) {
  const float q = x * r;
  return __builtin_fmaf(__builtin_fmaf(-q, h, x), r, q);
}

// One Runge-Kutta combination u* = u0 + dt * (0 + c[0] k[0] + ... ) with the
// reference's term order (zero coefficients left out by the host); k arrays
// are batch-sized like the state (instance s at s * n * n).
struct RkTerms {
  const float* ku[4];
  const float* kv[4];
  float c[4];
  int n;
};

// Pointers of one FVM instance for one step.
struct FvmBuffers {
  const float* u;        // state in (n, n)
  const float* v;
  float* us;             // u*, v* scratch
  float* vs;
  float2* spec;          // packed half spectrum, (n, n/2)
  float* u_out;          // state out
  float* v_out;
  const float* fx;       // forcing column f_x(y_j), n floats (may be null)
  // Spectrum multiplier of the Poisson column kernel, entry (p, c) of an
  // (n, n/2 + 1) array: lam_mode 0 a table lam, row-major; 1 the table lam
  // stored column by column (c * n + p); 2 Lambda^+ / n^2 computed from
  // lam_xy (x eigenvalues, n doubles, then y eigenvalues, n/2 + 1).
  const float* lam;
  const double* lam_xy;
  int lam_mode;
  const float2* tw;      // twiddles exp(-2 pi i k / n), n entries
  // Batches: instance k of a launch (blockIdx.y = k) is slots[k], or k when
  // slots is null; fields of instance s start at s * n * n (spectrum:
  // s * n * n / 2).
  const int* slots;
  int count;  // instances in this launch (gridDim.y); <= 0 means 1
  float* rhs;  // CG variant: div u* (n, n), single instance
  const float* q;  // CG variant: pressure correction (n, n), single instance
  const RkTerms* rk;  // Runge-Kutta combination (device memory) of rows-forward mode 5
  size_t inst;        // instance index of this block (set by fvm_instance)
};

// Conjugate-gradient pressure solve (cg.hip), one instance.
struct CgBuffers {
  float* x;            // solution q
  float* r;            // residual
  float* p[2];         // search directions (ping-pong by iteration parity)
  float* ap;           // A p
  float* part_rr[2];   // per-block partial sums of r.r (by parity)
  float* part_pap;     // per-block partial sums of p.Ap
  float* part_bb;      // per-block partial sums of b.b
  int* status;         // [0] done flag, [1] iterations run, [2] grid-barrier arrivals, [3] barrier timeout
};
int cg_blocks(int n);
int cg_launch_init(const FvmParams& p, const float* rhs, const CgBuffers& c, float rtol, float atol,
                   int maxiter, hipStream_t s);
// The whole solve (after cg_launch_init / cg_launch_init_helmholtz) as one
// cooperative launch; JCFD_ERR_UNSUPPORTED (no error recorded) when the
// device cannot keep all its workgroups resident.
int cg_launch_persistent(const FvmParams& p, const CgBuffers& c, float rtol, float atol, int maxiter,
                         bool helm, float coef, hipStream_t s);
// One CG iteration k (two kernels); a no-op once converged.
int cg_launch_iteration(const FvmParams& p, const CgBuffers& c, int k, float rtol, float atol,
                        int maxiter, hipStream_t s);
// The implicit diffusion system (I - coef L) x = b on the same buffers:
// x0 = b, then iterations as above.
int cg_launch_init_helmholtz(const FvmParams& p, const float* b, const CgBuffers& c, float coef,
                             hipStream_t s);
int cg_launch_iteration_helmholtz(const FvmParams& p, const CgBuffers& c, int k, float rtol, float atol,
                                  int maxiter, float coef, hipStream_t s);
int fvm_launch_adjoint_explicit(const FvmParams& p, const float* u, const float* v, const float* lu,
                                const float* lv, float* ou, float* ov, hipStream_t s);
int fvm_launch_curl_vjp(const FvmParams& p, const float* wb, float* ub, float* vb, hipStream_t s);
int fvm_launch_forward_rhs(const FvmParams& p, const FvmBuffers& b, hipStream_t s);
int fvm_launch_project_q(const FvmParams& p, const FvmBuffers& b, hipStream_t s);

inline int nb(const FvmBuffers& b) { return b.count > 0 ? b.count : 1; }

// Workgroups are dispatched round-robin over the 8 XCDs (each with its own
// L2): workgroup b of a row of nb runs on XCD b % 8.  This maps it to tile
// (b % 8) * (nb / 8) + b / 8 so every XCD gets a contiguous run of tiles, and
// neighbouring tiles -- which share halo rows or cache lines -- share an L2.
// (blockIdx.y rows of nb workgroups each keep the same XCD pattern when nb is
// a multiple of 8; otherwise the identity is kept.)
__device__ __forceinline__ int xcd_tile(int b, int nb) {
  return (nb % 8 != 0) ? b : (b % 8) * (nb / 8) + b / 8;
}

// p, known to point at global memory, as a generic pointer the compiler can
// trace back to the global address space.  Pointers read from memory (work
// list descriptors) are otherwise generic to the compiler: its flat
// accesses count on the LDS wait counter too, so every wait on them waits
// for all outstanding memory operations.
template <typename T>
__device__ __forceinline__ T* as_global(T* p) {
  using G = __attribute__((address_space(1))) T;
  G* gp = (G*)reinterpret_cast<uintptr_t>(p);
  return (T*)gp;
}

// Per-instance view of FvmBuffers for the instance this block works on.
__device__ __forceinline__ FvmBuffers fvm_instance(FvmBuffers b, int n) {
  const size_t s = b.slots ? static_cast<size_t>(b.slots[blockIdx.y]) : blockIdx.y;
  const size_t st = s * n * n;
  if (b.u) b.u += st;
  if (b.v) b.v += st;
  if (b.us) b.us += st;
  if (b.vs) b.vs += st;
  if (b.u_out) b.u_out += st;
  if (b.v_out) b.v_out += st;
  if (b.spec) b.spec += st / 2;
  b.inst = s;
  return b;
}

// Kernel launchers (fvm.hip).  All are stream-ordered and asynchronous.
int fvm_launch_step(const FvmParams& p, const FvmBuffers& b, hipStream_t s);
// Implicit diffusion solve in place on b.u_out / b.v_out (every instance of b)
// with the multiplier table `mult` ((n, n/2 + 1) floats, 1/n^2 folded in,
// stored column by column).
int fvm_launch_diffuse(const FvmParams& p, const FvmBuffers& b, const float* mult, hipStream_t s);
// Runge-Kutta pieces (every instance of b, default tile shapes):
// explicit terms of (b.u, b.v) into (b.us, b.vs) rows of all instances, and
// the combination b.u + dt * sum(b.rk) -> (b.us, b.vs) + its div spectrum,
// Poisson columns and velocity correction into (b.u_out, b.v_out).
int fvm_launch_rk_explicit(const FvmParams& p, const FvmBuffers& b, hipStream_t s);
int fvm_launch_rk_project(const FvmParams& p, const FvmBuffers& b, hipStream_t s);
// One kernel class of the step (0 rows-forward, 1 Poisson columns, 2 rows-project).
int fvm_launch_kernel(const FvmParams& p, const FvmBuffers& b, int which, hipStream_t s);
int fvm_launch_explicit_terms(const FvmParams& p, const float* u, const float* v,
                              float* du, float* dv, const float* fx, hipStream_t s);
// lam_xy: the plan's Laplacian eigenvalues (x, then y), for Lambda^+.
int fvm_launch_projection(const FvmParams& p, const float* u, const float* v, float* u_out,
                          float* v_out, float2* spec, const double* lam_xy, const float2* tw,
                          hipStream_t s);
int fvm_launch_divergence(const FvmParams& p, const float* u, const float* v, float* out,
                          hipStream_t s);
int fvm_launch_curl(const FvmParams& p, const float* u, const float* v, float* out, hipStream_t s);
// interpolation.linear along one axis (interpolation.py:36-64): a shift by f
// (offset difference an integer, blend = 0) or fw * shift(f) + cw * shift(c)
struct SlAxis {
  int blend, f, c;
  float fw, cw;
};
struct SlParams {
  SlAxis u0, u1, v0, v1;  // u and v to c's offset, axis 0 then axis 1
  float ox, oy, hx, hy, dt;
};
int fvm_launch_semilagrangian(int n, const float* c, const float* u, const float* v, float* out,
                              const SlParams& q, hipStream_t s);
int fvm_launch_filter(const FvmParams& p, const float* x, const float* multiplier, float* out,
                      float2* spec, const float2* tw, hipStream_t s);
bool fvm_size_supported(int n);
// Default tile shapes and radix of size p->n into p.
int fvm_default_tiles(FvmParams* p);
// Whether kernel class `which` (0 rows-forward, 1 Poisson columns,
// 2 rows-project) is built with tile `tile` for size n.
bool fvm_tile_supported(int n, int which, int tile);

// Mixed-size group step (fvm.hip): one launch per kernel class, one
// workgroup per work item
// (phase 0 rows-forward, 1 Poisson columns, 2 rows-project) advances every
// instance of a device-resident work list, whatever its grid size.
struct FvmGroupDesc {
  FvmParams p;
  FvmBuffers b;  // pointers already offset to the instance; u_out == u, v_out == v
};
struct FvmGroupItem {
  int desc;  // index into the descriptor table
  int tile;  // row block (phases 0, 2) or column tile (phase 1)
};
struct FvmGroupQueue {
  const FvmGroupDesc* descs;
  const FvmGroupItem* items[3];  // work list of each phase
  const int* n_items;            // device int[3]
};
constexpr int kFvmGroupMaxN = 512;
bool fvm_group_size_supported(int n);
// Work items per instance of grid size n in `phase`.
int fvm_group_tiles(int n, int phase);
// Launch one phase: grid = the phase's item count (one workgroup per item).
int fvm_group_launch(int phase, int max_n, const FvmGroupQueue& q, int grid, hipStream_t s);

// Learned interpolation (ml.hip).  The tower's output layer carries the
// 8 interpolations' 16 stencil coefficients each (constraint folded in).
constexpr int kLiCoefChannels = 128;
struct LiConvArgs {
  int n;
  const float* in;    // (batch, cin, n, n); the first layer: u (batch, n, n)
  const float* in_v;  // the first layer: v (batch, n, n); otherwise null
  const float* w;     // MFMA operand order (li_conv_weight_layout)
  const float* b;     // [cout]
  float* out;         // (batch, cout, n, n)
};
struct LiExplicitArgs {
  const float* coef;  // (launch instances, 128, n, n)
  const float* u;
  const float* v;
  float* ku;          // explicit terms out, (batch, n, n)
  float* kv;
  const float* fx;    // forcing column (null without forcing)
  const int* slots;   // instance of launch row blockIdx.y, or null = blockIdx.y
  // first coefficient channel of the interpolations each face flux uses:
  // u: x face (advected, velocity), y face (advected, velocity); then v
  int key[8];
};
bool li_conv_supported(int cin, int cout, bool relu);
int li_launch_conv(const LiConvArgs& a, int cin, int cout, bool relu, int batch, hipStream_t s);
// Host: weights [tap = 3 kh + kw][cinp][cout] (cinp = input channels padded
// to a multiple of 4) -> the device layout li_conv3x3 reads (same size).
void li_conv_weight_layout(const float* w, int cinp, int cout, float* frag);
int li_launch_explicit(const FvmParams& p, const LiExplicitArgs& a, int batch, hipStream_t s);
int li_launch_interpolate(int n, const float* coef, const float* x, int key, float* out, int batch,
                          hipStream_t s);

// Device-side parameters of a pseudospectral stage (spectral.hip).  The
// wavenumber tables and multipliers follow the reference's fp32 values;
// beta, g = gamma dt and mu are the stage's low-storage CN-RK scalars.
struct NsParams {
  const float* kx;        // fftfreq(n, dx), n values
  const float* ky;        // rfftfreq(n, dy), n/2+1 values
  const float2* tw;       // twiddles exp(-2 pi i k / n)
  const int* force_idx;   // (p, c) pairs of the forcing-curl entries
  const float2* force_val;
  int n_force;
  float two_pi;           // fp32(2 pi)
  float neg_four_pi_sq;   // fp32((2 pi i)^2)
  float inv_nn;           // 1 / n^2 (exact)
  float nu, drag;
  int smooth;
  int filt_row_lo, filt_row_hi, filt_cols;
  float beta, g, mu;
  int read_h;
  const int* slots;       // instance of launch row blockIdx.y, or null = blockIdx.y
  // Step-kernel tile shapes and FFT radix schedule (host side; see
  // ns_default_tiles / ns_tile_supported).
  int w_cols, rb_rows, radix;
};

struct NsBuffers {
  const float2* u_in;  // state in, (batch, n, n/2+1) words, internal layout
  float2* u_out;       // state out, internal layout
  float2* h;           // low-storage RK register, internal layout
  float2* g;           // 4 packed spectra per instance, (batch, 4, n x n/2), tiled
  float2* pbuf;        // packed product spectrum, (batch, n x n/2), tiled
};

// Mixed-size spectral group (spectral.hip): one launch per kernel class of
// the step covers a work list of (instance, tile) items over instances of
// any built grid size, one workgroup per item; each item runs the per-size
// kernels' tile body.
// Descriptors are per (instance, stage): the member plan's stage parameters
// and the instance's buffers (internal state layout).
struct NsGroupDesc {
  NsParams q;
  int n;
  float2* ws;    // state, internal layout
  float2* h;     // low-storage RK register
  float2* g;     // four packed spectra
  float2* pbuf;  // packed product spectrum
};
struct NsGroupItem {
  int desc;
  int tile;  // column tile (list 0) or row block (list 1)
};
struct NsGroupQueue {
  const NsGroupDesc* descs;     // [instance][stage], JCFD_MAX_RK_STAGES per instance
  const NsGroupItem* items[2];  // 0: column tiles, 1: row blocks
  const int* n_items;           // device int[2]
};
// Kernel classes of a group step (as the per-size step): 0 inverse columns,
// 1 rows, 2 forward columns + update + next inverse, 3 forward columns +
// final update.
constexpr int kNsGroupClasses = 4;
bool ns_group_size_supported(int n);
int ns_group_tiles(int n, int list);
int ns_group_launch(int cls, int max_n, const NsGroupQueue& q, int stage, int grid, hipStream_t s);
bool ns_size_supported(int n);
// Default tile shapes and radix for size n into q.
int ns_default_tiles(NsParams* q, int n);
// Whether step kernel class `which` (0 cols-inverse, 1 rows, 2/3 cols-forward)
// is built with tile `tile` (columns per tile; rows per workgroup for 1).
bool ns_tile_supported(int n, int which, int tile);
// The step kernels keep the state (and h, G, P) in an internal layout of the
// same size as the external (n, n/2 + 1) one: column tiles of 4 words, each
// a contiguous block, then the Nyquist column.  ns_launch_state_convert maps
// `count` instances (slots[i], or i) between the two.
int ns_launch_state_convert(int n, bool to_internal, const float2* src, float2* dst, const int* slots,
                            int count, hipStream_t s);
// One step of an nstages-stage low-storage CN-RK tableau (stages[k] holds
// stage k's scalars); b.u_in / b.u_out / b.h in the internal state layout.
int ns_launch_step(int n, const NsParams* stages, int nstages, const NsBuffers& b, int batch,
                   hipStream_t s);
// Explicit terms of w (internal state layout) into out (external layout).
int ns_launch_explicit_terms(int n, const NsParams& q, const float2* w, float2* out, float2* g,
                             float2* pbuf, int batch, hipStream_t s);
int ns_launch_rfft2(int n, const NsParams& q, const float* x, float2* out, float2* pbuf, int batch,
                    hipStream_t s);
// One kernel class of the step: 0 cols-inverse, 1 rows, 2 cols-forward with
// update + next inverse, 3 cols-forward with the final update.
int ns_launch_kernel(int n, int which, const NsParams* stages, int nstages, const NsBuffers& b,
                     int batch, hipStream_t s);
int ns_launch_irfft2(int n, const NsParams& q, const float2* in, float* x, float2* pbuf, int batch,
                     hipStream_t s);

}  // namespace jcfd
