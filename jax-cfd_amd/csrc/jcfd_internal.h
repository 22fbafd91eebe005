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
  float inv_dx_sq, inv_dy_sq, lap_sum;  // (1/dx)^2, (1/dy)^2 and their fp32 sum
  float dt;          // time step
  float dt_over_dx, dt_over_dy;         // Courant factors dt/h (rounded from fp64)
  float nu;          // viscosity / density
  float rho;         // density
  float lin;         // linear forcing coefficient
  int has_diffusion;
  int has_forcing;
  int convect;       // JCFD_CONVECT_*
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
  const float* lam;      // Lambda^+ / n^2, (n, n/2 + 1) floats
  const float2* tw;      // twiddles exp(-2 pi i k / n), n entries
};

// Kernel launchers (fvm.hip).  All are stream-ordered and asynchronous.
int fvm_launch_step(const FvmParams& p, const FvmBuffers& b, hipStream_t s);
// One kernel class of the step (0 rows-forward, 1 Poisson columns, 2 rows-project).
int fvm_launch_kernel(const FvmParams& p, const FvmBuffers& b, int which, hipStream_t s);
int fvm_launch_explicit_terms(const FvmParams& p, const float* u, const float* v,
                              float* du, float* dv, const float* fx, hipStream_t s);
int fvm_launch_projection(const FvmParams& p, const float* u, const float* v, float* u_out,
                          float* v_out, float2* spec, const float* lam, const float2* tw,
                          hipStream_t s);
int fvm_launch_divergence(const FvmParams& p, const float* u, const float* v, float* out,
                          hipStream_t s);
int fvm_launch_curl(const FvmParams& p, const float* u, const float* v, float* out, hipStream_t s);
int fvm_launch_filter(const FvmParams& p, const float* x, const float* multiplier, float* out,
                      float2* spec, const float2* tw, hipStream_t s);
bool fvm_size_supported(int n);

}  // namespace jcfd
