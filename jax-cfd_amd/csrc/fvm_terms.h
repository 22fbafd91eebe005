// Per-cell FVM terms shared by the step kernels (fvm.hip) and the
// learned-interpolation kernels (ml.hip): the LDS row window and the
// diffusion + forcing part of the explicit terms, in the reference's
// operation order.
#pragma once
#include <hip/hip_runtime.h>

#include "jcfd_internal.h"

#pragma clang fp contract(off)

namespace jcfd {

// Rows of an LDS staging window (local row lr; the kernel fixes which
// global row that is, e.g. i0 - 3 + lr for the step kernels).
template <int N>
struct Window {
  const float* su;  // u rows
  const float* sv;  // v rows
  __device__ __forceinline__ float U(int lr, int j) const { return su[lr * N + (j & (N - 1))]; }
  __device__ __forceinline__ float V(int lr, int j) const { return sv[lr * N + (j & (N - 1))]; }
};

// f / rho; x / 1 is x exactly, so the common rho = 1 skips the division.
__device__ __forceinline__ float over_rho(float f, float rho) { return rho == 1.0f ? f : f / rho; }

// k += nu/rho lap + forcing/rho for one component, in the reference's order
// (equations.py:102-110, finite_differences.py:127-133 for the Laplacian).
// V_COMP selects v (no forcing column) instead of u.
template <bool V_COMP, typename WIN>
__device__ __forceinline__ float add_diffusion_forcing(const WIN& w, int lr, int j,
                                                       const FvmParams& p, float fx_j, float k) {
  auto c = [&](int r, int jj) { return V_COMP ? w.V(r, jj) : w.U(r, jj); };
  if (p.has_diffusion) {
    const float cc = c(lr, j);
    const float lap = ((-2.f * cc) * p.lap_sum + (c(lr - 1, j) + c(lr + 1, j)) * p.inv_dx_sq) +
                      (c(lr, j - 1) + c(lr, j + 1)) * p.inv_dy_sq;
    k = k + p.nu * lap;
  }
  if (p.has_forcing) {
    const float f = p.lin * c(lr, j) + (V_COMP ? 0.f : fx_j);
    k = k + over_rho(f, p.rho);
  }
  return k;
}

}  // namespace jcfd
