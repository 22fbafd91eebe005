// Hand-written adjoint of the FVM forward-Euler step (advect_van_leer).
//
// Forward: u_{s+1} = P(u_s + dt k(u_s)) with k = convection + nu lap + forcing
// (jax_cfd/base/equations.py:102-114, time_stepping.py:88-104) and P the
// projection (pressure.py:181-198).  P is self-adjoint (P = I - G L^+ D with
// D^T = -G for periodic differences and L^+ symmetric), so the reverse step
// is
//   lambda* = P lambda_{s+1}               (the forward projection kernels)
//   lambda_s = lambda* + dt J(u_s)^T lambda*
// and this file provides J^T: for every input cell it gathers, in a fixed
// order, the contributions of every Van Leer face flux that reads it
// (advection.py:210-342, the forward stencil of fvm.hip), each flux's local
// gradient evaluated with the differentiation rules of jax.grad on the
// reference code: sign' = 0; abs'(x) = +1 for x >= 0 and -1 otherwise;
// where(c, a, b) sends the cotangent to the selected branch only (so at
// den = 0 the guarded denominator is the constant 1 and d phi / d den =
// |num|).  The diffusion part is nu * lap (symmetric), the linear forcing
// part lin * I.  No atomics: results are deterministic.
#include <hip/hip_runtime.h>

#include "jcfd_internal.h"

#pragma clang fp contract(off)

namespace jcfd {
namespace {

constexpr int kThreads = 256;

__device__ __forceinline__ float sgnf(float x) { return (x > 0.f) ? 1.f : ((x < 0.f) ? -1.f : 0.f); }
// jax: d|x|/dx = select(x >= 0, 1, -1)
__device__ __forceinline__ float dabs(float x) { return x >= 0.f ? 1.f : -1.f; }

// Gradient of the Van Leer face flux (vl_flux in fvm.hip) with respect to its
// eight inputs: in = {A, B, c_pos, c_neg, cm2, cm1, c0, cp1}, U = .5 A + .5 B.
__device__ __forceinline__ void vl_flux_grad(const float (&in)[8], float h, float (&g)[8]) {
  const float U = 0.5f * in[0] + 0.5f * in[1];
  const bool pos = U > 0.f;
  const float cp = in[2], cn = in[3], cm2 = in[4], cm1 = in[5], c0 = in[6], cp1 = in[7];
  const float num = pos ? (cm1 - cm2) : (cp1 - c0);
  const float den = c0 - cm1;
  const float aden = fabsf(den), anum = fabsf(num);
  const bool guard = aden > 0.f;
  const float D = guard ? (aden + anum) : 1.0f;
  const float q = den / D;
  const float s = sgnf(den) + sgnf(num);
  const float phi = (num * s) * q;
  const float a = fabsf(U);
  const float pre = (0.5f * (1.f - h * a)) * a;
  // d/dU
  const float dpre_dU = dabs(U) * (0.5f - h * a);
  const float dF_dU = (pos ? cp : cn) + dpre_dU * phi;
  // through phi
  const float dD_dnum = guard ? dabs(num) : 0.f;
  const float dD_dden = guard ? dabs(den) : 0.f;
  const float dq_dnum = -den * dD_dnum / (D * D);
  const float dq_dden = 1.f / D - den * dD_dden / (D * D);
  const float dF_dnum = pre * (s * q + (num * s) * dq_dnum);
  const float dF_dden = pre * ((num * s) * dq_dden);
  const float np = pos ? 1.f : 0.f, nn = pos ? 0.f : 1.f;
  g[0] = 0.5f * dF_dU;
  g[1] = 0.5f * dF_dU;
  g[2] = pos ? U : 0.f;
  g[3] = pos ? 0.f : U;
  g[4] = -np * dF_dnum;
  g[5] = np * dF_dnum - dF_dden;
  g[6] = -nn * dF_dnum + dF_dden;
  g[7] = nn * dF_dnum;
}

struct Slot {
  int ro, co, f;  // row / column offset from the flux position, field (0 u, 1 v)
};

// Input slots of the four flux families relative to the flux position (a, b),
// exactly as fvm.hip's advect_van_leer face fluxes (face_fx_u .. face_fy_v) read them.
__constant__ Slot kSlots[4][8] = {
    // Fx_u(a, b): U = .5 u[a-1] + .5 u[a]; c_pos u[a-1], c_neg u[a]; limiter u[a-2..a+1]
    {{-1, 0, 0}, {0, 0, 0}, {-1, 0, 0}, {0, 0, 0}, {-2, 0, 0}, {-1, 0, 0}, {0, 0, 0}, {1, 0, 0}},
    // Fy_u(a, b): V = .5 v[a-1, b] + .5 v[a, b]; upwind u[a-1, b], u[a-1, b+1];
    // limiter u[a, b-1..b+2]
    {{-1, 0, 1}, {0, 0, 1}, {-1, 0, 0}, {-1, 1, 0}, {0, -1, 0}, {0, 0, 0}, {0, 1, 0}, {0, 2, 0}},
    // Fx_v(a, b): U = .5 u[a, b-1] + .5 u[a, b]; upwind v[a, b-1], v[a+1, b-1];
    // limiter v[a-1..a+2, b]
    {{0, -1, 0}, {0, 0, 0}, {0, -1, 1}, {1, -1, 1}, {-1, 0, 1}, {0, 0, 1}, {1, 0, 1}, {2, 0, 1}},
    // Fy_v(a, b): V = .5 v[a, b-1] + .5 v[a, b]; upwind v[a, b-1], v[a, b];
    // limiter v[a, b-2..b+1]
    {{0, -1, 1}, {0, 0, 1}, {0, -1, 1}, {0, 0, 1}, {0, -2, 1}, {0, -1, 1}, {0, 0, 1}, {0, 1, 1}},
};

template <int N>
__device__ __forceinline__ float at(const float* f, int i, int j) {
  return f[(i & (N - 1)) * N + (j & (N - 1))];
}

// Cotangent of flux family T at (a, b) from the cotangents of k_u, k_v.
template <int N>
__device__ __forceinline__ float flux_bar(int t, int a, int b, const float* lu, const float* lv,
                                          const FvmParams& p) {
  switch (t) {
    case 0: return over_step(at<N>(lu, a, b) - at<N>(lu, a - 1, b), p.dx, p.rcp_dx);
    case 1: return over_step(at<N>(lu, a - 1, b + 1) - at<N>(lu, a - 1, b), p.dy, p.rcp_dy);
    case 2: return over_step(at<N>(lv, a + 1, b - 1) - at<N>(lv, a, b - 1), p.dx, p.rcp_dx);
    default: return over_step(at<N>(lv, a, b) - at<N>(lv, a, b - 1), p.dy, p.rcp_dy);
  }
}

// (J^T lambda) at (i, j) for field `f` (convection part only).
template <int N>
__device__ float convection_vjp(int f, int i, int j, const float* u, const float* v, const float* lu,
                                const float* lv, const FvmParams& p) {
  float acc = 0.f;
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const float h = (t == 0 || t == 2) ? p.dt_over_dx : p.dt_over_dy;
#pragma unroll
    for (int s = 0; s < 8; ++s) {
      const Slot sl = kSlots[t][s];
      if (sl.f != f) continue;
      const int a = i - sl.ro, b = j - sl.co;  // flux position reading (i, j) in slot s
      float in[8], g[8];
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        const Slot q = kSlots[t][k];
        in[k] = at<N>(q.f == 0 ? u : v, a + q.ro, b + q.co);
      }
      vl_flux_grad(in, h, g);
      acc = acc + flux_bar<N>(t, a, b, lu, lv, p) * g[s];
    }
  }
  return acc;
}

// lambda_out = lambda + dt * J(u)^T lambda  (lambda = P lambda_{s+1} already).
template <int N>
__global__ __launch_bounds__(kThreads) void fvm_adjoint_explicit(FvmParams p, const float* __restrict__ u,
                                                                 const float* __restrict__ v,
                                                                 const float* __restrict__ lu,
                                                                 const float* __restrict__ lv,
                                                                 float* __restrict__ out_u,
                                                                 float* __restrict__ out_v) {
  const int e = blockIdx.x * kThreads + threadIdx.x;
  if (e >= N * N) return;
  const int i = e / N, j = e % N;
  float gu = convection_vjp<N>(0, i, j, u, v, lu, lv, p);
  float gv = convection_vjp<N>(1, i, j, u, v, lu, lv, p);
  if (p.has_diffusion) {
    const float lapu = ((-2.f * lu[e]) * p.lap_sum + (at<N>(lu, i - 1, j) + at<N>(lu, i + 1, j)) * p.inv_dx_sq) +
                       (at<N>(lu, i, j - 1) + at<N>(lu, i, j + 1)) * p.inv_dy_sq;
    const float lapv = ((-2.f * lv[e]) * p.lap_sum + (at<N>(lv, i - 1, j) + at<N>(lv, i + 1, j)) * p.inv_dx_sq) +
                       (at<N>(lv, i, j - 1) + at<N>(lv, i, j + 1)) * p.inv_dy_sq;
    gu = gu + p.nu * lapu;
    gv = gv + p.nu * lapv;
  }
  if (p.has_forcing) {
    gu = gu + (p.lin / p.rho) * lu[e];
    gv = gv + (p.lin / p.rho) * lv[e];
  }
  out_u[e] = lu[e] + p.dt * gu;
  out_v[e] = lv[e] + p.dt * gv;
}

// Cotangent of curl_2d (finite_differences.py:186-193): w = dv/dx - du/dy with
// forward differences -> u_bar[a,b] = (w_bar[a,b] - w_bar[a,b-1]) / dy,
// v_bar[a,b] = (w_bar[a-1,b] - w_bar[a,b]) / dx.
template <int N>
__global__ __launch_bounds__(kThreads) void fvm_curl_vjp(FvmParams p, const float* __restrict__ wb,
                                                         float* __restrict__ ub, float* __restrict__ vb) {
  const int e = blockIdx.x * kThreads + threadIdx.x;
  if (e >= N * N) return;
  const int i = e / N, j = e % N;
  ub[e] = over_step(wb[e] - at<N>(wb, i, j - 1), p.dy, p.rcp_dy);
  vb[e] = over_step(at<N>(wb, i - 1, j) - wb[e], p.dx, p.rcp_dx);
}

template <int N>
int adjoint_impl(const FvmParams& p, const float* u, const float* v, const float* lu, const float* lv,
                 float* ou, float* ov, hipStream_t s) {
  hipLaunchKernelGGL((fvm_adjoint_explicit<N>), dim3((N * N + kThreads - 1) / kThreads), dim3(kThreads),
                     0, s, p, u, v, lu, lv, ou, ov);
  return check_hip(hipGetLastError(), "fvm_adjoint_explicit");
}

template <int N>
int curl_vjp_impl(const FvmParams& p, const float* wb, float* ub, float* vb, hipStream_t s) {
  hipLaunchKernelGGL((fvm_curl_vjp<N>), dim3((N * N + kThreads - 1) / kThreads), dim3(kThreads), 0, s,
                     p, wb, ub, vb);
  return check_hip(hipGetLastError(), "fvm_curl_vjp");
}

}  // namespace

int fvm_launch_adjoint_explicit(const FvmParams& p, const float* u, const float* v, const float* lu,
                                const float* lv, float* ou, float* ov, hipStream_t s) {
  switch (p.n) {
    case 32: return adjoint_impl<32>(p, u, v, lu, lv, ou, ov, s);
    case 64: return adjoint_impl<64>(p, u, v, lu, lv, ou, ov, s);
    case 128: return adjoint_impl<128>(p, u, v, lu, lv, ou, ov, s);
    case 256: return adjoint_impl<256>(p, u, v, lu, lv, ou, ov, s);
    case 512: return adjoint_impl<512>(p, u, v, lu, lv, ou, ov, s);
    case 1024: return adjoint_impl<1024>(p, u, v, lu, lv, ou, ov, s);
    default: return fail(JCFD_ERR_UNSUPPORTED, "adjoint grid size not built");
  }
}

int fvm_launch_curl_vjp(const FvmParams& p, const float* wb, float* ub, float* vb, hipStream_t s) {
  switch (p.n) {
    case 32: return curl_vjp_impl<32>(p, wb, ub, vb, s);
    case 64: return curl_vjp_impl<64>(p, wb, ub, vb, s);
    case 128: return curl_vjp_impl<128>(p, wb, ub, vb, s);
    case 256: return curl_vjp_impl<256>(p, wb, ub, vb, s);
    case 512: return curl_vjp_impl<512>(p, wb, ub, vb, s);
    case 1024: return curl_vjp_impl<1024>(p, wb, ub, vb, s);
    default: return fail(JCFD_ERR_UNSUPPORTED, "curl vjp grid size not built");
  }
}

}  // namespace jcfd
