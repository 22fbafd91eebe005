// Conjugate-gradient pressure solve for the FVM step (gfx950).
//
// solve_cg (jax_cfd/base/pressure.py:67-112) calls
// jax.scipy.sparse.linalg.cg(laplacian, div v, x0 = 0, tol, atol): the
// unpreconditioned CG recurrence
//   alpha = gamma / <p, Ap>;  x += alpha p;  r -= alpha Ap;
//   gamma' = <r, r>;  beta = gamma' / gamma;  p = r + beta p,
// run while gamma > max(tol^2 <b, b>, atol^2) and k < maxiter, with A the
// periodic 5-point Laplacian of finite_differences.laplacian
// (finite_differences.py:127-133).
//
// Each iteration is two kernels over row blocks:
//   cg_direction  p_k = r_k + beta_k p_{k-1} (rows +-1 halo recomputed in
//                 LDS), Ap = L p_k, per-block partial <p, Ap>; convergence is
//                 tested here against the reduced <r_k, r_k>.
//   cg_update     x += alpha p_k, r -= alpha Ap, per-block partial <r, r>.
// Dot products are reduced in a fixed order (block tree, then every block
// reduces the partials array itself), so runs are deterministic.  Once the
// test fails, block 0 raises a flag and every later kernel returns at once,
// so the host can enqueue iterations in chunks and read the flag between
// chunks.
#include <hip/hip_runtime.h>

#include "jcfd_internal.h"

#pragma clang fp contract(off)

namespace jcfd {
namespace {

constexpr int kThreads = 256;

__device__ __forceinline__ float block_sum(float v, float* red) {
  const int tid = threadIdx.x;
  red[tid] = v;
  __syncthreads();
  for (int s = kThreads / 2; s > 0; s >>= 1) {
    if (tid < s) red[tid] = red[tid] + red[tid + s];
    __syncthreads();
  }
  const float out = red[0];
  __syncthreads();
  return out;
}

// Sum of the per-block partials, same order in every block.
__device__ __forceinline__ float sum_partials(const float* part, int nb, float* red) {
  float v = 0.f;
  for (int i = threadIdx.x; i < nb; i += kThreads) v = v + part[i];
  return block_sum(v, red);
}

__device__ __forceinline__ float tol2(const FvmParams& p, float bs, float rtol, float atol) {
  const float a = (rtol * rtol) * bs;
  const float b = atol * atol;
  return a > b ? a : b;
}

template <int N, int RB>
__global__ __launch_bounds__(kThreads) void cg_init(const float* __restrict__ rhs, CgBuffers c) {
  __shared__ float red[kThreads];
  const int i0 = blockIdx.x * RB;
  float rr = 0.f;
  for (int e = threadIdx.x; e < RB * N; e += kThreads) {
    const int idx = i0 * N + e;
    const float b = rhs[idx];
    c.x[idx] = 0.f;
    c.r[idx] = b - 0.f;  // r0 = b - A x0 with x0 = 0
    c.p[0][idx] = 0.f;
    rr = rr + b * b;
  }
  const float s = block_sum(rr, red);
  if (threadIdx.x == 0) {
    c.part_rr[0][blockIdx.x] = s;
    c.part_bb[blockIdx.x] = s;
    if (blockIdx.x == 0) {
      c.status[0] = 0;
      c.status[1] = 0;
    }
  }
}

template <int N, int RB>
__global__ __launch_bounds__(kThreads) void cg_direction(FvmParams p, CgBuffers c, int k, float rtol,
                                                         float atol, int maxiter) {
  __shared__ float red[kThreads];
  __shared__ float pt[(RB + 2) * N];  // p_k rows i0-1 .. i0+RB
  if (c.status[0]) return;
  constexpr int NB = N / RB;
  const int i0 = blockIdx.x * RB;
  const float gamma = sum_partials(c.part_rr[k & 1], NB, red);
  const float bs = sum_partials(c.part_bb, NB, red);
  if (!(gamma > tol2(p, bs, rtol, atol) && k < maxiter)) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      c.status[0] = 1;
      c.status[1] = k;
    }
    return;
  }
  const float beta = (k == 0) ? 0.f : gamma / sum_partials(c.part_rr[(k + 1) & 1], NB, red);
  const float* pold = c.p[k & 1];
  float* pnew = c.p[(k + 1) & 1];
  for (int e = threadIdx.x; e < (RB + 2) * N; e += kThreads) {
    const int lr = e / N, j = e % N;
    const int gi = (i0 - 1 + lr) & (N - 1);
    const float v = c.r[gi * N + j] + beta * pold[gi * N + j];
    pt[e] = v;
    if (lr >= 1 && lr <= RB) pnew[gi * N + j] = v;
  }
  __syncthreads();
  float pap = 0.f;
  for (int e = threadIdx.x; e < RB * N; e += kThreads) {
    const int r = e / N, j = e % N;
    const float* row = pt + (r + 1) * N;
    const float pc = row[j];
    const float lap = ((-2.f * pc) * p.lap_sum + (row[j - N] + row[j + N]) * p.inv_dx_sq) +
                      (row[(j - 1) & (N - 1)] + row[(j + 1) & (N - 1)]) * p.inv_dy_sq;
    c.ap[(i0 + r) * N + j] = lap;
    pap = pap + pc * lap;
  }
  const float s = block_sum(pap, red);
  if (threadIdx.x == 0) c.part_pap[blockIdx.x] = s;
}

template <int N, int RB>
__global__ __launch_bounds__(kThreads) void cg_update(CgBuffers c, int k) {
  __shared__ float red[kThreads];
  if (c.status[0]) return;
  constexpr int NB = N / RB;
  const int i0 = blockIdx.x * RB;
  const float gamma = sum_partials(c.part_rr[k & 1], NB, red);
  const float pap = sum_partials(c.part_pap, NB, red);
  const float alpha = gamma / pap;
  const float* pk = c.p[(k + 1) & 1];
  float rr = 0.f;
  for (int e = threadIdx.x; e < RB * N; e += kThreads) {
    const int idx = i0 * N + e;
    c.x[idx] = c.x[idx] + alpha * pk[idx];
    const float rn = c.r[idx] - alpha * c.ap[idx];
    c.r[idx] = rn;
    rr = rr + rn * rn;
  }
  const float s = block_sum(rr, red);
  if (threadIdx.x == 0) c.part_rr[(k + 1) & 1][blockIdx.x] = s;
}

template <int N> struct CgRows { static constexpr int RB = N >= 512 ? N / 256 : 1; };

template <int N>
struct CgOps {
  static constexpr int RB = CgRows<N>::RB;
  static int init(const float* rhs, const CgBuffers& c, hipStream_t s) {
    hipLaunchKernelGGL((cg_init<N, RB>), dim3(N / RB), dim3(kThreads), 0, s, rhs, c);
    return check_hip(hipGetLastError(), "cg_init");
  }
  static int iteration(const FvmParams& p, const CgBuffers& c, int k, float rtol, float atol,
                       int maxiter, hipStream_t s) {
    hipLaunchKernelGGL((cg_direction<N, RB>), dim3(N / RB), dim3(kThreads), 0, s, p, c, k, rtol,
                       atol, maxiter);
    hipLaunchKernelGGL((cg_update<N, RB>), dim3(N / RB), dim3(kThreads), 0, s, c, k);
    return check_hip(hipGetLastError(), "cg iteration");
  }
};

}  // namespace

int cg_blocks(int n) { return n >= 512 ? 256 : n; }

int cg_launch_init(const FvmParams& p, const float* rhs, const CgBuffers& c, float, float, int,
                   hipStream_t s) {
  switch (p.n) {
    case 32: return CgOps<32>::init(rhs, c, s);
    case 64: return CgOps<64>::init(rhs, c, s);
    case 128: return CgOps<128>::init(rhs, c, s);
    case 256: return CgOps<256>::init(rhs, c, s);
    case 512: return CgOps<512>::init(rhs, c, s);
    case 1024: return CgOps<1024>::init(rhs, c, s);
    default: return fail(JCFD_ERR_UNSUPPORTED, "CG grid size not built");
  }
}

int cg_launch_iteration(const FvmParams& p, const CgBuffers& c, int k, float rtol, float atol,
                        int maxiter, hipStream_t s) {
  switch (p.n) {
    case 32: return CgOps<32>::iteration(p, c, k, rtol, atol, maxiter, s);
    case 64: return CgOps<64>::iteration(p, c, k, rtol, atol, maxiter, s);
    case 128: return CgOps<128>::iteration(p, c, k, rtol, atol, maxiter, s);
    case 256: return CgOps<256>::iteration(p, c, k, rtol, atol, maxiter, s);
    case 512: return CgOps<512>::iteration(p, c, k, rtol, atol, maxiter, s);
    case 1024: return CgOps<1024>::iteration(p, c, k, rtol, atol, maxiter, s);
    default: return fail(JCFD_ERR_UNSUPPORTED, "CG grid size not built");
  }
}

}  // namespace jcfd
