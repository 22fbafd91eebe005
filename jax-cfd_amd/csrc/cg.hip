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
// Each iteration is two steps over row blocks, as two kernels or, where the
// device can hold all the row blocks at once, as one cooperative launch for
// the whole solve with a grid barrier between the steps (cg_persistent):
//   cg_direction  p_k = r_k + beta_k p_{k-1} (rows +-1 halo recomputed in
//                 LDS), Ap = L p_k, per-block partial <p, Ap>; convergence is
//                 tested here against the reduced <r_k, r_k>.
//   cg_update     x += alpha p_k, r -= alpha Ap, per-block partial <r, r>.
// The same two kernels solve the implicit diffusion system of
// diffusion.solve_cg (jax_cfd/base/diffusion.py:137-167), one velocity
// component at a time: A x = x - coef L x with coef = fp32(dt viscosity),
// x0 = b = the component, so r0 = b - A b (cg_init_helmholtz).
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

constexpr int kWaves = kThreads / 64;
static_assert(kWaves == 4, "block sums combine four wavefronts");

// Workgroup sums of K values at once: a butterfly across each wavefront's 64
// lanes (every lane ends with the same bits: each step adds the same two
// numbers in both lanes of a pair), then the four wavefront sums combined
// as (w0 + w1) + (w2 + w3).  One barrier pair for all K sums; the order is
// fixed, so every workgroup gets the same result from the same inputs.
// red holds K * kWaves floats.
template <int K>
__device__ __forceinline__ void block_sums(float (&v)[K], float* red) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < K; ++q) {
    float x = v[q];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x = x + __shfl_xor(x, o, 64);
    v[q] = x;
  }
  if (lane == 0) {
#pragma unroll
    for (int q = 0; q < K; ++q) red[q * kWaves + wave] = v[q];
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < K; ++q)
    v[q] = (red[q * kWaves] + red[q * kWaves + 1]) + (red[q * kWaves + 2] + red[q * kWaves + 3]);
  __syncthreads();  // red may be written again
}

__device__ __forceinline__ float block_sum(float v, float* red) {
  float a[1] = {v};
  block_sums<1>(a, red);
  return a[0];
}

__device__ __forceinline__ float tol2(const FvmParams& p, float bs, float rtol, float atol) {
  const float a = (rtol * rtol) * bs;
  const float b = atol * atol;
  return a > b ? a : b;
}

template <int N, int RB>
__global__ __launch_bounds__(kThreads) void cg_init(const float* __restrict__ rhs, CgBuffers c) {
  __shared__ float red[kWaves];
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
      c.status[2] = 0;  // grid-barrier arrivals of the persistent solve
      c.status[3] = 0;  // grid-barrier timeout
    }
  }
}

// periodic 5-point Laplacian at (i, j) of an N x N field, summed as
// finite_differences.laplacian does (finite_differences.py:127-133)
__device__ __forceinline__ float lap5(const FvmParams& p, float centre, float up, float down, float left,
                                      float right) {
  return ((-2.f * centre) * p.lap_sum + (up + down) * p.inv_dx_sq) + (left + right) * p.inv_dy_sq;
}

// x0 = b, r0 = b - (b - coef L b), p_{-1} = 0, partial <r0, r0> and <b, b>
template <int N, int RB>
__global__ __launch_bounds__(kThreads) void cg_init_helmholtz(FvmParams p, const float* __restrict__ b,
                                                              CgBuffers c, float coef) {
  __shared__ float red[2 * kWaves];
  const int i0 = blockIdx.x * RB;
  float sums[2] = {0.f, 0.f};  // <r, r>, <b, b>
  for (int e = threadIdx.x; e < RB * N; e += kThreads) {
    const int i = i0 + e / N, j = e % N;
    const int idx = i * N + j;
    const float bc = b[idx];
    const float lap = lap5(p, bc, b[((i - 1) & (N - 1)) * N + j], b[((i + 1) & (N - 1)) * N + j],
                           b[i * N + ((j - 1) & (N - 1))], b[i * N + ((j + 1) & (N - 1))]);
    const float r = bc - (bc - coef * lap);
    c.x[idx] = bc;
    c.r[idx] = r;
    c.p[0][idx] = 0.f;
    sums[0] = sums[0] + r * r;
    sums[1] = sums[1] + bc * bc;
  }
  block_sums<2>(sums, red);
  if (threadIdx.x == 0) {
    c.part_rr[0][blockIdx.x] = sums[0];
    c.part_bb[blockIdx.x] = sums[1];
    if (blockIdx.x == 0) {
      c.status[0] = 0;
      c.status[1] = 0;
      c.status[2] = 0;
      c.status[3] = 0;
    }
  }
}

// One direction step; HELM selects A = I - coef L over A = L.
template <int N, int RB, bool HELM>
__device__ __forceinline__ bool direction_body(const FvmParams& p, const CgBuffers& c, int k, float rtol,
                                               float atol, int maxiter, float coef, float* red,
                                               float* pt) {
  if (c.status[0]) return true;
  constexpr int NB = N / RB;
  constexpr int TOT = (RB + 2) * N, PT = (TOT + kThreads - 1) / kThreads;
  const int i0 = blockIdx.x * RB;
  const float* pold = c.p[k & 1];
  float* pnew = c.p[(k + 1) & 1];
  // r and p_{k-1} over rows i0-1 .. i0+RB, loaded before the reductions so
  // the two memory latencies overlap (indices clamped: loads unconditional)
  float rv[PT], pv[PT];
#pragma unroll
  for (int q = 0; q < PT; ++q) {
    const int e = min(static_cast<int>(threadIdx.x) + q * kThreads, TOT - 1);
    const int gi = (i0 - 1 + e / N) & (N - 1);
    rv[q] = c.r[gi * N + e % N];
    pv[q] = pold[gi * N + e % N];
  }
  // gamma_k = <r_k, r_k>, <b, b> and gamma_{k-1}, reduced together (the
  // last is unused, and part_rr of the other parity unset, at k = 0)
  float sums[3] = {0.f, 0.f, 0.f};
  for (int e = threadIdx.x; e < NB; e += kThreads) {
    sums[0] = sums[0] + c.part_rr[k & 1][e];
    sums[1] = sums[1] + c.part_bb[e];
    if (k > 0) sums[2] = sums[2] + c.part_rr[(k + 1) & 1][e];
  }
  block_sums<3>(sums, red);
  const float gamma = sums[0], bs = sums[1];
  if (!(gamma > tol2(p, bs, rtol, atol) && k < maxiter)) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      c.status[0] = 1;
      c.status[1] = k;
    }
    return true;
  }
  const float beta = (k == 0) ? 0.f : gamma / sums[2];
#pragma unroll
  for (int q = 0; q < PT; ++q) {
    const int e = static_cast<int>(threadIdx.x) + q * kThreads;
    if (TOT % kThreads != 0 && e >= TOT) break;
    const int lr = e / N, j = e % N;
    const int gi = (i0 - 1 + lr) & (N - 1);
    const float v = rv[q] + beta * pv[q];
    pt[e] = v;
    if (lr >= 1 && lr <= RB) pnew[gi * N + j] = v;
  }
  __syncthreads();
  float pap = 0.f;
  for (int e = threadIdx.x; e < RB * N; e += kThreads) {
    const int r = e / N, j = e % N;
    const float* row = pt + (r + 1) * N;
    const float pc = row[j];
    const float lap = lap5(p, pc, row[j - N], row[j + N], row[(j - 1) & (N - 1)], row[(j + 1) & (N - 1)]);
    const float ap = HELM ? pc - coef * lap : lap;
    c.ap[(i0 + r) * N + j] = ap;
    pap = pap + pc * ap;
  }
  const float s = block_sum(pap, red);
  if (threadIdx.x == 0) c.part_pap[blockIdx.x] = s;
  return false;
}

template <int N, int RB>
__global__ __launch_bounds__(kThreads) void cg_direction(FvmParams p, CgBuffers c, int k, float rtol,
                                                         float atol, int maxiter) {
  __shared__ float red[3 * kWaves];
  __shared__ float pt[(RB + 2) * N];  // p_k rows i0-1 .. i0+RB
  direction_body<N, RB, false>(p, c, k, rtol, atol, maxiter, 0.f, red, pt);
}

template <int N, int RB>
__global__ __launch_bounds__(kThreads) void cg_direction_helmholtz(FvmParams p, CgBuffers c, int k,
                                                                   float rtol, float atol, int maxiter,
                                                                   float coef) {
  __shared__ float red[3 * kWaves];
  __shared__ float pt[(RB + 2) * N];
  direction_body<N, RB, true>(p, c, k, rtol, atol, maxiter, coef, red, pt);
}

template <int N, int RB>
__device__ __forceinline__ void update_body(const CgBuffers& c, int k, float* red) {
  constexpr int NB = N / RB;
  constexpr int TOT = RB * N, PU = (TOT + kThreads - 1) / kThreads;
  const int i0 = blockIdx.x * RB;
  const float* pk = c.p[(k + 1) & 1];
  // this block's x, p_k, r, Ap, loaded before the reductions (clamped)
  float xv[PU], pv[PU], rv[PU], av[PU];
#pragma unroll
  for (int q = 0; q < PU; ++q) {
    const int idx = i0 * N + min(static_cast<int>(threadIdx.x) + q * kThreads, TOT - 1);
    xv[q] = c.x[idx];
    pv[q] = pk[idx];
    rv[q] = c.r[idx];
    av[q] = c.ap[idx];
  }
  float sums[2] = {0.f, 0.f};  // gamma_k, <p, Ap>
  for (int e = threadIdx.x; e < NB; e += kThreads) {
    sums[0] = sums[0] + c.part_rr[k & 1][e];
    sums[1] = sums[1] + c.part_pap[e];
  }
  block_sums<2>(sums, red);
  const float alpha = sums[0] / sums[1];
  float rr = 0.f;
#pragma unroll
  for (int q = 0; q < PU; ++q) {
    const int e = static_cast<int>(threadIdx.x) + q * kThreads;
    if (TOT % kThreads != 0 && e >= TOT) break;
    const int idx = i0 * N + e;
    c.x[idx] = xv[q] + alpha * pv[q];
    const float rn = rv[q] - alpha * av[q];
    c.r[idx] = rn;
    rr = rr + rn * rn;
  }
  const float s = block_sum(rr, red);
  if (threadIdx.x == 0) c.part_rr[(k + 1) & 1][blockIdx.x] = s;
}

template <int N, int RB>
__global__ __launch_bounds__(kThreads) void cg_update(CgBuffers c, int k) {
  __shared__ float red[2 * kWaves];
  if (c.status[0]) return;
  update_body<N, RB>(c, k, red);
}

// Polls of a grid barrier before a workgroup gives up (seconds of waiting;
// a barrier of a healthy launch completes in microseconds).
constexpr unsigned kBarrierSpinLimit = 1u << 24;

// Grid-wide barrier of a cooperative launch (every workgroup resident at
// once): the arrival counter grows by one per workgroup and barrier, and the
// workgroups of barrier number b wait until it reaches target = b *
// gridDim.x.  The agent-scope release of the arrival and the acquire after
// the wait make every workgroup's earlier global writes visible to all the
// others (the caches of every XCD included), as a kernel boundary would.
// Bounded: a workgroup that polls kBarrierSpinLimit times raises the timeout
// flag and returns false, so a broken launch ends with an error, not a hang.
__device__ __forceinline__ bool grid_barrier(unsigned* count, int* timeout, unsigned target, int* s_ok) {
  __syncthreads();
  if (threadIdx.x == 0) {
    int ok = 1;
    __hip_atomic_fetch_add(count, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    unsigned spins = 0;
    // wrap-safe: the count may pass 2^32 in very long solves
    while (static_cast<int>(target - __hip_atomic_load(count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) > 0) {
      __builtin_amdgcn_s_sleep(2);
      if (++spins > kBarrierSpinLimit) {
        __hip_atomic_store(timeout, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        ok = 0;
        break;
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    *s_ok = ok;
  }
  __syncthreads();
  return *s_ok != 0;
}

// The whole solve in one cooperative launch: the iterations of cg_direction
// and cg_update with a grid barrier in place of each kernel boundary.  Every
// workgroup computes the same reduced gamma (same partials, same order), so
// all of them stop at the same iteration; the per-workgroup arithmetic is
// that of the two-kernel form, so results are bitwise the same.
template <int N, int RB, bool HELM>
__global__ __launch_bounds__(kThreads) void cg_persistent(FvmParams p, CgBuffers c, float rtol, float atol,
                                                          int maxiter, float coef) {
  __shared__ float red[3 * kWaves];
  __shared__ float pt[(RB + 2) * N];
  __shared__ int s_ok;
  unsigned* count = reinterpret_cast<unsigned*>(&c.status[2]);
  unsigned target = 0;
  for (int k = 0;; ++k) {
    if (direction_body<N, RB, HELM>(p, c, k, rtol, atol, maxiter, coef, red, pt)) break;
    target += gridDim.x;
    if (!grid_barrier(count, &c.status[3], target, &s_ok)) break;
    update_body<N, RB>(c, k, red);
    target += gridDim.x;
    if (!grid_barrier(count, &c.status[3], target, &s_ok)) break;
  }
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
  static int init_helmholtz(const FvmParams& p, const float* b, const CgBuffers& c, float coef,
                            hipStream_t s) {
    hipLaunchKernelGGL((cg_init_helmholtz<N, RB>), dim3(N / RB), dim3(kThreads), 0, s, p, b, c, coef);
    return check_hip(hipGetLastError(), "cg_init_helmholtz");
  }
  // The whole solve as one cooperative launch (N / RB workgroups, all
  // resident).  JCFD_ERR_UNSUPPORTED when the device cannot hold them.
  static int persistent(const FvmParams& p, const CgBuffers& c, float rtol, float atol, int maxiter,
                        bool helm, float coef, hipStream_t s) {
    const void* fn = helm ? reinterpret_cast<const void*>(&cg_persistent<N, RB, true>)
                          : reinterpret_cast<const void*>(&cg_persistent<N, RB, false>);
    int dev = 0, coop = 0, cus = 0, per_cu = 0, rc;
    if ((rc = check_hip(hipGetDevice(&dev), "hipGetDevice")) ||
        (rc = check_hip(hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev),
                        "cooperative launch attribute")) ||
        (rc = check_hip(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev), "CU count")) ||
        (rc = check_hip(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, kThreads, 0), "occupancy")))
      return rc;
    if (!coop || per_cu * cus < N / RB) return JCFD_ERR_UNSUPPORTED;
    FvmParams pp = p;
    CgBuffers cc = c;
    void* args[] = {&pp, &cc, &rtol, &atol, &maxiter, &coef};
    return check_hip(hipLaunchCooperativeKernel(fn, dim3(N / RB), dim3(kThreads), args, 0, s),
                     "cg_persistent");
  }
  static int iteration_helmholtz(const FvmParams& p, const CgBuffers& c, int k, float rtol, float atol,
                                 int maxiter, float coef, hipStream_t s) {
    hipLaunchKernelGGL((cg_direction_helmholtz<N, RB>), dim3(N / RB), dim3(kThreads), 0, s, p, c, k,
                       rtol, atol, maxiter, coef);
    hipLaunchKernelGGL((cg_update<N, RB>), dim3(N / RB), dim3(kThreads), 0, s, c, k);
    return check_hip(hipGetLastError(), "cg helmholtz iteration");
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

int cg_launch_persistent(const FvmParams& p, const CgBuffers& c, float rtol, float atol, int maxiter,
                         bool helm, float coef, hipStream_t s) {
  switch (p.n) {
    case 32: return CgOps<32>::persistent(p, c, rtol, atol, maxiter, helm, coef, s);
    case 64: return CgOps<64>::persistent(p, c, rtol, atol, maxiter, helm, coef, s);
    case 128: return CgOps<128>::persistent(p, c, rtol, atol, maxiter, helm, coef, s);
    case 256: return CgOps<256>::persistent(p, c, rtol, atol, maxiter, helm, coef, s);
    case 512: return CgOps<512>::persistent(p, c, rtol, atol, maxiter, helm, coef, s);
    case 1024: return CgOps<1024>::persistent(p, c, rtol, atol, maxiter, helm, coef, s);
    default: return fail(JCFD_ERR_UNSUPPORTED, "CG grid size not built");
  }
}

int cg_launch_init_helmholtz(const FvmParams& p, const float* b, const CgBuffers& c, float coef,
                             hipStream_t s) {
  switch (p.n) {
    case 32: return CgOps<32>::init_helmholtz(p, b, c, coef, s);
    case 64: return CgOps<64>::init_helmholtz(p, b, c, coef, s);
    case 128: return CgOps<128>::init_helmholtz(p, b, c, coef, s);
    case 256: return CgOps<256>::init_helmholtz(p, b, c, coef, s);
    case 512: return CgOps<512>::init_helmholtz(p, b, c, coef, s);
    case 1024: return CgOps<1024>::init_helmholtz(p, b, c, coef, s);
    default: return fail(JCFD_ERR_UNSUPPORTED, "CG grid size not built");
  }
}

int cg_launch_iteration_helmholtz(const FvmParams& p, const CgBuffers& c, int k, float rtol, float atol,
                                  int maxiter, float coef, hipStream_t s) {
  switch (p.n) {
    case 32: return CgOps<32>::iteration_helmholtz(p, c, k, rtol, atol, maxiter, coef, s);
    case 64: return CgOps<64>::iteration_helmholtz(p, c, k, rtol, atol, maxiter, coef, s);
    case 128: return CgOps<128>::iteration_helmholtz(p, c, k, rtol, atol, maxiter, coef, s);
    case 256: return CgOps<256>::iteration_helmholtz(p, c, k, rtol, atol, maxiter, coef, s);
    case 512: return CgOps<512>::iteration_helmholtz(p, c, k, rtol, atol, maxiter, coef, s);
    case 1024: return CgOps<1024>::iteration_helmholtz(p, c, k, rtol, atol, maxiter, coef, s);
    default: return fail(JCFD_ERR_UNSUPPORTED, "CG grid size not built");
  }
}

}  // namespace jcfd
