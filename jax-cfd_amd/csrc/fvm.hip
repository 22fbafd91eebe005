// FVM time step on the staggered MAC grid: hand-written kernels for gfx950.
//
// One forward-Euler step of semi_implicit_navier_stokes
// (jax_cfd/base/equations.py:120-151, time_stepping.py:59-118) with
// advect_van_leer convection (advection.py:210-342), explicit diffusion
// (diffusion.py:35-37), Kolmogorov + linear forcing (forcings.py:63-178) and
// the periodic solve_fast_diag projection (pressure.py:115-198) runs as three
// kernels over one instance:
//
//   A  fvm_rows_forward   rows [i0, i0+RB): stage u, v rows i0-3..i0+RB+1 in
//                         LDS, evaluate the explicit terms and the Euler
//                         update u* = u + dt k (rows i0-1.. for the backward
//                         difference), write u*, v*, form div u* and take the
//                         packed real FFT of each row -> spectrum rows.
//   B  fvm_poisson_cols   tiles of W spectrum columns: forward column FFT,
//                         multiply by Lambda^+ / n^2 (the DC/Nyquist pair in
//                         word 0 is unpacked, scaled separately and repacked),
//                         inverse column FFT.
//   C  fvm_rows_project   rows [i0, i0+RB]: inverse packed real FFT -> q, then
//                         u = u* - (q[i+1,j] - q[i,j]) / dx,
//                         v = v* - (q[i,j+1] - q[i,j]) / dy.
//
// Layout: u, v are (n, n) fp32, C-contiguous, axis 1 (y) fastest, offsets
// (1, .5) and (.5, 1).  Threads run along y so every HBM access is
// coalesced.  The arithmetic keeps the reference's operation order (no FMA
// contraction in this file) so fp32 results track jax-cfd's to round-off.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "jcfd_fft.h"
#include "jcfd_internal.h"

#pragma clang fp contract(off)

namespace jcfd {
namespace {

constexpr int kThreads = 256;

__device__ __forceinline__ float sgnf(float x) { return (x > 0.f) ? 1.f : ((x < 0.f) ? -1.f : x); }

// Van Leer flux correction (advection.py:284-334 for one face):
// pre = 0.5 (1 - (dt/h)|U|)|U|;  phi = num (sgn den + sgn num) den / D,
// D = |den| + |num| if |den| > 0 else 1.
__device__ __forceinline__ float vl_correction(float U, float num_pos, float num_neg, float den,
                                               float dt_over_h) {
  const float num = (U > 0.f) ? num_pos : num_neg;
  const float aden = fabsf(den);
  const float dnn = (aden > 0.f) ? (aden + fabsf(num)) : 1.0f;
  const float inv = den / dnn;
  const float phi = (num * (sgnf(den) + sgnf(num))) * inv;
  const float av = fabsf(U);
  const float pre = (0.5f * (1.f - dt_over_h * av)) * av;
  return pre * phi;
}

// Upwind flux + correction for a face with face velocity U, upwind values
// c_pos (U > 0) / c_neg (U <= 0) and limiter stencil (cm2, cm1, c0, cp1) around
// the face: num_pos = cm1 - cm2, num_neg = cp1 - c0, den = c0 - cm1.
__device__ __forceinline__ float vl_flux(float U, float c_pos, float c_neg, float cm2, float cm1,
                                         float c0, float cp1, float dt_over_h) {
  const float up = (U > 0.f) ? U * c_pos : U * c_neg;
  return up + vl_correction(U, cm1 - cm2, cp1 - c0, c0 - cm1, dt_over_h);
}

// Flux through the face between c0 = c[k] and cp1 = c[k+1] for
// advect_van_leer_using_limiters (advection.py:387-395): advect_general with
// linear velocity interpolation (advection.py:34-116) and
// apply_tvd_limiter(lax_wendroff, van_leer_limiter) (interpolation.py:96-303).
__device__ __forceinline__ float vll_flux(float cm1, float c0, float cp1, float cp2, float U,
                                          float dt_over_h) {
  const bool pos = U > 0.f;
  const float low = pos ? c0 : cp1;                       // upwind
  const float cour = dt_over_h * U;                       // Lax-Wendroff
  const float d = cp1 - c0;
  const float high = pos ? c0 + (0.5f * (1.f - cour)) * d : cp1 - (0.5f * (1.f + cour)) * d;
  const float den = (d != 0.f) ? d : 1.0f;                // safe_div
  const float r = pos ? (c0 - cm1) / den : (cp2 - cp1) / den;
  const float onep = 1.f + r;
  const float phi = (r > 0.f) ? (2.f * r) / ((onep != 0.f) ? onep : 1.0f) : 0.f;  // van_leer_limiter
  const float x = low - (low - high) * phi;
  return x * U;
}

// Rows of the LDS staging window: local row lr holds global row i0 - 3 + lr.
template <int N>
struct Window {
  const float* su;  // u rows
  const float* sv;  // v rows
  __device__ __forceinline__ float U(int lr, int j) const { return su[lr * N + (j & (N - 1))]; }
  __device__ __forceinline__ float V(int lr, int j) const { return sv[lr * N + (j & (N - 1))]; }
};

// Explicit terms k_u, k_v at (local row lr, column j); see advection.py:210-342
// for the face fluxes (the limiter stencil of the cross-direction faces is
// read one cell over along the component's staggered axis, as the reference's
// loop variable at advection.py:291 makes it), equations.py:102-110 for the
// sum, finite_differences.py:127-133 for the Laplacian.
template <int N>
__device__ __forceinline__ void add_diffusion_forcing(const Window<N>& w, int lr, int j,
                                                      const FvmParams& p, float fx_j, float& ku,
                                                      float& kv);

template <int N>
__device__ __forceinline__ void vll_advection_at(const Window<N>& w, int lr, int j,
                                                 const FvmParams& p, float& adv_u, float& adv_v) {
  const float hx = p.dt_over_dx, hy = p.dt_over_dy;
  // u (offset (1, .5)): x faces between rows a, a+1 with U = .5 u[a] + .5 u[a+1];
  // y faces between columns b, b+1 with V = .5 v[i, b] + .5 v[i+1, b].
  auto fx_u = [&](int a) {
    const float Uf = 0.5f * w.U(a, j) + 0.5f * w.U(a + 1, j);
    return vll_flux(w.U(a - 1, j), w.U(a, j), w.U(a + 1, j), w.U(a + 2, j), Uf, hx);
  };
  auto fy_u = [&](int b) {
    const float Vf = 0.5f * w.V(lr, b) + 0.5f * w.V(lr + 1, b);
    return vll_flux(w.U(lr, b - 1), w.U(lr, b), w.U(lr, b + 1), w.U(lr, b + 2), Vf, hy);
  };
  adv_u = -((fx_u(lr) - fx_u(lr - 1)) / p.dx + (fy_u(j) - fy_u(j - 1)) / p.dy);
  // v (offset (.5, 1)): x faces with U = .5 u[a, j] + .5 u[a, j+1]; y faces
  // with V = .5 v[i, b] + .5 v[i, b+1].
  auto fx_v = [&](int a) {
    const float Uf = 0.5f * w.U(a, j) + 0.5f * w.U(a, j + 1);
    return vll_flux(w.V(a - 1, j), w.V(a, j), w.V(a + 1, j), w.V(a + 2, j), Uf, hx);
  };
  auto fy_v = [&](int b) {
    const float Vf = 0.5f * w.V(lr, b) + 0.5f * w.V(lr, b + 1);
    return vll_flux(w.V(lr, b - 1), w.V(lr, b), w.V(lr, b + 1), w.V(lr, b + 2), Vf, hy);
  };
  adv_v = -((fx_v(lr) - fx_v(lr - 1)) / p.dx + (fy_v(j) - fy_v(j - 1)) / p.dy);
}

template <int N>
__device__ __forceinline__ void explicit_terms_at(const Window<N>& w, int lr, int j,
                                                  const FvmParams& p, float fx_j, float& ku,
                                                  float& kv) {
  const float hx = p.dt_over_dx, hy = p.dt_over_dy;
  if (p.convect == JCFD_CONVECT_VAN_LEER_LIMITERS) {
    vll_advection_at<N>(w, lr, j, p, ku, kv);
    add_diffusion_forcing<N>(w, lr, j, p, fx_j, ku, kv);
    return;
  }
  // ---- u component (offset (1, .5)) ----
  // x faces at a = i and a = i + 1 (local rows lr, lr + 1)
  auto fx_u = [&](int a) {
    const float Uf = 0.5f * w.U(a - 1, j) + 0.5f * w.U(a, j);
    return vl_flux(Uf, w.U(a - 1, j), w.U(a, j), w.U(a - 2, j), w.U(a - 1, j), w.U(a, j),
                   w.U(a + 1, j), hx);
  };
  // y faces at (a, b) = (i + 1, j) and (i + 1, j - 1)
  auto fy_u = [&](int a, int b) {
    const float Vf = 0.5f * w.V(a - 1, b) + 0.5f * w.V(a, b);
    return vl_flux(Vf, w.U(a - 1, b), w.U(a - 1, b + 1), w.U(a, b - 1), w.U(a, b), w.U(a, b + 1),
                   w.U(a, b + 2), hy);
  };
  const float dfx_u = (fx_u(lr + 1) - fx_u(lr)) / p.dx;
  const float dfy_u = (fy_u(lr + 1, j) - fy_u(lr + 1, j - 1)) / p.dy;
  const float adv_u = -(dfx_u + dfy_u);

  // ---- v component (offset (.5, 1)) ----
  // x faces at (a, b) = (i, j + 1) and (i - 1, j + 1)
  auto fx_v = [&](int a, int b) {
    const float Uf = 0.5f * w.U(a, b - 1) + 0.5f * w.U(a, b);
    return vl_flux(Uf, w.V(a, b - 1), w.V(a + 1, b - 1), w.V(a - 1, b), w.V(a, b), w.V(a + 1, b),
                   w.V(a + 2, b), hx);
  };
  // y faces at b = j + 1 and b = j (row i)
  auto fy_v = [&](int b) {
    const float Vf = 0.5f * w.V(lr, b - 1) + 0.5f * w.V(lr, b);
    return vl_flux(Vf, w.V(lr, b - 1), w.V(lr, b), w.V(lr, b - 2), w.V(lr, b - 1), w.V(lr, b),
                   w.V(lr, b + 1), hy);
  };
  const float dfx_v = (fx_v(lr, j + 1) - fx_v(lr - 1, j + 1)) / p.dx;
  const float dfy_v = (fy_v(j + 1) - fy_v(j)) / p.dy;
  const float adv_v = -(dfx_v + dfy_v);

  ku = adv_u;
  kv = adv_v;
  add_diffusion_forcing<N>(w, lr, j, p, fx_j, ku, kv);
}

// k += nu/rho lap + forcing/rho, in the reference's order (equations.py:102-110).
template <int N>
__device__ __forceinline__ void add_diffusion_forcing(const Window<N>& w, int lr, int j,
                                                      const FvmParams& p, float fx_j, float& ku,
                                                      float& kv) {
  if (p.has_diffusion) {
    const float uc = w.U(lr, j), vc = w.V(lr, j);
    const float lap_u = ((-2.f * uc) * p.lap_sum + (w.U(lr - 1, j) + w.U(lr + 1, j)) * p.inv_dx_sq) +
                        (w.U(lr, j - 1) + w.U(lr, j + 1)) * p.inv_dy_sq;
    const float lap_v = ((-2.f * vc) * p.lap_sum + (w.V(lr - 1, j) + w.V(lr + 1, j)) * p.inv_dx_sq) +
                        (w.V(lr, j - 1) + w.V(lr, j + 1)) * p.inv_dy_sq;
    ku = ku + p.nu * lap_u;
    kv = kv + p.nu * lap_v;
  }
  if (p.has_forcing) {
    const float fu = p.lin * w.U(lr, j) + fx_j;
    const float fv = p.lin * w.V(lr, j) + 0.f;
    ku = ku + fu / p.rho;
    kv = kv + fv / p.rho;
  }
}

template <int N, int ROWS>
__device__ __forceinline__ void stage_rows(float* su, float* sv, const float* __restrict__ u,
                                           const float* __restrict__ v, int row0, int tid) {
  for (int e = tid; e < ROWS * N; e += kThreads) {
    const int lr = e / N, j = e % N;
    const int gi = (row0 + lr) & (N - 1);
    su[e] = u[gi * N + j];
    sv[e] = v[gi * N + j];
  }
}

// LDS layouts of the three step kernels, carved from one byte buffer so the
// same bodies serve the per-size kernels (static LDS) and the mixed-size
// group kernels (dynamic LDS sized for the largest grid present).  float2
// arrays come first so every array stays 8-byte aligned.
// Arrays a MODE does not use get size 0.
template <int N, int RB, int MODE>
struct RowsForwardLds {
  static constexpr int H = N / 2, WR = RB + 5;
  static constexpr bool kFft = MODE == 0 || MODE == 1 || MODE == 3;   // row FFT of div / field
  static constexpr bool kWin = MODE == 0 || MODE == 2 || MODE == 4;   // stencil window
  static constexpr bool kStar = MODE == 0 || MODE == 1 || MODE == 4;  // u*, v* rows
  static constexpr int zb = 0;                                  // float2[RB * H]: div rows, then FFTs
  static constexpr int twn = zb + (kFft ? RB * H * 8 : 0);      // float2[N]
  static constexpr int twh = twn + (kFft ? N * 8 : 0);          // float2[H]
  static constexpr int su = twh + (kFft ? H * 8 : 0);           // float[WR * N]: u rows i0-3 .. i0+RB+1
  static constexpr int sv = su + (kWin ? WR * N * 4 : 0);       // float[WR * N]
  static constexpr int sus = sv + (kWin ? WR * N * 4 : 0);      // float[(RB + 1) * N]: u* rows i0-1 ..
  static constexpr int svs = sus + (kStar ? (RB + 1) * N * 4 : 0);  // float[RB * N]: v* rows i0 ..
  static constexpr int bytes = svs + (kStar ? RB * N * 4 : 0);
};

template <int N, int W>
struct PoissonLds {
  static constexpr int buf = 0;               // float2[N * W]
  static constexpr int twn = buf + N * W * 8;  // float2[N]
  static constexpr int bytes = twn + N * 8;
};

template <int N, int RB, int MODE>
struct RowsProjectLds {
  static constexpr int H = N / 2;
  static constexpr bool kFft = MODE != 2;
  static constexpr int qb = 0;                          // float2[(RB + 1) * H]
  static constexpr int twn = qb + (RB + 1) * H * 8;     // float2[N]
  static constexpr int twh = twn + (kFft ? N * 8 : 0);  // float2[H]
  static constexpr int bytes = twh + (kFft ? H * 8 : 0);
};

template <typename T>
__device__ __forceinline__ T* lds_at(char* smem, int offset) {
  return reinterpret_cast<T*>(smem + offset);
}

// Rows [i0, i0 + RB) of one instance (b already points at the instance).
// MODE 0: full step (write u*, v*, spectrum of div u*).
// MODE 1: projection only (spectrum of div u for the given u, v).
// MODE 2: explicit terms only (write k_u, k_v to us, vs).
// MODE 3: spectrum of the field b.u itself (spectral filtering).
// MODE 4: CG right-hand side (write div u* to b.rhs, no FFT).
template <int N, int RB, int MODE>
__device__ __forceinline__ void rows_forward_body(const FvmParams& p, const FvmBuffers& b, int i0,
                                                  int tid, char* smem) {
  constexpr int H = N / 2;
  constexpr int WR = RB + 5;
  using L = RowsForwardLds<N, RB, MODE>;
  float* su = lds_at<float>(smem, L::su);
  float* sv = lds_at<float>(smem, L::sv);
  float* sus = lds_at<float>(smem, L::sus);
  float* svs = lds_at<float>(smem, L::svs);
  float2* zb = lds_at<float2>(smem, L::zb);
  float2* twn = lds_at<float2>(smem, L::twn);
  float2* twh = lds_at<float2>(smem, L::twh);

  if constexpr (MODE == 3) {
    float* zf3 = reinterpret_cast<float*>(zb);
    for (int e = tid; e < RB * N; e += kThreads) zf3[e] = b.u[(i0 + e / N) * N + e % N];
  } else if constexpr (MODE == 1) {
    // u* = u: stage rows i0-1 .. i0+RB-1 of u and rows i0 .. of v directly.
    for (int e = tid; e < (RB + 1) * N; e += kThreads) {
      const int lr = e / N, j = e % N;
      sus[e] = b.u[((i0 - 1 + lr) & (N - 1)) * N + j];
      if (lr < RB) svs[e] = b.v[((i0 + lr) & (N - 1)) * N + j];
    }
  } else {
    stage_rows<N, WR>(su, sv, b.u, b.v, i0 - 3, tid);
  }
  if constexpr (L::kFft) {
    load_twiddles(twn, b.tw, N, 1, tid, kThreads);
    load_twiddles(twh, b.tw, H, 2, tid, kThreads);
  }
  __syncthreads();

  if constexpr (L::kWin) {
    Window<N> w{su, sv};
    // rows i0-1 .. i0+RB-1 -> local rows 2 .. RB+2 of the window
    const int first = (MODE == 2) ? 1 : 0;
    for (int e = tid + first * N; e < (RB + 1) * N; e += kThreads) {
      const int r = e / N, j = e % N;  // r = 0 -> row i0 - 1
      const int lr = r + 2;
      const float fxj = p.has_forcing ? b.fx[j] : 0.f;
      float ku, kv;
      explicit_terms_at<N>(w, lr, j, p, fxj, ku, kv);
      const int gi = (i0 - 1 + r) & (N - 1);
      if constexpr (MODE == 2) {
        b.us[gi * N + j] = ku;
        b.vs[gi * N + j] = kv;
      } else {
        const float ust = w.U(lr, j) + p.dt * ku;
        sus[e] = ust;
        if (r >= 1) {
          const float vst = w.V(lr, j) + p.dt * kv;
          svs[e - N] = vst;
          b.us[gi * N + j] = ust;
          b.vs[gi * N + j] = vst;
        }
      }
    }
    if constexpr (MODE == 2) return;
    __syncthreads();
  }

  if constexpr (MODE != 3) {
    // div u* (finite_differences.py:136-143): (0 + (u*[i] - u*[i-1]) / dx) + (v*[j] - v*[j-1]) / dy
    float* zf = reinterpret_cast<float*>(zb);
    for (int e = tid; e < RB * N; e += kThreads) {
      const int r = e / N, j = e % N;
      const float ddx = (sus[(r + 1) * N + j] - sus[r * N + j]) / p.dx;
      const float ddy = (svs[r * N + j] - svs[r * N + ((j - 1) & (N - 1))]) / p.dy;
      zf[e] = ddx + ddy;
      if constexpr (MODE == 4) b.rhs[(i0 + r) * N + j] = ddx + ddy;
    }
    if constexpr (MODE == 4) return;
    __syncthreads();
  }
  lds_fft<H, RB, kThreads, 1, H, false, false>(zb, twh, tid);
  for (int e = tid; e < RB * H; e += kThreads) {
    const int r = e / H, k = e % H;
    b.spec[(i0 + r) * H + k] = r2c_combine(zb + r * H, k, H, twn[k]);
  }
}

template <int N, int RB, int MODE>
__global__ __launch_bounds__(kThreads) void fvm_rows_forward(FvmParams p, FvmBuffers b) {
  __shared__ __align__(16) char smem[RowsForwardLds<N, RB, MODE>::bytes];
  rows_forward_body<N, RB, MODE>(p, fvm_instance(b, N), blockIdx.x * RB, threadIdx.x, smem);
}

// Column pass of the Poisson solve on the tile of W packed-spectrum columns
// starting at c0 (b already points at the instance).
template <int N, int W>
__device__ __forceinline__ void poisson_cols_body(const FvmBuffers& b, float scale, int c0, int tid,
                                                  char* smem) {
  constexpr int H = N / 2;
  constexpr int PER = (N * W + kThreads - 1) / kThreads;
  using L = PoissonLds<N, W>;
  float2* buf = lds_at<float2>(smem, L::buf);
  float2* twn = lds_at<float2>(smem, L::twn);
  for (int e = tid; e < N * W; e += kThreads) {
    const int pp = e / W, c = e % W;
    buf[e] = b.spec[pp * H + c0 + c];
  }
  load_twiddles(twn, b.tw, N, 1, tid, kThreads);
  __syncthreads();
  lds_fft<N, W, kThreads, W, 1, true, false>(buf, twn, tid);
  float2 nv[PER];
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    const int e = tid + q * kThreads;
    if (e < N * W) {
      const int pp = e / W, c = e % W, cg = c0 + c;
      const float2 y = buf[e];
      if (cg == 0) {
        // word 0 packs the DC (real) and Nyquist (imaginary) columns.
        const float2 ym = cconj(buf[((N - pp) & (N - 1)) * W + c]);
        const float2 f0 = make_float2(0.5f * (y.x + ym.x), 0.5f * (y.y + ym.y));
        const float2 d = csub(y, ym);
        const float2 fh = make_float2(0.5f * d.y, -0.5f * d.x);
        const float l0 = b.lam[pp * (H + 1)] * scale;
        const float lh = b.lam[pp * (H + 1) + H] * scale;
        const float2 g0 = cscale(f0, l0), gh = cscale(fh, lh);
        nv[q] = make_float2(g0.x - gh.y, g0.y + gh.x);  // g0 + i gh
      } else {
        nv[q] = cscale(y, b.lam[pp * (H + 1) + cg] * scale);
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    const int e = tid + q * kThreads;
    if (e < N * W) buf[e] = nv[q];
  }
  __syncthreads();
  lds_fft<N, W, kThreads, W, 1, true, true>(buf, twn, tid);
  for (int e = tid; e < N * W; e += kThreads) {
    const int pp = e / W, c = e % W;
    b.spec[pp * H + c0 + c] = buf[e];
  }
}

template <int N, int W>
__global__ __launch_bounds__(kThreads) void fvm_poisson_cols(FvmBuffers b, float scale) {
  __shared__ __align__(16) char smem[PoissonLds<N, W>::bytes];
  poisson_cols_body<N, W>(fvm_instance(b, N), scale, blockIdx.x * W, threadIdx.x, smem);
}

// Inverse row transforms of q and the velocity correction for rows
// [i0, i0 + RB) (b already points at the instance).
// MODE 0: velocity correction; MODE 1: write q itself to u_out;
// MODE 2: velocity correction from the real q of the CG solve (no FFT).
template <int N, int RB, int MODE>
__device__ __forceinline__ void rows_project_body(const FvmParams& p, const FvmBuffers& b, int i0,
                                                  int tid, char* smem) {
  constexpr int H = N / 2;
  constexpr int T = RB + 1;
  using L = RowsProjectLds<N, RB, MODE>;
  float2* qb = lds_at<float2>(smem, L::qb);
  float2* twn = lds_at<float2>(smem, L::twn);
  float2* twh = lds_at<float2>(smem, L::twh);
  if constexpr (MODE == 2) {
    float* qf = reinterpret_cast<float*>(qb);
    for (int e = tid; e < T * N; e += kThreads) qf[e] = b.q[((i0 + e / N) & (N - 1)) * N + e % N];
    __syncthreads();
  } else {
    load_twiddles(twn, b.tw, N, 1, tid, kThreads);
    load_twiddles(twh, b.tw, H, 2, tid, kThreads);
    __syncthreads();
    for (int e = tid; e < T * H; e += kThreads) {
      const int r = e / H, k = e % H;
      const int gi = (i0 + r) & (N - 1);
      const float2 pk = b.spec[gi * H + k];
      const float2 pmk = b.spec[gi * H + ((H - k) & (H - 1))];
      qb[e] = c2r_combine(pk, pmk, k, twn[k]);
    }
    __syncthreads();
    lds_fft<H, T, kThreads, 1, H, false, true>(qb, twh, tid);
  }
  const float* q = reinterpret_cast<const float*>(qb);
  if constexpr (MODE == 1) {
    for (int e = tid; e < RB * N; e += kThreads) b.u_out[i0 * N + e] = q[e];
    return;
  }
  for (int e = tid; e < RB * N; e += kThreads) {
    const int r = e / N, j = e % N;
    const int gi = i0 + r;
    const float qc = q[r * N + j];
    const float gx = (q[(r + 1) * N + j] + (-qc)) / p.dx;
    const float gy = (q[r * N + ((j + 1) & (N - 1))] + (-qc)) / p.dy;
    b.u_out[gi * N + j] = b.us[gi * N + j] - gx;
    b.v_out[gi * N + j] = b.vs[gi * N + j] - gy;
  }
}

template <int N, int RB, int MODE>
__global__ __launch_bounds__(kThreads) void fvm_rows_project(FvmParams p, FvmBuffers b) {
  __shared__ __align__(16) char smem[RowsProjectLds<N, RB, MODE>::bytes];
  rows_project_body<N, RB, MODE>(p, fvm_instance(b, N), blockIdx.x * RB, threadIdx.x, smem);
}

// curl_2d (finite_differences.py:186-193): forward differences, offset (1, 1).
__global__ void fvm_curl_kernel(FvmParams p, const float* __restrict__ u,
                                const float* __restrict__ v, float* __restrict__ out) {
  const int n = p.n;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n * n) return;
  const int i = e / n, j = e % n;
  const int ip = (i + 1) % n, jp = (j + 1) % n;
  const float dvdx = (v[ip * n + j] + (-v[e])) / p.dx;
  const float dudy = (u[i * n + jp] + (-u[e])) / p.dy;
  out[e] = dvdx - dudy;
}

__global__ void fvm_divergence_kernel(FvmParams p, const float* __restrict__ u,
                                      const float* __restrict__ v, float* __restrict__ out) {
  const int n = p.n;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n * n) return;
  const int i = e / n, j = e % n;
  const int im = (i - 1 + n) % n, jm = (j - 1 + n) % n;
  const float ddx = (u[e] - u[im * n + j]) / p.dx;
  const float ddy = (v[e] - v[i * n + jm]) / p.dy;
  out[e] = ddx + ddy;
}

// Per-size tile shapes: RB rows per row block, W columns per column tile.
template <int N> struct Tiles;
template <> struct Tiles<32> { static constexpr int RB = 4, W = 8; };
template <> struct Tiles<64> { static constexpr int RB = 4, W = 8; };
template <> struct Tiles<128> { static constexpr int RB = 4, W = 8; };
template <> struct Tiles<256> { static constexpr int RB = 4, W = 8; };
template <> struct Tiles<512> { static constexpr int RB = 2, W = 8; };
template <> struct Tiles<1024> { static constexpr int RB = 1, W = 4; };

template <int N>
int launch_step(const FvmParams& p, const FvmBuffers& b, hipStream_t s) {
  constexpr int RB = Tiles<N>::RB, W = Tiles<N>::W;
  hipLaunchKernelGGL((fvm_rows_forward<N, RB, 0>), dim3(N / RB, nb(b)), dim3(kThreads), 0, s, p, b);
  hipLaunchKernelGGL((fvm_poisson_cols<N, W>), dim3(N / 2 / W, nb(b)), dim3(kThreads), 0, s, b, 1.0f);
  hipLaunchKernelGGL((fvm_rows_project<N, RB, 0>), dim3(N / RB, nb(b)), dim3(kThreads), 0, s, p, b);
  return check_hip(hipGetLastError(), "fvm step launch");
}

template <int N>
int launch_one(const FvmParams& p, const FvmBuffers& b, int which, hipStream_t s) {
  constexpr int RB = Tiles<N>::RB, W = Tiles<N>::W;
  switch (which) {
    case 0:
      hipLaunchKernelGGL((fvm_rows_forward<N, RB, 0>), dim3(N / RB, nb(b)), dim3(kThreads), 0, s, p, b);
      break;
    case 1:
      hipLaunchKernelGGL((fvm_poisson_cols<N, W>), dim3(N / 2 / W, nb(b)), dim3(kThreads), 0, s, b, 1.0f);
      break;
    default:
      hipLaunchKernelGGL((fvm_rows_project<N, RB, 0>), dim3(N / RB, nb(b)), dim3(kThreads), 0, s, p, b);
  }
  return check_hip(hipGetLastError(), "fvm kernel launch");
}

template <int N>
int launch_explicit(const FvmParams& p, const FvmBuffers& b, hipStream_t s) {
  constexpr int RB = Tiles<N>::RB;
  hipLaunchKernelGGL((fvm_rows_forward<N, RB, 2>), dim3(N / RB, nb(b)), dim3(kThreads), 0, s, p, b);
  return check_hip(hipGetLastError(), "fvm explicit-terms launch");
}

template <int N>
int launch_projection(const FvmParams& p, const FvmBuffers& b, hipStream_t s) {
  constexpr int RB = Tiles<N>::RB, W = Tiles<N>::W;
  hipLaunchKernelGGL((fvm_rows_forward<N, RB, 1>), dim3(N / RB, nb(b)), dim3(kThreads), 0, s, p, b);
  hipLaunchKernelGGL((fvm_poisson_cols<N, W>), dim3(N / 2 / W, nb(b)), dim3(kThreads), 0, s, b, 1.0f);
  hipLaunchKernelGGL((fvm_rows_project<N, RB, 0>), dim3(N / RB, nb(b)), dim3(kThreads), 0, s, p, b);
  return check_hip(hipGetLastError(), "fvm projection launch");
}

template <int N>
int launch_filter(const FvmParams& p, const FvmBuffers& b, hipStream_t s) {
  constexpr int RB = Tiles<N>::RB, W = Tiles<N>::W;
  const float scale = 1.0f / (static_cast<float>(N) * static_cast<float>(N));
  hipLaunchKernelGGL((fvm_rows_forward<N, RB, 3>), dim3(N / RB, nb(b)), dim3(kThreads), 0, s, p, b);
  hipLaunchKernelGGL((fvm_poisson_cols<N, W>), dim3(N / 2 / W, nb(b)), dim3(kThreads), 0, s, b, scale);
  hipLaunchKernelGGL((fvm_rows_project<N, RB, 1>), dim3(N / RB, nb(b)), dim3(kThreads), 0, s, p, b);
  return check_hip(hipGetLastError(), "spectral filter launch");
}

template <template <int> class F, typename... Args>
int dispatch_n(int n, Args&&... args) {
  switch (n) {
    case 32: return F<32>::run(args...);
    case 64: return F<64>::run(args...);
    case 128: return F<128>::run(args...);
    case 256: return F<256>::run(args...);
    case 512: return F<512>::run(args...);
    case 1024: return F<1024>::run(args...);
    default: return fail(JCFD_ERR_UNSUPPORTED, "grid size not supported: " + std::to_string(n));
  }
}

template <int N> struct StepOp {
  static int run(const FvmParams& p, const FvmBuffers& b, hipStream_t s) { return launch_step<N>(p, b, s); }
};
template <int N> struct ExplicitOp {
  static int run(const FvmParams& p, const FvmBuffers& b, hipStream_t s) { return launch_explicit<N>(p, b, s); }
};
template <int N> struct ForwardRhsOp {
  static int run(const FvmParams& p, const FvmBuffers& b, hipStream_t s) {
    constexpr int RB = Tiles<N>::RB;
    hipLaunchKernelGGL((fvm_rows_forward<N, RB, 4>), dim3(N / RB), dim3(kThreads), 0, s, p, b);
    return check_hip(hipGetLastError(), "fvm forward (CG rhs)");
  }
};
template <int N> struct ProjectQOp {
  static int run(const FvmParams& p, const FvmBuffers& b, hipStream_t s) {
    constexpr int RB = Tiles<N>::RB;
    hipLaunchKernelGGL((fvm_rows_project<N, RB, 2>), dim3(N / RB), dim3(kThreads), 0, s, p, b);
    return check_hip(hipGetLastError(), "fvm project (CG q)");
  }
};
template <int N> struct OneOp {
  static int run(const FvmParams& p, const FvmBuffers& b, int which, hipStream_t s) {
    return launch_one<N>(p, b, which, s);
  }
};
template <int N> struct FilterOp {
  static int run(const FvmParams& p, const FvmBuffers& b, hipStream_t s) { return launch_filter<N>(p, b, s); }
};
template <int N> struct ProjOp {
  static int run(const FvmParams& p, const FvmBuffers& b, hipStream_t s) { return launch_projection<N>(p, b, s); }
};

// ---- mixed-size group step -------------------------------------------------
// Each phase is one persistent launch: a grid of as many workgroups as the
// device holds at once walks the phase's work list (instance, tile) with a
// fixed stride, so one launch covers every instance of every grid size and
// the launch arguments never depend on the list's contents (a captured graph
// stays valid as instances come and go).  The host sorts the list largest
// grid first, so the stride deals every workgroup a similar mix.  There is
// no shared work counter: device-scope atomics on one address serialise
// across the XCDs and cost more than they balance.  Each item runs the same
// tile body as the per-size kernels, so an instance's arithmetic -- and its
// result, bit for bit -- does not depend on what it is grouped with.
template <int N, int PHASE>
__device__ __forceinline__ void group_tile(const FvmGroupDesc& d, int tile, int tid, char* smem) {
  constexpr int RB = Tiles<N>::RB, W = Tiles<N>::W;
  if constexpr (PHASE == 0)
    rows_forward_body<N, RB, 0>(d.p, d.b, tile * RB, tid, smem);
  else if constexpr (PHASE == 1)
    poisson_cols_body<N, W>(d.b, 1.0f, tile * W, tid, smem);
  else
    rows_project_body<N, RB, 0>(d.p, d.b, tile * RB, tid, smem);
}

template <int PHASE>
__global__ __launch_bounds__(kThreads) void fvm_group_phase(FvmGroupQueue q) {
  extern __shared__ float4 fvm_group_smem[];
  char* smem = reinterpret_cast<char*>(fvm_group_smem);
  const int tid = threadIdx.x;
  const int total = q.n_items[PHASE];
  for (int it = blockIdx.x; it < total; it += gridDim.x) {
    const FvmGroupItem w = q.items[PHASE][it];
    const FvmGroupDesc d = q.descs[w.desc];
    // A fresh copy of the thread index per item: keeps the compiler from
    // hoisting every per-thread address of every tile body out of this loop
    // (which would hold them all in registers across items).
    int t = tid;
    asm volatile("" : "+v"(t));
    switch (d.p.n) {
      case 32: group_tile<32, PHASE>(d, w.tile, t, smem); break;
      case 64: group_tile<64, PHASE>(d, w.tile, t, smem); break;
      case 128: group_tile<128, PHASE>(d, w.tile, t, smem); break;
      case 256: group_tile<256, PHASE>(d, w.tile, t, smem); break;
      case 512: group_tile<512, PHASE>(d, w.tile, t, smem); break;
      default: break;  // rejected on the host
    }
    __syncthreads();  // LDS is reused by the next item
  }
}

template <int N>
constexpr int group_lds_bytes(int phase) {
  return phase == 0   ? RowsForwardLds<N, Tiles<N>::RB, 0>::bytes
         : phase == 1 ? PoissonLds<N, Tiles<N>::W>::bytes
                      : RowsProjectLds<N, Tiles<N>::RB, 0>::bytes;
}

int group_lds(int phase, int max_n) {
  int bytes = 0;
  if (max_n >= 32) bytes = std::max(bytes, group_lds_bytes<32>(phase));
  if (max_n >= 64) bytes = std::max(bytes, group_lds_bytes<64>(phase));
  if (max_n >= 128) bytes = std::max(bytes, group_lds_bytes<128>(phase));
  if (max_n >= 256) bytes = std::max(bytes, group_lds_bytes<256>(phase));
  if (max_n >= 512) bytes = std::max(bytes, group_lds_bytes<512>(phase));
  return bytes;
}

template <int N>
int group_tiles_n(int phase) {
  return phase == 1 ? N / 2 / Tiles<N>::W : N / Tiles<N>::RB;
}


}  // namespace

bool fvm_size_supported(int n) {
  return n == 32 || n == 64 || n == 128 || n == 256 || n == 512 || n == 1024;
}

bool fvm_group_size_supported(int n) { return fvm_size_supported(n) && n <= kFvmGroupMaxN; }

int fvm_group_tiles(int n, int phase) {
  switch (n) {
    case 32: return group_tiles_n<32>(phase);
    case 64: return group_tiles_n<64>(phase);
    case 128: return group_tiles_n<128>(phase);
    case 256: return group_tiles_n<256>(phase);
    case 512: return group_tiles_n<512>(phase);
    default: return 0;
  }
}

// The grid is the number of workgroups of the phase that are resident at
// once (occupancy for the LDS of the largest grid present, times the CU
// count).  It does not depend on the work list.
const void* group_kernel(int phase) {
  switch (phase) {
    case 0: return reinterpret_cast<const void*>(&fvm_group_phase<0>);
    case 1: return reinterpret_cast<const void*>(&fvm_group_phase<1>);
    default: return reinterpret_cast<const void*>(&fvm_group_phase<2>);
  }
}

int fvm_group_grid(int phase, int max_n, int* grid) {
  int dev = 0, cus = 0, per_cu = 0;
  int rc;
  if ((rc = check_hip(hipGetDevice(&dev), "hipGetDevice")) ||
      (rc = check_hip(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev),
                      "CU count")) ||
      (rc = check_hip(hipOccupancyMaxActiveBlocksPerMultiprocessor(
                          &per_cu, group_kernel(phase), kThreads, group_lds(phase, max_n)),
                      "group occupancy")))
    return rc;
  *grid = std::max(1, per_cu) * std::max(1, cus);
  return JCFD_OK;
}

int fvm_group_launch(int phase, int max_n, const FvmGroupQueue& q, int grid, hipStream_t s) {
  if (!fvm_group_size_supported(max_n))
    return fail(JCFD_ERR_UNSUPPORTED, "group grid size not supported: " + std::to_string(max_n));
  const size_t lds = static_cast<size_t>(group_lds(phase, max_n));
  switch (phase) {
    case 0: hipLaunchKernelGGL(fvm_group_phase<0>, dim3(grid), dim3(kThreads), lds, s, q); break;
    case 1: hipLaunchKernelGGL(fvm_group_phase<1>, dim3(grid), dim3(kThreads), lds, s, q); break;
    default: hipLaunchKernelGGL(fvm_group_phase<2>, dim3(grid), dim3(kThreads), lds, s, q);
  }
  return check_hip(hipGetLastError(), "fvm group launch");
}

int fvm_launch_step(const FvmParams& p, const FvmBuffers& b, hipStream_t s) {
  return dispatch_n<StepOp>(p.n, p, b, s);
}

int fvm_launch_kernel(const FvmParams& p, const FvmBuffers& b, int which, hipStream_t s) {
  return dispatch_n<OneOp>(p.n, p, b, which, s);
}

int fvm_launch_forward_rhs(const FvmParams& p, const FvmBuffers& b, hipStream_t s) {
  return dispatch_n<ForwardRhsOp>(p.n, p, b, s);
}

int fvm_launch_project_q(const FvmParams& p, const FvmBuffers& b, hipStream_t s) {
  return dispatch_n<ProjectQOp>(p.n, p, b, s);
}

int fvm_launch_explicit_terms(const FvmParams& p, const float* u, const float* v, float* du,
                              float* dv, const float* fx, hipStream_t s) {
  FvmBuffers b{};
  b.u = u;
  b.v = v;
  b.us = du;
  b.vs = dv;
  b.fx = fx;
  return dispatch_n<ExplicitOp>(p.n, p, b, s);
}

int fvm_launch_projection(const FvmParams& p, const float* u, const float* v, float* u_out,
                          float* v_out, float2* spec, const float* lam, const float2* tw,
                          hipStream_t s) {
  FvmBuffers b{};
  b.u = u;
  b.v = v;
  b.us = const_cast<float*>(u);
  b.vs = const_cast<float*>(v);
  b.u_out = u_out;
  b.v_out = v_out;
  b.spec = spec;
  b.lam = lam;
  b.tw = tw;
  return dispatch_n<ProjOp>(p.n, p, b, s);
}

int fvm_launch_filter(const FvmParams& p, const float* x, const float* multiplier, float* out,
                      float2* spec, const float2* tw, hipStream_t s) {
  FvmBuffers b{};
  b.u = x;
  b.u_out = out;
  b.spec = spec;
  b.lam = multiplier;
  b.tw = tw;
  return dispatch_n<FilterOp>(p.n, p, b, s);
}

int fvm_launch_curl(const FvmParams& p, const float* u, const float* v, float* out, hipStream_t s) {
  const int total = p.n * p.n;
  hipLaunchKernelGGL(fvm_curl_kernel, dim3((total + 255) / 256), dim3(256), 0, s, p, u, v, out);
  return check_hip(hipGetLastError(), "curl launch");
}

int fvm_launch_divergence(const FvmParams& p, const float* u, const float* v, float* out,
                          hipStream_t s) {
  const int total = p.n * p.n;
  hipLaunchKernelGGL(fvm_divergence_kernel, dim3((total + 255) / 256), dim3(256), 0, s, p, u, v, out);
  return check_hip(hipGetLastError(), "divergence launch");
}

}  // namespace jcfd
