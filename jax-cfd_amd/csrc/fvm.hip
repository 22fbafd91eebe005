// FVM time step on the staggered MAC grid: hand-written kernels for gfx950.
//
// One forward-Euler step of semi_implicit_navier_stokes
// (jax_cfd/base/equations.py:120-151, time_stepping.py:59-118) with
// advect_van_leer convection (advection.py:210-342; or
// advect_van_leer_using_limiters, 387-395, convect_linear, 184-207, or
// advect_upwind, 126-130, chosen by a template argument of the stencil
// bodies), explicit diffusion
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
// contraction in this file; quotients by the grid steps are the correctly
// rounded ones, see over_step) so fp32 results track jax-cfd's to round-off.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "jcfd_fft.h"
#include "jcfd_internal.h"
#include "fvm_terms.h"

#pragma clang fp contract(off)

namespace jcfd {
namespace {

constexpr int kThreads = 256;

__device__ __forceinline__ float sgnf(float x) { return (x > 0.f) ? 1.f : ((x < 0.f) ? -1.f : x); }

// Radix schedule of the packed row transforms (length n/2).  A row kernel
// transforms a handful of rows per workgroup (4 per marching batch, RB or
// RB + 1 per row tile), often one per wavefront: radix-8 passes then keep
// n/16 butterflies per transform busy, half a wavefront's lanes or fewer
// for n <= 512, where radix-4 passes (n/8 butterflies) issue fewer
// instructions.  A function of n and the plan's radix only, so every tile
// shape (and every group shape) gives the same bits.
template <int N, int R>
constexpr int kRowRadix = (R == 8 && N <= 512) ? 4 : R;

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

// Byte offset into a workgroup's LDS buffer as a typed pointer.
template <typename T>
__device__ __forceinline__ T* lds_at(char* smem, int offset) {
  return reinterpret_cast<T*>(smem + offset);
}

// The same rows held in a ring of K slots: local row a lives in slot a % K.
template <int N, int K>
struct RingWindow {
  const float* su;
  const float* sv;
  __device__ __forceinline__ float U(int a, int j) const { return su[(a & (K - 1)) * N + (j & (N - 1))]; }
  __device__ __forceinline__ float V(int a, int j) const { return sv[(a & (K - 1)) * N + (j & (N - 1))]; }
};

// Face fluxes of the convection term, each evaluated once per face.  With
// the indexing below a cell at local window row lr, column j reads
//   adv_u = -((FX(lr + 1, j) - FX(lr, j)) / dx + (FY(lr, j) - FY(lr, j - 1)) / dy)
//   adv_v = -((GX(lr + 1, j) - GX(lr, j)) / dx + (GY(lr, j) - GY(lr, j - 1)) / dy)
// for u (offset (1, .5)) and v (offset (.5, 1)).  advect_van_leer
// (advection.py:210-342): upwind flux plus the limited correction, whose
// stencil for the cross-direction faces is read one cell over along the
// component's staggered axis, as the reference's loop variable at
// advection.py:291 makes it.  advect_van_leer_using_limiters
// (advection.py:387-395): advect_general with linear face velocities and
// the TVD Lax-Wendroff / van_leer_limiter face value.  The operation order
// of each face and of the differences is the reference's.
template <int CONV, typename WIN>
__device__ __forceinline__ float face_fx_u(const WIN& w, int a, int j, const FvmParams& p) {
  const float h = p.dt_over_dx;
  if constexpr (CONV == JCFD_CONVECT_LINEAR) {  // u and its velocity both at the face
    const float Uf = 0.5f * w.U(a - 1, j) + 0.5f * w.U(a, j);
    return Uf * Uf;
  }
  if constexpr (CONV == JCFD_CONVECT_UPWIND) {  // upwind u times the face velocity
    const float Uf = 0.5f * w.U(a - 1, j) + 0.5f * w.U(a, j);
    return ((Uf > 0.f) ? w.U(a - 1, j) : w.U(a, j)) * Uf;
  }
  if constexpr (CONV == JCFD_CONVECT_VAN_LEER_LIMITERS) {
    const int c = a - 1;  // face between rows c and c + 1
    const float Uf = 0.5f * w.U(c, j) + 0.5f * w.U(c + 1, j);
    return vll_flux(w.U(c - 1, j), w.U(c, j), w.U(c + 1, j), w.U(c + 2, j), Uf, h);
  }
  const float Uf = 0.5f * w.U(a - 1, j) + 0.5f * w.U(a, j);
  return vl_flux(Uf, w.U(a - 1, j), w.U(a, j), w.U(a - 2, j), w.U(a - 1, j), w.U(a, j), w.U(a + 1, j), h);
}

template <int CONV, typename WIN>
__device__ __forceinline__ float face_fy_u(const WIN& w, int lr, int b, const FvmParams& p) {
  const float h = p.dt_over_dy;
  if constexpr (CONV == JCFD_CONVECT_LINEAR) {  // (u at the face) * (v at the face)
    const float cf = 0.5f * w.U(lr, b) + 0.5f * w.U(lr, b + 1);
    const float Vf = 0.5f * w.V(lr, b) + 0.5f * w.V(lr + 1, b);
    return cf * Vf;
  }
  if constexpr (CONV == JCFD_CONVECT_UPWIND) {
    const float Vf = 0.5f * w.V(lr, b) + 0.5f * w.V(lr + 1, b);
    return ((Vf > 0.f) ? w.U(lr, b) : w.U(lr, b + 1)) * Vf;
  }
  if constexpr (CONV == JCFD_CONVECT_VAN_LEER_LIMITERS) {
    const float Vf = 0.5f * w.V(lr, b) + 0.5f * w.V(lr + 1, b);
    return vll_flux(w.U(lr, b - 1), w.U(lr, b), w.U(lr, b + 1), w.U(lr, b + 2), Vf, h);
  }
  const int a = lr + 1;
  const float Vf = 0.5f * w.V(a - 1, b) + 0.5f * w.V(a, b);
  return vl_flux(Vf, w.U(a - 1, b), w.U(a - 1, b + 1), w.U(a, b - 1), w.U(a, b), w.U(a, b + 1),
                 w.U(a, b + 2), h);
}

template <int CONV, typename WIN>
__device__ __forceinline__ float face_fx_v(const WIN& w, int a, int j, const FvmParams& p) {
  const float h = p.dt_over_dx;
  const int c = a - 1;
  if constexpr (CONV == JCFD_CONVECT_LINEAR) {  // (v at the face) * (u at the face)
    const float cf = 0.5f * w.V(c, j) + 0.5f * w.V(c + 1, j);
    const float Uf = 0.5f * w.U(c, j) + 0.5f * w.U(c, j + 1);
    return cf * Uf;
  }
  if constexpr (CONV == JCFD_CONVECT_UPWIND) {
    const float Uf = 0.5f * w.U(c, j) + 0.5f * w.U(c, j + 1);
    return ((Uf > 0.f) ? w.V(c, j) : w.V(c + 1, j)) * Uf;
  }
  if constexpr (CONV == JCFD_CONVECT_VAN_LEER_LIMITERS) {
    const float Uf = 0.5f * w.U(c, j) + 0.5f * w.U(c, j + 1);
    return vll_flux(w.V(c - 1, j), w.V(c, j), w.V(c + 1, j), w.V(c + 2, j), Uf, h);
  }
  const int b = j + 1;
  const float Uf = 0.5f * w.U(c, b - 1) + 0.5f * w.U(c, b);
  return vl_flux(Uf, w.V(c, b - 1), w.V(c + 1, b - 1), w.V(c - 1, b), w.V(c, b), w.V(c + 1, b),
                 w.V(c + 2, b), h);
}

template <int CONV, typename WIN>
__device__ __forceinline__ float face_fy_v(const WIN& w, int lr, int j, const FvmParams& p) {
  const float h = p.dt_over_dy;
  if constexpr (CONV == JCFD_CONVECT_LINEAR) {  // v and its velocity both at the face
    const float Vf = 0.5f * w.V(lr, j) + 0.5f * w.V(lr, j + 1);
    return Vf * Vf;
  }
  if constexpr (CONV == JCFD_CONVECT_UPWIND) {
    const float Vf = 0.5f * w.V(lr, j) + 0.5f * w.V(lr, j + 1);
    return ((Vf > 0.f) ? w.V(lr, j) : w.V(lr, j + 1)) * Vf;
  }
  if constexpr (CONV == JCFD_CONVECT_VAN_LEER_LIMITERS) {
    const float Vf = 0.5f * w.V(lr, j) + 0.5f * w.V(lr, j + 1);
    return vll_flux(w.V(lr, j - 1), w.V(lr, j), w.V(lr, j + 1), w.V(lr, j + 2), Vf, h);
  }
  const int b = j + 1;
  const float Vf = 0.5f * w.V(lr, b - 1) + 0.5f * w.V(lr, b);
  return vl_flux(Vf, w.V(lr, b - 1), w.V(lr, b), w.V(lr, b - 2), w.V(lr, b - 1), w.V(lr, b),
                 w.V(lr, b + 1), h);
}

// Rows row0 .. row0 + ROWS - 1 (periodic) of u and v into LDS, all loads of
// a chunk issued before its first LDS write.
template <int N, int ROWS>
__device__ __forceinline__ void stage_rows(float* su, float* sv, const float* __restrict__ u,
                                           const float* __restrict__ v, int row0, int tid) {
  staged<ROWS * N, 8, kThreads>(
      tid,
      [&](int e) {
        const int gi = (row0 + e / N) & (N - 1), j = e % N;
        return make_float2(u[gi * N + j], v[gi * N + j]);
      },
      [&](int e, float2 x) {
        su[e] = x.x;
        sv[e] = x.y;
      });
}

// LDS layouts of the three step kernels, carved from one byte buffer so the
// same bodies serve the per-size kernels (static LDS) and the mixed-size
// group kernels (dynamic LDS sized for the largest grid present).  float2
// arrays come first so every array stays 8-byte aligned.
// Arrays a MODE does not use get size 0.
template <int N, int RB, int MODE>
struct RowsForwardLds {
  static constexpr int H = N / 2, WR = RB + 5;
  static constexpr bool kFft = MODE == 0 || MODE == 1 || MODE == 3 || MODE == 5;  // row FFT
  static constexpr bool kWin = MODE == 0 || MODE == 2 || MODE == 4;                // stencil window
  static constexpr bool kStar = MODE == 0 || MODE == 1 || MODE == 4 || MODE == 5;  // u*, v* rows
  // Window rows of the cells whose explicit terms are formed: u at LO_U..HI
  // (global rows i0-1.. for the backward difference of div u*, or i0.. for
  // MODE 2), v at 3..HI (global rows i0..i0+RB-1).
  static constexpr int LO_U = MODE == 2 ? 3 : 2, HI = RB + 2;
  static constexpr int NU = HI - LO_U + 1, NV = RB;
  // Rows of the face-flux area: x faces + y faces of one component, or the
  // u*, v* rows afterwards.
  static constexpr int FR = kWin ? (2 * NU + 1) : (kStar ? 2 * RB + 1 : 0);
  static constexpr int twn = 0;                              // float2[N]
  static constexpr int twh = twn + (kFft ? N * 8 : 0);       // float2[H]
  // float2[RB * H]: div rows, then their FFTs; shares the window's space when
  // there is a window (the window is dead by then)
  static constexpr int zb = twh + (kFft ? H * 8 : 0);
  static constexpr int su = zb + ((kFft && !kWin) ? RB * H * 8 : 0);  // float[WR * N]: rows i0-3 ..
  static constexpr int sv = su + (kWin ? WR * N * 4 : 0);              // float[WR * N]
  static constexpr int flux = sv + (kWin ? WR * N * 4 : 0);            // float[FR * N]
  static constexpr int bytes = flux + FR * N * 4;
  static_assert(!(kFft && kWin) || RB * H * 8 <= 2 * WR * N * 4, "zb must fit in the window");
};

// Marching variant of kernel A (full step only), for n >= 256: a workgroup
// walks RB rows i0-1 .. i0+RB-1 one at a time over a ring of 8 staged rows of
// u and v (each row read once per workgroup, plus a 4-row halo), fetching
// the next row into registers while the current row is computed.  x-face
// fluxes carry over from row to row in registers; y-face fluxes and v* of
// the current row go through LDS for the column neighbours; div rows collect
// FB at a time and are transformed together.  Same per-cell arithmetic as
// rows_forward_body, so the same results bit for bit.
template <int N, int RB>
struct MarchLds {
  static constexpr int H = N / 2, K = 8, FB = 4;
  static constexpr int twn = 0;                  // float2[N]
  static constexpr int twh = twn + N * 8;        // float2[H]
  static constexpr int zb = twh + H * 8;         // float2[FB * H]: div rows, then FFTs
  static constexpr int su = zb + FB * H * 8;     // float[K * N]: ring of u rows
  static constexpr int sv = su + K * N * 4;      // float[K * N]: ring of v rows
  static constexpr int fy = sv + K * N * 4;      // float[N]: y-face fluxes of u, current row
  static constexpr int gy = fy + N * 4;          // float[N]: y-face fluxes of v, current row
  static constexpr int vs = gy + N * 4;          // float[N]: v* of the current row
  static constexpr int bytes = vs + N * 4;
};

template <int N, int RB, int R = 8, int CONV = JCFD_CONVECT_VAN_LEER>
__device__ __forceinline__ void rows_forward_march_body(const FvmParams& p, const FvmBuffers& b,
                                                        int i0, int tid, char* smem) {
  using L = MarchLds<N, RB>;
  constexpr int H = N / 2, K = L::K, FB = L::FB, CPT = N / kThreads;
  static_assert(N % kThreads == 0 && RB % FB == 0, "march shape");
  float* su = lds_at<float>(smem, L::su);
  float* sv = lds_at<float>(smem, L::sv);
  float* fy = lds_at<float>(smem, L::fy);
  float* gy = lds_at<float>(smem, L::gy);
  float* vs = lds_at<float>(smem, L::vs);
  float2* zb = lds_at<float2>(smem, L::zb);
  float* zf = reinterpret_cast<float*>(zb);
  float2* twn = lds_at<float2>(smem, L::twn);
  float2* twh = lds_at<float2>(smem, L::twh);
  const RingWindow<N, K> w{su, sv};
  // ring row a holds global row i0 - 3 + a
  auto grow = [&](int a) { return (i0 - 3 + a) & (N - 1); };
  TwiddleStage<N, 1, kThreads> tsn;
  TwiddleStage<H, 2, kThreads> tsh;
  tsn.load(b.tw, tid);
  tsh.load(b.tw, tid);
  float fx_prev[CPT], gx_prev[CPT], us_prev[CPT], fxj[CPT];
#pragma unroll
  for (int c = 0; c < CPT; ++c) fxj[c] = p.has_forcing ? b.fx[tid + c * kThreads] : 0.f;
  staged<5 * N, 8, kThreads>(
      tid,
      [&](int e) {
        const int g = grow(e / N) * N + e % N;
        return make_float2(b.u[g], b.v[g]);
      },
      [&](int e, float2 x) {
        const int a = e / N, j = e % N;
        su[(a & (K - 1)) * N + j] = x.x;
        sv[(a & (K - 1)) * N + j] = x.y;
      });
  tsn.store(twn, tid);
  tsh.store(twh, tid);
  __syncthreads();
#pragma unroll
  for (int c = 0; c < CPT; ++c) {
    const int j = tid + c * kThreads;
    fx_prev[c] = face_fx_u<CONV>(w, 2, j, p);  // x face below the first u row (i0 - 1)
    gx_prev[c] = face_fx_v<CONV>(w, 3, j, p);  // x face below the first v row (i0)
    us_prev[c] = 0.f;
  }
  // step s: cells of ring row a = s + 2 (global row i0 - 1 + s); u for every
  // s, v and div for s >= 1.  Reads ring rows a - 1 .. a + 2.
  for (int s = 0; s <= RB; ++s) {
    const int a = s + 2;
    const bool vrow = s >= 1;
    const bool more = s < RB;
    float pu[CPT], pv[CPT];
    if (more) {
#pragma unroll
      for (int c = 0; c < CPT; ++c) {
        const int j = tid + c * kThreads;
        pu[c] = b.u[grow(a + 3) * N + j];
        pv[c] = b.v[grow(a + 3) * N + j];
      }
    }
#pragma unroll
    for (int c = 0; c < CPT; ++c) {
      const int j = tid + c * kThreads;
      fy[j] = face_fy_u<CONV>(w, a, j, p);
      if (vrow) gy[j] = face_fy_v<CONV>(w, a, j, p);
    }
    __syncthreads();
    float us_cur[CPT];
#pragma unroll
    for (int c = 0; c < CPT; ++c) {
      const int j = tid + c * kThreads;
      const int jm = (j - 1) & (N - 1);
      const float fx_next = face_fx_u<CONV>(w, a + 1, j, p);
      const float adv_u = -(over_step(fx_next - fx_prev[c], p.dx, p.rcp_dx) +
                            over_step(fy[j] - fy[jm], p.dy, p.rcp_dy));
      const float ku = add_diffusion_forcing<false>(w, a, j, p, fxj[c], adv_u);
      us_cur[c] = w.U(a, j) + p.dt * ku;
      fx_prev[c] = fx_next;
      if (vrow) {
        const float gx_next = face_fx_v<CONV>(w, a + 1, j, p);
        const float adv_v = -(over_step(gx_next - gx_prev[c], p.dx, p.rcp_dx) +
                              over_step(gy[j] - gy[jm], p.dy, p.rcp_dy));
        const float kv = add_diffusion_forcing<true>(w, a, j, p, 0.f, adv_v);
        const float vst = w.V(a, j) + p.dt * kv;
        gx_prev[c] = gx_next;
        vs[j] = vst;
        const int gi = i0 - 1 + s;
        b.us[gi * N + j] = us_cur[c];
        b.vs[gi * N + j] = vst;
      }
    }
    __syncthreads();  // v* row complete; y-face rows read
    if (vrow) {
      const int d = s - 1;  // div row i0 + d
#pragma unroll
      for (int c = 0; c < CPT; ++c) {
        const int j = tid + c * kThreads;
        // div u* (finite_differences.py:136-143)
        const float ddx = over_step(us_cur[c] - us_prev[c], p.dx, p.rcp_dx);
        const float ddy = over_step(vs[j] - vs[(j - 1) & (N - 1)], p.dy, p.rcp_dy);
        zf[(d % FB) * N + j] = ddx + ddy;
      }
    }
#pragma unroll
    for (int c = 0; c < CPT; ++c) us_prev[c] = us_cur[c];
    if (more) {
#pragma unroll
      for (int c = 0; c < CPT; ++c) {
        const int j = tid + c * kThreads;
        su[((a + 3) & (K - 1)) * N + j] = pu[c];
        sv[((a + 3) & (K - 1)) * N + j] = pv[c];
      }
    }
    __syncthreads();
    if (vrow && (s - 1) % FB == FB - 1) {
      const int d0 = s - FB;  // first div row of this batch
      lds_fft_waves<H, FB, kThreads, 1, H, false, kRowRadix<N, R>>(zb, twh, tid);
      for (int e = tid; e < FB * H; e += kThreads) {
        const int r = e / H, k = e % H;
        b.spec[(i0 + d0 + r) * H + k] = r2c_combine(zb + r * H, k, H, twn[k]);
      }
      // zb is written again only after the next step's first barrier
    }
  }
}

// The convection scheme is a template argument of the stencil bodies, chosen
// once per workgroup here (a run-time choice per face costs the VALU-bound
// stencil ~15 %: the extra branches split the face computations apart).
#define JCFD_CONVECT_DISPATCH(CALL)                                   \
  switch (p.convect) {                                               \
    case JCFD_CONVECT_VAN_LEER_LIMITERS: CALL(JCFD_CONVECT_VAN_LEER_LIMITERS); break; \
    case JCFD_CONVECT_LINEAR: CALL(JCFD_CONVECT_LINEAR); break;      \
    case JCFD_CONVECT_UPWIND: CALL(JCFD_CONVECT_UPWIND); break;      \
    default: CALL(JCFD_CONVECT_VAN_LEER); break;                     \
  }

template <int N, int RB, int R = 8>
__global__ __launch_bounds__(kThreads) void fvm_rows_forward_march(FvmParams p, FvmBuffers b) {
  __shared__ __align__(16) char smem[MarchLds<N, RB>::bytes];
  const FvmBuffers bi = fvm_instance(b, N);
  const int i0 = xcd_tile(blockIdx.x, gridDim.x) * RB;
#define JCFD_MARCH(C) rows_forward_march_body<N, RB, R, C>(p, bi, i0, threadIdx.x, smem)
  JCFD_CONVECT_DISPATCH(JCFD_MARCH)
#undef JCFD_MARCH
}

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
  static constexpr int twh = qb + (RB + 1) * H * 8;     // float2[H]
  static constexpr int bytes = twh + (kFft ? H * 8 : 0);
};

// Rows [i0, i0 + RB) of one instance (b already points at the instance).
// MODE 0: full step (write u*, v*, spectrum of div u*).
// MODE 1: projection only (spectrum of div u for the given u, v).
// MODE 2: explicit terms only (write k_u, k_v to us, vs).
// MODE 3: spectrum of the field b.u itself (spectral filtering).
// MODE 4: CG right-hand side (write div u* to b.rhs, no FFT).
// MODE 5: Runge-Kutta combination u* = u0 + dt * (0 + sum c_t k_t) with
//         u0 = (b.u, b.v) and the terms of *b.rk (time_stepping.py:91-101);
//         write u*, v*, spectrum of div u*.
template <int N, int RB, int MODE, int R = 8, int CONV = JCFD_CONVECT_VAN_LEER>
__device__ __forceinline__ void rows_forward_body(const FvmParams& p, const FvmBuffers& b, int i0,
                                                  int tid, char* smem) {
  constexpr int H = N / 2;
  constexpr int WR = RB + 5;
  using L = RowsForwardLds<N, RB, MODE>;
  constexpr int LO_U = L::LO_U, NU = L::NU, NV = L::NV;
  float* su = lds_at<float>(smem, L::su);
  float* sv = lds_at<float>(smem, L::sv);
  float* flux = lds_at<float>(smem, L::flux);
  float* sus = flux;                                   // u* rows i0-1 .. i0+RB-1
  float* svs = flux + (RB + 1) * N;                    // v* rows i0 .. i0+RB-1
  float2* zb = lds_at<float2>(smem, L::kWin ? L::su : L::zb);
  float2* twn = lds_at<float2>(smem, L::twn);
  float2* twh = lds_at<float2>(smem, L::twh);

  // Twiddle and forcing loads are issued first and written to LDS after the
  // field rows' loads (a wait for a load also waits for every earlier one).
  [[maybe_unused]] TwiddleStage<N, 1, kThreads> tsn;
  [[maybe_unused]] TwiddleStage<H, 2, kThreads> tsh;
  if constexpr (L::kFft) {
    tsn.load(b.tw, tid);
    tsh.load(b.tw, tid);
  }
  constexpr int PU = (NU * N + kThreads - 1) / kThreads;
  [[maybe_unused]] float fxr[L::kWin ? PU : 1];
  if constexpr (L::kWin) {
#pragma unroll
    for (int k = 0; k < PU; ++k) fxr[k] = p.has_forcing ? b.fx[(tid + k * kThreads) % N] : 0.f;
  }
  if constexpr (MODE == 3) {
    float* zf3 = reinterpret_cast<float*>(zb);
    const float* ur = b.u + static_cast<size_t>(i0) * N;
    staged<RB * N, 16, kThreads>(tid, [&](int e) { return ur[e]; }, [&](int e, float x) { zf3[e] = x; });
  } else if constexpr (MODE == 1) {
    // u* = u: stage rows i0-1 .. i0+RB-1 of u and rows i0 .. of v directly.
    staged<(RB + 1) * N, 8, kThreads>(
        tid,
        [&](int e) {
          const int lr = e / N, j = e % N;
          return make_float2(b.u[((i0 - 1 + lr) & (N - 1)) * N + j],
                             lr < RB ? b.v[((i0 + lr) & (N - 1)) * N + j] : 0.f);
        },
        [&](int e, float2 x) {
          sus[e] = x.x;
          if (e < RB * N) svs[e] = x.y;
        });
  } else if constexpr (MODE == 5) {
    const RkTerms& t = *b.rk;
    const size_t off = b.inst * static_cast<size_t>(N) * N;
    const float dt = p.dt;
    for (int e = tid; e < (RB + 1) * N; e += kThreads) {
      const int lr = e / N, j = e % N;
      const int gu = ((i0 - 1 + lr) & (N - 1)) * N + j;
      float acc = 0.f;
      for (int k = 0; k < t.n; ++k) acc = acc + t.c[k] * as_global(t.ku[k])[off + gu];
      const float ust = b.u[gu] + dt * acc;
      sus[e] = ust;
      if (lr >= 1) b.us[gu] = ust;
      if (lr < RB) {
        const int gv = ((i0 + lr) & (N - 1)) * N + j;
        float accv = 0.f;
        for (int k = 0; k < t.n; ++k) accv = accv + t.c[k] * as_global(t.kv[k])[off + gv];
        const float vst = b.v[gv] + dt * accv;
        svs[e] = vst;
        b.vs[gv] = vst;
      }
    }
  } else {
    stage_rows<N, WR>(su, sv, b.u, b.v, i0 - 3, tid);
  }
  if constexpr (L::kFft) {
    tsn.store(twn, tid);
    tsh.store(twh, tid);
  }
  __syncthreads();

  if constexpr (L::kWin) {
    const Window<N> w{su, sv};
    constexpr int PV = (NV * N + kThreads - 1) / kThreads;
    float ru[PU], rv[PV];  // explicit terms (MODE 2) or u*, v* of this thread's cells
    // ---- u: x faces of rows LO_U .. HI + 1, y faces of rows LO_U .. HI
    for (int e = tid; e < (NU + 1) * N; e += kThreads) flux[e] = face_fx_u<CONV>(w, LO_U + e / N, e % N, p);
    for (int e = tid; e < NU * N; e += kThreads)
      flux[(NU + 1) * N + e] = face_fy_u<CONV>(w, LO_U + e / N, e % N, p);
    __syncthreads();
#pragma unroll
    for (int k = 0; k < PU; ++k) {
      const int e = tid + k * kThreads;
      if (e < NU * N) {
        const int r = e / N, j = e % N, lr = LO_U + r;
        const float* fy = flux + (NU + 1) * N + r * N;
        const float adv = -(over_step(flux[(r + 1) * N + j] - flux[r * N + j], p.dx, p.rcp_dx) +
                            over_step(fy[j] - fy[(j - 1) & (N - 1)], p.dy, p.rcp_dy));
        const float ku = add_diffusion_forcing<false>(w, lr, j, p, fxr[k], adv);
        ru[k] = (MODE == 2) ? ku : w.U(lr, j) + p.dt * ku;
      }
    }
    __syncthreads();
    // ---- v: x faces of rows 3 .. HI + 1, y faces of rows 3 .. HI
    for (int e = tid; e < (NV + 1) * N; e += kThreads) flux[e] = face_fx_v<CONV>(w, 3 + e / N, e % N, p);
    for (int e = tid; e < NV * N; e += kThreads)
      flux[(NV + 1) * N + e] = face_fy_v<CONV>(w, 3 + e / N, e % N, p);
    __syncthreads();
#pragma unroll
    for (int k = 0; k < PV; ++k) {
      const int e = tid + k * kThreads;
      if (e < NV * N) {
        const int r = e / N, j = e % N, lr = 3 + r;
        const float* fy = flux + (NV + 1) * N + r * N;
        const float adv = -(over_step(flux[(r + 1) * N + j] - flux[r * N + j], p.dx, p.rcp_dx) +
                            over_step(fy[j] - fy[(j - 1) & (N - 1)], p.dy, p.rcp_dy));
        const float kv = add_diffusion_forcing<true>(w, lr, j, p, 0.f, adv);
        rv[k] = (MODE == 2) ? kv : w.V(lr, j) + p.dt * kv;
      }
    }
    if constexpr (MODE == 2) {
#pragma unroll
      for (int k = 0; k < PU; ++k) {
        const int e = tid + k * kThreads;
        if (e < NU * N) b.us[(i0 + e / N) * N + e % N] = ru[k];
      }
#pragma unroll
      for (int k = 0; k < PV; ++k) {
        const int e = tid + k * kThreads;
        if (e < NV * N) b.vs[(i0 + e / N) * N + e % N] = rv[k];
      }
      return;
    }
    __syncthreads();  // the flux area becomes the u*, v* rows
#pragma unroll
    for (int k = 0; k < PU; ++k) {
      const int e = tid + k * kThreads;
      if (e < NU * N) {
        sus[e] = ru[k];
        if (e >= N) b.us[(i0 - 1) * N + e] = ru[k];  // rows i0 .. (row i0-1 is another block's)
      }
    }
#pragma unroll
    for (int k = 0; k < PV; ++k) {
      const int e = tid + k * kThreads;
      if (e < NV * N) {
        svs[e] = rv[k];
        b.vs[i0 * N + e] = rv[k];
      }
    }
    __syncthreads();
  }

  if constexpr (MODE != 3) {
    // div u* (finite_differences.py:136-143): (0 + (u*[i] - u*[i-1]) / dx) + (v*[j] - v*[j-1]) / dy
    float* zf = reinterpret_cast<float*>(zb);
    for (int e = tid; e < RB * N; e += kThreads) {
      const int r = e / N, j = e % N;
      const float ddx = over_step(sus[(r + 1) * N + j] - sus[r * N + j], p.dx, p.rcp_dx);
      const float ddy = over_step(svs[r * N + j] - svs[r * N + ((j - 1) & (N - 1))], p.dy, p.rcp_dy);
      zf[e] = ddx + ddy;
      if constexpr (MODE == 4) b.rhs[(i0 + r) * N + j] = ddx + ddy;
    }
    if constexpr (MODE == 4) return;
    __syncthreads();
  }
  lds_fft_waves<H, RB, kThreads, 1, H, false, kRowRadix<N, R>>(zb, twh, tid);
  for (int e = tid; e < RB * H; e += kThreads) {
    const int r = e / H, k = e % H;
    b.spec[(i0 + r) * H + k] = r2c_combine(zb + r * H, k, H, twn[k]);
  }
}

template <int N, int RB, int MODE, int R = 8>
__global__ __launch_bounds__(kThreads) void fvm_rows_forward(FvmParams p, FvmBuffers b) {
  __shared__ __align__(16) char smem[RowsForwardLds<N, RB, MODE>::bytes];
  const FvmBuffers bi = fvm_instance(b, N);
  const int i0 = xcd_tile(blockIdx.x, gridDim.x) * RB;
  if constexpr (RowsForwardLds<N, RB, MODE>::kWin) {
#define JCFD_ROWS(C) rows_forward_body<N, RB, MODE, R, C>(p, bi, i0, threadIdx.x, smem)
    JCFD_CONVECT_DISPATCH(JCFD_ROWS)
#undef JCFD_ROWS
  } else {
    rows_forward_body<N, RB, MODE, R>(p, bi, i0, threadIdx.x, smem);
  }
}

// Entry (p, c) of the column pass's spectrum multiplier (see lam_mode).  For
// Lambda^+ it is computed as the host tabulated it -- fp64 eigenvalue sum,
// pseudo-inverse where |s| > 10 eps(fp32) (fast_diagonalization.py:199-266
// with array_utils.py:168-173), / n^2, rounded to fp32; every operation is
// IEEE, so the bits equal the table's -- which saves the kernel a table read
// per spectrum word.
constexpr double kLamCutoff = static_cast<double>(10.0f * 1.1920928955078125e-07f);
template <int N>
__device__ __forceinline__ float lam_value(const FvmBuffers& b, int p, int c) {
  if (b.lam_mode == 2) {
    const double s = b.lam_xy[p] + b.lam_xy[N + c];
    const double inv = (fabs(s) > kLamCutoff) ? 1.0 / s : 0.0;
    return static_cast<float>(inv / (static_cast<double>(N) * N));
  }
  return b.lam_mode == 1 ? b.lam[c * N + p] : b.lam[p * (N / 2 + 1) + c];
}

// Column pass of the Poisson solve on the tile of W packed-spectrum columns
// starting at c0 (b already points at the instance).
template <int N, int W, int R = 8>
__device__ __forceinline__ void poisson_cols_body(const FvmBuffers& b, float scale, int c0, int tid,
                                                  char* smem) {
  constexpr int H = N / 2;
  constexpr int PER = (N * W + kThreads - 1) / kThreads;
  using L = PoissonLds<N, W>;
  float2* buf = lds_at<float2>(smem, L::buf);
  float2* twn = lds_at<float2>(smem, L::twn);
  TwiddleStage<N, 1, kThreads> tsn;
  tsn.load(b.tw, tid);
  staged<N * W, 8, kThreads>(
      tid, [&](int e) { return b.spec[(e / W) * H + c0 + e % W]; },
      [&](int e, float2 x) { buf[e] = x; });
  // the multipliers of this thread's elements arrive during the transform
  float lam[PER];
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    const int e0 = tid + q * kThreads;
    const int e = e0 < N * W ? e0 : N * W - 1;
    lam[q] = lam_value<N>(b, e / W, c0 + e % W);
  }
  tsn.store(twn, tid);
  __syncthreads();
  lds_fft<N, W, kThreads, W, 1, true, false, R>(buf, twn, tid);
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
        const float l0 = lam[q] * scale;
        const float lh = lam_value<N>(b, pp, H) * scale;
        const float2 g0 = cscale(f0, l0), gh = cscale(fh, lh);
        nv[q] = make_float2(g0.x - gh.y, g0.y + gh.x);  // g0 + i gh
      } else {
        nv[q] = cscale(y, lam[q] * scale);
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
  lds_fft<N, W, kThreads, W, 1, true, true, R>(buf, twn, tid);
  staged<N * W, 16, kThreads>(
      tid, [&](int e) { return buf[e]; },
      [&](int e, float2 x) { b.spec[(e / W) * H + c0 + e % W] = x; });
}

template <int N, int W, int R = 8>
__global__ __launch_bounds__(kThreads) void fvm_poisson_cols(FvmBuffers b, float scale) {
  __shared__ __align__(16) char smem[PoissonLds<N, W>::bytes];
  poisson_cols_body<N, W, R>(fvm_instance(b, N), scale, xcd_tile(blockIdx.x, gridDim.x) * W,
                             threadIdx.x, smem);
}

// Inverse row transforms of q and the velocity correction for rows
// [i0, i0 + RB) (b already points at the instance).
// MODE 0: velocity correction; MODE 1: write q itself to u_out;
// MODE 2: velocity correction from the real q of the CG solve (no FFT);
// MODE 3: u_out = u + q (the implicit diffusion solve).
template <int N, int RB, int MODE, int R = 8>
__device__ __forceinline__ void rows_project_body(const FvmParams& p, const FvmBuffers& b, int i0,
                                                  int tid, char* smem) {
  constexpr int H = N / 2;
  constexpr int T = RB + 1;
  using L = RowsProjectLds<N, RB, MODE>;
  constexpr int PR = (RB * N + kThreads - 1) / kThreads;
  float2* qb = lds_at<float2>(smem, L::qb);
  float2* twh = lds_at<float2>(smem, L::twh);
  // (u*, v*) (MODES 0, 2) or u (MODE 3) at this thread's cells: loaded
  // before the transform, used after it
  [[maybe_unused]] float pa[PR], pb[PR];
  auto prefetch = [&]() {
#pragma unroll
    for (int k = 0; k < PR; ++k) {
      const int e0 = tid + k * kThreads;
      const int e = i0 * N + (e0 < RB * N ? e0 : RB * N - 1);
      if constexpr (MODE == 0 || MODE == 2) {
        pa[k] = b.us[e];
        pb[k] = b.vs[e];
      } else if constexpr (MODE == 3) {
        pa[k] = b.u[e];
      }
    }
  };
  if constexpr (MODE == 2) {
    float* qf = reinterpret_cast<float*>(qb);
    staged<T * N, 16, kThreads>(
        tid, [&](int e) { return b.q[((i0 + e / N) & (N - 1)) * N + e % N]; },
        [&](int e, float x) { qf[e] = x; });
    prefetch();
    __syncthreads();
  } else {
    TwiddleStage<H, 2, kThreads> tsh;
    tsh.load(b.tw, tid);
    // the combine's twiddle exp(-2 pi i k / n) is entry k of the table
    staged<T * H, 8, kThreads>(
        tid,
        [&](int e) {
          const int k = e % H;
          const float2* row = b.spec + ((i0 + e / H) & (N - 1)) * H;
          return F2Triple{row[k], row[(H - k) & (H - 1)], b.tw[k]};
        },
        [&](int e, F2Triple x) { qb[e] = c2r_combine(x.a, x.b, e % H, x.t); });
    prefetch();
    tsh.store(twh, tid);
    __syncthreads();
    lds_fft_waves<H, T, kThreads, 1, H, true, kRowRadix<N, R>>(qb, twh, tid);
  }
  const float* q = reinterpret_cast<const float*>(qb);
#pragma unroll
  for (int k = 0; k < PR; ++k) {
    const int e = tid + k * kThreads;
    if (e < RB * N) {
      if constexpr (MODE == 1) {
        b.u_out[i0 * N + e] = q[e];
      } else if constexpr (MODE == 3) {
        // implicit diffusion: u + (1 - nu dt L)^-1 (nu dt L) u (diffusion.py:198-208)
        b.u_out[i0 * N + e] = pa[k] + q[e];
      } else {
        const int r = e / N, j = e % N;
        const int gi = i0 + r;
        const float qc = q[r * N + j];
        const float gx = over_step(q[(r + 1) * N + j] + (-qc), p.dx, p.rcp_dx);
        const float gy = over_step(q[r * N + ((j + 1) & (N - 1))] + (-qc), p.dy, p.rcp_dy);
        b.u_out[gi * N + j] = pa[k] - gx;
        b.v_out[gi * N + j] = pb[k] - gy;
      }
    }
  }
}

template <int N, int RB, int MODE, int R = 8>
__global__ __launch_bounds__(kThreads) void fvm_rows_project(FvmParams p, FvmBuffers b) {
  __shared__ __align__(16) char smem[RowsProjectLds<N, RB, MODE>::bytes];
  rows_project_body<N, RB, MODE, R>(p, fvm_instance(b, N), xcd_tile(blockIdx.x, gridDim.x) * RB,
                                    threadIdx.x, smem);
}

// curl_2d (finite_differences.py:186-193): forward differences, offset (1, 1).
__global__ void fvm_curl_kernel(FvmParams p, const float* __restrict__ u,
                                const float* __restrict__ v, float* __restrict__ out) {
  const int n = p.n;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n * n) return;
  const int i = e / n, j = e % n;
  const int ip = (i + 1) % n, jp = (j + 1) % n;
  const float dvdx = over_step(v[ip * n + j] + (-v[e]), p.dx, p.rcp_dx);
  const float dudy = over_step(u[i * n + jp] + (-u[e]), p.dy, p.rcp_dy);
  out[e] = dvdx - dudy;
}

__global__ void fvm_divergence_kernel(FvmParams p, const float* __restrict__ u,
                                      const float* __restrict__ v, float* __restrict__ out) {
  const int n = p.n;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n * n) return;
  const int i = e / n, j = e % n;
  const int im = (i - 1 + n) % n, jm = (j - 1 + n) % n;
  const float ddx = over_step(u[e] - u[im * n + j], p.dx, p.rcp_dx);
  const float ddy = over_step(v[e] - v[i * n + jm], p.dy, p.rcp_dy);
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

// ---- tunable step kernels ---------------------------------------------------
// The three step kernels are built for a menu of tile shapes and both radix
// schedules; a plan's FvmParams carries its choice (rb_fwd, w_cols, rb_proj,
// radix).  Tile shapes only regroup the same per-cell / per-column
// arithmetic, so they never change a result; the radix schedule changes the
// FFTs' rounding.
constexpr int kTileLdsMax = 96 * 1024;

// rows_forward tiles 1..8 stage RB + 5 rows at once; 16 and 32 march (n >= 256).
template <int N, int RB>
constexpr bool fwd_ok() {
  if constexpr (RB >= 16)
    return N >= kThreads && N <= 512 && RB <= N / 8 && MarchLds<N, RB>::bytes <= kTileLdsMax;
  else
    return RB <= N / 4 && RowsForwardLds<N, RB, 0>::bytes <= kTileLdsMax;
}
template <int N, int W>
constexpr bool cols_ok() {
  return W <= N / 2 && PoissonLds<N, W>::bytes <= kTileLdsMax;
}
template <int N, int RB>
constexpr bool proj_ok() {
  return RB <= N / 4 && RowsProjectLds<N, RB, 0>::bytes <= kTileLdsMax;
}

// which: 0 rows-forward (tile = rows per block), 1 Poisson columns (tile =
// columns per tile), 2 rows-project (tile = rows per block).  Returns
// JCFD_ERR_UNSUPPORTED for a shape not on the menu of size N.
template <int N, int T, int R>
int launch_variant(int which, const FvmParams& p, const FvmBuffers& b, hipStream_t s) {
  if (which == 0) {
    if constexpr (fwd_ok<N, T>()) {
      if constexpr (T >= 16)
        hipLaunchKernelGGL((fvm_rows_forward_march<N, T, R>), dim3(N / T, nb(b)), dim3(kThreads), 0, s, p, b);
      else
        hipLaunchKernelGGL((fvm_rows_forward<N, T, 0, R>), dim3(N / T, nb(b)), dim3(kThreads), 0, s, p, b);
      return JCFD_OK;
    }
  } else if (which == 1) {
    if constexpr (T <= 16 && cols_ok<N, T>()) {
      hipLaunchKernelGGL((fvm_poisson_cols<N, T, R>), dim3(N / 2 / T, nb(b)), dim3(kThreads), 0, s, b, 1.0f);
      return JCFD_OK;
    }
  } else {
    if constexpr (T <= 8 && proj_ok<N, T>()) {
      hipLaunchKernelGGL((fvm_rows_project<N, T, 0, R>), dim3(N / T, nb(b)), dim3(kThreads), 0, s, p, b);
      return JCFD_OK;
    }
  }
  return JCFD_ERR_UNSUPPORTED;
}

template <int N, int R>
int launch_tile(int which, int tile, const FvmParams& p, const FvmBuffers& b, hipStream_t s) {
  switch (tile) {
    case 1: return launch_variant<N, 1, R>(which, p, b, s);
    case 2: return launch_variant<N, 2, R>(which, p, b, s);
    case 4: return launch_variant<N, 4, R>(which, p, b, s);
    case 8: return launch_variant<N, 8, R>(which, p, b, s);
    case 16: return which == 2 ? JCFD_ERR_UNSUPPORTED : launch_variant<N, 16, R>(which, p, b, s);
    case 32: return which == 0 ? launch_variant<N, 32, R>(which, p, b, s) : JCFD_ERR_UNSUPPORTED;
    default: return JCFD_ERR_UNSUPPORTED;
  }
}

template <int N>
int launch_one(const FvmParams& p, const FvmBuffers& b, int which, hipStream_t s) {
  const int tile = which == 0 ? p.rb_fwd : (which == 1 ? p.w_cols : p.rb_proj);
  const int rc = p.radix == 4 ? launch_tile<N, 4>(which, tile, p, b, s)
                              : launch_tile<N, 8>(which, tile, p, b, s);
  if (rc) return fail(rc, "fvm tile shape not built for n = " + std::to_string(N));
  return check_hip(hipGetLastError(), "fvm kernel launch");
}

template <int N>
int launch_step(const FvmParams& p, const FvmBuffers& b, hipStream_t s) {
  int rc;
  for (int k = 0; k < 3; ++k)
    if ((rc = launch_one<N>(p, b, k, s))) return rc;
  return JCFD_OK;
}

template <int N>
bool tile_on_menu(int which, int tile) {
  auto row = [](int t) {
    switch (t) {
      case 1: return fwd_ok<N, 1>();
      case 2: return fwd_ok<N, 2>();
      case 4: return fwd_ok<N, 4>();
      case 8: return fwd_ok<N, 8>();
      case 16: return fwd_ok<N, 16>();
      case 32: return fwd_ok<N, 32>();
      default: return false;
    }
  };
  auto proj = [](int t) {
    switch (t) {
      case 1: return proj_ok<N, 1>();
      case 2: return proj_ok<N, 2>();
      case 4: return proj_ok<N, 4>();
      case 8: return proj_ok<N, 8>();
      default: return false;
    }
  };
  auto col = [](int t) {
    switch (t) {
      case 1: return cols_ok<N, 1>();
      case 2: return cols_ok<N, 2>();
      case 4: return cols_ok<N, 4>();
      case 8: return cols_ok<N, 8>();
      case 16: return cols_ok<N, 16>();
      default: return false;
    }
  };
  return which == 0 ? row(tile) : (which == 1 ? col(tile) : proj(tile));
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

// In place on b.u_out and b.v_out (every instance of b): u <- u + irfft2(m .
// rfft2(u)) with m = D / n^2 the implicit-diffusion multiplier.
template <int N>
int launch_diffuse(const FvmParams& p, const FvmBuffers& b, const float* mult, hipStream_t s) {
  constexpr int RB = Tiles<N>::RB, W = Tiles<N>::W;
  for (int c = 0; c < 2; ++c) {
    FvmBuffers f = b;
    float* field = c == 0 ? b.u_out : b.v_out;
    f.u = field;
    f.u_out = field;
    f.lam = mult;
    f.lam_mode = 1;  // the plan's multiplier table, column by column
    hipLaunchKernelGGL((fvm_rows_forward<N, RB, 3>), dim3(N / RB, nb(f)), dim3(kThreads), 0, s, p, f);
    hipLaunchKernelGGL((fvm_poisson_cols<N, W>), dim3(N / 2 / W, nb(f)), dim3(kThreads), 0, s, f, 1.0f);
    hipLaunchKernelGGL((fvm_rows_project<N, RB, 3>), dim3(N / RB, nb(f)), dim3(kThreads), 0, s, p, f);
  }
  return check_hip(hipGetLastError(), "implicit diffusion launch");
}

template <int N>
int launch_rk_explicit(const FvmParams& p, const FvmBuffers& b, hipStream_t s) {
  constexpr int RB = Tiles<N>::RB;
  hipLaunchKernelGGL((fvm_rows_forward<N, RB, 2>), dim3(N / RB, nb(b)), dim3(kThreads), 0, s, p, b);
  return check_hip(hipGetLastError(), "rk explicit-terms launch");
}

template <int N>
int launch_rk_project(const FvmParams& p, const FvmBuffers& b, hipStream_t s) {
  constexpr int RB = Tiles<N>::RB, W = Tiles<N>::W;
  hipLaunchKernelGGL((fvm_rows_forward<N, RB, 5>), dim3(N / RB, nb(b)), dim3(kThreads), 0, s, p, b);
  hipLaunchKernelGGL((fvm_poisson_cols<N, W>), dim3(N / 2 / W, nb(b)), dim3(kThreads), 0, s, b, 1.0f);
  hipLaunchKernelGGL((fvm_rows_project<N, RB, 0>), dim3(N / RB, nb(b)), dim3(kThreads), 0, s, p, b);
  return check_hip(hipGetLastError(), "rk projection launch");
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
template <int N> struct DiffuseOp {
  static int run(const FvmParams& p, const FvmBuffers& b, const float* mult, hipStream_t s) {
    return launch_diffuse<N>(p, b, mult, s);
  }
};
template <int N> struct RkExplicitOp {
  static int run(const FvmParams& p, const FvmBuffers& b, hipStream_t s) { return launch_rk_explicit<N>(p, b, s); }
};
template <int N> struct RkProjectOp {
  static int run(const FvmParams& p, const FvmBuffers& b, hipStream_t s) { return launch_rk_project<N>(p, b, s); }
};
template <int N> struct ProjOp {
  static int run(const FvmParams& p, const FvmBuffers& b, hipStream_t s) { return launch_projection<N>(p, b, s); }
};

// ---- mixed-size group step -------------------------------------------------
// Each phase is one launch over the phase's work list (instance, tile), one
// workgroup per item, so one launch covers every instance of every grid
// size.  The host sorts the list largest grid first.  There is no shared
// work counter: device-scope atomics on one address serialise across the
// XCDs and cost more than they balance.  Each item runs the same tile body
// as the per-size kernels, so an instance's arithmetic -- and its result,
// bit for bit -- does not depend on what it is grouped with.
// Tile shapes of the group's items, sized so that no item is much longer
// than a workgroup's share of a mixed group: the marching stencil for
// n >= 256 (as a tuned batch of those sizes uses) with 4096 cells per item
// (16 rows of 256, 8 rows of 512), the default row blocks below that, and
// Poisson column tiles of about 2048 words (n * W).  Every shape gives the
// same results bit for bit.
template <int N>
struct GroupTiles {
  static constexpr bool kMarch = N >= 256;
  static constexpr int RB_FWD = kMarch ? 4096 / N : Tiles<N>::RB;
  static constexpr int W = 2048 / N < N / 2 ? 2048 / N : N / 2;
  static constexpr int RB_PROJ = Tiles<N>::RB;
};

template <int N, int PHASE>
__device__ __forceinline__ void group_tile(const FvmGroupDesc& d, int tile, int tid, char* smem) {
  using G = GroupTiles<N>;
  if constexpr (PHASE == 0) {
    const FvmParams& p = d.p;
    if constexpr (G::kMarch) {
#define JCFD_GMARCH(C) rows_forward_march_body<N, G::RB_FWD, 8, C>(d.p, d.b, tile * G::RB_FWD, tid, smem)
      JCFD_CONVECT_DISPATCH(JCFD_GMARCH)
#undef JCFD_GMARCH
    } else {
#define JCFD_GROWS(C) rows_forward_body<N, G::RB_FWD, 0, 8, C>(d.p, d.b, tile * G::RB_FWD, tid, smem)
      JCFD_CONVECT_DISPATCH(JCFD_GROWS)
#undef JCFD_GROWS
    }
  } else if constexpr (PHASE == 1) {
    poisson_cols_body<N, G::W>(d.b, 1.0f, tile * G::W, tid, smem);
  } else {
    rows_project_body<N, G::RB_PROJ, 0>(d.p, d.b, tile * G::RB_PROJ, tid, smem);
  }
}

// One workgroup per item (the host launches exactly the phase's item
// count): the hardware's workgroup dispatch keeps the CUs full, as for the
// per-size kernels, with no loop state held in registers (a persistent
// fixed-stride walk over the list measured 7-15 % slower).
template <int PHASE>
__global__ __launch_bounds__(kThreads) void fvm_group_phase(FvmGroupQueue q) {
  extern __shared__ float4 fvm_group_smem[];
  char* smem = reinterpret_cast<char*>(fvm_group_smem);
  const int it = blockIdx.x;
  if (it >= q.n_items[PHASE]) return;
  const FvmGroupItem w = q.items[PHASE][it];
  FvmGroupDesc d = q.descs[w.desc];
  // the descriptor's pointers are all global memory
  d.b.u = as_global(d.b.u);
  d.b.v = as_global(d.b.v);
  d.b.us = as_global(d.b.us);
  d.b.vs = as_global(d.b.vs);
  d.b.spec = as_global(d.b.spec);
  d.b.u_out = as_global(d.b.u_out);
  d.b.v_out = as_global(d.b.v_out);
  d.b.fx = as_global(d.b.fx);
  d.b.lam = as_global(d.b.lam);
  d.b.lam_xy = as_global(d.b.lam_xy);
  d.b.tw = as_global(d.b.tw);
  d.b.slots = as_global(d.b.slots);
  d.b.rhs = as_global(d.b.rhs);
  d.b.q = as_global(d.b.q);
  d.b.rk = as_global(d.b.rk);
  const int t = threadIdx.x;
  switch (d.p.n) {
    case 32: group_tile<32, PHASE>(d, w.tile, t, smem); break;
    case 64: group_tile<64, PHASE>(d, w.tile, t, smem); break;
    case 128: group_tile<128, PHASE>(d, w.tile, t, smem); break;
    case 256: group_tile<256, PHASE>(d, w.tile, t, smem); break;
    case 512: group_tile<512, PHASE>(d, w.tile, t, smem); break;
    default: break;  // rejected on the host
  }
}

template <int N>
constexpr int group_lds_bytes(int phase) {
  using G = GroupTiles<N>;
  if (phase == 0)
    return G::kMarch ? MarchLds<N, G::RB_FWD>::bytes : RowsForwardLds<N, G::RB_FWD, 0>::bytes;
  return phase == 1 ? PoissonLds<N, G::W>::bytes : RowsProjectLds<N, G::RB_PROJ, 0>::bytes;
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
  using G = GroupTiles<N>;
  return phase == 0 ? N / G::RB_FWD : phase == 1 ? N / 2 / G::W : N / G::RB_PROJ;
}


}  // namespace

bool fvm_size_supported(int n) {
  return n == 32 || n == 64 || n == 128 || n == 256 || n == 512 || n == 1024;
}

bool fvm_group_size_supported(int n) { return fvm_size_supported(n) && n <= kFvmGroupMaxN; }

template <int N> struct DefaultTilesOp {
  static int run(FvmParams* p) {
    p->rb_fwd = Tiles<N>::RB;
    p->w_cols = Tiles<N>::W;
    p->rb_proj = Tiles<N>::RB;
    p->radix = 8;
    return JCFD_OK;
  }
};
template <int N> struct TileMenuOp {
  static int run(int which, int tile, bool* ok) {
    *ok = tile_on_menu<N>(which, tile);
    return JCFD_OK;
  }
};

int fvm_default_tiles(FvmParams* p) { return dispatch_n<DefaultTilesOp>(p->n, p); }

bool fvm_tile_supported(int n, int which, int tile) {
  bool ok = false;
  return dispatch_n<TileMenuOp>(n, which, tile, &ok) == JCFD_OK && ok;
}

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

int fvm_launch_diffuse(const FvmParams& p, const FvmBuffers& b, const float* mult, hipStream_t s) {
  return dispatch_n<DiffuseOp>(p.n, p, b, mult, s);
}

int fvm_launch_rk_explicit(const FvmParams& p, const FvmBuffers& b, hipStream_t s) {
  return dispatch_n<RkExplicitOp>(p.n, p, b, s);
}

int fvm_launch_rk_project(const FvmParams& p, const FvmBuffers& b, hipStream_t s) {
  return dispatch_n<RkProjectOp>(p.n, p, b, s);
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
                          float* v_out, float2* spec, const double* lam_xy, const float2* tw,
                          hipStream_t s) {
  FvmBuffers b{};
  b.u = u;
  b.v = v;
  b.us = const_cast<float*>(u);
  b.vs = const_cast<float*>(v);
  b.u_out = u_out;
  b.v_out = v_out;
  b.spec = spec;
  b.lam_xy = lam_xy;
  b.lam_mode = 2;  // Lambda^+ from the plan's eigenvalues
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

// advect_step_semilagrangian (advection.py:345-382) of a periodic N x N field
// c at offset (ox, oy): u and v linearly interpolated to c's offset
// (interpolation.py:36-93, axis 0 then axis 1), departure points
// x - dt u on c's mesh, their fractional indices x / h - o, then bilinear
// interpolation with periodic wrap in the term order of
// jax.scipy.ndimage.map_coordinates(order=1, mode='wrap'): weights
// (1 - f, f) per axis, contributions (lo,lo), (lo,up), (up,lo), (up,up)
// summed left to right, each weight product formed before the value.
__device__ __forceinline__ int wrap_index(int k, int n) { return ((k % n) + n) % n; }

// one axis of interpolation.linear at (i, j) of a field whose value at
// (i, j) along the other axis is given by `at`
template <typename At>
__device__ __forceinline__ float lin_axis(const SlAxis& a, At at) {
  if (!a.blend) return at(a.f);
  return a.fw * at(a.f) + a.cw * at(a.c);
}

__device__ __forceinline__ float lin_interp(const float* f, int n, int i, int j, const SlAxis& a0,
                                            const SlAxis& a1) {
  auto along0 = [&](int dj) {
    const int jj = wrap_index(j + dj, n);
    return lin_axis(a0, [&](int di) { return f[wrap_index(i + di, n) * n + jj]; });
  };
  return lin_axis(a1, along0);
}

__global__ __launch_bounds__(256) void fvm_semilagrangian_kernel(int n, const float* __restrict__ c,
                                                                 const float* __restrict__ u,
                                                                 const float* __restrict__ v,
                                                                 float* __restrict__ out, SlParams q) {
  const int e = blockIdx.x * 256 + threadIdx.x;
  if (e >= n * n) return;
  const int i = e / n, j = e % n;
  const float ui = lin_interp(u, n, i, j, q.u0, q.u1);
  const float vi = lin_interp(v, n, i, j, q.v0, q.v1);
  const float x = 0.f + (static_cast<float>(i) + q.ox) * q.hx;  // grid.mesh(offset), lower 0
  const float y = 0.f + (static_cast<float>(j) + q.oy) * q.hy;
  const float xi = (x - q.dt * ui) / q.hx - q.ox;
  const float yi = (y - q.dt * vi) / q.hy - q.oy;
  const float fx = floorf(xi), fy = floorf(yi);
  const float wx1 = xi - fx, wy1 = yi - fy;
  const float wx0 = 1.f - wx1, wy0 = 1.f - wy1;
  const int i0 = static_cast<int>(fx), j0 = static_cast<int>(fy);
  const int a0 = wrap_index(i0, n), a1 = wrap_index(i0 + 1, n);
  const int b0 = wrap_index(j0, n), b1 = wrap_index(j0 + 1, n);
  const float r = (((wx0 * wy0) * c[a0 * n + b0] + (wx0 * wy1) * c[a0 * n + b1]) +
                   (wx1 * wy0) * c[a1 * n + b0]) +
                  (wx1 * wy1) * c[a1 * n + b1];
  out[e] = r;
}

int fvm_launch_semilagrangian(int n, const float* c, const float* u, const float* v, float* out,
                              const SlParams& q, hipStream_t s) {
  const int total = n * n;
  hipLaunchKernelGGL(fvm_semilagrangian_kernel, dim3((total + 255) / 256), dim3(256), 0, s, n, c, u, v,
                     out, q);
  return check_hip(hipGetLastError(), "semilagrangian launch");
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
