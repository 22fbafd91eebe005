// Pseudospectral vorticity step (NavierStokes2D + a low-storage Runge-Kutta /
// Crank-Nicolson stepper: crank_nicolson_rk4, _rk3 or any tableau) for gfx950.
//
// State: the complex64 rfft2 of the vorticity, (n, n/2+1) per instance
// (jax_cfd/spectral/equations.py:129-190).  One step
// (spectral/time_stepping.py:117-170; CN-RK4, 186-197, has five) is stages of
//   F = explicit_terms(u);  h = F + beta_k h;
//   u = (u + gamma_k dt h + mu_k L u) / (1 - mu_k L),  L = nu lap - drag,
// where explicit_terms = filter * rfftn(-(d_x w * v_x + d_y w * v_y)) + forcing
// curl, with v from vorticity_to_velocity (spectral/utils.py:154-189).
//
// Kernels (all launched with blockIdx.y = instance of a same-size batch):
//   ns_cols_inverse   column tile of W packed columns: build the four spectra
//                     d_x w, d_y w, v_x, v_y (scaled by 1/n^2), inverse column
//                     FFT each, write them packed -> G[4].
//   ns_rows           RB rows: four inverse packed real row FFTs, the
//                     product -(gx vx + gy vy), forward packed real row FFT -> P.
//   ns_cols_forward   column tile: forward column FFT of P, unpack the DC/
//                     Nyquist word, 2/3 filter, forcing curl, then either write
//                     F (explicit-terms seam) or the RK/CN update of (u, h),
//                     optionally followed directly by the next stage's
//                     ns_cols_inverse work on the same tile.
//
// Packed half spectra: word 0 of a row holds (X[0], X[n/2]) of a real row;
// the column pass of word 0 therefore carries two Hermitian columns, which
// are separated with Y(p) = A(p) + i B(p), A = (Y(p) + conj Y(-p))/2,
// B = (Y(p) - conj Y(-p))/(2i).  On the way in, (A, B) are symmetrised the
// same way, which reproduces irfftn's dropping of the imaginary parts of the
// DC and Nyquist columns after the axis-0 inverse transform.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>

#include "jcfd_fft.h"
#include "jcfd_internal.h"

#pragma clang fp contract(off)

namespace jcfd {
namespace {

constexpr int kThreads = 256;

// Internal layout of packed half spectra (n rows x n/2 words: the four
// derivative spectra G and the product spectrum P): column tiles of kTw<N>
// words, each tile one contiguous block of n rows x kTw words.  A column
// kernel's tile is then a contiguous block, and a row kernel's rows cover
// whole cache lines in every tile (4 rows x 4 words = 128 bytes), where the
// row-major layout left 32-byte pieces of 2 KB rows to both.  From 512 on
// the column kernels run two-column tiles: tiles of 2 words make each of
// their rows a whole 16-byte piece of one tile (with 4-word tiles two
// workgroups wrote the halves of every 32-byte piece; the update-and-inverse
// kernel at 64 x 512^2 ran 149 instead of 133 us).
template <int N>
constexpr int kTw = N >= 512 ? 2 : 4;
template <int N>
__device__ __forceinline__ int tix(int p, int c) {
  return ((c / kTw<N>) * N + p) * kTw<N> + (c % kTw<N>);
}
// Internal state layout (n rows x (n/2 + 1) words, the same size as the
// external rfft2 layout): columns 0 .. n/2 - 1 tiled as above, then column
// n/2 (the Nyquist column) as n consecutive words.
template <int N>
__device__ __forceinline__ int six(int p, int c) {
  return c < N / 2 ? tix<N>(p, c) : (N / 2) * N + p;
}

// Per-element multipliers of the four derivative spectra, following the
// reference's fp32/complex64 operation order.
struct NsMult {
  float ax;   // 2 pi kx  (d_x multiplier is i ax)
  float ay;   // 2 pi ky
  float psi;  // -1 / laplace'  (laplace' = (2 pi i)^2 (kx^2 + ky^2), [0,0] := 1)
};

__device__ __forceinline__ NsMult ns_mult(const NsParams& q, float kx, float ky, bool origin) {
  NsMult m;
  m.ax = q.two_pi * kx;
  m.ay = q.two_pi * ky;
  const float s = kx * kx + ky * ky;
  const float lap = origin ? 1.0f : q.neg_four_pi_sq * s;
  m.psi = -1.0f / lap;
  return m;
}

// i a * z
__device__ __forceinline__ float2 mul_i(float a, float2 z) { return make_float2(-a * z.y, a * z.x); }

// Spectrum m in {0: d_x w, 1: d_y w, 2: v_x, 3: v_y, 4: w itself} at (p, c).
__device__ __forceinline__ float2 ns_spectrum(int m, const NsMult& k, float2 w) {
  switch (m) {
    case 0: return mul_i(k.ax, w);
    case 1: return mul_i(k.ay, w);
    case 2: return mul_i(k.ay, make_float2(k.psi * w.x, k.psi * w.y));
    case 3: return mul_i(-k.ax, make_float2(k.psi * w.x, k.psi * w.y));
    default: return w;
  }
}

// Spectra built and inverse-transformed per round of the inverse column
// work: a round's spectra sit side by side as NSP * W columns of one LDS
// tile, so one set of FFT passes (and barriers) serves all of them.  NSP is
// the largest of 4, 2, 1 whose round fits kNsRoundWords complex words.
constexpr int kNsMaxSpectraPerRound = 4;
constexpr int kNsRoundWords = 2048;
template <int N, int W>
struct NsRound {
  static constexpr int value = (kNsMaxSpectraPerRound >= 4 && 4 * N * W <= kNsRoundWords) ? 4
                               : (kNsMaxSpectraPerRound >= 2 && 2 * N * W <= kNsRoundWords) ? 2
                                                                                            : 1;
};

template <int N, int W>
struct ColShared {
  float2 buf[N * W * NsRound<N, W>::value];  // work tile [p][c] (inverse rounds: [p][spectrum][c])
  float2 w0[N];       // tile 0: vorticity column 0
  float2 wn[N];       // tile 0: vorticity column N/2
  float2 tw[N];
  float kx[N];        // fftfreq
  float ky[N / 2 + 1];  // rfftfreq
};

// Elements of a column tile are owned by threads as e = tid + k * kThreads;
// each thread keeps its vorticity values in registers across the four spectra
// (only tile 0's DC/Nyquist columns, which pair rows p and -p, also go through
// LDS).
template <int N, int W>
struct WTile {
  static constexpr int PER = (N * W + kThreads - 1) / kThreads;
  float2 w[PER];
};

// Build spectrum m of the tile into sm.buf (scaled by 1/n^2), symmetrising and
// packing the DC/Nyquist pair into local column 0 of tile 0.  psi[k] is
// -1/laplacian at this thread's element k (columns other than 0; one
// division per element for all four spectra); 2 pi kx, 2 pi ky come from
// the LDS wavenumber tables.
template <int N, int W, int NSP = 1>
__device__ __forceinline__ void build_spectrum(ColShared<N, W>& sm, const WTile<N, W>& wt,
                                               const float (&psi)[WTile<N, W>::PER],
                                               const NsParams& q, int m, int slot, int c0, int tid) {
  constexpr int H = N / 2;
#pragma unroll
  for (int k = 0; k < WTile<N, W>::PER; ++k) {
    const int e = tid + k * kThreads;
    if (e >= N * W) break;
    const int p = e / W, c = e % W, cg = c0 + c;
    float2 out;
    if (cg != 0) {
      NsMult mk;
      mk.ax = q.two_pi * sm.kx[p];
      mk.ay = q.two_pi * sm.ky[cg];
      mk.psi = psi[k];
      out = ns_spectrum(m, mk, wt.w[k]);
    } else {
      const int pm = (N - p) & (N - 1);
      const float2 a0 = ns_spectrum(m, ns_mult(q, sm.kx[p], sm.ky[0], p == 0), sm.w0[p]);
      const float2 b0 = ns_spectrum(m, ns_mult(q, sm.kx[pm], sm.ky[0], pm == 0), sm.w0[pm]);
      const float2 ah = ns_spectrum(m, ns_mult(q, sm.kx[p], sm.ky[H], false), sm.wn[p]);
      const float2 bh = ns_spectrum(m, ns_mult(q, sm.kx[pm], sm.ky[H], false), sm.wn[pm]);
      const float2 s0 = make_float2(0.5f * (a0.x + b0.x), 0.5f * (a0.y - b0.y));
      const float2 sh = make_float2(0.5f * (ah.x + bh.x), 0.5f * (ah.y - bh.y));
      out = make_float2(s0.x - sh.y, s0.y + sh.x);  // s0 + i sh
    }
    sm.buf[p * (NSP * W) + slot * W + c] = make_float2(out.x * q.inv_nn, out.y * q.inv_nn);
  }
}

// Inverse column transforms of the tile -> G[0..NSPEC): the four derivative
// spectra (NSPEC = 4) or the field itself (NSPEC = 1).  sm.w0 / sm.wn must be
// visible (barrier) before the call.
template <int N, int W, int NSPEC = 4, int R = 8>
__device__ __forceinline__ void cols_inverse_body(ColShared<N, W>& sm, const WTile<N, W>& wt,
                                                  const NsParams& q, float2* __restrict__ g,
                                                  int c0, int tid) {
  constexpr int H = N / 2;
  constexpr int PER = WTile<N, W>::PER;
  const size_t gstride = static_cast<size_t>(N) * H;
  float psi[PER];
#pragma unroll
  for (int k = 0; k < PER; ++k) {
    const int e = tid + k * kThreads;
    const int p = (e < N * W ? e : N * W - 1) / W, cg = c0 + e % W;
    psi[k] = (NSPEC == 4 && cg != 0) ? ns_mult(q, sm.kx[p], sm.ky[cg], false).psi : 0.f;
  }
  constexpr int NSP = NSPEC < NsRound<N, W>::value ? NSPEC : NsRound<N, W>::value;
  static_assert(NSPEC % NSP == 0, "whole rounds");
#pragma unroll 1
  for (int m0 = 0; m0 < NSPEC; m0 += NSP) {
#pragma unroll
    for (int j = 0; j < NSP; ++j)
      build_spectrum<N, W, NSP>(sm, wt, psi, q, NSPEC == 1 ? 4 : m0 + j, j, c0, tid);
    __syncthreads();
    lds_fft<N, NSP * W, kThreads, NSP * W, 1, true, true, R>(sm.buf, sm.tw, tid);
#pragma unroll
    for (int j = 0; j < NSP; ++j) {
      float2* gm = g + (m0 + j) * gstride;
      float2 o[PER];
#pragma unroll
      for (int k = 0; k < PER; ++k) {
        const int e = tid + k * kThreads;
        if ((N * W) % kThreads == 0 || e < N * W) o[k] = sm.buf[(e / W) * (NSP * W) + j * W + e % W];
      }
#pragma unroll
      for (int k = 0; k < PER; ++k) {
        const int e = tid + k * kThreads;
        if ((N * W) % kThreads == 0 || e < N * W) gm[tix<N>(e / W, c0 + e % W)] = o[k];
      }
    }
    __syncthreads();
  }
}

// The state tile (internal layout when TILED, else the external rfft2
// layout) into registers; tile 0 also stages columns 0 and N/2 in LDS.
template <int N, int W, bool TILED>
__device__ __forceinline__ void load_w_tile(ColShared<N, W>& sm, WTile<N, W>& wt, const NsParams& q,
                                            const float2* __restrict__ w, int c0, int tid) {
  constexpr int H1 = N / 2 + 1;
  auto at = [&](int p, int c) { return TILED ? six<N>(p, c) : p * H1 + c; };
#pragma unroll
  for (int k = 0; k < WTile<N, W>::PER; ++k) {
    const int e = tid + k * kThreads;
    if (e >= N * W) break;
    const int p = e / W, c = e % W;
    wt.w[k] = w[at(p, c0 + c)];
  }
  if (c0 == 0) {
    constexpr int PN = (N + kThreads - 1) / kThreads;
    float2 a[PN], b[PN];
#pragma unroll
    for (int k = 0; k < PN; ++k) {
      const int p = tid + k * kThreads;
      if (N % kThreads == 0 || p < N) {
        a[k] = w[at(p, 0)];
        b[k] = w[at(p, N / 2)];
      }
    }
#pragma unroll
    for (int k = 0; k < PN; ++k) {
      const int p = tid + k * kThreads;
      if (N % kThreads == 0 || p < N) {
        sm.w0[p] = a[k];
        sm.wn[p] = b[k];
      }
    }
  }
}

// Wavenumber tables staged through registers (load() issues the global
// loads, store() writes LDS), like TwiddleStage.
template <int N>
struct WavenumberStage {
  static constexpr int PX = (N + kThreads - 1) / kThreads;
  static constexpr int PY = (N / 2 + 1 + kThreads - 1) / kThreads;
  float x[PX], y[PY];
  __device__ __forceinline__ void load(const NsParams& q, int tid) {
#pragma unroll
    for (int k = 0; k < PX; ++k) {
      const int i = tid + k * kThreads;
      x[k] = q.kx[N % kThreads == 0 || i < N ? i : N - 1];
    }
#pragma unroll
    for (int k = 0; k < PY; ++k) {
      const int i = tid + k * kThreads;
      y[k] = q.ky[i <= N / 2 ? i : N / 2];
    }
  }
  __device__ __forceinline__ void store(float* kx, float* ky, int tid) const {
#pragma unroll
    for (int k = 0; k < PX; ++k) {
      const int i = tid + k * kThreads;
      if (N % kThreads == 0 || i < N) kx[i] = x[k];
    }
#pragma unroll
    for (int k = 0; k < PY; ++k) {
      const int i = tid + k * kThreads;
      if (i <= N / 2) ky[i] = y[k];
    }
  }
};


// Minimum waves per SIMD asked of the column kernels: 5 for the 512-point
// tiles of one or two columns, the most their LDS (inverse rounds of two or
// four spectra) allows; other shapes keep the compiler's choice (wider
// tiles would spill).
template <int N, int W>
struct NsColWaves {
  static constexpr int value = (N == 512 && W <= 2) ? 5 : 1;
};

// One column tile: the state's W columns from c0 -> the inverse transforms of
// the four spectra (NSPEC = 4) or of the field (NSPEC = 1); w and g already
// point at the instance.
template <int N, int W, int NSPEC, int R = 8>
__device__ __forceinline__ void cols_inverse_tile(ColShared<N, W>& sm, const NsParams& q,
                                                  const float2* __restrict__ w, float2* __restrict__ g,
                                                  int c0, int tid) {
  WTile<N, W> wt;
  TwiddleStage<N, 1, kThreads> tw;
  WavenumberStage<N> wn;
  tw.load(q.tw, tid);
  wn.load(q, tid);
  // the step (NSPEC = 4) reads the internal state, irfft2 (NSPEC = 1) the
  // caller's array
  load_w_tile<N, W, NSPEC == 4>(sm, wt, q, w, c0, tid);
  tw.store(sm.tw, tid);
  wn.store(sm.kx, sm.ky, tid);
  __syncthreads();
  cols_inverse_body<N, W, NSPEC, R>(sm, wt, q, g, c0, tid);
}

template <int N, int W, int NSPEC, int R = 8>
__global__ __launch_bounds__(kThreads, (NsColWaves<N, W>::value)) void ns_cols_inverse(NsParams q, const float2* __restrict__ w,
                                                            float2* __restrict__ g) {
  __shared__ ColShared<N, W> sm;
  const int c0 = xcd_tile(blockIdx.x, gridDim.x) * W;
  const size_t inst = q.slots ? static_cast<size_t>(q.slots[blockIdx.y]) : blockIdx.y;
  cols_inverse_tile<N, W, NSPEC, R>(sm, q, w + inst * static_cast<size_t>(N) * (N / 2 + 1),
                                    g + inst * NSPEC * static_cast<size_t>(N) * (N / 2), c0, threadIdx.x);
}

// Packed real row transforms of a batch of real (N, N) fields.
template <int N, int RB>
__global__ __launch_bounds__(kThreads) void rows_r2c(NsParams q, const float* __restrict__ x,
                                                     float2* __restrict__ pout) {
  constexpr int H = N / 2;
  __shared__ float2 zb[RB * H];
  __shared__ float2 twn[N];
  __shared__ float2 twh[H];
  const int tid = threadIdx.x;
  const int i0 = xcd_tile(blockIdx.x, gridDim.x) * RB;
  const size_t inst = q.slots ? static_cast<size_t>(q.slots[blockIdx.y]) : blockIdx.y;
  x += inst * static_cast<size_t>(N) * N;
  pout += inst * static_cast<size_t>(N) * H;
  TwiddleStage<N, 1, kThreads> ta;
  TwiddleStage<H, 2, kThreads> tb;
  ta.load(q.tw, tid);
  tb.load(q.tw, tid);
  const float* xr = x + static_cast<size_t>(i0) * N;
  float* zf = reinterpret_cast<float*>(zb);
  staged<RB * N, 16, kThreads>(tid, [&](int e) { return xr[e]; }, [&](int e, float v) { zf[e] = v; });
  ta.store(twn, tid);
  tb.store(twh, tid);
  __syncthreads();
  lds_fft_waves<H, RB, kThreads, 1, H, false>(zb, twh, tid);
  for (int e = tid; e < RB * H; e += kThreads) {
    const int r = e / H, k = e % H;
    pout[tix<N>(i0 + r, k)] = r2c_combine(zb + r * H, k, H, twn[k]);
  }
}

template <int N, int RB>
__global__ __launch_bounds__(kThreads) void rows_c2r(NsParams q, const float2* __restrict__ pin,
                                                     float* __restrict__ x) {
  constexpr int H = N / 2;
  __shared__ float2 zb[RB * H];
  __shared__ float2 twh[H];
  const int tid = threadIdx.x;
  const int i0 = xcd_tile(blockIdx.x, gridDim.x) * RB;
  const size_t inst = q.slots ? static_cast<size_t>(q.slots[blockIdx.y]) : blockIdx.y;
  x += inst * static_cast<size_t>(N) * N;
  pin += inst * static_cast<size_t>(N) * H;
  TwiddleStage<H, 2, kThreads> tb;
  tb.load(q.tw, tid);
  staged<RB * H, 8, kThreads>(
      tid,
      [&](int e) {
        const int r = e / H, k = e % H;
        return F2Triple{pin[tix<N>(i0 + r, k)], pin[tix<N>(i0 + r, (H - k) & (H - 1))], q.tw[k]};
      },
      [&](int e, F2Triple v) { zb[e] = c2r_combine(v.a, v.b, e % H, v.t); });
  tb.store(twh, tid);
  __syncthreads();
  lds_fft_waves<H, RB, kThreads, 1, H, true>(zb, twh, tid);
  const float* zf = reinterpret_cast<const float*>(zb);
  for (int e = tid; e < RB * N; e += kThreads) x[static_cast<size_t>(i0) * N + e] = zf[e];
}

template <int N, int RB>
struct RowsShared {
  float2 rows[4 * RB * (N / 2)];  // transform t = m * RB + r
  float2 twn[N];
  float2 twh[N / 2];
};

// One row block: rows i0 .. i0+RB-1 of the four spectra G -> the product
// spectrum rows of P; g and pout already point at the instance.
template <int N, int RB, int R = 8>
__device__ __forceinline__ void rows_tile(RowsShared<N, RB>& sm, const NsParams& q,
                                          const float2* __restrict__ g, float2* __restrict__ pout,
                                          int i0, int tid) {
  constexpr int H = N / 2;
  const size_t gstride = static_cast<size_t>(N) * H;
  float2* rows = sm.rows;
  // the product spectrum overwrites the first transform's rows (gx, element
  // by element: each element is read and written by its owner only)
  float2* prod = sm.rows;
  float2* twn = sm.twn;
  float2* twh = sm.twh;
  // Load the four spectra's rows in chunks (all loads of a chunk issued
  // together); the twiddle tables go to LDS while the first chunk is in
  // flight.  Words are taken in mirror pairs (kk, H - kk), kk < H/2: both
  // outputs of a pair need both words, so each word is read once.  Element
  // e walks the first half of the tiled layout in memory order: word w of
  // the tile, then row r (the RB rows of a tile are contiguous), then the
  // tile, then the spectrum m; word H/2 (its own mirror) goes alone.
  constexpr int HH = H / 2;
  constexpr int TW = kTw<N>;
  static_assert(HH % TW == 0, "whole tiles in each half row");
  constexpr int TOT = 4 * RB * HH;
  constexpr int PER = (TOT + kThreads - 1) / kThreads;
  constexpr int C = PER < 8 ? PER : 8;
  static_assert(4 * RB <= kThreads, "one thread per middle word");
  TwiddleStage<N, 1, kThreads> ta;
  TwiddleStage<H, 2, kThreads> tb;
  ta.load(q.tw, tid);
  tb.load(q.tw, tid);
  float2 mid = make_float2(0.f, 0.f);
  if (tid < 4 * RB) mid = g[(tid / RB) * gstride + tix<N>(i0 + tid % RB, HH)];
#pragma unroll 1
  for (int k0 = 0; k0 < PER; k0 += C) {
    float2 a[C], b[C];
#pragma unroll
    for (int k = 0; k < C; ++k) {
      const int e = tid + (k0 + k) * kThreads;
      if (e < TOT) {
        const int r = (e / TW) % RB, kk = ((e / (TW * RB)) % (HH / TW)) * TW + e % TW;
        const float2* gm = g + (e / (RB * HH)) * gstride;
        a[k] = gm[tix<N>(i0 + r, kk)];
        b[k] = gm[tix<N>(i0 + r, (H - kk) & (H - 1))];
      }
    }
    if (k0 == 0) {
      ta.store(twn, tid);
      tb.store(twh, tid);
      __syncthreads();
      if (tid < 4 * RB) rows[tid * H + HH] = c2r_combine(mid, mid, HH, twn[HH]);
    }
#pragma unroll
    for (int k = 0; k < C; ++k) {
      const int e = tid + (k0 + k) * kThreads;
      if (e < TOT) {
        const int r = (e / TW) % RB, kk = ((e / (TW * RB)) % (HH / TW)) * TW + e % TW;
        float2* row = rows + ((e / (RB * HH)) * RB + r) * H;
        row[kk] = c2r_combine(a[k], b[k], kk, twn[kk]);
        if (kk != 0) row[H - kk] = c2r_combine(b[k], a[k], H - kk, twn[H - kk]);
      }
    }
  }
  __syncthreads();
  lds_fft_waves<H, 4 * RB, kThreads, 1, H, true, R>(rows, twh, tid);
  // advection = -(grad_x * vx + grad_y * vy)  (spectral/equations.py:170)
  const float* gx = reinterpret_cast<const float*>(rows);
  const float* gy = gx + RB * N;
  const float* vx = gy + RB * N;
  const float* vy = vx + RB * N;
  float* pr = reinterpret_cast<float*>(prod);
  staged<RB * N, 8, kThreads>(
      tid, [&](int e) { return -(gx[e] * vx[e] + gy[e] * vy[e]); },
      [&](int e, float v) { pr[e] = v; });
  __syncthreads();
  // The product's RB forward transforms: with radix-8 passes a transform
  // has H / 8 butterflies, half a wavefront's lanes or fewer for n <= 512;
  // radix-4 passes (H / 4 butterflies) issue fewer instructions per
  // wavefront there.  A function of n and the plan's radix only, so every
  // row tile shape still gives the same bits.
  constexpr int RF = (R == 8 && N <= 512) ? 4 : R;
  lds_fft_waves<H, RB, kThreads, 1, H, false, RF>(prod, twh, tid);
  // element e in the tiled layout's memory order, as for the loads
  auto row_of = [](int e) { return (e / TW) % RB; };
  auto word_of = [](int e) { return (e / (TW * RB)) * TW + e % TW; };
  staged<RB * H, 8, kThreads>(
      tid, [&](int e) { return r2c_combine(prod + row_of(e) * H, word_of(e), H, twn[word_of(e)]); },
      [&](int e, float2 v) { pout[tix<N>(i0 + row_of(e), word_of(e))] = v; });
}


template <int N, int RB, int R = 8>
__global__ __launch_bounds__(kThreads) void ns_rows(NsParams q, const float2* __restrict__ g,
                                                    float2* __restrict__ pout) {
  __shared__ RowsShared<N, RB> sm;
  const size_t inst = q.slots ? static_cast<size_t>(q.slots[blockIdx.y]) : blockIdx.y;
  const size_t gstride = static_cast<size_t>(N) * (N / 2);
  rows_tile<N, RB, R>(sm, q, g + inst * 4 * gstride, pout + inst * gstride,
                      xcd_tile(blockIdx.x, gridDim.x) * RB, threadIdx.x);
}

// External rfft2 layout <-> internal state layout of one instance per
// blockIdx.y (slots[blockIdx.y] when given); one word per thread, walked in
// the internal order (contiguous internal side).
template <int N, bool TO_INTERNAL>
__global__ __launch_bounds__(kThreads) void ns_state_convert(const float2* __restrict__ src,
                                                             float2* __restrict__ dst,
                                                             const int* __restrict__ slots) {
  constexpr int H = N / 2, H1 = H + 1, TOT = N * H1;
  const int t = blockIdx.x * kThreads + threadIdx.x;
  if (t >= TOT) return;
  const size_t inst = slots ? static_cast<size_t>(slots[blockIdx.y]) : blockIdx.y;
  int p, c;
  if (t < N * H) {
    constexpr int TW = kTw<N>;
    const int tile = t / (N * TW), r = t % (N * TW);
    p = r / TW;
    c = tile * TW + r % TW;
  } else {
    p = t - N * H;
    c = H;
  }
  const size_t base = inst * TOT;
  if (TO_INTERNAL)
    dst[base + t] = src[base + static_cast<size_t>(p) * H1 + c];
  else
    dst[base + static_cast<size_t>(p) * H1 + c] = src[base + t];
}

constexpr int kForceRegs = 4;  // forcing-curl entries kept in registers

// 2/3 rule mask (spectral/utils.py:128-134).
__device__ __forceinline__ float ns_filter(const NsParams& q, int p, int c) {
  if (!q.smooth) return 1.0f;
  const bool row = (p < q.filt_row_lo) || (p >= q.filt_row_hi);
  return (row && c < q.filt_cols) ? 1.0f : 0.0f;
}

// MODE 0: write F (explicit terms) to u_out in the external layout.
// MODE 1: RK/CN update of (u, h) (internal state layout).
// MODE 2: update, then the next stage's inverse column work on the new tile.
// One column tile of the forward column pass: P's W columns from c0 ->
// explicit terms (MODE 0), the RK/CN update (MODE 1), or the update followed
// by the next stage's inverse work (MODE 2); all pointers already point at
// the instance.
template <int N, int W, int MODE, int R = 8>
__device__ __forceinline__ void cols_forward_tile(ColShared<N, W>& sm, const NsParams& q,
                                                  const float2* __restrict__ pin, const float2* u_in,
                                                  float2* u_out, float2* h, float2* __restrict__ g,
                                                  int c0, int tid) {
  constexpr int H = N / 2;
  constexpr int H1 = H + 1;
  constexpr int PER = (N * W + kThreads - 1) / kThreads;

  // Every global load of the kernel's first half is issued up front: the
  // tile of P, the state (u, h) at this thread's elements for the update
  // (each element is read and written by its owner only, so reading u and h
  // before the transform is safe also when u_in == u_out), the twiddles.
  // Issue order matters: a wait for a load also waits for every load issued
  // before it, so the tables (needed first) go first and (u, h) (needed
  // after the transform) last.
  TwiddleStage<N, 1, kThreads> tws;
  [[maybe_unused]] WavenumberStage<N> wns;
  tws.load(q.tw, tid);
  if constexpr (MODE != 0) wns.load(q, tid);
  float2 pv[PER];
#pragma unroll
  for (int k = 0; k < PER; ++k) {
    const int e = tid + k * kThreads;
    if ((N * W) % kThreads == 0 || e < N * W) pv[k] = pin[tix<N>(e / W, c0 + e % W)];
  }
  [[maybe_unused]] float2 u_pre[MODE == 0 ? 1 : PER], h_pre[MODE == 0 ? 1 : PER];
  if constexpr (MODE != 0) {
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const int e = tid + k * kThreads;
      if ((N * W) % kThreads == 0 || e < N * W) {
        const int idx = six<N>(e / W, c0 + e % W);
        u_pre[k] = u_in[idx];
        h_pre[k] = q.read_h ? h[idx] : make_float2(0.f, 0.f);
      }
    }
  }
  tws.store(sm.tw, tid);
  if constexpr (MODE != 0) wns.store(sm.kx, sm.ky, tid);
#pragma unroll
  for (int k = 0; k < PER; ++k) {
    const int e = tid + k * kThreads;
    if ((N * W) % kThreads == 0 || e < N * W) sm.buf[e] = pv[k];
  }
  __syncthreads();
  lds_fft<N, W, kThreads, W, 1, true, false, R>(sm.buf, sm.tw, tid);

  // h = F + beta h;  u = (u + g h + mu L u) / (1 - mu L), with the old
  // (u, h) at (p, cg) given.
  auto update_with = [&](int p, int cg, float2 fv, float2 uo, float2 ho) -> float2 {
    const int idx = six<N>(p, cg);
    float2 hv = fv;
    if (q.read_h) hv = make_float2(fv.x + q.beta * ho.x, fv.y + q.beta * ho.y);
    h[idx] = hv;
    const float kx = sm.kx[p], ky = sm.ky[cg];
    const float lap = q.neg_four_pi_sq * (kx * kx + ky * ky);
    const float L = q.nu * lap - q.drag;
    const float2 lu = make_float2(L * uo.x, L * uo.y);
    const float2 x = make_float2((uo.x + q.g * hv.x) + q.mu * lu.x, (uo.y + q.g * hv.y) + q.mu * lu.y);
    const float inv = 1.0f / (1.0f - q.mu * L);
    const float2 un = make_float2(inv * x.x, inv * x.y);
    u_out[idx] = un;
    return un;
  };
  // Column N/2 (tile 0 only): (u, h) read here.
  auto update_nyquist = [&](int p, float2 fv) -> float2 {
    const int idx = six<N>(p, H);
    const float2 ho = q.read_h ? h[idx] : make_float2(0.f, 0.f);
    return update_with(p, H, fv, u_in[idx], ho);
  };

  // The forcing-curl entries, read once (the element loop's stores could
  // alias them as far as the compiler knows, so it would reload them for
  // every element); longer lists are read from memory in the loop.
  int fpr[kForceRegs], fcr[kForceRegs];
  float2 fvr[kForceRegs];
#pragma unroll
  for (int t = 0; t < kForceRegs; ++t) {
    const int tt = t < q.n_force ? t : 0;
    fpr[t] = q.n_force > 0 ? q.force_idx[2 * tt] : -1;
    fcr[t] = q.n_force > 0 ? q.force_idx[2 * tt + 1] : -1;
    fvr[t] = q.n_force > 0 ? q.force_val[tt] : make_float2(0.f, 0.f);
  }

  // Per element: F at (p, cg) (filter, then the sparse, time-constant forcing
  // curl entries), then the update (or F itself for MODE 0); tile 0 also
  // handles column N/2.  One pass, so no array of F values stays live.
  [[maybe_unused]] WTile<N, W> wt;
#pragma unroll
  for (int k = 0; k < PER; ++k) {
    const int e = tid + k * kThreads;
    if (e < N * W) {
      const int p = e / W, c = e % W, cg = c0 + c;
      const float2 y = sm.buf[e];
      float2 f, fh = make_float2(0.f, 0.f);
      if (cg == 0) {
        const float2 ym = cconj(sm.buf[((N - p) & (N - 1)) * W]);
        const float2 a = make_float2(0.5f * (y.x + ym.x), 0.5f * (y.y + ym.y));
        const float2 d = csub(y, ym);
        const float2 b = make_float2(0.5f * d.y, -0.5f * d.x);
        const float m0 = ns_filter(q, p, 0), mh = ns_filter(q, p, H);
        f = make_float2(a.x * m0, a.y * m0);
        fh = make_float2(b.x * mh, b.y * mh);
      } else {
        const float m = ns_filter(q, p, cg);
        f = make_float2(y.x * m, y.y * m);
      }
      if (q.n_force <= kForceRegs) {
#pragma unroll
        for (int t = 0; t < kForceRegs; ++t) {
          if (t < q.n_force && fpr[t] == p && fcr[t] == cg) f = cadd(f, fvr[t]);
          if (t < q.n_force && cg == 0 && fpr[t] == p && fcr[t] == H) fh = cadd(fh, fvr[t]);
        }
      } else {
        for (int t = 0; t < q.n_force; ++t) {
          const int fp = q.force_idx[2 * t], fc = q.force_idx[2 * t + 1];
          if (fp == p && fc == cg) f = cadd(f, q.force_val[t]);
          if (cg == 0 && fp == p && fc == H) fh = cadd(fh, q.force_val[t]);
        }
      }
      if constexpr (MODE == 0) {
        u_out[p * H1 + cg] = f;
        if (cg == 0) u_out[p * H1 + H] = fh;
      } else {
        const float2 un = update_with(p, cg, f, u_pre[k], h_pre[k]);
        if constexpr (MODE == 2) {
          wt.w[k] = un;
          if (cg == 0) {
            sm.w0[p] = un;
            sm.wn[p] = update_nyquist(p, fh);
          }
        } else {
          if (cg == 0) update_nyquist(p, fh);
        }
      }
    }
  }
  if constexpr (MODE == 2) {
    __syncthreads();  // sm.buf read by all; sm.w0 / sm.wn written
    cols_inverse_body<N, W, 4, R>(sm, wt, q, g, c0, tid);
  }
}


template <int N, int W, int MODE, int R = 8>
__global__ __launch_bounds__(kThreads, (NsColWaves<N, W>::value)) void ns_cols_forward(NsParams q, const float2* __restrict__ pin,
                                                            const float2* u_in, float2* u_out,
                                                            float2* h, float2* __restrict__ g) {
  __shared__ ColShared<N, W> sm;
  const size_t inst = q.slots ? static_cast<size_t>(q.slots[blockIdx.y]) : blockIdx.y;
  const size_t ext = static_cast<size_t>(N) * (N / 2 + 1);
  cols_forward_tile<N, W, MODE, R>(sm, q, pin + inst * static_cast<size_t>(N) * (N / 2),
                                   u_in ? u_in + inst * ext : nullptr, u_out + inst * ext,
                                   h ? h + inst * ext : nullptr,
                                   g ? g + inst * 4 * static_cast<size_t>(N) * (N / 2) : nullptr,
                                   xcd_tile(blockIdx.x, gridDim.x) * W, threadIdx.x);
}

// ---- mixed-size group step ---------------------------------------------------
// Every item is 1024 complex words of work whatever its grid size (column
// tiles of W = 1024 / n columns, row blocks of 1024 / n rows; capped by the
// sizes' limits), so workgroups of every size take about the same time.  Tile shapes regroup the same per-column and
// per-row arithmetic: results are bit-identical to the per-size kernels'.
template <int N>
struct NsGroupShape {
  static constexpr int W = (1024 / N < N / 2) ? 1024 / N : N / 2;
  static constexpr int RB = (1024 / N < N / 4) ? 1024 / N : N / 4;
};

template <int N, int CLS>
__device__ __forceinline__ void ns_group_tile(const NsGroupDesc& d, int tile, int tid, char* smem) {
  using S = NsGroupShape<N>;
  if constexpr (CLS == 1) {
    rows_tile<N, S::RB, 8>(*reinterpret_cast<RowsShared<N, S::RB>*>(smem), d.q, d.g, d.pbuf,
                           tile * S::RB, tid);
  } else {
    auto& sm = *reinterpret_cast<ColShared<N, S::W>*>(smem);
    if constexpr (CLS == 0)
      cols_inverse_tile<N, S::W, 4, 8>(sm, d.q, d.ws, d.g, tile * S::W, tid);
    else
      cols_forward_tile<N, S::W, CLS == 2 ? 2 : 1, 8>(sm, d.q, d.pbuf, d.ws, d.ws, d.h, d.g,
                                                       tile * S::W, tid);
  }
}

// One workgroup per item (the host launches exactly the list's item count):
// the hardware's workgroup dispatch keeps the CUs full, as for the per-size
// kernels, with no loop state held in registers.
template <int CLS>
__global__ __launch_bounds__(kThreads) void ns_group_phase(const NsGroupDesc* __restrict__ descs,
                                                           const NsGroupItem* __restrict__ items,
                                                           const int* __restrict__ n_items, int stage) {
  extern __shared__ float4 ns_group_smem[];
  char* smem = reinterpret_cast<char*>(ns_group_smem);
  const int it = blockIdx.x;
  if (it >= *n_items) return;
  const NsGroupItem w = items[it];
  // the descriptor's pointers are all global memory
  NsGroupDesc d = descs[static_cast<size_t>(w.desc) * JCFD_MAX_RK_STAGES + stage];
  d.ws = as_global(d.ws);
  d.h = as_global(d.h);
  d.g = as_global(d.g);
  d.pbuf = as_global(d.pbuf);
  d.q.kx = as_global(d.q.kx);
  d.q.ky = as_global(d.q.ky);
  d.q.tw = as_global(d.q.tw);
  d.q.force_idx = as_global(d.q.force_idx);
  d.q.force_val = as_global(d.q.force_val);
  const int t = threadIdx.x;
  switch (d.n) {
    case 32: ns_group_tile<32, CLS>(d, w.tile, t, smem); break;
    case 64: ns_group_tile<64, CLS>(d, w.tile, t, smem); break;
    case 128: ns_group_tile<128, CLS>(d, w.tile, t, smem); break;
    case 256: ns_group_tile<256, CLS>(d, w.tile, t, smem); break;
    case 512: ns_group_tile<512, CLS>(d, w.tile, t, smem); break;
    case 1024: ns_group_tile<1024, CLS>(d, w.tile, t, smem); break;
    default: break;  // rejected on the host
  }
}

template <int N>
constexpr int ns_group_lds_n(int cls) {
  using S = NsGroupShape<N>;
  return cls == 1 ? static_cast<int>(sizeof(RowsShared<N, S::RB>))
                  : static_cast<int>(sizeof(ColShared<N, S::W>));
}

// dynamic LDS stays within the 64 KiB a launch gets without raising the
// kernel's limit
static_assert(ns_group_lds_n<1024>(0) <= 65536 && ns_group_lds_n<1024>(1) <= 65536 &&
                  ns_group_lds_n<512>(0) <= 65536 && ns_group_lds_n<512>(1) <= 65536 &&
                  ns_group_lds_n<64>(0) <= 65536 && ns_group_lds_n<64>(1) <= 65536,
              "group LDS");

int ns_group_lds(int cls, int max_n) {
  int bytes = 0;
  if (max_n >= 32) bytes = std::max(bytes, ns_group_lds_n<32>(cls));
  if (max_n >= 64) bytes = std::max(bytes, ns_group_lds_n<64>(cls));
  if (max_n >= 128) bytes = std::max(bytes, ns_group_lds_n<128>(cls));
  if (max_n >= 256) bytes = std::max(bytes, ns_group_lds_n<256>(cls));
  if (max_n >= 512) bytes = std::max(bytes, ns_group_lds_n<512>(cls));
  if (max_n >= 1024) bytes = std::max(bytes, ns_group_lds_n<1024>(cls));
  return bytes;
}


template <int N> struct NsTiles;
template <> struct NsTiles<32> { static constexpr int W = 8, RB = 2; };
template <> struct NsTiles<64> { static constexpr int W = 8, RB = 2; };
template <> struct NsTiles<128> { static constexpr int W = 8, RB = 2; };
template <> struct NsTiles<256> { static constexpr int W = 8, RB = 2; };
template <> struct NsTiles<512> { static constexpr int W = 8, RB = 1; };
template <> struct NsTiles<1024> { static constexpr int W = 4, RB = 1; };

template <int N>
struct NsOps {
  static constexpr int W = NsTiles<N>::W, RB = NsTiles<N>::RB;
  static int cols_inverse(const NsParams& q, const float2* w, float2* g, int batch, hipStream_t s) {
    hipLaunchKernelGGL((ns_cols_inverse<N, W, 4>), dim3(N / 2 / W, batch), dim3(kThreads), 0, s, q, w, g);
    return check_hip(hipGetLastError(), "ns_cols_inverse");
  }
  static int rfft2(const NsParams& q, const float* x, float2* out, float2* pbuf, int batch,
                   hipStream_t s) {
    NsParams plain = q;
    plain.smooth = 0;
    plain.n_force = 0;
    hipLaunchKernelGGL((rows_r2c<N, RB>), dim3(N / RB, batch), dim3(kThreads), 0, s, q, x, pbuf);
    hipLaunchKernelGGL((ns_cols_forward<N, W, 0>), dim3(N / 2 / W, batch), dim3(kThreads), 0, s, plain,
                       pbuf, nullptr, out, nullptr, nullptr);
    return check_hip(hipGetLastError(), "rfft2");
  }
  static int irfft2(const NsParams& q, const float2* in, float* x, float2* pbuf, int batch,
                    hipStream_t s) {
    hipLaunchKernelGGL((ns_cols_inverse<N, W, 1>), dim3(N / 2 / W, batch), dim3(kThreads), 0, s, q, in, pbuf);
    hipLaunchKernelGGL((rows_c2r<N, RB>), dim3(N / RB, batch), dim3(kThreads), 0, s, q, pbuf, x);
    return check_hip(hipGetLastError(), "irfft2");
  }
  static int rows(const NsParams& q, const float2* g, float2* pbuf, int batch, hipStream_t s) {
    hipLaunchKernelGGL((ns_rows<N, RB>), dim3(N / RB, batch), dim3(kThreads), 0, s, q, g, pbuf);
    return check_hip(hipGetLastError(), "ns_rows");
  }
  // Forward column transforms of P -> explicit terms F (no update).
  static int explicit_cols(const NsParams& q, const float2* pbuf, float2* out, int batch, hipStream_t s) {
    hipLaunchKernelGGL((ns_cols_forward<N, W, 0>), dim3(N / 2 / W, batch), dim3(kThreads), 0, s, q, pbuf,
                       nullptr, out, nullptr, nullptr);
    return check_hip(hipGetLastError(), "ns_cols_forward");
  }
};

// ---- tunable step kernels ---------------------------------------------------
// The step's column kernels are built for a menu of tile widths W and its row
// kernel for a menu of rows per workgroup RB, each with both radix schedules;
// NsParams carries a plan's choice (w_cols, rb_rows, radix).  Tile shapes
// regroup the same per-column / per-row arithmetic and never change results.
constexpr int kNsLdsMax = 96 * 1024;

template <int N, int W>
constexpr bool ns_cols_ok() {
  return W <= N / 2 && N * W <= 16 * kThreads && sizeof(ColShared<N, W>) <= kNsLdsMax;
}
template <int N, int RB>
constexpr bool ns_rows_ok() {
  return RB <= N / 4 && (4 * RB * (N / 2) + N + N / 2) * 8 <= kNsLdsMax;
}

// which: 0 cols-inverse, 1 rows, 2 cols-forward + update + next inverse,
// 3 cols-forward + final update.  JCFD_ERR_UNSUPPORTED off the menu.
template <int N, int T, int R>
int ns_variant(int which, const NsParams& q, const float2* a, const float2* u_in, float2* u_out,
               float2* h, float2* g, int batch, hipStream_t s) {
  if (which == 1) {
    if constexpr (ns_rows_ok<N, T>()) {
      hipLaunchKernelGGL((ns_rows<N, T, R>), dim3(N / T, batch), dim3(kThreads), 0, s, q, a, g);
      return JCFD_OK;
    }
    return JCFD_ERR_UNSUPPORTED;
  }
  if constexpr (ns_cols_ok<N, T>()) {
    const dim3 grid(N / 2 / T, batch);
    if (which == 0)
      hipLaunchKernelGGL((ns_cols_inverse<N, T, 4, R>), grid, dim3(kThreads), 0, s, q, a, g);
    else if (which == 2)
      hipLaunchKernelGGL((ns_cols_forward<N, T, 2, R>), grid, dim3(kThreads), 0, s, q, a, u_in, u_out, h, g);
    else
      hipLaunchKernelGGL((ns_cols_forward<N, T, 1, R>), grid, dim3(kThreads), 0, s, q, a, u_in, u_out, h, g);
    return JCFD_OK;
  }
  return JCFD_ERR_UNSUPPORTED;
}

template <int N, int R>
int ns_tile(int which, int tile, const NsParams& q, const float2* a, const float2* u_in, float2* u_out,
            float2* h, float2* g, int batch, hipStream_t s) {
  switch (tile) {
    case 1: return ns_variant<N, 1, R>(which, q, a, u_in, u_out, h, g, batch, s);
    case 2: return ns_variant<N, 2, R>(which, q, a, u_in, u_out, h, g, batch, s);
    case 4: return ns_variant<N, 4, R>(which, q, a, u_in, u_out, h, g, batch, s);
    case 8: return which == 1 ? JCFD_ERR_UNSUPPORTED : ns_variant<N, 8, R>(which, q, a, u_in, u_out, h, g, batch, s);
    case 16: return which == 1 ? JCFD_ERR_UNSUPPORTED : ns_variant<N, 16, R>(which, q, a, u_in, u_out, h, g, batch, s);
    default: return JCFD_ERR_UNSUPPORTED;
  }
}

// One tuned launch.  For rows (which 1) `a` is G and the product spectrum is
// written to g; for the column kernels `a` is the state (0) or P (2, 3).
template <int N>
int ns_launch(int which, const NsParams& q, const float2* a, const float2* u_in, float2* u_out,
              float2* h, float2* g, int batch, hipStream_t s) {
  const int tile = which == 1 ? q.rb_rows : q.w_cols;
  const int rc = q.radix == 4 ? ns_tile<N, 4>(which, tile, q, a, u_in, u_out, h, g, batch, s)
                              : ns_tile<N, 8>(which, tile, q, a, u_in, u_out, h, g, batch, s);
  if (rc) return fail(rc, "spectral tile shape not built for n = " + std::to_string(N));
  return check_hip(hipGetLastError(), "spectral step launch");
}

template <int N>
int ns_step_impl(const NsParams* stages, int nstages, const NsBuffers& b, int batch, hipStream_t s) {
  int rc = ns_launch<N>(0, stages[0], b.u_in, nullptr, nullptr, nullptr, b.g, batch, s);
  for (int k = 0; k < nstages && rc == JCFD_OK; ++k) {
    rc = ns_launch<N>(1, stages[k], b.g, nullptr, nullptr, nullptr, b.pbuf, batch, s);
    if (rc) break;
    const float2* u_src = (k == 0) ? b.u_in : b.u_out;
    rc = ns_launch<N>(k + 1 < nstages ? 2 : 3, stages[k], b.pbuf, u_src, b.u_out, b.h, b.g, batch, s);
  }
  return rc;
}

template <int N>
bool ns_tile_on_menu(int which, int tile) {
  switch (tile) {
    case 1: return which == 1 ? ns_rows_ok<N, 1>() : ns_cols_ok<N, 1>();
    case 2: return which == 1 ? ns_rows_ok<N, 2>() : ns_cols_ok<N, 2>();
    case 4: return which == 1 ? ns_rows_ok<N, 4>() : ns_cols_ok<N, 4>();
    case 8: return which != 1 && ns_cols_ok<N, 8>();
    case 16: return which != 1 && ns_cols_ok<N, 16>();
    default: return false;
  }
}

template <int N>
int ns_explicit_impl(const NsParams& q, const float2* w, float2* out, float2* g, float2* pbuf,
                     int batch, hipStream_t s) {
  int rc = NsOps<N>::cols_inverse(q, w, g, batch, s);
  if (!rc) rc = NsOps<N>::rows(q, g, pbuf, batch, s);
  if (!rc) rc = NsOps<N>::explicit_cols(q, pbuf, out, batch, s);
  return rc;
}

}  // namespace

bool ns_group_size_supported(int n) { return ns_size_supported(n); }

int ns_group_tiles(int n, int list) {
  switch (n) {
    case 32: return list == 0 ? 16 / NsGroupShape<32>::W : 32 / NsGroupShape<32>::RB;
    case 64: return list == 0 ? 32 / NsGroupShape<64>::W : 64 / NsGroupShape<64>::RB;
    case 128: return list == 0 ? 64 / NsGroupShape<128>::W : 128 / NsGroupShape<128>::RB;
    case 256: return list == 0 ? 128 / NsGroupShape<256>::W : 256 / NsGroupShape<256>::RB;
    case 512: return list == 0 ? 256 / NsGroupShape<512>::W : 512 / NsGroupShape<512>::RB;
    case 1024: return list == 0 ? 512 / NsGroupShape<1024>::W : 1024 / NsGroupShape<1024>::RB;
    default: return 0;
  }
}

int ns_group_launch(int cls, int max_n, const NsGroupQueue& q, int stage, int grid, hipStream_t s) {
  if (!ns_group_size_supported(max_n))
    return fail(JCFD_ERR_UNSUPPORTED, "spectral group grid size not supported: " + std::to_string(max_n));
  const size_t lds = static_cast<size_t>(ns_group_lds(cls, max_n));
  switch (cls) {
    case 0: hipLaunchKernelGGL(ns_group_phase<0>, dim3(grid), dim3(kThreads), lds, s, q.descs, q.items[0], q.n_items, stage); break;
    case 1: hipLaunchKernelGGL(ns_group_phase<1>, dim3(grid), dim3(kThreads), lds, s, q.descs, q.items[1], q.n_items + 1, stage); break;
    case 2: hipLaunchKernelGGL(ns_group_phase<2>, dim3(grid), dim3(kThreads), lds, s, q.descs, q.items[0], q.n_items, stage); break;
    default: hipLaunchKernelGGL(ns_group_phase<3>, dim3(grid), dim3(kThreads), lds, s, q.descs, q.items[0], q.n_items, stage);
  }
  return check_hip(hipGetLastError(), "spectral group launch");
}

bool ns_size_supported(int n) {
  return n == 32 || n == 64 || n == 128 || n == 256 || n == 512 || n == 1024;
}

namespace {
template <int N>
int ns_one_impl(int which, const NsParams* stages, int nstages, const NsBuffers& b, int batch,
                hipStream_t s) {
  switch (which) {
    case 0: return ns_launch<N>(0, stages[0], b.u_in, nullptr, nullptr, nullptr, b.g, batch, s);
    case 1: return ns_launch<N>(1, stages[0], b.g, nullptr, nullptr, nullptr, b.pbuf, batch, s);
    case 2: return ns_launch<N>(2, stages[nstages > 1 ? 1 : 0], b.pbuf, b.u_out, b.u_out, b.h, b.g,
                                batch, s);
    default: return ns_launch<N>(3, stages[nstages - 1], b.pbuf, b.u_out, b.u_out, b.h, b.g, batch, s);
  }
}

template <int N> struct NsDefaultOp {
  static void run(NsParams* q) {
    q->w_cols = NsTiles<N>::W;
    q->rb_rows = NsTiles<N>::RB;
    q->radix = 8;
  }
};
}  // namespace

int ns_default_tiles(NsParams* q, int n) {
  switch (n) {
    case 32: NsDefaultOp<32>::run(q); return JCFD_OK;
    case 64: NsDefaultOp<64>::run(q); return JCFD_OK;
    case 128: NsDefaultOp<128>::run(q); return JCFD_OK;
    case 256: NsDefaultOp<256>::run(q); return JCFD_OK;
    case 512: NsDefaultOp<512>::run(q); return JCFD_OK;
    case 1024: NsDefaultOp<1024>::run(q); return JCFD_OK;
    default: return fail(JCFD_ERR_UNSUPPORTED, "spectral grid size not built");
  }
}

bool ns_tile_supported(int n, int which, int tile) {
  switch (n) {
    case 32: return ns_tile_on_menu<32>(which, tile);
    case 64: return ns_tile_on_menu<64>(which, tile);
    case 128: return ns_tile_on_menu<128>(which, tile);
    case 256: return ns_tile_on_menu<256>(which, tile);
    case 512: return ns_tile_on_menu<512>(which, tile);
    case 1024: return ns_tile_on_menu<1024>(which, tile);
    default: return false;
  }
}

int ns_launch_kernel(int n, int which, const NsParams* stages, int nstages, const NsBuffers& b,
                     int batch, hipStream_t s) {
  if (nstages < 1) return fail(JCFD_ERR_INVALID_ARGUMENT, "at least one stage");
  switch (n) {
    case 32: return ns_one_impl<32>(which, stages, nstages, b, batch, s);
    case 64: return ns_one_impl<64>(which, stages, nstages, b, batch, s);
    case 128: return ns_one_impl<128>(which, stages, nstages, b, batch, s);
    case 256: return ns_one_impl<256>(which, stages, nstages, b, batch, s);
    case 512: return ns_one_impl<512>(which, stages, nstages, b, batch, s);
    case 1024: return ns_one_impl<1024>(which, stages, nstages, b, batch, s);
    default: return fail(JCFD_ERR_UNSUPPORTED, "spectral grid size not built");
  }
}

int ns_launch_step(int n, const NsParams* stages, int nstages, const NsBuffers& b, int batch,
                   hipStream_t s) {
  if (nstages < 1) return fail(JCFD_ERR_INVALID_ARGUMENT, "at least one stage");
  switch (n) {
    case 32: return ns_step_impl<32>(stages, nstages, b, batch, s);
    case 64: return ns_step_impl<64>(stages, nstages, b, batch, s);
    case 128: return ns_step_impl<128>(stages, nstages, b, batch, s);
    case 256: return ns_step_impl<256>(stages, nstages, b, batch, s);
    case 512: return ns_step_impl<512>(stages, nstages, b, batch, s);
    case 1024: return ns_step_impl<1024>(stages, nstages, b, batch, s);
    default: return fail(JCFD_ERR_UNSUPPORTED, "spectral grid size not built");
  }
}

int ns_launch_rfft2(int n, const NsParams& q, const float* x, float2* out, float2* pbuf, int batch,
                    hipStream_t s) {
  switch (n) {
    case 32: return NsOps<32>::rfft2(q, x, out, pbuf, batch, s);
    case 64: return NsOps<64>::rfft2(q, x, out, pbuf, batch, s);
    case 128: return NsOps<128>::rfft2(q, x, out, pbuf, batch, s);
    case 256: return NsOps<256>::rfft2(q, x, out, pbuf, batch, s);
    case 512: return NsOps<512>::rfft2(q, x, out, pbuf, batch, s);
    case 1024: return NsOps<1024>::rfft2(q, x, out, pbuf, batch, s);
    default: return fail(JCFD_ERR_UNSUPPORTED, "spectral grid size not built");
  }
}

int ns_launch_irfft2(int n, const NsParams& q, const float2* in, float* x, float2* pbuf, int batch,
                     hipStream_t s) {
  switch (n) {
    case 32: return NsOps<32>::irfft2(q, in, x, pbuf, batch, s);
    case 64: return NsOps<64>::irfft2(q, in, x, pbuf, batch, s);
    case 128: return NsOps<128>::irfft2(q, in, x, pbuf, batch, s);
    case 256: return NsOps<256>::irfft2(q, in, x, pbuf, batch, s);
    case 512: return NsOps<512>::irfft2(q, in, x, pbuf, batch, s);
    case 1024: return NsOps<1024>::irfft2(q, in, x, pbuf, batch, s);
    default: return fail(JCFD_ERR_UNSUPPORTED, "spectral grid size not built");
  }
}

namespace {
template <int N>
int ns_convert_n(bool to_internal, const float2* src, float2* dst, const int* slots, int count,
                 hipStream_t s) {
  const dim3 grid((N * (N / 2 + 1) + kThreads - 1) / kThreads, count);
  if (to_internal)
    hipLaunchKernelGGL((ns_state_convert<N, true>), grid, dim3(kThreads), 0, s, src, dst, slots);
  else
    hipLaunchKernelGGL((ns_state_convert<N, false>), grid, dim3(kThreads), 0, s, src, dst, slots);
  return check_hip(hipGetLastError(), "ns_state_convert");
}
}  // namespace

int ns_launch_state_convert(int n, bool to_internal, const float2* src, float2* dst, const int* slots,
                            int count, hipStream_t s) {
  switch (n) {
    case 32: return ns_convert_n<32>(to_internal, src, dst, slots, count, s);
    case 64: return ns_convert_n<64>(to_internal, src, dst, slots, count, s);
    case 128: return ns_convert_n<128>(to_internal, src, dst, slots, count, s);
    case 256: return ns_convert_n<256>(to_internal, src, dst, slots, count, s);
    case 512: return ns_convert_n<512>(to_internal, src, dst, slots, count, s);
    case 1024: return ns_convert_n<1024>(to_internal, src, dst, slots, count, s);
    default: return fail(JCFD_ERR_UNSUPPORTED, "spectral grid size not built");
  }
}

int ns_launch_explicit_terms(int n, const NsParams& q, const float2* w, float2* out, float2* g,
                             float2* pbuf, int batch, hipStream_t s) {
  switch (n) {
    case 32: return ns_explicit_impl<32>(q, w, out, g, pbuf, batch, s);
    case 64: return ns_explicit_impl<64>(q, w, out, g, pbuf, batch, s);
    case 128: return ns_explicit_impl<128>(q, w, out, g, pbuf, batch, s);
    case 256: return ns_explicit_impl<256>(q, w, out, g, pbuf, batch, s);
    case 512: return ns_explicit_impl<512>(q, w, out, g, pbuf, batch, s);
    case 1024: return ns_explicit_impl<1024>(q, w, out, g, pbuf, batch, s);
    default: return fail(JCFD_ERR_UNSUPPORTED, "spectral grid size not built");
  }
}

}  // namespace jcfd
