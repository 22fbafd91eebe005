// Workgroup-level FFT building blocks for gfx950 (device code only).
//
// Transforms live in LDS and are done in place with Stockham auto-sort passes
// (radix 8, then 4, then 2; or at most 4, a schedule the tile tuner may pick):
// each pass loads R values per butterfly into
// registers, waits on a barrier, applies twiddles and an R-point DFT in
// registers, and writes the results back in auto-sorted order.  A workgroup
// runs T independent transforms of length N at once; element e of transform t
// is at buf[e * ES + t * TS], so the same code serves
//   * row transforms      (transform-major: ES = 1, TS = row pitch), and
//   * column-tile transforms (element-major: ES = tile width, TS = 1).
// Butterflies are enumerated so that neighbouring lanes touch neighbouring
// LDS words in both layouts (TFAST picks transform-index-fastest).
//
// Real transforms of length n use the half-length complex transform of
// z[m] = x[2m] + i x[2m+1] (a real row in LDS reinterpreted as float2 is
// already z) plus one combine step.  The "packed" half spectrum stores the
// n/2 + 1 bins of a real row in n/2 complex words: word 0 holds
// (X[0], X[n/2]) (both real), words 1..n/2-1 hold X[1..n/2-1].
//
// Conventions: forward = exp(-2 pi i jk/N), unnormalised; inverse =
// exp(+2 pi i jk/N), unnormalised (callers fold 1/N into their multipliers).
// Twiddle tables hold exp(-2 pi i k/N), k in [0, N), rounded from fp64.
#pragma once
#include <hip/hip_runtime.h>

// No multiply-add contraction: whether the backend fuses a product into an
// add depends on the surrounding code after inlining, so contracted helpers
// could round differently from one kernel (or tile shape) to the next.
#pragma clang fp contract(off)

namespace jcfd {

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
// Complex product with two explicit fused multiply-adds (2 VALU operations
// fewer than 4 products and 2 sums).  Explicit fmaf is never re-associated,
// so every kernel still rounds it the same way.
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(__builtin_fmaf(a.x, b.x, -(a.y * b.y)), __builtin_fmaf(a.x, b.y, a.y * b.x));
}
__device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
// multiply by -i (forward) or +i (inverse)
template <bool INV>
__device__ __forceinline__ float2 mul_mi(float2 a) {
  return INV ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
}

template <bool INV>
__device__ __forceinline__ void dft2(float2& a, float2& b) {
  float2 t = a;
  a = cadd(t, b);
  b = csub(t, b);
}

template <bool INV>
__device__ __forceinline__ void dft4(float2& v0, float2& v1, float2& v2, float2& v3) {
  float2 t0 = cadd(v0, v2), t1 = csub(v0, v2);
  float2 t2 = cadd(v1, v3), t3 = mul_mi<INV>(csub(v1, v3));
  v0 = cadd(t0, t2);
  v2 = csub(t0, t2);
  v1 = cadd(t1, t3);
  v3 = csub(t1, t3);
}

template <bool INV>
__device__ __forceinline__ void dft8(float2 (&v)[8]) {
  constexpr float h = 0.70710678118654752440f;
  dft4<INV>(v[0], v[2], v[4], v[6]);  // even half -> E0..E3 in v0,v2,v4,v6
  dft4<INV>(v[1], v[3], v[5], v[7]);  // odd half  -> O0..O3 in v1,v3,v5,v7
  // W8^s, s = 1..3 (forward: exp(-i pi s / 4))
  float2 o1 = v[3], o2 = v[5], o3 = v[7];
  o1 = INV ? make_float2(h * (o1.x - o1.y), h * (o1.x + o1.y))
           : make_float2(h * (o1.x + o1.y), h * (o1.y - o1.x));
  o2 = mul_mi<INV>(o2);
  o3 = INV ? make_float2(-h * (o3.x + o3.y), h * (o3.x - o3.y))
           : make_float2(h * (o3.y - o3.x), -h * (o3.x + o3.y));
  float2 e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6], o0 = v[1];
  v[0] = cadd(e0, o0);
  v[4] = csub(e0, o0);
  v[1] = cadd(e1, o1);
  v[5] = csub(e1, o1);
  v[2] = cadd(e2, o2);
  v[6] = csub(e2, o2);
  v[3] = cadd(e3, o3);
  v[7] = csub(e3, o3);
}

template <int R, bool INV>
__device__ __forceinline__ void dft_radix(float2 (&v)[R]) {
  if constexpr (R == 2) {
    dft2<INV>(v[0], v[1]);
  } else if constexpr (R == 4) {
    dft4<INV>(v[0], v[1], v[2], v[3]);
  } else {
    static_assert(R == 8, "radix");
    dft8<INV>(v);
  }
}

// One in-place Stockham pass of radix R with NS = product of earlier radices.
template <int N, int R, int NS, int T, int NT, int ES, int TS, bool TFAST, bool INV>
__device__ __forceinline__ void stockham_pass(float2* buf, const float2* tw, int tid) {
  constexpr int NB = N / R;
  constexpr int TOT = T * NB;
  constexpr int BPT = (TOT + NT - 1) / NT;
  constexpr int TWS = N / (NS * R);
  float2 v[BPT][R];
#pragma unroll
  for (int b = 0; b < BPT; ++b) {
    const int g = tid + b * NT;
    if (TOT % NT == 0 || g < TOT) {
      const int t = TFAST ? g % T : g / NB;
      const int j = TFAST ? g / T : g % NB;
#pragma unroll
      for (int r = 0; r < R; ++r) v[b][r] = buf[(j + r * NB) * ES + t * TS];
    }
  }
  __syncthreads();
#pragma unroll
  for (int b = 0; b < BPT; ++b) {
    const int g = tid + b * NT;
    if (TOT % NT == 0 || g < TOT) {
      const int t = TFAST ? g % T : g / NB;
      const int j = TFAST ? g / T : g % NB;
      const int k = j % NS;
      if constexpr (NS > 1) {
#pragma unroll
        for (int r = 1; r < R; ++r) {
          float2 w = tw[r * k * TWS];
          if (INV) w.y = -w.y;
          v[b][r] = cmul(v[b][r], w);
        }
      }
      float2 x[R];
#pragma unroll
      for (int r = 0; r < R; ++r) x[r] = v[b][r];
      dft_radix<R, INV>(x);
      const int base = (j / NS) * NS * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) buf[(base + r * NS) * ES + t * TS] = x[r];
    }
  }
  __syncthreads();
}

// Radix schedule: the largest radix <= MAXR (8 or 4) that divides what is
// left, largest first (N = 512, MAXR 8: 8 8 8; MAXR 4: 4 4 4 4 2).
template <int N, int NS, int T, int NT, int ES, int TS, bool TFAST, bool INV, int MAXR>
__device__ __forceinline__ void fft_passes(float2* buf, const float2* tw, int tid) {
  if constexpr (NS < N) {
    constexpr int REM = N / NS;
    constexpr int R = (MAXR >= 8 && REM % 8 == 0) ? 8 : ((REM % 4 == 0) ? 4 : 2);
    stockham_pass<N, R, NS, T, NT, ES, TS, TFAST, INV>(buf, tw, tid);
    fft_passes<N, NS * R, T, NT, ES, TS, TFAST, INV, MAXR>(buf, tw, tid);
  }
}

// T in-place complex transforms of length N in LDS.  Callers must have a
// barrier between writing `buf` and calling this; it ends with a barrier.
template <int N, int T, int NT, int ES, int TS, bool TFAST, bool INV, int MAXR = 8>
__device__ __forceinline__ void lds_fft(float2* buf, const float2* tw, int tid) {
  static_assert((N & (N - 1)) == 0 && N >= 2, "power-of-two length");
  static_assert(MAXR == 8 || MAXR == 4, "radix schedule");
  fft_passes<N, 1, T, NT, ES, TS, TFAST, INV, MAXR>(buf, tw, tid);
}

// ---- one transform per wavefront -------------------------------------------
// The same Stockham passes, with wavefront w running transforms w, w + NW, ...
// (its lanes take their (transform, butterfly) pairs): a pass depends only on
// its own wavefront's
// earlier LDS writes, so passes are separated by a wavefront-level fence
// instead of a workgroup barrier.  Arithmetic per transform is unchanged, so
// results are bit-identical to lds_fft.  Best when elements of a transform
// are contiguous (ES = 1), so a wavefront's lanes touch consecutive words.

// Orders this wavefront's LDS accesses (all of its lanes) before the ones
// that follow.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int N, int R, int NS, int T, int NW, int ES, int TS, bool INV>
__device__ __forceinline__ void stockham_pass_wave(float2* buf, const float2* tw, int wave, int lane) {
  constexpr int NB = N / R;
  constexpr int TPW = (T + NW - 1) / NW;  // transforms per wavefront: t = wave + a * NW
  constexpr int ITEMS = TPW * NB;         // (transform, butterfly) pairs of the wavefront
  constexpr int IPL = (ITEMS + 63) / 64;  // pairs per lane
  constexpr int TWS = N / (NS * R);
  float2 v[IPL][R];
#pragma unroll
  for (int i = 0; i < IPL; ++i) {
    const int g = lane + 64 * i, a = g / NB, j = g % NB, t = wave + a * NW;
    if ((ITEMS % 64 == 0 || g < ITEMS) && (T % NW == 0 || t < T)) {
#pragma unroll
      for (int r = 0; r < R; ++r) v[i][r] = buf[(j + r * NB) * ES + t * TS];
    }
  }
  wave_lds_sync();
#pragma unroll
  for (int i = 0; i < IPL; ++i) {
    const int g = lane + 64 * i, a = g / NB, j = g % NB, t = wave + a * NW;
    if ((ITEMS % 64 == 0 || g < ITEMS) && (T % NW == 0 || t < T)) {
      const int k = j % NS;
      if constexpr (NS > 1) {
#pragma unroll
        for (int r = 1; r < R; ++r) {
          float2 w = tw[r * k * TWS];
          if (INV) w.y = -w.y;
          v[i][r] = cmul(v[i][r], w);
        }
      }
      float2 x[R];
#pragma unroll
      for (int r = 0; r < R; ++r) x[r] = v[i][r];
      dft_radix<R, INV>(x);
      const int base = (j / NS) * NS * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) buf[(base + r * NS) * ES + t * TS] = x[r];
    }
  }
  wave_lds_sync();
}

template <int N, int NS, int T, int NW, int ES, int TS, bool INV, int MAXR>
__device__ __forceinline__ void fft_passes_wave(float2* buf, const float2* tw, int wave, int lane) {
  if constexpr (NS < N) {
    constexpr int REM = N / NS;
    constexpr int R = (MAXR >= 8 && REM % 8 == 0) ? 8 : ((REM % 4 == 0) ? 4 : 2);
    stockham_pass_wave<N, R, NS, T, NW, ES, TS, INV>(buf, tw, wave, lane);
    fft_passes_wave<N, NS * R, T, NW, ES, TS, INV, MAXR>(buf, tw, wave, lane);
  }
}

// T in-place complex transforms of length N in LDS by NT / 64 wavefronts,
// same element placement and results as lds_fft.  Callers must have a
// barrier between writing `buf` and calling this; it ends with a barrier.
template <int N, int T, int NT, int ES, int TS, bool INV, int MAXR = 8>
__device__ __forceinline__ void lds_fft_waves(float2* buf, const float2* tw, int tid) {
  static_assert((N & (N - 1)) == 0 && N >= 2, "power-of-two length");
  static_assert(MAXR == 8 || MAXR == 4, "radix schedule");
  static_assert(NT % 64 == 0, "whole wavefronts");
  fft_passes_wave<N, 1, T, NT / 64, ES, TS, INV, MAXR>(buf, tw, tid / 64, tid % 64);
  __syncthreads();
}

// A length-N twiddle table (entries k * STRIDE of the source table) staged
// through registers: load() issues all of a thread's global loads, store()
// writes them to LDS.  Splitting the two lets a kernel keep the loads in
// flight together with its data loads (a loop with a run-time bound waits on
// each load before the next is issued).
template <int N, int STRIDE, int NT>
struct TwiddleStage {
  static constexpr int PER = (N + NT - 1) / NT;
  float2 v[PER];
  // Loads are unconditional (surplus threads read entry N - 1): a load under
  // a branch makes the compiler's wait before the first use of any later
  // load a wait for all of them.
  __device__ __forceinline__ void load(const float2* __restrict__ src, int tid) {
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const int i = tid + k * NT;
      v[k] = src[(N % NT == 0 || i < N ? i : N - 1) * STRIDE];
    }
  }
  __device__ __forceinline__ void store(float2* dst, int tid) const {
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const int i = tid + k * NT;
      if (N % NT == 0 || i < N) dst[i] = v[k];
    }
  }
};

// Stage TOT elements through registers, up to CH per thread at a time
// (element e = tid + k * NT): every load of a chunk is issued before the
// chunk's first write, so a chunk costs one memory latency.  load() is
// called unconditionally (surplus lanes load element TOT - 1), store() only
// for e < TOT.
template <int TOT, int CH, int NT, typename Load, typename Store>
__device__ __forceinline__ void staged(int tid, Load&& load, Store&& store) {
  constexpr int PER = (TOT + NT - 1) / NT;
  constexpr int C = PER < CH ? PER : CH;
  // every e of every chunk is < TOT only when the threads tile TOT exactly
  // AND the chunks tile PER exactly; otherwise the last chunk's surplus
  // lanes must be clamped (e.g. TOT = 51 * NT in chunks of 16)
  constexpr bool kExact = TOT % NT == 0 && PER % C == 0;
  using V = decltype(load(0));
#pragma unroll 1
  for (int k0 = 0; k0 < PER; k0 += C) {
    V v[C];
#pragma unroll
    for (int k = 0; k < C; ++k) {
      const int e = tid + (k0 + k) * NT;
      v[k] = load(kExact || e < TOT ? e : TOT - 1);
    }
#pragma unroll
    for (int k = 0; k < C; ++k) {
      const int e = tid + (k0 + k) * NT;
      if (e < TOT) store(e, v[k]);
    }
  }
}

struct F2Triple {
  float2 a, b, t;
};


// After an in-place length-n/2 forward transform of z (row in LDS), the
// packed half-spectrum word k of the real row: needs Z[k] and Z[n/2 - k].
// twn = exp(-2 pi i k / n) for this k.
__device__ __forceinline__ float2 r2c_combine(const float2* Z, int k, int half, float2 twn) {
  if (k == 0) {
    float2 z0 = Z[0];
    return make_float2(z0.x + z0.y, z0.x - z0.y);
  }
  const float2 a = Z[k];
  const float2 b = cconj(Z[half - k]);
  // Ze = (a + b)/2, Zo = (a - b)/(2i); X = Ze + W^k Zo
  const float2 ze = make_float2(0.5f * (a.x + b.x), 0.5f * (a.y + b.y));
  const float2 d = csub(a, b);
  const float2 zo = make_float2(0.5f * d.y, -0.5f * d.x);
  return cadd(ze, cmul(twn, zo));
}

// Inverse combine: from packed half-spectrum words P[k], P[n/2 - k] build
// Z[k] whose length-n/2 inverse transform interleaves the (unnormalised)
// real output.  twn = exp(-2 pi i k / n).
__device__ __forceinline__ float2 c2r_combine(float2 pk, float2 pmk, int k, float2 twn) {
  float2 a, b;
  if (k == 0) {
    a = make_float2(pk.x, 0.f);  // X[0], imaginary part dropped
    b = make_float2(pk.y, 0.f);  // conj(X[n/2]), real
  } else {
    a = pk;
    b = cconj(pmk);
  }
  const float2 e = cadd(a, b);
  const float2 o = cmul(csub(a, b), cconj(twn));
  return make_float2(e.x - o.y, e.y + o.x);  // E + i O
}

}  // namespace jcfd
