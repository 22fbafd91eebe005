// Learned-interpolation FVM convection (jax_cfd/ml/interpolations.py:31-121,
// FusedLearnedInterpolation under advections.modular_self_advection) for
// gfx950.
//
// The reference runs one CNN tower (towers.py:144-262: periodic 3x3
// convolutions, relu between them) over the stacked velocity components,
// turns each interpolation's slice of the output into polynomially
// constrained stencil coefficients, bias + logits . nullspace
// (layers.py:451-544, 586-642), and interpolates u and v to the control-volume
// faces with them (4 x 4 patches, layers_util.py:324-447) before
// advect_general (base/advection.py:34-116).  Here:
//
//   li_conv3x3    one periodic 3x3 convolution layer as an implicit GEMM on
//                 the matrix cores (v_mfma_f32_16x16x4_f32, fp32 in, fp32
//                 accumulate): D[co][px] = sum over (tap, ci) of
//                 W[tap][ci][co] * x[ci][px + tap], + bias, relu.  A
//                 workgroup owns 4 rows x 32 columns of pixels (one row per
//                 wavefront) and every output channel; the input tile with
//                 its one-cell periodic halo goes through LDS in chunks of
//                 16 channels, double-buffered (the next chunk loaded during
//                 the current chunk's nine taps).  The weights are stored in
//                 MFMA operand order, so each wavefront reads its A operands
//                 straight from global memory (cache-resident: every
//                 workgroup reads the same few hundred KB), one vector load
//                 per lane for four k-steps, the next step's loads in flight
//                 during the current step's products; one barrier per
//                 chunk.  The tower's output layer is the
//                 constraint folded in: its 128 channels are the 8
//                 interpolations' 16 stencil coefficients each (weights
//                 W . nullspace, bias b . nullspace + constraint bias, formed
//                 in fp64 on the host), so no logits are materialised.
//   li_explicit   per row block: the 8 interpolations (16-term dot products
//                 of coefficients and 4 x 4 patches of u or v from LDS), the
//                 face fluxes aligned_c * aligned_v, their divergence, then
//                 diffusion and forcing -- the explicit terms k_u, k_v that
//                 the forward-Euler combination kernel (rows-forward mode 5)
//                 turns into u*, projected as every FVM step is.
//   li_interpolate one interpolation of one field (the interpolators of
//                 FusedLearnedInterpolation as a testing seam).
//
// Activations are planar per instance: (batch, channels, n, n) fp32.
#include <hip/hip_runtime.h>

#include <type_traits>
#include <utility>

#include "fvm_terms.h"
#include "jcfd_fft.h"
#include "jcfd_internal.h"

#pragma clang fp contract(off)

namespace jcfd {
namespace {

constexpr int kThreads = 256;
typedef float floatx4 __attribute__((ext_vector_type(4)));

// Output tile of a convolution workgroup and its LDS halo tile.
constexpr int kConvRows = 4;                         // one row per wavefront
constexpr int kConvCols = 32;                        // two 16-pixel MFMA blocks
constexpr int kHaloW = kConvCols + 2;
constexpr int kHaloPlane = (kConvRows + 2) * kHaloW;  // 204 words
constexpr int kHaloPitch = 208;  // plane pitch: 4 channels of a k-step land 16 banks apart

// Weight layout of a convolution layer in MFMA operand order.  K-steps of 4
// input channels are grouped KG at a time (one KG-wide vector per lane);
// fragment (tap, group g, channel block cb) holds, for lane l = i + 16 k
// (the A[i][k] lane of v_mfma_f32_16x16x4_f32), the KG weights
// W[tap][ci = 4 (g KG + kk) + k][co = 16 cb + i], kk = 0 .. KG-1, at
// float offset (((tap GPT + g) NCB + cb) 64 + l) KG + kk.
template <int CINP, int COUT>
struct ConvFrag {
  static constexpr int KS = CINP / 4;         // k-steps
  static constexpr int KG = KS < 4 ? KS : 4;  // k-steps per weight vector
  static constexpr int GPT = KS / KG;         // groups per tap
  static constexpr int NCB = COUT / 16;       // output-channel blocks
  static constexpr int STEPS = 9 * GPT;       // (tap, group) pairs, tap-major
};
template <int KG>
using ConvAVec = std::conditional_t<KG == 4, float4, float>;
template <int KG>
__device__ __forceinline__ float conv_acomp(const ConvAVec<KG>& v, int kk) {
  if constexpr (KG == 4) return kk == 0 ? v.x : (kk == 1 ? v.y : (kk == 2 ? v.z : v.w));
  else return v;
}

// Calls f(integral_constant<int, S>) for S = B .. E-1 (a loop the compiler
// sees fully unrolled, with compile-time step numbers).
template <int B, int E, typename F>
__device__ __forceinline__ void static_for(F&& f) {
  if constexpr (B < E) {
    f(std::integral_constant<int, B>{});
    static_for<B + 1, E>(f);
  }
}

// v_mfma_f32_16x16x4_f32 operand layout (checked on gfx950): A[i][k] in lane
// i + 16 k, B[k][j] in lane j + 16 k, D[i][j] in lane j + 16 (i / 4),
// register i % 4.  Here i = output channel, j = pixel, k = input channel.
//
// The input channels go in chunks of 4 KG (16 for 16 or more channels):
// the halo tile of one chunk sits in one of two LDS buffers while the next
// chunk's tile is loaded into registers, so only the first chunk's load is
// exposed.  Each chunk runs its 9 taps (step s = 9 chunk + tap, weight
// fragment tap GPT + chunk); every accumulator therefore sums its products
// chunk by chunk, tap by tap within a chunk, channels in order within a
// tap.  The steps are unrolled at compile time so that each weight wait
// counts exactly the loads issued after it: a chunk's halo loads follow
// the weight prefetch of its chunk's first step and are waited on two
// steps later (loads and their waits share one counter, in issue order).
template <int CINP, int COUT, bool RELU, bool FIRST>
__global__ __launch_bounds__(kThreads) void li_conv3x3(LiConvArgs a) {
  static_assert(CINP % 4 == 0 && COUT % 16 == 0, "MFMA blocking");
  using F = ConvFrag<CINP, COUT>;
  constexpr int NCB = F::NCB, KG = F::KG, GPT = F::GPT, STEPS = F::STEPS;
  constexpr int CH = 4 * KG;                  // channels per chunk
  constexpr int CTOT = CH * kHaloPlane;       // halo words per chunk
  constexpr int HPT = (CTOT + kThreads - 1) / kThreads;
  constexpr int NBUF = GPT > 1 ? 2 : 1;
  using AV = ConvAVec<KG>;
  __shared__ float shalo[NBUF * CH * kHaloPitch];
  const int n = a.n;
  const int tid = threadIdx.x, wave = tid / 64, lane = tid % 64;
  const int tiles_x = n / kConvCols;
  const int t = xcd_tile(blockIdx.x, gridDim.x);
  const int i0 = (t / tiles_x) * kConvRows, j0 = (t % tiles_x) * kConvCols;
  const size_t plane = static_cast<size_t>(n) * n;
  const size_t inst = blockIdx.y;
  const float* in = a.in + inst * (FIRST ? 1 : CINP) * plane;
  const float* in_v = FIRST ? a.in_v + inst * plane : nullptr;
  // fragment (tap GPT + chunk, cb) of this lane at wf[(fragment * NCB + cb) * 64]
  const AV* wf = reinterpret_cast<const AV*>(a.w) + lane;
  auto load = [&](int s, AV (&w)[NCB]) {
    const int sc = s < STEPS ? s : STEPS - 1;  // unconditional: the last prefetch repeats
    const int fr = (sc % 9) * GPT + sc / 9;
#pragma unroll
    for (int cb = 0; cb < NCB; ++cb) w[cb] = wf[(fr * NCB + cb) * 64];
  };
  // halo tile of chunk g (rows i0-1 .. i0+4, columns j0-1 .. j0+32,
  // periodic) through registers
  float hreg[HPT];
  auto halo_load = [&](int g) {
#pragma unroll
    for (int k = 0; k < HPT; ++k) {
      const int e0 = tid + k * kThreads;
      const int e = (CTOT % kThreads == 0 || e0 < CTOT) ? e0 : CTOT - 1;
      const int c = g * CH + e / kHaloPlane, r = (e % kHaloPlane) / kHaloW, cc = e % kHaloW;
      const size_t gi = static_cast<size_t>((i0 - 1 + r) & (n - 1)) * n + ((j0 - 1 + cc) & (n - 1));
      if constexpr (FIRST) hreg[k] = c == 0 ? in[gi] : (c == 1 ? in_v[gi] : 0.f);
      else hreg[k] = in[c * plane + gi];
    }
  };
  auto halo_store = [&](int buf) {
    float* dst = shalo + buf * CH * kHaloPitch;
#pragma unroll
    for (int k = 0; k < HPT; ++k) {
      const int e = tid + k * kThreads;
      if (CTOT % kThreads == 0 || e < CTOT) dst[(e / kHaloPlane) * kHaloPitch + e % kHaloPlane] = hreg[k];
    }
  };
  // the first step's weights go out first, then chunk 0's halo
  AV wa[NCB], wb[NCB];
  load(0, wa);
  halo_load(0);
  halo_store(0);
  __syncthreads();

  floatx4 acc[NCB][2];
#pragma unroll
  for (int cb = 0; cb < NCB; ++cb) acc[cb][0] = acc[cb][1] = floatx4{0.f, 0.f, 0.f, 0.f};
  const int kq = lane / 16, lp = lane % 16;
  auto step = [&](int s, const AV (&w)[NCB]) {
    const int tap = s % 9, g = s / 9;
    const float* bin = shalo + (g % NBUF) * CH * kHaloPitch + (wave + tap / 3) * kHaloW + tap % 3 + lp;
#pragma unroll
    for (int kk = 0; kk < KG; ++kk) {
      const int ci = kk * 4 + kq;
      const float b0 = bin[ci * kHaloPitch];
      const float b1 = bin[ci * kHaloPitch + 16];
#pragma unroll
      for (int cb = 0; cb < NCB; ++cb) {
        const float av = conv_acomp<KG>(w[cb], kk);
        acc[cb][0] = __builtin_amdgcn_mfma_f32_16x16x4f32(av, b0, acc[cb][0], 0, 0, 0);
        acc[cb][1] = __builtin_amdgcn_mfma_f32_16x16x4f32(av, b1, acc[cb][1], 0, 0, 0);
      }
    }
  };
  // step s computes with one register set while the other receives step
  // s + 1's weights (the sets alternate, so no copy waits for a load)
  auto run = [&](auto sc, AV (&cur)[NCB], AV (&nxt)[NCB]) {
    constexpr int s = decltype(sc)::value;
    load(s + 1, nxt);
    if constexpr (s % 9 == 0 && s / 9 + 1 < GPT) halo_load(s / 9 + 1);
    step(s, cur);
    if constexpr (s % 9 == 8 && s / 9 + 1 < GPT) {
      halo_store((s / 9 + 1) % NBUF);  // its buffer's last reader was chunk s/9 - 1
      __syncthreads();
    }
  };
  static_for<0, STEPS>([&](auto sc) {
    if constexpr (decltype(sc)::value % 2 == 0) run(sc, wa, wb);
    else run(sc, wb, wa);
  });

  // out[co][i0 + wave][j0 + px] = act(acc + b[co])
  float* out = a.out + inst * COUT * plane + static_cast<size_t>(i0 + wave) * n + j0;
#pragma unroll
  for (int cb = 0; cb < NCB; ++cb) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int co = cb * 16 + 4 * kq + r;
      const float bias = a.b[co];
#pragma unroll
      for (int pb = 0; pb < 2; ++pb) {
        float x = acc[cb][pb][r] + bias;
        if constexpr (RELU) x = fmaxf(x, 0.f);
        out[co * plane + pb * 16 + lp] = x;
      }
    }
  }
}

// The 16-term dot product of one interpolation at cell (gi, j): coefficient
// s = 4 (a + 1) + (b + 1) multiplies x[gi + a, j + b], a, b in {-1 .. 2}
// (extract_patches' channel order).  Summed as NumPy's pairwise reduction
// of 16 terms: r_k = t_k + t_{k+8}, then ((r0+r1)+(r2+r3))+((r4+r5)+(r6+r7)).
// `s` holds LDS rows, local row lr = gi - row0; `c` points at the
// interpolation's first coefficient plane at (gi, j).
template <int N>
__device__ __forceinline__ float li_dot(const float* __restrict__ c, size_t plane,
                                        const float* s, int lr, int j) {
  float t[16];
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    const int da = q / 4 - 1, db = q % 4 - 1;
    t[q] = c[q * plane] * s[(lr + da) * N + ((j + db) & (N - 1))];
  }
  float r[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) r[k] = t[k] + t[k + 8];
  return ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]));
}

// Explicit terms of the learned self-advection step for rows [i0, i0+RB):
// k = (-div(aligned_c * aligned_v) + nu lap) + f, per component, with the
// fluxes of advect_general (advection.py:81-116, _advect_aligned 34-78).
template <int N, int RB>
__global__ __launch_bounds__(kThreads) void li_explicit(FvmParams p, LiExplicitArgs a) {
  constexpr int ROWS = RB + 4;  // u, v rows i0-2 .. i0+RB+1
  __shared__ float su[ROWS * N], sv[ROWS * N];
  __shared__ float fxs[(RB + 1) * N], fys[RB * N];
  const int tid = threadIdx.x;
  const int i0 = xcd_tile(blockIdx.x, gridDim.x) * RB;
  const size_t plane = static_cast<size_t>(N) * N;
  const size_t inst = a.slots ? static_cast<size_t>(a.slots[blockIdx.y]) : blockIdx.y;
  const float* u = a.u + inst * plane;
  const float* v = a.v + inst * plane;
  const float* coef = a.coef + blockIdx.y * static_cast<size_t>(kLiCoefChannels) * plane;
  staged<ROWS * N, 8, kThreads>(
      tid,
      [&](int e) {
        const size_t g = static_cast<size_t>((i0 - 2 + e / N) & (N - 1)) * N + e % N;
        return make_float2(u[g], v[g]);
      },
      [&](int e, float2 x) {
        su[e] = x.x;
        sv[e] = x.y;
      });
  const Window<N> w{su, sv};
  __syncthreads();
  // face flux of one interpolation pair at local row lr (global gi):
  // aligned_c (coefficients at base kc, field sc) * aligned_v (kv, sv_)
  auto flux = [&](int kc, const float* fc, int kv, const float* fv, int lr, int j) {
    const size_t at = static_cast<size_t>((i0 - 2 + lr) & (N - 1)) * N + j;
    const float cval = li_dot<N>(coef + kc * plane + at, plane, fc, lr, j);
    const float vval = li_dot<N>(coef + kv * plane + at, plane, fv, lr, j);
    return cval * vval;
  };
#pragma unroll 1
  for (int comp = 0; comp < 2; ++comp) {
    const float* sc = comp == 0 ? su : sv;  // the advected component
    const int* key = a.key + 4 * comp;      // x face: (c, velocity), y face: (c, velocity)
    const float* vx = su;                   // velocity of the x faces: u
    const float* vy = sv;                   // of the y faces: v
    // x faces of rows i0-1 .. i0+RB-1 (local rows 1 .. RB+1), y faces of rows i0 ..
    for (int e = tid; e < (RB + 1) * N; e += kThreads)
      fxs[e] = flux(key[0], sc, key[1], vx, 1 + e / N, e % N);
    for (int e = tid; e < RB * N; e += kThreads)
      fys[e] = flux(key[2], sc, key[3], vy, 2 + e / N, e % N);
    __syncthreads();
    float* kout = (comp == 0 ? a.ku : a.kv) + inst * plane;
    for (int e = tid; e < RB * N; e += kThreads) {
      const int r = e / N, j = e % N;
      const float adv = -(over_step(fxs[(r + 1) * N + j] - fxs[r * N + j], p.dx, p.rcp_dx) +
                          over_step(fys[r * N + j] - fys[r * N + ((j - 1) & (N - 1))], p.dy, p.rcp_dy));
      const float k = comp == 0 ? add_diffusion_forcing<false>(w, r + 2, j, p, p.has_forcing ? a.fx[j] : 0.f, adv)
                                : add_diffusion_forcing<true>(w, r + 2, j, p, 0.f, adv);
      kout[static_cast<size_t>(i0 + r) * N + j] = k;
    }
    __syncthreads();
  }
}

// out = interpolation `key` (coefficient planes key .. key+15) applied to x.
template <int N>
__global__ __launch_bounds__(kThreads) void li_interpolate(const float* __restrict__ coef,
                                                           const float* __restrict__ x, int key,
                                                           float* __restrict__ out) {
  constexpr int RB = kThreads / N > 0 ? kThreads / N : 1;
  constexpr int ROWS = RB + 3;
  __shared__ float sx[ROWS * N];
  const int tid = threadIdx.x;
  const int i0 = blockIdx.x * RB;
  const size_t plane = static_cast<size_t>(N) * N;
  x += blockIdx.y * plane;
  coef += blockIdx.y * static_cast<size_t>(kLiCoefChannels) * plane + key * plane;
  out += blockIdx.y * plane;
  for (int e = tid; e < ROWS * N; e += kThreads)
    sx[e] = x[static_cast<size_t>((i0 - 1 + e / N) & (N - 1)) * N + e % N];
  __syncthreads();
  for (int e = tid; e < RB * N; e += kThreads) {
    const int r = e / N, j = e % N;
    const size_t at = static_cast<size_t>(i0 + r) * N + j;
    out[at] = li_dot<N>(coef + at, plane, sx, r + 1, j);
  }
}

// ---- launchers ------------------------------------------------------------
template <int CINP, int COUT, bool RELU, bool FIRST>
int conv_launch(const LiConvArgs& a, int batch, hipStream_t s) {
  const dim3 grid((a.n / kConvCols) * (a.n / kConvRows), batch);
  hipLaunchKernelGGL((li_conv3x3<CINP, COUT, RELU, FIRST>), grid, dim3(kThreads), 0, s, a);
  return check_hip(hipGetLastError(), "li_conv3x3");
}

template <int COUT>
int conv_first(const LiConvArgs& a, bool relu, int batch, hipStream_t s) {
  return relu ? conv_launch<4, COUT, true, true>(a, batch, s) : conv_launch<4, COUT, false, true>(a, batch, s);
}

template <int CINP>
int conv_hidden(const LiConvArgs& a, int cout, bool relu, int batch, hipStream_t s) {
  if (relu && cout == CINP) return conv_launch<CINP, CINP, true, false>(a, batch, s);
  if (!relu && cout == kLiCoefChannels) return conv_launch<CINP, kLiCoefChannels, false, false>(a, batch, s);
  return fail(JCFD_ERR_UNSUPPORTED, "convolution " + std::to_string(CINP) + " -> " + std::to_string(cout) +
                                        (relu ? " (relu)" : "") + " not built");
}

template <int N>
int explicit_launch(const FvmParams& p, const LiExplicitArgs& a, int batch, hipStream_t s) {
  constexpr int RB = N <= 256 ? 4 : 2;
  hipLaunchKernelGGL((li_explicit<N, RB>), dim3(N / RB, batch), dim3(kThreads), 0, s, p, a);
  return check_hip(hipGetLastError(), "li_explicit");
}

template <int N>
int interpolate_launch(const float* coef, const float* x, int key, float* out, int batch, hipStream_t s) {
  constexpr int RB = kThreads / N > 0 ? kThreads / N : 1;
  hipLaunchKernelGGL((li_interpolate<N>), dim3(N / RB, batch), dim3(kThreads), 0, s, coef, x, key, out);
  return check_hip(hipGetLastError(), "li_interpolate");
}

}  // namespace

void li_conv_weight_layout(const float* w, int cinp, int cout, float* frag) {
  // the ConvFrag layout: k-step groups of up to 4, fragments tap-major
  const int ks = cinp / 4, kg = ks < 4 ? ks : 4, gpt = ks / kg, ncb = cout / 16;
  for (int tap = 0; tap < 9; ++tap)
    for (int g = 0; g < gpt; ++g)
      for (int cb = 0; cb < ncb; ++cb)
        for (int l = 0; l < 64; ++l)
          for (int kk = 0; kk < kg; ++kk) {
            const int ci = (g * kg + kk) * 4 + l / 16, co = cb * 16 + l % 16;
            frag[(((static_cast<size_t>(tap) * gpt + g) * ncb + cb) * 64 + l) * kg + kk] =
                w[(static_cast<size_t>(tap) * cinp + ci) * cout + co];
          }
}

bool li_conv_supported(int cin, int cout, bool relu) {
  const bool hidden_out = relu && (cout == 16 || cout == 32 || cout == 64);
  const bool coef_out = !relu && cout == kLiCoefChannels;
  if (cin == 2) return hidden_out || coef_out;
  if (cin == 16 || cin == 32 || cin == 64) return (relu && cout == cin) || coef_out;
  return false;
}

int li_launch_conv(const LiConvArgs& a, int cin, int cout, bool relu, int batch, hipStream_t s) {
  if (a.n < kConvCols || (a.n & (a.n - 1)) != 0)
    return fail(JCFD_ERR_UNSUPPORTED, "learned interpolation needs n a power of two >= 32");
  if (!li_conv_supported(cin, cout, relu))
    return fail(JCFD_ERR_UNSUPPORTED, "convolution " + std::to_string(cin) + " -> " +
                                          std::to_string(cout) + " not built");
  if (cin == 2) {
    switch (cout) {
      case 16: return conv_first<16>(a, relu, batch, s);
      case 32: return conv_first<32>(a, relu, batch, s);
      case 64: return conv_first<64>(a, relu, batch, s);
      default: return conv_first<kLiCoefChannels>(a, relu, batch, s);
    }
  }
  switch (cin) {
    case 16: return conv_hidden<16>(a, cout, relu, batch, s);
    case 32: return conv_hidden<32>(a, cout, relu, batch, s);
    default: return conv_hidden<64>(a, cout, relu, batch, s);
  }
}

int li_launch_explicit(const FvmParams& p, const LiExplicitArgs& a, int batch, hipStream_t s) {
  switch (p.n) {
    case 32: return explicit_launch<32>(p, a, batch, s);
    case 64: return explicit_launch<64>(p, a, batch, s);
    case 128: return explicit_launch<128>(p, a, batch, s);
    case 256: return explicit_launch<256>(p, a, batch, s);
    case 512: return explicit_launch<512>(p, a, batch, s);
    case 1024: return explicit_launch<1024>(p, a, batch, s);
    default: return fail(JCFD_ERR_UNSUPPORTED, "learned interpolation grid size not built");
  }
}

int li_launch_interpolate(int n, const float* coef, const float* x, int key, float* out, int batch,
                          hipStream_t s) {
  switch (n) {
    case 32: return interpolate_launch<32>(coef, x, key, out, batch, s);
    case 64: return interpolate_launch<64>(coef, x, key, out, batch, s);
    case 128: return interpolate_launch<128>(coef, x, key, out, batch, s);
    case 256: return interpolate_launch<256>(coef, x, key, out, batch, s);
    case 512: return interpolate_launch<512>(coef, x, key, out, batch, s);
    case 1024: return interpolate_launch<1024>(coef, x, key, out, batch, s);
    default: return fail(JCFD_ERR_UNSUPPORTED, "learned interpolation grid size not built");
  }
}

}  // namespace jcfd
