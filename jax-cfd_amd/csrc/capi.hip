// C ABI of libjcfd_amd: argument validation, plans, and the step loops.
// Kernels live in fvm.hip / spectral.hip; this file only orchestrates.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "jcfd.h"
#include "jcfd_internal.h"
#include "jcfd_profile.h"

namespace jcfd {

namespace {
thread_local std::string g_last_error;
}

void set_error(const std::string& msg) { g_last_error = msg; }

int fail(int code, const std::string& msg) {
  set_error(msg);
  return code;
}

int check_hip(hipError_t e, const char* what) {
  if (e == hipSuccess) return JCFD_OK;
  return fail(JCFD_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

namespace {

constexpr int kDLFloat = 2;
constexpr int kDLComplex = 5;
constexpr int kDLROCM = 10;

// Validate a C-contiguous 2-D tensor of the given dtype and shape on `device`.
int check_tensor(const jcfd_tensor* t, const char* name, int code, int bits, int64_t d0, int64_t d1,
                 int device, void** ptr) {
  if (t == nullptr) return fail(JCFD_ERR_INVALID_ARGUMENT, std::string(name) + ": null tensor");
  if (t->device.device_type != kDLROCM)
    return fail(JCFD_ERR_INVALID_ARGUMENT,
                std::string(name) + ": expected a ROCm device tensor (device_type 10), got " +
                    std::to_string(t->device.device_type));
  if (device >= 0 && t->device.device_id != device)
    return fail(JCFD_ERR_INVALID_ARGUMENT, std::string(name) + ": tensor on device " +
                                               std::to_string(t->device.device_id) +
                                               ", plan on device " + std::to_string(device));
  if (t->dtype.code != code || t->dtype.bits != bits || t->dtype.lanes != 1)
    return fail(JCFD_ERR_INVALID_ARGUMENT,
                std::string(name) + ": wrong dtype (expected " +
                    (code == kDLFloat ? "float32" : "complex64") + ")");
  if (t->ndim != 2 || t->shape == nullptr || t->shape[0] != d0 || t->shape[1] != d1)
    return fail(JCFD_ERR_INVALID_ARGUMENT, std::string(name) + ": expected shape (" +
                                               std::to_string(d0) + ", " + std::to_string(d1) + ")");
  if (t->strides != nullptr && !(t->strides[1] == 1 && t->strides[0] == d1))
    return fail(JCFD_ERR_INVALID_ARGUMENT, std::string(name) + ": must be C-contiguous");
  if (t->data == nullptr) return fail(JCFD_ERR_INVALID_ARGUMENT, std::string(name) + ": null data");
  *ptr = static_cast<char*>(t->data) + t->byte_offset;
  return JCFD_OK;
}

// A batch of `batch` (d1, d2) fields: shape (batch, d1, d2), or (d1, d2) when
// batch == 1.  C-contiguous.
int check_batched(const jcfd_tensor* t, const char* name, int code, int bits, int batch, int64_t d1,
                  int64_t d2, int device, void** ptr) {
  if (t == nullptr) return fail(JCFD_ERR_INVALID_ARGUMENT, std::string(name) + ": null tensor");
  if (t->ndim == 2 && batch == 1) return check_tensor(t, name, code, bits, d1, d2, device, ptr);
  if (t->device.device_type != kDLROCM)
    return fail(JCFD_ERR_INVALID_ARGUMENT, std::string(name) + ": expected a ROCm device tensor");
  if (device >= 0 && t->device.device_id != device)
    return fail(JCFD_ERR_INVALID_ARGUMENT, std::string(name) + ": tensor on another device");
  if (t->dtype.code != code || t->dtype.bits != bits || t->dtype.lanes != 1)
    return fail(JCFD_ERR_INVALID_ARGUMENT, std::string(name) + ": wrong dtype (expected " +
                                               (code == kDLFloat ? "float32" : "complex64") + ")");
  if (t->ndim != 3 || t->shape == nullptr || t->shape[0] != batch || t->shape[1] != d1 ||
      t->shape[2] != d2)
    return fail(JCFD_ERR_INVALID_ARGUMENT, std::string(name) + ": expected shape (" +
                                               std::to_string(batch) + ", " + std::to_string(d1) +
                                               ", " + std::to_string(d2) + ")");
  if (t->strides != nullptr && !(t->strides[2] == 1 && t->strides[1] == d2 && t->strides[0] == d1 * d2))
    return fail(JCFD_ERR_INVALID_ARGUMENT, std::string(name) + ": must be C-contiguous");
  if (t->data == nullptr) return fail(JCFD_ERR_INVALID_ARGUMENT, std::string(name) + ": null data");
  *ptr = static_cast<char*>(t->data) + t->byte_offset;
  return JCFD_OK;
}

int check_spectral(const jcfd_tensor* t, const char* name, int batch, int n, int device, void** ptr) {
  return check_batched(t, name, kDLComplex, 64, batch, n, n / 2 + 1, device, ptr);
}

template <typename T>
int device_alloc(T** p, size_t count, int64_t* total) {
  hipError_t e = hipMalloc(reinterpret_cast<void**>(p), count * sizeof(T));
  if (e != hipSuccess) {
    *p = nullptr;
    (void)hipGetLastError();
    return fail(JCFD_ERR_OUT_OF_MEMORY, "hipMalloc of " + std::to_string(count * sizeof(T)) +
                                            " bytes failed: " + hipGetErrorString(e));
  }
  *total += static_cast<int64_t>(count * sizeof(T));
  return JCFD_OK;
}

// exp(-2 pi i k / n), k in [0, n), rounded from fp64.
std::vector<float2> twiddles(int n) {
  std::vector<float2> tw(n);
  for (int k = 0; k < n; ++k) {
    const double a = -2.0 * M_PI * static_cast<double>(k) / static_cast<double>(n);
    tw[k] = make_float2(static_cast<float>(std::cos(a)), static_cast<float>(std::sin(a)));
  }
  return tw;
}

// Eigenvalues (2 cos(2 pi p / n) - 2) / h^2 of the periodic 5-point
// Laplacian per axis, in fp64 (fast_diagonalization.py:199-266 with
// array_utils.py:168-173): x for p in [0, n), then y for c in [0, n/2]
// (the rfft layout of the last axis).  The Poisson column kernel forms
// Lambda^+ / n^2 from their sums, as a table of it would hold it.
std::vector<double> laplacian_eigenvalues(int n, double dx, double dy) {
  std::vector<double> e(static_cast<size_t>(n) + n / 2 + 1);
  for (int p = 0; p < n; ++p)
    e[p] = (-2.0 / (dx * dx)) + 2.0 * std::cos(2.0 * M_PI * p / n) / (dx * dx);
  for (int c = 0; c <= n / 2; ++c)
    e[n + c] = (-2.0 / (dy * dy)) + 2.0 * std::cos(2.0 * M_PI * c / n) / (dy * dy);
  return e;
}

// D / n^2 for the implicit diffusion solve (diffusion.py:180-185 through
// fast_diagonalization.py:196-225): D = (dt nu) s / (1 - (dt nu) s) with s the
// summed periodic Laplacian eigenvalues, in fp64, rfft layout on the last axis;
// stored column by column like Lambda^+.
std::vector<float> diffusion_multiplier(int n, double dx, double dy, double nu, double dt) {
  const int h1 = n / 2 + 1;
  std::vector<float> m(static_cast<size_t>(n) * h1);
  const double dt_nu = dt * nu;
  for (int p = 0; p < n; ++p) {
    const double lx = (-2.0 / (dx * dx)) + 2.0 * std::cos(2.0 * M_PI * p / n) / (dx * dx);
    for (int c = 0; c < h1; ++c) {
      const double ly = (-2.0 / (dy * dy)) + 2.0 * std::cos(2.0 * M_PI * c / n) / (dy * dy);
      const double x = dt_nu * (lx + ly);
      m[static_cast<size_t>(c) * n + p] = static_cast<float>(x / (1.0 - x) / (static_cast<double>(n) * n));
    }
  }
  return m;
}

// Replays `chunk` steps captured once into a hipGraph.  Capture happens on a
// plan-owned non-blocking stream (default priority); events order it after
// the caller's prior work and the caller's later work after it.  The graph
// bakes in its buffer addresses, so it is rebuilt when `key` (the output
// state pointer) or the chunk length changes.
// Captured multi-step launch sequences, replayed with hipGraphLaunch.  Up to
// kMaxGraphs are kept per plan (least recently used dropped), keyed by the
// buffers and slot list a sequence was captured with and its step count, so
// a plan that alternates between a few buffer sets captures each once.
struct GraphCache {
  static constexpr int kMaxGraphs = 32;
  struct Key {
    const void* k;
    const void* k2;
    const void* k4;
    int k3;
    int chunk;
    bool operator==(const Key& o) const {
      return k == o.k && k2 == o.k2 && k4 == o.k4 && k3 == o.k3 && chunk == o.chunk;
    }
  };
  struct Entry {
    Key key;
    hipGraphExec_t exec;
    uint64_t used;
  };
  bool enabled = true;
  hipStream_t stream = nullptr;
  hipEvent_t ev_in = nullptr, ev_out = nullptr;
  std::vector<Entry> entries;
  uint64_t tick = 0;

  void release() {
    for (Entry& e : entries) (void)hipGraphExecDestroy(e.exec);
    entries.clear();
    if (ev_in) (void)hipEventDestroy(ev_in);
    if (ev_out) (void)hipEventDestroy(ev_out);
    if (stream) (void)hipStreamDestroy(stream);
    ev_in = ev_out = nullptr;
    stream = nullptr;
  }

  // Replays `reps` times the chunk (chunk_len steps) that launch(s) enqueues
  // on stream s, capturing it first if its key is not cached.
  template <typename F>
  int run(const void* k, const void* k2, const void* k4, int k3, int chunk_len, int64_t reps,
          hipStream_t caller, F&& launch) {
    const Key key{k, k2, k4, k3, chunk_len};
    Entry* hit = nullptr;
    for (Entry& e : entries)
      if (e.key == key) hit = &e;
    if (hit == nullptr) {
      int rc = capture(key, launch);
      if (rc) return rc;
      hit = &entries.back();
    }
    hit->used = ++tick;
    int rc;
    if ((rc = check_hip(hipEventRecord(ev_in, caller), "event record")) ||
        (rc = check_hip(hipStreamWaitEvent(stream, ev_in, 0), "stream wait")))
      return rc;
    for (int64_t r = 0; r < reps; ++r)
      if ((rc = check_hip(hipGraphLaunch(hit->exec, stream), "graph launch"))) return rc;
    if ((rc = check_hip(hipEventRecord(ev_out, stream), "event record")) ||
        (rc = check_hip(hipStreamWaitEvent(caller, ev_out, 0), "stream wait")))
      return rc;
    return JCFD_OK;
  }

 private:
  template <typename F>
  int capture(const Key& key, F&& launch) {
    int rc = JCFD_OK;
    if (stream == nullptr) {
      if ((rc = check_hip(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "graph stream")) ||
          (rc = check_hip(hipEventCreateWithFlags(&ev_in, hipEventDisableTiming), "graph event")) ||
          (rc = check_hip(hipEventCreateWithFlags(&ev_out, hipEventDisableTiming), "graph event")))
        return rc;
    }
    if (entries.size() >= static_cast<size_t>(kMaxGraphs)) {  // drop the least recently used
      size_t lru = 0;
      for (size_t i = 1; i < entries.size(); ++i)
        if (entries[i].used < entries[lru].used) lru = i;
      (void)hipGraphExecDestroy(entries[lru].exec);
      entries.erase(entries.begin() + static_cast<std::ptrdiff_t>(lru));
    }
    hipGraph_t graph = nullptr;
    if ((rc = check_hip(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal), "begin capture")))
      return rc;
    const int lrc = launch(stream);
    const hipError_t ec = hipStreamEndCapture(stream, &graph);
    if (lrc) {
      if (graph) (void)hipGraphDestroy(graph);
      return lrc;
    }
    if ((rc = check_hip(ec, "end capture"))) return rc;
    hipGraphExec_t exec = nullptr;
    rc = check_hip(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0), "graph instantiate");
    (void)hipGraphDestroy(graph);
    if (rc) return rc;
    entries.push_back(Entry{key, exec, 0});
    return JCFD_OK;
  }
};

// Steps per captured graph.  Shorter runs (and a long run's remainder) launch
// directly: replaying a graph goes through the cache's own stream and two
// event hand-offs, which cost a server of many pools more device time than
// the host saves on a short chunk (measured on the mixed request stream).
constexpr int kGraphChunk = 32;

// nsteps repetitions of step(stream): whole chunks of kGraphChunk through the
// graph cache when it is enabled (keyed by k..k4), the rest directly on s.
template <typename Step>
int run_steps(GraphCache& graphs, const void* k, const void* k2, const void* k4, int k3,
              int64_t nsteps, hipStream_t s, Step&& step) {
  auto chunk_of = [&](int len) {
    return [&step, len](hipStream_t gs) {
      for (int i = 0; i < len; ++i) {
        const int r = step(gs);
        if (r) return r;
      }
      return static_cast<int>(JCFD_OK);
    };
  };
  int64_t remaining = nsteps;
  int rc;
  if (graphs.enabled && remaining >= kGraphChunk) {
    if ((rc = graphs.run(k, k2, k4, k3, kGraphChunk, remaining / kGraphChunk, s, chunk_of(kGraphChunk))))
      return rc;
    remaining %= kGraphChunk;
  }
  for (int64_t i = 0; i < remaining; ++i)
    if ((rc = step(s))) return rc;
  return JCFD_OK;
}

}  // namespace
}  // namespace jcfd

using namespace jcfd;

struct jcfd_fvm_plan {
  jcfd_fvm_config cfg;
  FvmParams params;
  int device = 0;
  int batch = 1;
  // CG pressure solve (pressure == JCFD_PRESSURE_CG)
  CgBuffers cg{};
  float* rhs = nullptr;
  int* status_host = nullptr;  // pinned, 4 ints
  bool cg_persistent = false;  // one cooperative launch per CG solve (opt-in: slower, see the header)
  int cg_maxiter = 0;
  int cg_last_iterations = 0;
  // CG implicit diffusion (diffusion == JCFD_DIFFUSION_IMPLICIT_CG), same buffers
  float diff_coef = 0.f;  // fp32(dt * viscosity)
  int diff_maxiter = 0;
  int diff_last_iterations[2] = {0, 0};
  float* us = nullptr;
  float* vs = nullptr;
  float2* spec = nullptr;
  double* lam_xy = nullptr;  // Laplacian eigenvalues: x (n), then y (n/2 + 1)
  float2* tw = nullptr;
  float* fx = nullptr;
  float* dlam = nullptr;  // implicit diffusion multiplier / n^2 (diffusion == implicit)
  // Runge-Kutta tableau steps (rk_stages > 1 or b != [1]): stage slopes k_i,
  // the projected stage state, and the combinations (stages 1.., final).
  int rk_stages = 0;
  float* kbuf = nullptr;   // [stage][u, v][batch][n][n]
  float* sbuf = nullptr;   // [u, v][batch][n][n]
  RkTerms* rk_terms = nullptr;  // device, rk_stages entries
  int64_t bytes = 0;
  GraphCache graphs;
};

struct jcfd_ns2d_plan {
  jcfd_ns2d_config cfg;
  int device = 0;
  int n = 0, batch = 1;
  NsParams stages[JCFD_MAX_RK_STAGES];
  int n_stages = 0;
  float* kx = nullptr;
  float* ky = nullptr;
  float2* tw = nullptr;
  int* force_idx = nullptr;
  float2* force_val = nullptr;
  float2* h = nullptr;
  float2* g = nullptr;
  float2* pbuf = nullptr;
  float2* ws = nullptr;  // the state in the step kernels' internal layout
  int64_t bytes = 0;
  GraphCache graphs;
};

namespace {
// Carpenter-Kennedy CN-RK4 coefficients (spectral/time_stepping.py:192-194).
const double kCnAlpha[6] = {0, 0.1496590219993, 0.3704009573644, 0.6222557631345, 0.9582821306748, 1};
const double kCnBeta[5] = {0, -0.4178904745, -1.192151694643, -1.697784692471, -1.514183444257};
const double kCnGamma[5] = {0.1496590219993, 0.3792103129999, 0.8229550293869, 0.6994504559488,
                            0.1530572479681};

constexpr int kDLInt = 0;

// A 1-D int32 device tensor of distinct slot indices in [0, batch); the
// values are the caller's contract (the kernels index with them directly),
// only the shape, dtype and device are checked here.
int check_slots(const jcfd_tensor* t, int batch, int device, void** ptr, int* count) {
  if (t == nullptr) return fail(JCFD_ERR_INVALID_ARGUMENT, "slots: null tensor");
  if (t->device.device_type != kDLROCM || (device >= 0 && t->device.device_id != device))
    return fail(JCFD_ERR_INVALID_ARGUMENT, "slots: must be on the plan's ROCm device");
  if (t->dtype.code != kDLInt || t->dtype.bits != 32 || t->dtype.lanes != 1)
    return fail(JCFD_ERR_INVALID_ARGUMENT, "slots: dtype must be int32");
  if (t->ndim != 1 || t->shape == nullptr || t->shape[0] < 0 || t->shape[0] > batch)
    return fail(JCFD_ERR_INVALID_ARGUMENT, "slots: must be 1-D with at most `batch` entries");
  if (t->strides != nullptr && t->shape[0] > 1 && t->strides[0] != 1)
    return fail(JCFD_ERR_INVALID_ARGUMENT, "slots: must be contiguous");
  *count = static_cast<int>(t->shape[0]);
  *ptr = static_cast<char*>(t->data) + t->byte_offset;
  return JCFD_OK;
}

FvmBuffers fvm_buffers(const jcfd_fvm_plan* p) {
  FvmBuffers b{};
  b.us = p->us;
  b.vs = p->vs;
  b.spec = p->spec;
  b.fx = p->fx;
  b.lam_xy = p->lam_xy;
  b.lam_mode = 2;
  b.tw = p->tw;
  return b;
}

int cg_alloc(jcfd_fvm_plan* p) {
  const size_t nn = static_cast<size_t>(p->params.n) * p->params.n;
  const size_t nb = static_cast<size_t>(cg_blocks(p->params.n));
  CgBuffers& c = p->cg;
  int rc;
  if ((rc = device_alloc(&c.x, nn, &p->bytes)) || (rc = device_alloc(&c.r, nn, &p->bytes)) ||
      (rc = device_alloc(&c.p[0], nn, &p->bytes)) || (rc = device_alloc(&c.p[1], nn, &p->bytes)) ||
      (rc = device_alloc(&c.ap, nn, &p->bytes)) || (rc = device_alloc(&p->rhs, nn, &p->bytes)) ||
      (rc = device_alloc(&c.part_rr[0], nb, &p->bytes)) ||
      (rc = device_alloc(&c.part_rr[1], nb, &p->bytes)) ||
      (rc = device_alloc(&c.part_pap, nb, &p->bytes)) || (rc = device_alloc(&c.part_bb, nb, &p->bytes)) ||
      (rc = device_alloc(&c.status, 4, &p->bytes)))
    return rc;
  if ((rc = check_hip(hipHostMalloc(reinterpret_cast<void**>(&p->status_host), 4 * sizeof(int),
                                    hipHostMallocDefault),
                      "pinned status")))
    return rc;
  p->cg_maxiter = p->cfg.cg_maxiter > 0 ? p->cfg.cg_maxiter : static_cast<int>(10 * nn);
  return JCFD_OK;
}

constexpr int kCgChunk = 64;  // iterations enqueued between convergence checks

// Enqueues CG iterations (iterate(k, s)) in chunks and reads the convergence
// flag between chunks, so this synchronises the stream.  The first chunk is
// `first` iterations long, the rest kCgChunk.  Returns the iteration count in
// *iterations.
template <typename Iterate>
int cg_iterate(jcfd_fvm_plan* p, int maxiter, int first, int* iterations, hipStream_t s,
               Iterate iterate) {
  int rc = JCFD_OK;
  for (int k = 0, chunk = first; rc == JCFD_OK; k += chunk, chunk = kCgChunk) {
    for (int i = 0; i < chunk && rc == JCFD_OK; ++i) rc = iterate(k + i, s);
    if (rc) break;
    if ((rc = check_hip(hipMemcpyAsync(p->status_host, p->cg.status, 2 * sizeof(int),
                                       hipMemcpyDeviceToHost, s), "cg status")) ||
        (rc = check_hip(hipStreamSynchronize(s), "cg status sync")))
      break;
    if (p->status_host[0]) {
      *iterations = p->status_host[1];
      break;
    }
    if (k + chunk > maxiter) {
      *iterations = maxiter;
      break;
    }
  }
  return rc;
}

// The solve as one cooperative launch (cg_launch_persistent) when enabled
// and the device holds it: reads the status once at the end.  Returns
// JCFD_ERR_UNSUPPORTED, with nothing launched, when the caller should run
// the chunked two-kernel form instead.
int cg_persistent_solve(jcfd_fvm_plan* p, float rtol, float atol, int maxiter, bool helm, float coef,
                        int* iterations, hipStream_t s) {
  if (!p->cg_persistent) return JCFD_ERR_UNSUPPORTED;
  int rc = cg_launch_persistent(p->params, p->cg, rtol, atol, maxiter, helm, coef, s);
  if (rc) return rc;
  if ((rc = check_hip(hipMemcpyAsync(p->status_host, p->cg.status, 4 * sizeof(int), hipMemcpyDeviceToHost,
                                     s), "cg status")) ||
      (rc = check_hip(hipStreamSynchronize(s), "cg status sync")))
    return rc;
  if (p->status_host[3] || !p->status_host[0])
    return fail(JCFD_ERR_DEVICE, "persistent CG: a grid barrier timed out");
  *iterations = p->status_host[1];
  return JCFD_OK;
}

// CG on p->rhs into p->cg.x (solve_cg).
int cg_solve(jcfd_fvm_plan* p, hipStream_t s) {
  const float rtol = static_cast<float>(p->cfg.cg_rtol), atol = static_cast<float>(p->cfg.cg_atol);
  int rc = cg_launch_init(p->params, p->rhs, p->cg, rtol, atol, p->cg_maxiter, s);
  if (rc) return rc;
  rc = cg_persistent_solve(p, rtol, atol, p->cg_maxiter, false, 0.f, &p->cg_last_iterations, s);
  if (rc != JCFD_ERR_UNSUPPORTED) return rc;
  return cg_iterate(p, p->cg_maxiter, kCgChunk, &p->cg_last_iterations, s, [&](int k, hipStream_t st) {
    return cg_launch_iteration(p->params, p->cg, k, rtol, atol, p->cg_maxiter, st);
  });
}

// diffusion.solve_cg (diffusion.py:137-167) in place on both components of
// an instance.  The system is close to the identity for stable time steps,
// so the first chunk is short.
constexpr int kDiffusionFirstChunk = 8;

int cg_diffuse(jcfd_fvm_plan* p, float* u, float* v, hipStream_t s) {
  const float rtol = static_cast<float>(p->cfg.diff_cg_rtol);
  const float atol = static_cast<float>(p->cfg.diff_cg_atol);
  const size_t bytes = sizeof(float) * static_cast<size_t>(p->params.n) * p->params.n;
  float* comps[2] = {u, v};
  for (int m = 0; m < 2; ++m) {
    int rc = cg_launch_init_helmholtz(p->params, comps[m], p->cg, p->diff_coef, s);
    if (rc == JCFD_OK) {
      rc = cg_persistent_solve(p, rtol, atol, p->diff_maxiter, true, p->diff_coef,
                               &p->diff_last_iterations[m], s);
      if (rc == JCFD_ERR_UNSUPPORTED)
        rc = cg_iterate(p, p->diff_maxiter, kDiffusionFirstChunk, &p->diff_last_iterations[m], s,
                        [&](int k, hipStream_t st) {
                          return cg_launch_iteration_helmholtz(p->params, p->cg, k, rtol, atol,
                                                               p->diff_maxiter, p->diff_coef, st);
                        });
    }
    if (rc == JCFD_OK)
      rc = check_hip(hipMemcpyAsync(comps[m], p->cg.x, bytes, hipMemcpyDeviceToDevice, s),
                     "copy diffused component");
    if (rc) return rc;
  }
  return JCFD_OK;
}

// The implicit diffusion solve of the plan in place on the step's output.
int fvm_diffuse_output(jcfd_fvm_plan* p, const FvmBuffers& b, hipStream_t s) {
  if (p->dlam != nullptr) return fvm_launch_diffuse(p->params, b, p->dlam, s);
  if (p->cfg.diffusion == JCFD_DIFFUSION_IMPLICIT_CG) return cg_diffuse(p, b.u_out, b.v_out, s);
  return JCFD_OK;
}

// nsteps steps: the first reads b.u/b.v, the rest run in place on
// b.u_out/b.v_out (replayed from a graph when enabled).
// One navier_stokes_rk step (time_stepping.py:83-104): k0 = F(u0); for each
// later stage, u_i = P(u0 + dt sum a k) and k_i = F(u_i); then
// P(u0 + dt sum b k) into b.u_out.
int fvm_rk_step(const jcfd_fvm_plan* p, const FvmBuffers& b, hipStream_t s) {
  const size_t field = static_cast<size_t>(p->params.n) * p->params.n * p->batch;
  auto ku = [&](int i) { return p->kbuf + (2 * i) * field; };
  auto kv = [&](int i) { return p->kbuf + (2 * i + 1) * field; };
  FvmBuffers e = b;
  e.us = ku(0);
  e.vs = kv(0);
  int rc = fvm_launch_rk_explicit(p->params, e, s);
  for (int i = 1; i < p->rk_stages && rc == JCFD_OK; ++i) {
    FvmBuffers c = b;
    c.rk = p->rk_terms + (i - 1);
    c.u_out = p->sbuf;
    c.v_out = p->sbuf + field;
    if ((rc = fvm_launch_rk_project(p->params, c, s))) break;
    FvmBuffers e2 = b;
    e2.u = p->sbuf;
    e2.v = p->sbuf + field;
    e2.us = ku(i);
    e2.vs = kv(i);
    rc = fvm_launch_rk_explicit(p->params, e2, s);
  }
  if (rc) return rc;
  FvmBuffers c = b;
  c.rk = p->rk_terms + (p->rk_stages - 1);
  return fvm_launch_rk_project(p->params, c, s);
}

// One step: the fast-diagonal step kernels (or the Runge-Kutta stages), then
// (implicit diffusion plans) the diffusion solve in place on the output.
int fvm_full_step(jcfd_fvm_plan* p, const FvmBuffers& b, hipStream_t s) {
  int rc = p->rk_terms != nullptr ? fvm_rk_step(p, b, s) : fvm_launch_step(p->params, b, s);
  if (rc == JCFD_OK) rc = fvm_diffuse_output(p, b, s);
  return rc;
}

int fvm_run(jcfd_fvm_plan* p, FvmBuffers b, int64_t nsteps, hipStream_t s) {
  int rc;
  if (p->cfg.pressure == JCFD_PRESSURE_CG) {
    b.rhs = p->rhs;
    b.q = p->cg.x;
    for (int64_t step = 0; step < nsteps; ++step) {
      if ((rc = fvm_launch_forward_rhs(p->params, b, s)) || (rc = cg_solve(p, s)) ||
          (rc = fvm_launch_project_q(p->params, b, s)))
        return rc;
      if ((rc = fvm_diffuse_output(p, b, s))) return rc;
      b.u = b.u_out;
      b.v = b.v_out;
    }
    return JCFD_OK;
  }
  if (p->cfg.diffusion == JCFD_DIFFUSION_IMPLICIT_CG) {
    // the CG diffusion solve reads its convergence flag on the host: never captured
    for (int64_t step = 0; step < nsteps; ++step) {
      if ((rc = fvm_full_step(p, b, s))) return rc;
      b.u = b.u_out;
      b.v = b.v_out;
    }
    return JCFD_OK;
  }
  if ((rc = fvm_full_step(p, b, s))) return rc;
  b.u = b.u_out;
  b.v = b.v_out;
  return run_steps(p->graphs, b.u_out, b.v_out, b.slots, b.count, nsteps - 1, s,
                   [&](hipStream_t gs) { return fvm_full_step(p, b, gs); });
}
}  // namespace

extern "C" {

const char* jcfd_last_error(void) { return g_last_error.c_str(); }

const char* jcfd_version(void) { return "jcfd_amd 0.1 (gfx950)"; }

int jcfd_fvm_plan_create(const jcfd_fvm_config* config, jcfd_fvm_plan** plan) {
  if (config == nullptr || plan == nullptr) return fail(JCFD_ERR_INVALID_ARGUMENT, "null argument");
  *plan = nullptr;
  const jcfd_fvm_config& c = *config;
  if (!fvm_size_supported(c.n))
    return fail(JCFD_ERR_UNSUPPORTED,
                "FVM grid size " + std::to_string(c.n) + " not built (32..1024, powers of two)");
  if (!(c.dx > 0) || !(c.dy > 0) || !std::isfinite(c.dx) || !std::isfinite(c.dy))
    return fail(JCFD_ERR_INVALID_ARGUMENT, "grid steps must be positive and finite");
  if (!std::isfinite(c.dt) || !std::isfinite(c.density) || c.density == 0)
    return fail(JCFD_ERR_INVALID_ARGUMENT, "dt and density must be finite, density non-zero");
  if (c.convect != JCFD_CONVECT_VAN_LEER && c.convect != JCFD_CONVECT_VAN_LEER_LIMITERS &&
      c.convect != JCFD_CONVECT_LINEAR && c.convect != JCFD_CONVECT_UPWIND)
    return fail(JCFD_ERR_UNSUPPORTED,
                "convection must be advect_van_leer or advect_van_leer_using_limiters");
  if (c.pressure != JCFD_PRESSURE_FAST_DIAG && c.pressure != JCFD_PRESSURE_CG)
    return fail(JCFD_ERR_UNSUPPORTED, "pressure solve must be solve_fast_diag or solve_cg");
  if (c.pressure == JCFD_PRESSURE_CG && c.batch > 1)
    return fail(JCFD_ERR_UNSUPPORTED, "the CG pressure solve is built for one instance per plan");
  if (c.pressure == JCFD_PRESSURE_CG && (!std::isfinite(c.cg_rtol) || !std::isfinite(c.cg_atol)))
    return fail(JCFD_ERR_INVALID_ARGUMENT, "CG tolerances must be finite");
  if (c.diffusion != JCFD_DIFFUSION_EXPLICIT && c.diffusion != JCFD_DIFFUSION_IMPLICIT_FAST_DIAG &&
      c.diffusion != JCFD_DIFFUSION_IMPLICIT_CG)
    return fail(JCFD_ERR_UNSUPPORTED, "diffusion must be explicit, implicit fast-diagonal or implicit CG");
  if (c.diffusion != JCFD_DIFFUSION_EXPLICIT && (!c.has_viscosity || !std::isfinite(c.viscosity)))
    return fail(JCFD_ERR_INVALID_ARGUMENT, "implicit diffusion needs a finite viscosity");
  if (c.diffusion == JCFD_DIFFUSION_IMPLICIT_CG && c.batch > 1)
    return fail(JCFD_ERR_UNSUPPORTED, "the CG diffusion solve is built for one instance per plan");
  if (c.diffusion == JCFD_DIFFUSION_IMPLICIT_CG &&
      (!std::isfinite(c.diff_cg_rtol) || !std::isfinite(c.diff_cg_atol)))
    return fail(JCFD_ERR_INVALID_ARGUMENT, "CG diffusion tolerances must be finite");
  if (c.rk_stages < 0 || c.rk_stages > 4)
    return fail(JCFD_ERR_UNSUPPORTED, "Runge-Kutta tableaux of 1 to 4 stages are built");
  for (int k = 0; k < 6; ++k)
    if (!std::isfinite(c.rk_a[k])) return fail(JCFD_ERR_INVALID_ARGUMENT, "rk_a must be finite");
  for (int k = 0; k < 4; ++k)
    if (!std::isfinite(c.rk_b[k])) return fail(JCFD_ERR_INVALID_ARGUMENT, "rk_b must be finite");
  const bool generic_rk = c.rk_stages > 1 || (c.rk_stages == 1 && c.rk_b[0] != 1.0);
  if (generic_rk && (c.pressure != JCFD_PRESSURE_FAST_DIAG || c.diffusion != JCFD_DIFFUSION_EXPLICIT))
    return fail(JCFD_ERR_UNSUPPORTED,
                "Runge-Kutta tableaux are built with solve_fast_diag and explicit diffusion");

  auto* p = new (std::nothrow) jcfd_fvm_plan();
  if (p == nullptr) return fail(JCFD_ERR_OUT_OF_MEMORY, "host allocation failed");
  p->cfg = c;
  p->cfg.forcing_x = nullptr;
  p->graphs.enabled = false;  // one instance: kernels already run back to back
  if (hipGetDevice(&p->device) != hipSuccess) {
    delete p;
    return fail(JCFD_ERR_DEVICE, "hipGetDevice failed (no GPU?)");
  }
  FvmParams& q = p->params;
  const int n = c.n;
  q.n = n;
  q.dx = static_cast<float>(c.dx);
  q.dy = static_cast<float>(c.dy);
  // finite_differences.laplacian: scales = square(1 / array(step, float32)), summed in fp32
  const float ix = 1.0f / static_cast<float>(c.dx), iy = 1.0f / static_cast<float>(c.dy);
  q.inv_dx_sq = ix * ix;
  q.inv_dy_sq = iy * iy;
  q.rcp_dx = ix;  // x / dx as over_step(x, dx, rcp_dx)
  q.rcp_dy = iy;
  q.lap_sum = q.inv_dx_sq + q.inv_dy_sq;
  q.dt = static_cast<float>(c.dt);
  q.dt_over_dx = static_cast<float>(c.dt / c.dx);
  q.dt_over_dy = static_cast<float>(c.dt / c.dy);
  // implicit-diffusion plans leave the Laplacian out of the explicit terms
  q.has_diffusion = (c.has_viscosity && c.diffusion == JCFD_DIFFUSION_EXPLICIT) ? 1 : 0;
  q.nu = static_cast<float>(c.has_viscosity ? c.viscosity / c.density : 0.0);
  q.rho = static_cast<float>(c.density);
  q.has_forcing = c.has_forcing ? 1 : 0;
  q.lin = static_cast<float>(c.linear_coefficient);
  q.convect = c.convect;
  (void)fvm_default_tiles(&q);

  p->batch = c.batch > 1 ? c.batch : 1;
  p->graphs.enabled = p->batch > 1;
  const size_t nn = static_cast<size_t>(n) * n * p->batch;
  int rc = JCFD_OK;
  if ((rc = device_alloc(&p->us, nn, &p->bytes)) || (rc = device_alloc(&p->vs, nn, &p->bytes)) ||
      (rc = device_alloc(&p->spec, nn / 2, &p->bytes)) ||
      (rc = device_alloc(&p->lam_xy, static_cast<size_t>(n) + n / 2 + 1, &p->bytes)) ||
      (rc = device_alloc(&p->tw, static_cast<size_t>(n), &p->bytes)) ||
      (rc = device_alloc(&p->fx, static_cast<size_t>(n), &p->bytes)) ||
      ((c.pressure == JCFD_PRESSURE_CG || c.diffusion == JCFD_DIFFUSION_IMPLICIT_CG) &&
       (rc = cg_alloc(p))) ||
      (c.diffusion == JCFD_DIFFUSION_IMPLICIT_FAST_DIAG &&
       (rc = device_alloc(&p->dlam, static_cast<size_t>(n) * (n / 2 + 1), &p->bytes))) ||
      (generic_rk && ((rc = device_alloc(&p->kbuf, 2 * c.rk_stages * nn, &p->bytes)) ||
                      (rc = device_alloc(&p->sbuf, 2 * nn, &p->bytes)) ||
                      (rc = device_alloc(&p->rk_terms, static_cast<size_t>(c.rk_stages), &p->bytes))))) {
    jcfd_fvm_plan_destroy(p);
    return rc;
  }
  const std::vector<double> lam = laplacian_eigenvalues(n, c.dx, c.dy);
  const std::vector<float2> tw = twiddles(n);
  std::vector<float> fx(n, 0.0f);
  if (c.has_forcing && c.forcing_x != nullptr) std::memcpy(fx.data(), c.forcing_x, n * sizeof(float));
  if ((rc = check_hip(hipMemcpy(p->lam_xy, lam.data(), lam.size() * sizeof(double), hipMemcpyHostToDevice),
                      "copy Laplacian eigenvalues")) ||
      (rc = check_hip(hipMemcpy(p->tw, tw.data(), tw.size() * sizeof(float2), hipMemcpyHostToDevice),
                      "copy twiddles")) ||
      (rc = check_hip(hipMemcpy(p->fx, fx.data(), fx.size() * sizeof(float), hipMemcpyHostToDevice),
                      "copy forcing"))) {
    jcfd_fvm_plan_destroy(p);
    return rc;
  }
  if (generic_rk) {
    // terms of stage i = 1 .. stages-1 (row i-1 of a), then b; zero
    // coefficients are left out, as the reference's sums skip them
    p->rk_stages = c.rk_stages;
    std::vector<RkTerms> terms(c.rk_stages);
    auto fill = [&](RkTerms& t, const double* coef, int count) {
      t = RkTerms{};
      for (int j = 0; j < count; ++j) {
        if (coef[j] == 0.0) continue;
        t.ku[t.n] = p->kbuf + (2 * j) * nn;
        t.kv[t.n] = p->kbuf + (2 * j + 1) * nn;
        t.c[t.n] = static_cast<float>(coef[j]);
        ++t.n;
      }
    };
    int row = 0;
    for (int i = 1; i < c.rk_stages; ++i) {
      fill(terms[i - 1], c.rk_a + row, i);
      row += i;
    }
    fill(terms[c.rk_stages - 1], c.rk_b, c.rk_stages);
    if ((rc = check_hip(hipMemcpy(p->rk_terms, terms.data(), terms.size() * sizeof(RkTerms),
                                  hipMemcpyHostToDevice),
                        "copy Runge-Kutta terms"))) {
      jcfd_fvm_plan_destroy(p);
      return rc;
    }
  }
  if (c.diffusion == JCFD_DIFFUSION_IMPLICIT_CG) {
    // solve_cg's linear_op: u - dt * nu * laplacian(u), the Python product
    // dt * nu rounded once to float32
    p->diff_coef = static_cast<float>(c.dt * c.viscosity);
    p->diff_maxiter = c.diff_cg_maxiter > 0 ? c.diff_cg_maxiter : static_cast<int>(10 * n * n);
  }
  if (p->dlam != nullptr) {
    const std::vector<float> dm = diffusion_multiplier(n, c.dx, c.dy, c.viscosity, c.dt);
    if ((rc = check_hip(hipMemcpy(p->dlam, dm.data(), dm.size() * sizeof(float), hipMemcpyHostToDevice),
                        "copy diffusion multiplier"))) {
      jcfd_fvm_plan_destroy(p);
      return rc;
    }
  }
  *plan = p;
  return JCFD_OK;
}

int jcfd_fvm_plan_destroy(jcfd_fvm_plan* p) {
  if (p == nullptr) return JCFD_OK;
  for (void* ptr : {static_cast<void*>(p->us), static_cast<void*>(p->vs), static_cast<void*>(p->spec),
                    static_cast<void*>(p->lam_xy), static_cast<void*>(p->tw), static_cast<void*>(p->fx),
                    static_cast<void*>(p->dlam), static_cast<void*>(p->kbuf), static_cast<void*>(p->sbuf),
                    static_cast<void*>(p->rk_terms)})
    if (ptr != nullptr) (void)hipFree(ptr);
  for (void* ptr : {static_cast<void*>(p->cg.x), static_cast<void*>(p->cg.r),
                    static_cast<void*>(p->cg.p[0]), static_cast<void*>(p->cg.p[1]),
                    static_cast<void*>(p->cg.ap), static_cast<void*>(p->cg.part_rr[0]),
                    static_cast<void*>(p->cg.part_rr[1]), static_cast<void*>(p->cg.part_pap),
                    static_cast<void*>(p->cg.part_bb), static_cast<void*>(p->cg.status),
                    static_cast<void*>(p->rhs)})
    if (ptr != nullptr) (void)hipFree(ptr);
  if (p->status_host) (void)hipHostFree(p->status_host);
  p->graphs.release();
  delete p;
  return JCFD_OK;
}

int32_t jcfd_fvm_cg_iterations(const jcfd_fvm_plan* p) { return p ? p->cg_last_iterations : 0; }

int32_t jcfd_fvm_diffusion_cg_iterations(const jcfd_fvm_plan* p, int32_t component) {
  return (p && (component == 0 || component == 1)) ? p->diff_last_iterations[component] : 0;
}

int jcfd_fvm_plan_set_graphs(jcfd_fvm_plan* p, int32_t enabled) {
  if (p == nullptr) return fail(JCFD_ERR_INVALID_ARGUMENT, "null plan");
  p->graphs.enabled = enabled != 0;
  return JCFD_OK;
}

int jcfd_fvm_plan_set_cg_persistent(jcfd_fvm_plan* p, int32_t enabled) {
  if (p == nullptr) return fail(JCFD_ERR_INVALID_ARGUMENT, "null plan");
  p->cg_persistent = enabled != 0;
  return JCFD_OK;
}

int jcfd_fvm_plan_get_tiles(const jcfd_fvm_plan* p, jcfd_fvm_tiles* tiles) {
  if (p == nullptr || tiles == nullptr) return fail(JCFD_ERR_INVALID_ARGUMENT, "null argument");
  tiles->rows_forward = p->params.rb_fwd;
  tiles->poisson_cols = p->params.w_cols;
  tiles->rows_project = p->params.rb_proj;
  tiles->radix = p->params.radix;
  return JCFD_OK;
}

int jcfd_fvm_plan_set_tiles(jcfd_fvm_plan* p, const jcfd_fvm_tiles* tiles) {
  if (p == nullptr || tiles == nullptr) return fail(JCFD_ERR_INVALID_ARGUMENT, "null argument");
  const int n = p->params.n;
  if (tiles->radix != 8 && tiles->radix != 4)
    return fail(JCFD_ERR_UNSUPPORTED, "radix schedule must be 8 or 4");
  const int32_t t[3] = {tiles->rows_forward, tiles->poisson_cols, tiles->rows_project};
  const char* names[3] = {"rows_forward", "poisson_cols", "rows_project"};
  for (int k = 0; k < 3; ++k)
    if (!fvm_tile_supported(n, k, t[k]))
      return fail(JCFD_ERR_UNSUPPORTED, std::string(names[k]) + " tile " + std::to_string(t[k]) +
                                            " not built for n = " + std::to_string(n));
  p->params.rb_fwd = t[0];
  p->params.w_cols = t[1];
  p->params.rb_proj = t[2];
  p->params.radix = tiles->radix;
  p->graphs.release();  // captured launches carry the old shapes
  return JCFD_OK;
}

int64_t jcfd_fvm_plan_device_bytes(const jcfd_fvm_plan* p) { return p ? p->bytes : 0; }

int jcfd_fvm_step(jcfd_fvm_plan* p, const jcfd_tensor* u, const jcfd_tensor* v, jcfd_tensor* u_out,
                  jcfd_tensor* v_out, int64_t nsteps, void* stream) {
  if (p == nullptr) return fail(JCFD_ERR_INVALID_ARGUMENT, "null plan");
  if (nsteps < 0) return fail(JCFD_ERR_INVALID_ARGUMENT, "nsteps must be >= 0");
  const int n = p->params.n;
  void *pu, *pv, *pu2, *pv2;
  int rc;
  if ((rc = check_batched(u, "u", kDLFloat, 32, p->batch, n, n, p->device, &pu)) ||
      (rc = check_batched(v, "v", kDLFloat, 32, p->batch, n, n, p->device, &pv)) ||
      (rc = check_batched(u_out, "u_out", kDLFloat, 32, p->batch, n, n, p->device, &pu2)) ||
      (rc = check_batched(v_out, "v_out", kDLFloat, 32, p->batch, n, n, p->device, &pv2)))
    return rc;
  hipStream_t s = static_cast<hipStream_t>(stream);
  const size_t bytes = sizeof(float) * p->batch * static_cast<size_t>(n) * n;
  if (nsteps == 0) {
    if (pu2 != pu && (rc = check_hip(hipMemcpyAsync(pu2, pu, bytes, hipMemcpyDeviceToDevice, s), "copy")))
      return rc;
    if (pv2 != pv && (rc = check_hip(hipMemcpyAsync(pv2, pv, bytes, hipMemcpyDeviceToDevice, s), "copy")))
      return rc;
    return JCFD_OK;
  }
  FvmBuffers b = fvm_buffers(p);
  b.u = static_cast<const float*>(pu);
  b.v = static_cast<const float*>(pv);
  b.u_out = static_cast<float*>(pu2);
  b.v_out = static_cast<float*>(pv2);
  b.count = p->batch;
  return fvm_run(p, b, nsteps, s);
}

int jcfd_fvm_step_slots(jcfd_fvm_plan* p, jcfd_tensor* u, jcfd_tensor* v, const jcfd_tensor* slots,
                        int64_t nsteps, void* stream) {
  if (p == nullptr) return fail(JCFD_ERR_INVALID_ARGUMENT, "null plan");
  if (nsteps < 0) return fail(JCFD_ERR_INVALID_ARGUMENT, "nsteps must be >= 0");
  const int n = p->params.n;
  void *pu, *pv, *ps;
  int count = 0;
  int rc;
  if ((rc = check_batched(u, "u", kDLFloat, 32, p->batch, n, n, p->device, &pu)) ||
      (rc = check_batched(v, "v", kDLFloat, 32, p->batch, n, n, p->device, &pv)) ||
      (rc = check_slots(slots, p->batch, p->device, &ps, &count)))
    return rc;
  if (nsteps == 0 || count == 0) return JCFD_OK;
  FvmBuffers b = fvm_buffers(p);
  b.u = static_cast<const float*>(pu);
  b.v = static_cast<const float*>(pv);
  b.u_out = static_cast<float*>(pu);
  b.v_out = static_cast<float*>(pv);
  b.slots = static_cast<const int*>(ps);
  b.count = count;
  return fvm_run(p, b, nsteps, static_cast<hipStream_t>(stream));
}

}  // extern "C"

// ------------------------------------------------- allocation-free stepping
// Scratch of jcfd_fvm_step_noalloc: u*, v* and the packed divergence
// spectrum of every instance of one call (12 * batch * n^2 bytes), in a
// file-scope device array that every GPU the library runs on holds until the
// process ends.  Host code never uses the array's name as an address (on the
// host it is a stand-in symbol, not device memory): the device address comes
// from hipGetSymbolAddress, once per device, with that device current.
namespace {
constexpr size_t kNoallocScratchBytes = size_t(16) << 20;
constexpr int kNoallocMaxDevices = 64;
}  // namespace

static __device__ __align__(256) unsigned char scratch[kNoallocScratchBytes];

namespace {
std::mutex g_scratch_mu;
void* g_scratch_ptr[kNoallocMaxDevices] = {};

// Device address of `scratch` on `device` (looked up on the first call for
// that device, with it made current for the lookup).
int noalloc_scratch(int device, float** out) {
  if (device < 0 || device >= kNoallocMaxDevices)
    return fail(JCFD_ERR_UNSUPPORTED, "device id " + std::to_string(device) + " out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[device] == nullptr) {
    int current = -1;
    int rc;
    if ((rc = check_hip(hipGetDevice(&current), "hipGetDevice"))) return rc;
    if (current != device && (rc = check_hip(hipSetDevice(device), "hipSetDevice"))) return rc;
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (current != device) (void)hipSetDevice(current);
    if ((rc = check_hip(e, "hipGetSymbolAddress(scratch)"))) return rc;
    if (ptr == nullptr) return fail(JCFD_ERR_DEVICE, "hipGetSymbolAddress(scratch) gave a null address");
    g_scratch_ptr[device] = ptr;
  }
  *out = static_cast<float*>(g_scratch_ptr[device]);
  return JCFD_OK;
}
}  // namespace

extern "C" {

int jcfd_fvm_step_noalloc(jcfd_fvm_plan* p, const jcfd_tensor* u, const jcfd_tensor* v,
                          jcfd_tensor* u_out, jcfd_tensor* v_out, int64_t nsteps, void* stream) {
  if (p == nullptr) return fail(JCFD_ERR_INVALID_ARGUMENT, "null plan");
  if (nsteps < 0) return fail(JCFD_ERR_INVALID_ARGUMENT, "nsteps must be >= 0");
  if (p->cfg.pressure != JCFD_PRESSURE_FAST_DIAG || p->rk_terms != nullptr ||
      p->cfg.diffusion == JCFD_DIFFUSION_IMPLICIT_CG)
    return fail(JCFD_ERR_UNSUPPORTED,
                "jcfd_fvm_step_noalloc steps forward-Euler plans with solve_fast_diag "
                "(explicit or fast-diagonal implicit diffusion)");
  const int n = p->params.n;
  const size_t cells = static_cast<size_t>(p->batch) * n * n;
  const size_t need = 12 * cells;  // u*, v* (fp32) + packed spectrum (n^2 / 2 complex64)
  if (need > kNoallocScratchBytes)
    return fail(JCFD_ERR_INVALID_ARGUMENT,
                "jcfd_fvm_step_noalloc: batch " + std::to_string(p->batch) + " x " +
                    std::to_string(n) + "^2 needs " + std::to_string(need) +
                    " bytes of scratch, above the static maximum of " +
                    std::to_string(kNoallocScratchBytes) + " (12 * batch * n^2 <= 16 MiB)");
  void *pu, *pv, *pu2, *pv2;
  int rc;
  if ((rc = check_batched(u, "u", kDLFloat, 32, p->batch, n, n, p->device, &pu)) ||
      (rc = check_batched(v, "v", kDLFloat, 32, p->batch, n, n, p->device, &pv)) ||
      (rc = check_batched(u_out, "u_out", kDLFloat, 32, p->batch, n, n, p->device, &pu2)) ||
      (rc = check_batched(v_out, "v_out", kDLFloat, 32, p->batch, n, n, p->device, &pv2)))
    return rc;
  float* base = nullptr;
  if ((rc = noalloc_scratch(p->device, &base))) return rc;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (nsteps == 0) {
    const size_t bytes = sizeof(float) * cells;
    if (pu2 != pu && (rc = check_hip(hipMemcpyAsync(pu2, pu, bytes, hipMemcpyDeviceToDevice, s), "copy")))
      return rc;
    if (pv2 != pv && (rc = check_hip(hipMemcpyAsync(pv2, pv, bytes, hipMemcpyDeviceToDevice, s), "copy")))
      return rc;
    return JCFD_OK;
  }
  // Every scratch word a step reads was written earlier in the same step
  // (rows-forward writes u*, v* and the spectrum before the column and
  // projection kernels read them), so nothing needs clearing between calls.
  FvmBuffers b = fvm_buffers(p);
  b.us = base;
  b.vs = base + cells;
  b.spec = reinterpret_cast<float2*>(base + 2 * cells);
  b.u = static_cast<const float*>(pu);
  b.v = static_cast<const float*>(pv);
  b.u_out = static_cast<float*>(pu2);
  b.v_out = static_cast<float*>(pv2);
  b.count = p->batch;
  // direct launches only: capturing a graph would allocate
  for (int64_t k = 0; k < nsteps; ++k) {
    if ((rc = fvm_full_step(p, b, s))) return rc;
    b.u = b.u_out;
    b.v = b.v_out;
  }
  return JCFD_OK;
}

int jcfd_fvm_explicit_terms(jcfd_fvm_plan* p, const jcfd_tensor* u, const jcfd_tensor* v,
                            jcfd_tensor* du, jcfd_tensor* dv, void* stream) {
  if (p == nullptr) return fail(JCFD_ERR_INVALID_ARGUMENT, "null plan");
  const int n = p->params.n;
  void *pu, *pv, *pdu, *pdv;
  int rc;
  if ((rc = check_tensor(u, "u", kDLFloat, 32, n, n, p->device, &pu)) ||
      (rc = check_tensor(v, "v", kDLFloat, 32, n, n, p->device, &pv)) ||
      (rc = check_tensor(du, "du", kDLFloat, 32, n, n, p->device, &pdu)) ||
      (rc = check_tensor(dv, "dv", kDLFloat, 32, n, n, p->device, &pdv)))
    return rc;
  if (pdu == pu || pdu == pv || pdv == pu || pdv == pv)
    return fail(JCFD_ERR_INVALID_ARGUMENT, "explicit terms: outputs must not alias inputs");
  return fvm_launch_explicit_terms(p->params, static_cast<const float*>(pu),
                                   static_cast<const float*>(pv), static_cast<float*>(pdu),
                                   static_cast<float*>(pdv), p->fx, static_cast<hipStream_t>(stream));
}

int jcfd_fvm_projection(jcfd_fvm_plan* p, const jcfd_tensor* u, const jcfd_tensor* v,
                        jcfd_tensor* u_out, jcfd_tensor* v_out, void* stream) {
  if (p == nullptr) return fail(JCFD_ERR_INVALID_ARGUMENT, "null plan");
  const int n = p->params.n;
  void *pu, *pv, *pu2, *pv2;
  int rc;
  if ((rc = check_tensor(u, "u", kDLFloat, 32, n, n, p->device, &pu)) ||
      (rc = check_tensor(v, "v", kDLFloat, 32, n, n, p->device, &pv)) ||
      (rc = check_tensor(u_out, "u_out", kDLFloat, 32, n, n, p->device, &pu2)) ||
      (rc = check_tensor(v_out, "v_out", kDLFloat, 32, n, n, p->device, &pv2)))
    return rc;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (p->cfg.pressure == JCFD_PRESSURE_CG) {
    FvmBuffers b = fvm_buffers(p);
    b.us = static_cast<float*>(pu);
    b.vs = static_cast<float*>(pv);
    b.u_out = static_cast<float*>(pu2);
    b.v_out = static_cast<float*>(pv2);
    b.q = p->cg.x;
    if ((rc = fvm_launch_divergence(p->params, static_cast<const float*>(pu),
                                    static_cast<const float*>(pv), p->rhs, s)) ||
        (rc = cg_solve(p, s)))
      return rc;
    return fvm_launch_project_q(p->params, b, s);
  }
  return fvm_launch_projection(p->params, static_cast<const float*>(pu), static_cast<const float*>(pv),
                               static_cast<float*>(pu2), static_cast<float*>(pv2), p->spec, p->lam_xy,
                               p->tw, s);
}

int jcfd_fvm_diffuse(jcfd_fvm_plan* p, const jcfd_tensor* u, const jcfd_tensor* v, jcfd_tensor* u_out,
                     jcfd_tensor* v_out, void* stream) {
  if (p == nullptr) return fail(JCFD_ERR_INVALID_ARGUMENT, "null plan");
  if (p->dlam == nullptr && p->cfg.diffusion != JCFD_DIFFUSION_IMPLICIT_CG)
    return fail(JCFD_ERR_UNSUPPORTED, "jcfd_fvm_diffuse needs an implicit-diffusion plan");
  const int n = p->params.n;
  void *pu, *pv, *pu2, *pv2;
  int rc;
  if ((rc = check_tensor(u, "u", kDLFloat, 32, n, n, p->device, &pu)) ||
      (rc = check_tensor(v, "v", kDLFloat, 32, n, n, p->device, &pv)) ||
      (rc = check_tensor(u_out, "u_out", kDLFloat, 32, n, n, p->device, &pu2)) ||
      (rc = check_tensor(v_out, "v_out", kDLFloat, 32, n, n, p->device, &pv2)))
    return rc;
  hipStream_t s = static_cast<hipStream_t>(stream);
  const size_t bytes = sizeof(float) * static_cast<size_t>(n) * n;
  if (pu2 != pu && (rc = check_hip(hipMemcpyAsync(pu2, pu, bytes, hipMemcpyDeviceToDevice, s), "copy")))
    return rc;
  if (pv2 != pv && (rc = check_hip(hipMemcpyAsync(pv2, pv, bytes, hipMemcpyDeviceToDevice, s), "copy")))
    return rc;
  FvmBuffers b = fvm_buffers(p);
  b.u_out = static_cast<float*>(pu2);
  b.v_out = static_cast<float*>(pv2);
  return fvm_diffuse_output(p, b, s);
}

int jcfd_fvm_divergence(jcfd_fvm_plan* p, const jcfd_tensor* u, const jcfd_tensor* v,
                        jcfd_tensor* out, void* stream) {
  if (p == nullptr) return fail(JCFD_ERR_INVALID_ARGUMENT, "null plan");
  const int n = p->params.n;
  void *pu, *pv, *po;
  int rc;
  if ((rc = check_tensor(u, "u", kDLFloat, 32, n, n, p->device, &pu)) ||
      (rc = check_tensor(v, "v", kDLFloat, 32, n, n, p->device, &pv)) ||
      (rc = check_tensor(out, "out", kDLFloat, 32, n, n, p->device, &po)))
    return rc;
  return fvm_launch_divergence(p->params, static_cast<const float*>(pu), static_cast<const float*>(pv),
                               static_cast<float*>(po), static_cast<hipStream_t>(stream));
}

int jcfd_fvm_time_kernels(jcfd_fvm_plan* p, const jcfd_tensor* u, const jcfd_tensor* v,
                          int32_t reps, double* ms_per_launch, void* stream) {
  if (p == nullptr || ms_per_launch == nullptr)
    return fail(JCFD_ERR_INVALID_ARGUMENT, "null argument");
  if (reps <= 0) return fail(JCFD_ERR_INVALID_ARGUMENT, "reps must be positive");
  const int n = p->params.n;
  void *pu, *pv;
  int rc;
  if ((rc = check_batched(u, "u", kDLFloat, 32, p->batch, n, n, p->device, &pu)) ||
      (rc = check_batched(v, "v", kDLFloat, 32, p->batch, n, n, p->device, &pv)))
    return rc;
  hipStream_t s = static_cast<hipStream_t>(stream);
  FvmBuffers b{};
  b.count = p->batch;
  b.u = static_cast<const float*>(pu);
  b.v = static_cast<const float*>(pv);
  b.us = p->us;
  b.vs = p->vs;
  b.spec = p->spec;
  b.u_out = p->us;  // rows-project reads and writes the same element per thread
  b.v_out = p->vs;
  b.fx = p->fx;
  b.lam_xy = p->lam_xy;
  b.lam_mode = 2;
  b.tw = p->tw;
  hipEvent_t e0, e1;
  if ((rc = check_hip(hipEventCreate(&e0), "event")) ) return rc;
  if ((rc = check_hip(hipEventCreate(&e1), "event"))) {
    (void)hipEventDestroy(e0);
    return rc;
  }
  // one full step first so every scratch buffer holds valid data
  rc = fvm_launch_step(p->params, b, s);
  for (int k = 0; k < JCFD_FVM_KERNEL_COUNT && rc == JCFD_OK; ++k) {
    rc = check_hip(hipEventRecord(e0, s), "event record");
    for (int r = 0; r < reps && rc == JCFD_OK; ++r) rc = fvm_launch_kernel(p->params, b, k, s);
    if (rc == JCFD_OK) rc = check_hip(hipEventRecord(e1, s), "event record");
    if (rc == JCFD_OK) rc = check_hip(hipEventSynchronize(e1), "event sync");
    float ms = 0.f;
    if (rc == JCFD_OK) rc = check_hip(hipEventElapsedTime(&ms, e0, e1), "elapsed");
    ms_per_launch[k] = static_cast<double>(ms) / reps;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return rc;
}

int jcfd_fvm_kernel_bytes(const jcfd_fvm_plan* p, int64_t* bytes) {
  if (p == nullptr || bytes == nullptr) return fail(JCFD_ERR_INVALID_ARGUMENT, "null argument");
  // every instance of the plan: time_kernels launches them all
  const int64_t cells = static_cast<int64_t>(p->params.n) * p->params.n * p->batch;
  bytes[JCFD_FVM_KERNEL_ROWS_FORWARD] = 20 * cells;  // read u,v; write u*,v*; write packed spectrum
  bytes[JCFD_FVM_KERNEL_POISSON_COLS] = 8 * cells;   // read + write packed spectrum
  bytes[JCFD_FVM_KERNEL_ROWS_PROJECT] = 20 * cells;  // read spectrum, u*, v*; write u, v
  return JCFD_OK;
}

int jcfd_fvm_step_vjp(jcfd_fvm_plan* p, const jcfd_tensor* u, const jcfd_tensor* v,
                      const jcfd_tensor* lam_u, const jcfd_tensor* lam_v, jcfd_tensor* out_u,
                      jcfd_tensor* out_v, void* stream) {
  if (p == nullptr) return fail(JCFD_ERR_INVALID_ARGUMENT, "null plan");
  if (p->cfg.convect != JCFD_CONVECT_VAN_LEER || p->cfg.pressure != JCFD_PRESSURE_FAST_DIAG ||
      p->cfg.diffusion != JCFD_DIFFUSION_EXPLICIT || p->rk_terms != nullptr)
    return fail(JCFD_ERR_UNSUPPORTED,
                "the adjoint is built for advect_van_leer + explicit diffusion + solve_fast_diag");
  const int n = p->params.n;
  void *pu, *pv, *plu, *plv, *pou, *pov;
  int rc;
  if ((rc = check_tensor(u, "u", kDLFloat, 32, n, n, p->device, &pu)) ||
      (rc = check_tensor(v, "v", kDLFloat, 32, n, n, p->device, &pv)) ||
      (rc = check_tensor(lam_u, "lam_u", kDLFloat, 32, n, n, p->device, &plu)) ||
      (rc = check_tensor(lam_v, "lam_v", kDLFloat, 32, n, n, p->device, &plv)) ||
      (rc = check_tensor(out_u, "out_u", kDLFloat, 32, n, n, p->device, &pou)) ||
      (rc = check_tensor(out_v, "out_v", kDLFloat, 32, n, n, p->device, &pov)))
    return rc;
  for (void* o : {pou, pov})
    for (void* i : {pu, pv, plu, plv})
      if (o == i) return fail(JCFD_ERR_INVALID_ARGUMENT, "step vjp: outputs must not alias inputs");
  hipStream_t s = static_cast<hipStream_t>(stream);
  // lam* = P lam into the plan's u*/v* scratch, then the explicit-term transpose.
  if ((rc = fvm_launch_projection(p->params, static_cast<const float*>(plu),
                                  static_cast<const float*>(plv), p->us, p->vs, p->spec, p->lam_xy,
                                  p->tw, s)))
    return rc;
  return fvm_launch_adjoint_explicit(p->params, static_cast<const float*>(pu),
                                     static_cast<const float*>(pv), p->us, p->vs,
                                     static_cast<float*>(pou), static_cast<float*>(pov), s);
}

int jcfd_fvm_curl_vjp(jcfd_fvm_plan* p, const jcfd_tensor* w_bar, jcfd_tensor* u_bar,
                      jcfd_tensor* v_bar, void* stream) {
  if (p == nullptr) return fail(JCFD_ERR_INVALID_ARGUMENT, "null plan");
  const int n = p->params.n;
  void *pw, *pu, *pv;
  int rc;
  if ((rc = check_tensor(w_bar, "w_bar", kDLFloat, 32, n, n, p->device, &pw)) ||
      (rc = check_tensor(u_bar, "u_bar", kDLFloat, 32, n, n, p->device, &pu)) ||
      (rc = check_tensor(v_bar, "v_bar", kDLFloat, 32, n, n, p->device, &pv)))
    return rc;
  return fvm_launch_curl_vjp(p->params, static_cast<const float*>(pw), static_cast<float*>(pu),
                             static_cast<float*>(pv), static_cast<hipStream_t>(stream));
}

int jcfd_fvm_curl(jcfd_fvm_plan* p, const jcfd_tensor* u, const jcfd_tensor* v, jcfd_tensor* out,
                  void* stream) {
  if (p == nullptr) return fail(JCFD_ERR_INVALID_ARGUMENT, "null plan");
  const int n = p->params.n;
  void *pu, *pv, *po;
  int rc;
  if ((rc = check_tensor(u, "u", kDLFloat, 32, n, n, p->device, &pu)) ||
      (rc = check_tensor(v, "v", kDLFloat, 32, n, n, p->device, &pv)) ||
      (rc = check_tensor(out, "out", kDLFloat, 32, n, n, p->device, &po)))
    return rc;
  return fvm_launch_curl(p->params, static_cast<const float*>(pu), static_cast<const float*>(pv),
                         static_cast<float*>(po), static_cast<hipStream_t>(stream));
}

namespace {
// interpolation.py:36-64 for offset difference `delta` (Python floats there)
SlAxis sl_axis(double delta) {
  SlAxis a{};
  if (delta == std::floor(delta)) {
    a.f = static_cast<int>(delta);
    return a;
  }
  a.blend = 1;
  a.f = static_cast<int>(std::floor(delta));
  a.c = static_cast<int>(std::ceil(delta));
  const double fw = a.c - delta;
  a.fw = static_cast<float>(fw);
  a.cw = static_cast<float>(1.0 - fw);
  return a;
}
}  // namespace

int jcfd_fvm_advect_semilagrangian(jcfd_fvm_plan* p, const jcfd_tensor* c, double offset_x,
                                   double offset_y, const jcfd_tensor* u, const jcfd_tensor* v,
                                   double dt, jcfd_tensor* out, void* stream) {
  if (p == nullptr) return fail(JCFD_ERR_INVALID_ARGUMENT, "null plan");
  const int n = p->params.n;
  if (!std::isfinite(offset_x) || !std::isfinite(offset_y) || !std::isfinite(dt) ||
      std::fabs(offset_x) > 1e6 || std::fabs(offset_y) > 1e6)
    return fail(JCFD_ERR_INVALID_ARGUMENT, "offsets and dt must be finite");
  void *pc, *pu, *pv, *po;
  int rc;
  if ((rc = check_tensor(c, "c", kDLFloat, 32, n, n, p->device, &pc)) ||
      (rc = check_tensor(u, "u", kDLFloat, 32, n, n, p->device, &pu)) ||
      (rc = check_tensor(v, "v", kDLFloat, 32, n, n, p->device, &pv)) ||
      (rc = check_tensor(out, "out", kDLFloat, 32, n, n, p->device, &po)))
    return rc;
  if (po == pc || po == pu || po == pv)
    return fail(JCFD_ERR_INVALID_ARGUMENT, "out must not alias c, u or v");
  SlParams q{};
  q.u0 = sl_axis(offset_x - 1.0);  // u sits at (1, 0.5)
  q.u1 = sl_axis(offset_y - 0.5);
  q.v0 = sl_axis(offset_x - 0.5);  // v sits at (0.5, 1)
  q.v1 = sl_axis(offset_y - 1.0);
  q.ox = static_cast<float>(offset_x);
  q.oy = static_cast<float>(offset_y);
  q.hx = p->params.dx;
  q.hy = p->params.dy;
  q.dt = static_cast<float>(dt);
  return fvm_launch_semilagrangian(n, static_cast<const float*>(pc), static_cast<const float*>(pu),
                                   static_cast<const float*>(pv), static_cast<float*>(po), q,
                                   static_cast<hipStream_t>(stream));
}

int jcfd_fvm_spectral_filter(jcfd_fvm_plan* p, const jcfd_tensor* x, const jcfd_tensor* multiplier,
                             jcfd_tensor* out, void* stream) {
  if (p == nullptr) return fail(JCFD_ERR_INVALID_ARGUMENT, "null plan");
  const int n = p->params.n;
  void *px, *pm, *po;
  int rc;
  if ((rc = check_tensor(x, "x", kDLFloat, 32, n, n, p->device, &px)) ||
      (rc = check_tensor(multiplier, "multiplier", kDLFloat, 32, n, n / 2 + 1, p->device, &pm)) ||
      (rc = check_tensor(out, "out", kDLFloat, 32, n, n, p->device, &po)))
    return rc;
  return fvm_launch_filter(p->params, static_cast<const float*>(px), static_cast<const float*>(pm),
                           static_cast<float*>(po), p->spec, p->tw, static_cast<hipStream_t>(stream));
}

}  // extern "C"

// ------------------------------------------------- learned interpolation
struct jcfd_li_plan {
  jcfd_fvm_plan* fvm = nullptr;  // physics, projection and diffusion (forward Euler)
  int n = 0, batch = 1, device = 0;
  int n_layers = 0, hidden = 0, n_logits = 0;
  int cin[JCFD_LI_MAX_LAYERS] = {}, cout[JCFD_LI_MAX_LAYERS] = {};
  float* w[JCFD_LI_MAX_LAYERS] = {};  // device, li_conv_weight_layout of [tap][cin (first layer: 4)][cout]
  float* b[JCFD_LI_MAX_LAYERS] = {};
  float* w_logits = nullptr;  // the tower's own output layer, padded to 128 channels
  float* b_logits = nullptr;
  bool layer_set[JCFD_LI_MAX_LAYERS] = {};
  bool logits_set = false;
  float* act[2] = {nullptr, nullptr};  // (batch, hidden, n, n)
  float* coef = nullptr;               // (batch, 128, n, n)
  float* kbuf = nullptr;               // (2, batch, n, n): k_u, k_v
  RkTerms* rk = nullptr;               // u* = u + dt (0 + 1 k)
  int key[8] = {};
  int64_t bytes = 0;
};

namespace {
// The tower over (u, v) into `out` (the coefficient layer, or the logits
// layer when logits is set).
int li_tower(jcfd_li_plan* p, const float* u, const float* v, float* out, bool logits, hipStream_t s) {
  int rc = JCFD_OK;
  const float* in = u;
  for (int k = 0; k < p->n_layers && rc == JCFD_OK; ++k) {
    const bool last = k + 1 == p->n_layers;
    LiConvArgs a{};
    a.n = p->n;
    a.in = in;
    a.in_v = k == 0 ? v : nullptr;
    a.w = (last && logits) ? p->w_logits : p->w[k];
    a.b = (last && logits) ? p->b_logits : p->b[k];
    a.out = last ? out : p->act[k % 2];
    rc = li_launch_conv(a, p->cin[k], p->cout[k], !last, p->batch, s);
    in = a.out;
  }
  return rc;
}

int li_explicit(jcfd_li_plan* p, const float* u, const float* v, float* ku, float* kv, hipStream_t s) {
  int rc = li_tower(p, u, v, p->coef, false, s);
  if (rc) return rc;
  LiExplicitArgs a{};
  a.coef = p->coef;
  a.u = u;
  a.v = v;
  a.ku = ku;
  a.kv = kv;
  a.fx = p->fvm->params.has_forcing ? p->fvm->fx : nullptr;
  std::memcpy(a.key, p->key, sizeof(a.key));
  return li_launch_explicit(p->fvm->params, a, p->batch, s);
}

// One step: tower, explicit terms, u* = u + dt k with its projection, then
// (implicit-diffusion plans) the diffusion solve.
int li_step_once(jcfd_li_plan* p, const float* u, const float* v, float* u_out, float* v_out,
                 hipStream_t s) {
  const size_t field = static_cast<size_t>(p->n) * p->n * p->batch;
  int rc = li_explicit(p, u, v, p->kbuf, p->kbuf + field, s);
  if (rc) return rc;
  FvmBuffers c = fvm_buffers(p->fvm);
  c.u = u;
  c.v = v;
  c.u_out = u_out;
  c.v_out = v_out;
  c.count = p->batch;
  c.rk = p->rk;
  if ((rc = fvm_launch_rk_project(p->fvm->params, c, s))) return rc;
  return fvm_diffuse_output(p->fvm, c, s);
}

int li_ready(const jcfd_li_plan* p) {
  if (p == nullptr) return fail(JCFD_ERR_INVALID_ARGUMENT, "null plan");
  for (int k = 0; k < p->n_layers; ++k)
    if (!p->layer_set[k]) return fail(JCFD_ERR_INVALID_ARGUMENT, "tower layer " + std::to_string(k) + " not set");
  return JCFD_OK;
}

int li_upload_layer(const float* w, const float* b, int cin, int cin_dev, int cout, int cout_dev,
                    float* wd, float* bd) {
  // (3, 3, cin, cout) host -> [tap][cin_dev][cout_dev], zero padded, then
  // the kernel's operand order
  std::vector<float> wh(static_cast<size_t>(9) * cin_dev * cout_dev, 0.f), bh(cout_dev, 0.f);
  for (int t = 0; t < 9; ++t)
    for (int ci = 0; ci < cin; ++ci)
      for (int co = 0; co < cout; ++co)
        wh[(static_cast<size_t>(t) * cin_dev + ci) * cout_dev + co] =
            w[(static_cast<size_t>(t) * cin + ci) * cout + co];
  for (int co = 0; co < cout; ++co) bh[co] = b[co];
  for (float x : wh)
    if (!std::isfinite(x)) return fail(JCFD_ERR_INVALID_ARGUMENT, "weights must be finite");
  for (float x : bh)
    if (!std::isfinite(x)) return fail(JCFD_ERR_INVALID_ARGUMENT, "biases must be finite");
  std::vector<float> wf(wh.size());
  li_conv_weight_layout(wh.data(), cin_dev, cout_dev, wf.data());
  int rc;
  if ((rc = check_hip(hipMemcpy(wd, wf.data(), wf.size() * sizeof(float), hipMemcpyHostToDevice), "copy weights")))
    return rc;
  return check_hip(hipMemcpy(bd, bh.data(), bh.size() * sizeof(float), hipMemcpyHostToDevice), "copy biases");
}
}  // namespace

extern "C" {

int jcfd_li_plan_destroy(jcfd_li_plan* p) {
  if (p == nullptr) return JCFD_OK;
  for (int k = 0; k < JCFD_LI_MAX_LAYERS; ++k) {
    if (p->w[k]) (void)hipFree(p->w[k]);
    if (p->b[k]) (void)hipFree(p->b[k]);
  }
  for (void* ptr : {static_cast<void*>(p->w_logits), static_cast<void*>(p->b_logits),
                    static_cast<void*>(p->act[0]), static_cast<void*>(p->act[1]),
                    static_cast<void*>(p->coef), static_cast<void*>(p->kbuf), static_cast<void*>(p->rk)})
    if (ptr != nullptr) (void)hipFree(ptr);
  jcfd_fvm_plan_destroy(p->fvm);
  delete p;
  return JCFD_OK;
}

int jcfd_li_plan_create(const jcfd_li_config* config, jcfd_li_plan** plan) {
  if (config == nullptr || plan == nullptr) return fail(JCFD_ERR_INVALID_ARGUMENT, "null argument");
  *plan = nullptr;
  const jcfd_li_config& c = *config;
  const jcfd_fvm_config& f = c.fvm;
  if (f.n < 32) return fail(JCFD_ERR_UNSUPPORTED, "learned interpolation needs n >= 32");
  if (f.pressure != JCFD_PRESSURE_FAST_DIAG)
    return fail(JCFD_ERR_UNSUPPORTED, "learned interpolation is built with solve_fast_diag");
  if (f.diffusion == JCFD_DIFFUSION_IMPLICIT_CG)
    return fail(JCFD_ERR_UNSUPPORTED, "learned interpolation is built with explicit or fast-diagonal diffusion");
  if (f.rk_stages > 1 || (f.rk_stages == 1 && f.rk_b[0] != 1.0))
    return fail(JCFD_ERR_UNSUPPORTED, "learned interpolation is built with forward Euler");
  if (c.n_layers < 1 || c.n_layers > JCFD_LI_MAX_LAYERS)
    return fail(JCFD_ERR_UNSUPPORTED, "towers of 1 to " + std::to_string(JCFD_LI_MAX_LAYERS) +
                                          " convolutions are built");
  if (c.n_layers > 1 && c.hidden_channels != 16 && c.hidden_channels != 32 && c.hidden_channels != 64)
    return fail(JCFD_ERR_UNSUPPORTED, "hidden channels must be 16, 32 or 64");
  if (c.n_logits < 1 || c.n_logits > kLiCoefChannels)
    return fail(JCFD_ERR_INVALID_ARGUMENT, "the tower's output channels must be 1 .. 128");
  for (int k = 0; k < 8; ++k)
    if (c.interp_key[k] < 0 || c.interp_key[k] > kLiCoefChannels - 16 || c.interp_key[k] % 16 != 0)
      return fail(JCFD_ERR_INVALID_ARGUMENT, "interpolation keys must be multiples of 16 below 128");
  jcfd_fvm_config fc = f;
  fc.convect = JCFD_CONVECT_LINEAR;  // the convection is the learned one; the plan's kernels are unused
  fc.rk_stages = 0;
  auto* p = new (std::nothrow) jcfd_li_plan();
  if (p == nullptr) return fail(JCFD_ERR_OUT_OF_MEMORY, "host allocation failed");
  int rc = jcfd_fvm_plan_create(&fc, &p->fvm);
  if (rc) {
    delete p;
    return rc;
  }
  p->n = f.n;
  p->batch = p->fvm->batch;
  p->device = p->fvm->device;
  p->n_layers = c.n_layers;
  p->hidden = c.hidden_channels;
  p->n_logits = c.n_logits;
  std::memcpy(p->key, c.interp_key, sizeof(p->key));
  for (int k = 0; k < c.n_layers; ++k) {
    p->cin[k] = k == 0 ? 2 : c.hidden_channels;
    p->cout[k] = k + 1 == c.n_layers ? kLiCoefChannels : c.hidden_channels;
  }
  const size_t nn = static_cast<size_t>(p->n) * p->n * p->batch;
  for (int k = 0; k < c.n_layers && rc == JCFD_OK; ++k) {
    const int cin_dev = k == 0 ? 4 : p->cin[k];
    if ((rc = device_alloc(&p->w[k], static_cast<size_t>(9) * cin_dev * p->cout[k], &p->bytes))) break;
    rc = device_alloc(&p->b[k], static_cast<size_t>(p->cout[k]), &p->bytes);
  }
  const int cin_last = c.n_layers == 1 ? 4 : c.hidden_channels;
  if (rc == JCFD_OK &&
      ((rc = device_alloc(&p->w_logits, static_cast<size_t>(9) * cin_last * kLiCoefChannels, &p->bytes)) ||
       (rc = device_alloc(&p->b_logits, static_cast<size_t>(kLiCoefChannels), &p->bytes)) ||
       (c.n_layers > 1 && (rc = device_alloc(&p->act[0], nn * c.hidden_channels, &p->bytes))) ||
       (c.n_layers > 2 && (rc = device_alloc(&p->act[1], nn * c.hidden_channels, &p->bytes))) ||
       (rc = device_alloc(&p->coef, nn * kLiCoefChannels, &p->bytes)) ||
       (rc = device_alloc(&p->kbuf, 2 * nn, &p->bytes)) ||
       (rc = device_alloc(&p->rk, 1, &p->bytes)))) {
  }
  if (rc == JCFD_OK) {
    RkTerms t{};
    t.n = 1;
    t.c[0] = 1.0f;
    t.ku[0] = p->kbuf;
    t.kv[0] = p->kbuf + nn;
    rc = check_hip(hipMemcpy(p->rk, &t, sizeof(t), hipMemcpyHostToDevice), "copy combination");
  }
  if (rc) {
    jcfd_li_plan_destroy(p);
    return rc;
  }
  *plan = p;
  return JCFD_OK;
}

int64_t jcfd_li_plan_device_bytes(const jcfd_li_plan* p) {
  return p ? p->bytes + p->fvm->bytes : 0;
}

int jcfd_li_set_layer(jcfd_li_plan* p, int32_t layer, const float* w, const float* b) {
  if (p == nullptr || w == nullptr || b == nullptr) return fail(JCFD_ERR_INVALID_ARGUMENT, "null argument");
  if (layer < 0 || layer >= p->n_layers) return fail(JCFD_ERR_INVALID_ARGUMENT, "layer out of range");
  int rc = li_upload_layer(w, b, p->cin[layer], layer == 0 ? 4 : p->cin[layer], p->cout[layer],
                           p->cout[layer], p->w[layer], p->b[layer]);
  if (rc == JCFD_OK) p->layer_set[layer] = true;
  return rc;
}

int jcfd_li_set_logits_layer(jcfd_li_plan* p, const float* w, const float* b) {
  if (p == nullptr || w == nullptr || b == nullptr) return fail(JCFD_ERR_INVALID_ARGUMENT, "null argument");
  const int last = p->n_layers - 1;
  int rc = li_upload_layer(w, b, p->cin[last], last == 0 ? 4 : p->cin[last], p->n_logits,
                           kLiCoefChannels, p->w_logits, p->b_logits);
  if (rc == JCFD_OK) p->logits_set = true;
  return rc;
}

int jcfd_li_tower(jcfd_li_plan* p, const jcfd_tensor* u, const jcfd_tensor* v, jcfd_tensor* out,
                  int32_t coefficients, void* stream) {
  int rc = li_ready(p);
  if (rc) return rc;
  if (!coefficients && !p->logits_set) return fail(JCFD_ERR_INVALID_ARGUMENT, "logits layer not set");
  const int n = p->n;
  void *pu, *pv, *po;
  if ((rc = check_batched(u, "u", kDLFloat, 32, p->batch, n, n, p->device, &pu)) ||
      (rc = check_batched(v, "v", kDLFloat, 32, p->batch, n, n, p->device, &pv)) ||
      (rc = check_batched(out, "out", kDLFloat, 32, p->batch * kLiCoefChannels, n, n, p->device, &po)))
    return rc;
  return li_tower(p, static_cast<const float*>(pu), static_cast<const float*>(pv), static_cast<float*>(po),
                  !coefficients, static_cast<hipStream_t>(stream));
}

int jcfd_li_explicit_terms(jcfd_li_plan* p, const jcfd_tensor* u, const jcfd_tensor* v, jcfd_tensor* du,
                           jcfd_tensor* dv, void* stream) {
  int rc = li_ready(p);
  if (rc) return rc;
  const int n = p->n;
  void *pu, *pv, *pdu, *pdv;
  if ((rc = check_batched(u, "u", kDLFloat, 32, p->batch, n, n, p->device, &pu)) ||
      (rc = check_batched(v, "v", kDLFloat, 32, p->batch, n, n, p->device, &pv)) ||
      (rc = check_batched(du, "du", kDLFloat, 32, p->batch, n, n, p->device, &pdu)) ||
      (rc = check_batched(dv, "dv", kDLFloat, 32, p->batch, n, n, p->device, &pdv)))
    return rc;
  if (pdu == pu || pdu == pv || pdv == pu || pdv == pv)
    return fail(JCFD_ERR_INVALID_ARGUMENT, "explicit terms: outputs must not alias inputs");
  return li_explicit(p, static_cast<const float*>(pu), static_cast<const float*>(pv), static_cast<float*>(pdu),
                     static_cast<float*>(pdv), static_cast<hipStream_t>(stream));
}

int jcfd_li_interpolate(jcfd_li_plan* p, const jcfd_tensor* coefficients, const jcfd_tensor* x,
                        int32_t interpolation, jcfd_tensor* out, void* stream) {
  if (p == nullptr) return fail(JCFD_ERR_INVALID_ARGUMENT, "null plan");
  if (interpolation < 0 || interpolation >= 8)
    return fail(JCFD_ERR_INVALID_ARGUMENT, "interpolation index must be 0 .. 7");
  const int n = p->n;
  void *pc, *px, *po;
  int rc;
  if ((rc = check_batched(coefficients, "coefficients", kDLFloat, 32, p->batch * kLiCoefChannels, n, n,
                          p->device, &pc)) ||
      (rc = check_batched(x, "x", kDLFloat, 32, p->batch, n, n, p->device, &px)) ||
      (rc = check_batched(out, "out", kDLFloat, 32, p->batch, n, n, p->device, &po)))
    return rc;
  if (po == px) return fail(JCFD_ERR_INVALID_ARGUMENT, "interpolate: out must not alias x");
  return li_launch_interpolate(n, static_cast<const float*>(pc), static_cast<const float*>(px),
                               16 * interpolation, static_cast<float*>(po), p->batch,
                               static_cast<hipStream_t>(stream));
}

int jcfd_li_step(jcfd_li_plan* p, const jcfd_tensor* u, const jcfd_tensor* v, jcfd_tensor* u_out,
                 jcfd_tensor* v_out, int64_t nsteps, void* stream) {
  int rc = li_ready(p);
  if (rc) return rc;
  if (nsteps < 0) return fail(JCFD_ERR_INVALID_ARGUMENT, "nsteps must be >= 0");
  const int n = p->n;
  void *pu, *pv, *pu2, *pv2;
  if ((rc = check_batched(u, "u", kDLFloat, 32, p->batch, n, n, p->device, &pu)) ||
      (rc = check_batched(v, "v", kDLFloat, 32, p->batch, n, n, p->device, &pv)) ||
      (rc = check_batched(u_out, "u_out", kDLFloat, 32, p->batch, n, n, p->device, &pu2)) ||
      (rc = check_batched(v_out, "v_out", kDLFloat, 32, p->batch, n, n, p->device, &pv2)))
    return rc;
  hipStream_t s = static_cast<hipStream_t>(stream);
  const size_t bytes = sizeof(float) * p->batch * static_cast<size_t>(n) * n;
  if (nsteps == 0) {
    if (pu2 != pu && (rc = check_hip(hipMemcpyAsync(pu2, pu, bytes, hipMemcpyDeviceToDevice, s), "copy")))
      return rc;
    if (pv2 != pv && (rc = check_hip(hipMemcpyAsync(pv2, pv, bytes, hipMemcpyDeviceToDevice, s), "copy")))
      return rc;
    return JCFD_OK;
  }
  const float* cu = static_cast<const float*>(pu);
  const float* cv = static_cast<const float*>(pv);
  for (int64_t k = 0; k < nsteps; ++k) {
    if ((rc = li_step_once(p, cu, cv, static_cast<float*>(pu2), static_cast<float*>(pv2), s))) return rc;
    cu = static_cast<const float*>(pu2);
    cv = static_cast<const float*>(pv2);
  }
  return JCFD_OK;
}

int jcfd_li_time_kernels(jcfd_li_plan* p, const jcfd_tensor* u, const jcfd_tensor* v, int32_t reps,
                         double* ms_per_launch, void* stream) {
  int rc = li_ready(p);
  if (rc) return rc;
  if (ms_per_launch == nullptr || reps <= 0) return fail(JCFD_ERR_INVALID_ARGUMENT, "bad arguments");
  const int n = p->n;
  void *pu, *pv;
  if ((rc = check_batched(u, "u", kDLFloat, 32, p->batch, n, n, p->device, &pu)) ||
      (rc = check_batched(v, "v", kDLFloat, 32, p->batch, n, n, p->device, &pv)))
    return rc;
  hipStream_t s = static_cast<hipStream_t>(stream);
  const float* cu = static_cast<const float*>(pu);
  const float* cv = static_cast<const float*>(pv);
  const size_t field = static_cast<size_t>(n) * n * p->batch;
  hipEvent_t e0, e1;
  if ((rc = check_hip(hipEventCreate(&e0), "event"))) return rc;
  if ((rc = check_hip(hipEventCreate(&e1), "event"))) {
    (void)hipEventDestroy(e0);
    return rc;
  }
  // entries: each convolution, li_explicit, then the combination u + dt k
  // with its projection (rows-forward, Poisson columns, rows-project)
  rc = li_explicit(p, cu, cv, p->kbuf, p->kbuf + field, s);  // every buffer valid
  FvmBuffers c = fvm_buffers(p->fvm);
  c.u = cu;
  c.v = cv;
  c.u_out = p->fvm->us;  // rows-project reads and writes the same element per thread
  c.v_out = p->fvm->vs;
  c.count = p->batch;
  c.rk = p->rk;
  const int entries = p->n_layers + 2;
  for (int e = 0; e < entries && rc == JCFD_OK; ++e) {
    rc = check_hip(hipEventRecord(e0, s), "event record");
    for (int r = 0; r < reps && rc == JCFD_OK; ++r) {
      if (e < p->n_layers) {
        const bool last = e + 1 == p->n_layers;
        LiConvArgs a{};
        a.n = n;
        a.in = e == 0 ? cu : p->act[(e - 1) % 2];
        a.in_v = e == 0 ? cv : nullptr;
        a.w = p->w[e];
        a.b = p->b[e];
        a.out = last ? p->coef : p->act[e % 2];
        rc = li_launch_conv(a, p->cin[e], p->cout[e], !last, p->batch, s);
      } else if (e == p->n_layers) {
        LiExplicitArgs a{};
        a.coef = p->coef;
        a.u = cu;
        a.v = cv;
        a.ku = p->kbuf;
        a.kv = p->kbuf + field;
        a.fx = p->fvm->params.has_forcing ? p->fvm->fx : nullptr;
        std::memcpy(a.key, p->key, sizeof(a.key));
        rc = li_launch_explicit(p->fvm->params, a, p->batch, s);
      } else {
        rc = fvm_launch_rk_project(p->fvm->params, c, s);
      }
    }
    if (rc == JCFD_OK) rc = check_hip(hipEventRecord(e1, s), "event record");
    if (rc == JCFD_OK) rc = check_hip(hipEventSynchronize(e1), "event sync");
    float ms = 0.f;
    if (rc == JCFD_OK) rc = check_hip(hipEventElapsedTime(&ms, e0, e1), "elapsed");
    ms_per_launch[e] = static_cast<double>(ms) / reps;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return rc;
}

int jcfd_li_kernel_flops(const jcfd_li_plan* p, double* flops) {
  if (p == nullptr || flops == nullptr) return fail(JCFD_ERR_INVALID_ARGUMENT, "null argument");
  const double px = static_cast<double>(p->n) * p->n * p->batch;
  for (int k = 0; k < p->n_layers; ++k) {
    const int co = k + 1 == p->n_layers ? p->n_logits : p->cout[k];
    flops[k] = 2.0 * 9.0 * p->cin[k] * co * px;  // the tower's own multiply-adds
  }
  return JCFD_OK;
}

}  // extern "C"

// ------------------------------------------------------ mixed-size groups
struct jcfd_fvm_group {
  int device = 0;
  FvmGroupDesc* descs = nullptr;   // device descriptor table
  FvmGroupItem* items = nullptr;   // device work lists, phase k at items + k * cap_items
  int* d_counts = nullptr;         // device int[3]: items per phase
  int cap_descs = 0, cap_items = 0;
  char* staging = nullptr;         // pinned upload buffer
  size_t staging_bytes = 0;
  hipEvent_t staged = nullptr;     // recorded after the last upload
  bool staged_pending = false;
  std::vector<char> last;          // the table last uploaded
  int max_n = 0;
  int grid[3] = {0, 0, 0};
  int n_items[3] = {0, 0, 0};   // of the last step
  int64_t bytes[3] = {0, 0, 0};  // algorithmic bytes per phase launch, last step
  GraphCache graphs;
};

namespace {

void group_free_tables(jcfd_fvm_group* g) {
  for (void* ptr : {static_cast<void*>(g->descs), static_cast<void*>(g->items)})
    if (ptr != nullptr) (void)hipFree(ptr);
  g->descs = nullptr;
  g->items = nullptr;
  g->cap_descs = g->cap_items = 0;
}

// First sizes of a group's device work tables and pinned upload buffer:
// room for every slot of the default serving pools (960 instances, up to
// ~41k work items per list), so a serving loop never regrows them -- a
// regrow frees and reallocates device memory after a device-wide
// synchronisation, which stalls every stream of the process.
constexpr int kGroupMinDescs = 1024;
constexpr int kGroupMinItems = 65536;
constexpr size_t kGroupMinStaging = size_t(4) << 20;

// Grow the device tables to hold n_desc descriptors and n_item items per
// phase.  Tables in use by queued work are freed only after a device-wide
// synchronisation; the graph is keyed on their addresses and so recaptured.
int group_reserve(jcfd_fvm_group* g, int n_desc, int n_item) {
  if (n_desc <= g->cap_descs && n_item <= g->cap_items) return JCFD_OK;
  int rc;
  if (g->descs != nullptr && (rc = check_hip(hipDeviceSynchronize(), "group regrow sync"))) return rc;
  const int cd = std::max({n_desc, 2 * g->cap_descs, kGroupMinDescs});
  const int ci = std::max({n_item, 2 * g->cap_items, kGroupMinItems});
  group_free_tables(g);
  int64_t unused = 0;
  if ((rc = device_alloc(&g->descs, static_cast<size_t>(cd), &unused)) ||
      (rc = device_alloc(&g->items, static_cast<size_t>(3) * ci, &unused))) {
    group_free_tables(g);
    return rc;
  }
  g->cap_descs = cd;
  g->cap_items = ci;
  return JCFD_OK;
}

// Copy `table` (n_items[4] | descs | items phase 0 | 1 | 2) to the device
// through the pinned staging buffer, stream-ordered on s.
int group_upload(jcfd_fvm_group* g, const std::vector<char>& table, int n_desc, const int* counts,
                 hipStream_t s) {
  int rc;
  if (g->staged_pending && (rc = check_hip(hipEventSynchronize(g->staged), "staging reuse")))
    return rc;
  g->staged_pending = false;
  if (table.size() > g->staging_bytes) {
    if (g->staging != nullptr) (void)hipHostFree(g->staging);
    g->staging = nullptr;
    const size_t bytes = std::max({table.size(), 2 * g->staging_bytes, kGroupMinStaging});
    if ((rc = check_hip(hipHostMalloc(reinterpret_cast<void**>(&g->staging), bytes, hipHostMallocDefault),
                        "pinned group staging")))
      return rc;
    g->staging_bytes = bytes;
  }
  std::memcpy(g->staging, table.data(), table.size());
  size_t off = 4 * sizeof(int);
  if ((rc = check_hip(hipMemcpyAsync(g->d_counts, g->staging, 3 * sizeof(int), hipMemcpyHostToDevice, s),
                      "group upload")) ||
      (rc = check_hip(hipMemcpyAsync(g->descs, g->staging + off, n_desc * sizeof(FvmGroupDesc),
                                     hipMemcpyHostToDevice, s),
                      "group upload")))
    return rc;
  off += n_desc * sizeof(FvmGroupDesc);
  for (int k = 0; k < 3; ++k) {
    const size_t bytes = counts[k] * sizeof(FvmGroupItem);
    if (bytes > 0 &&
        (rc = check_hip(hipMemcpyAsync(g->items + static_cast<size_t>(k) * g->cap_items, g->staging + off,
                                       bytes, hipMemcpyHostToDevice, s),
                        "group upload")))
      return rc;
    off += bytes;
  }
  if ((rc = check_hip(hipEventRecord(g->staged, s), "staging event"))) return rc;
  g->staged_pending = true;
  return JCFD_OK;
}

FvmGroupQueue group_queue(const jcfd_fvm_group* g) {
  FvmGroupQueue q{};
  q.descs = g->descs;
  for (int k = 0; k < 3; ++k) q.items[k] = g->items + static_cast<size_t>(k) * g->cap_items;
  q.n_items = g->d_counts;
  return q;
}

}  // namespace

extern "C" {

int jcfd_fvm_group_create(jcfd_fvm_group** group) {
  if (group == nullptr) return fail(JCFD_ERR_INVALID_ARGUMENT, "null argument");
  *group = nullptr;
  auto* g = new (std::nothrow) jcfd_fvm_group();
  if (g == nullptr) return fail(JCFD_ERR_OUT_OF_MEMORY, "host allocation failed");
  int64_t unused = 0;
  int rc;
  if ((rc = check_hip(hipGetDevice(&g->device), "hipGetDevice")) ||
      (rc = device_alloc(&g->d_counts, 4, &unused)) ||
      (rc = check_hip(hipMemset(g->d_counts, 0, 4 * sizeof(int)), "zero counts")) ||
      (rc = check_hip(hipEventCreateWithFlags(&g->staged, hipEventDisableTiming), "staging event"))) {
    jcfd_fvm_group_destroy(g);
    return rc;
  }
  *group = g;
  return JCFD_OK;
}

int jcfd_fvm_group_destroy(jcfd_fvm_group* g) {
  if (g == nullptr) return JCFD_OK;
  if (g->staged_pending) (void)hipEventSynchronize(g->staged);
  g->graphs.release();
  group_free_tables(g);
  if (g->d_counts != nullptr) (void)hipFree(g->d_counts);
  if (g->staging != nullptr) (void)hipHostFree(g->staging);
  if (g->staged != nullptr) (void)hipEventDestroy(g->staged);
  delete g;
  return JCFD_OK;
}

int jcfd_fvm_group_set_graphs(jcfd_fvm_group* g, int32_t enabled) {
  if (g == nullptr) return fail(JCFD_ERR_INVALID_ARGUMENT, "null group");
  g->graphs.enabled = enabled != 0;
  return JCFD_OK;
}

int jcfd_fvm_group_step(jcfd_fvm_group* g, const jcfd_fvm_group_member* members, int32_t n_members,
                        int64_t nsteps, void* stream) {
  if (g == nullptr) return fail(JCFD_ERR_INVALID_ARGUMENT, "null group");
  if (n_members < 0 || (n_members > 0 && members == nullptr))
    return fail(JCFD_ERR_INVALID_ARGUMENT, "members: null array or negative count");
  if (nsteps < 0) return fail(JCFD_ERR_INVALID_ARGUMENT, "nsteps must be >= 0");
  // ---- validate and build the descriptor table
  std::vector<FvmGroupDesc> descs;
  std::vector<const void*> owners;   // (plan, slot) scratch and (field, slot) state must be unique
  int rc;
  for (int m = 0; m < n_members; ++m) {
    const jcfd_fvm_group_member& mb = members[m];
    jcfd_fvm_plan* p = mb.plan;
    const std::string who = "member " + std::to_string(m);
    if (p == nullptr) return fail(JCFD_ERR_INVALID_ARGUMENT, who + ": null plan");
    if (p->device != g->device)
      return fail(JCFD_ERR_INVALID_ARGUMENT, who + ": plan on another device than the group");
    if (p->cfg.pressure != JCFD_PRESSURE_FAST_DIAG)
      return fail(JCFD_ERR_UNSUPPORTED, who + ": groups step fast-diagonal plans only");
    if (p->params.radix != 8)
      return fail(JCFD_ERR_UNSUPPORTED, who + ": groups run the radix-8 FFT schedule only");
    // implicit-diffusion plans carry has_diffusion = 0 in their params (the
    // solve runs after the projection), so the group would silently step
    // them inviscid: reject every diffusion mode but the explicit one
    if (p->cfg.diffusion != JCFD_DIFFUSION_EXPLICIT || p->dlam != nullptr || p->rk_terms != nullptr)
      return fail(JCFD_ERR_UNSUPPORTED,
                  who + ": groups step forward-Euler explicit-diffusion plans only");
    const int n = p->params.n;
    if (!fvm_group_size_supported(n))
      return fail(JCFD_ERR_UNSUPPORTED, who + ": grid size " + std::to_string(n) +
                                            " not built for groups (32..512)");
    void *pu, *pv;
    if ((rc = check_batched(mb.u, "u", kDLFloat, 32, p->batch, n, n, p->device, &pu)) ||
        (rc = check_batched(mb.v, "v", kDLFloat, 32, p->batch, n, n, p->device, &pv)))
      return fail(rc, who + ": " + jcfd_last_error());
    if (mb.n_slots < 0 || mb.n_slots > p->batch || (mb.n_slots > 0 && mb.slots == nullptr))
      return fail(JCFD_ERR_INVALID_ARGUMENT, who + ": slots must hold 0..batch host int32 values");
    const size_t nn = static_cast<size_t>(n) * n;
    for (int k = 0; k < mb.n_slots; ++k) {
      const int s = mb.slots[k];
      if (s < 0 || s >= p->batch)
        return fail(JCFD_ERR_INVALID_ARGUMENT, who + ": slot " + std::to_string(s) + " out of range");
      FvmGroupDesc d{};
      d.p = p->params;
      d.b = fvm_buffers(p);
      d.b.us += s * nn;
      d.b.vs += s * nn;
      d.b.spec += s * nn / 2;
      float* u = static_cast<float*>(pu) + s * nn;
      float* v = static_cast<float*>(pv) + s * nn;
      d.b.u = u;
      d.b.v = v;
      d.b.u_out = u;
      d.b.v_out = v;
      d.b.count = 1;
      descs.push_back(d);
      owners.push_back(d.b.us);
      owners.push_back(u);
      owners.push_back(v);
    }
  }
  std::vector<const void*> sorted_owners = owners;
  std::sort(sorted_owners.begin(), sorted_owners.end());
  if (std::adjacent_find(sorted_owners.begin(), sorted_owners.end()) != sorted_owners.end())
    return fail(JCFD_ERR_INVALID_ARGUMENT,
                "group members repeat an instance (same plan slot or same field slot twice)");
  if (nsteps == 0 || descs.empty()) return JCFD_OK;
  // largest grids first: their items are the longest, so the tail stays short
  std::stable_sort(descs.begin(), descs.end(),
                   [](const FvmGroupDesc& a, const FvmGroupDesc& b) { return a.p.n > b.p.n; });
  const int n_desc = static_cast<int>(descs.size());
  const int max_n = descs.front().p.n;
  std::vector<FvmGroupItem> items[3];
  for (int k = 0; k < 3; ++k)
    for (int d = 0; d < n_desc; ++d)
      for (int t = 0, nt = fvm_group_tiles(descs[d].p.n, k); t < nt; ++t) items[k].push_back({d, t});
  const int counts[4] = {static_cast<int>(items[0].size()), static_cast<int>(items[1].size()),
                         static_cast<int>(items[2].size()), 0};
  std::vector<char> table(4 * sizeof(int) + n_desc * sizeof(FvmGroupDesc) +
                          (counts[0] + counts[1] + counts[2]) * sizeof(FvmGroupItem));
  {
    char* w = table.data();
    std::memcpy(w, counts, 4 * sizeof(int));
    w += 4 * sizeof(int);
    std::memcpy(w, descs.data(), n_desc * sizeof(FvmGroupDesc));
    w += n_desc * sizeof(FvmGroupDesc);
    for (int k = 0; k < 3; ++k) {
      std::memcpy(w, items[k].data(), counts[k] * sizeof(FvmGroupItem));
      w += counts[k] * sizeof(FvmGroupItem);
    }
  }
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int need_items = std::max({counts[0], counts[1], counts[2]});
  if (n_desc > g->cap_descs || need_items > g->cap_items) {
    if ((rc = group_reserve(g, n_desc, need_items))) return rc;
    g->last.clear();
  }
  if (table != g->last) {
    if ((rc = group_upload(g, table, n_desc, counts, s))) return rc;
    g->last = table;
  }
  g->max_n = max_n;
  for (int k = 0; k < 3; ++k) g->grid[k] = counts[k];  // one workgroup per item
  for (int k = 0; k < 3; ++k) g->n_items[k] = counts[k];
  g->bytes[0] = g->bytes[1] = g->bytes[2] = 0;
  for (const FvmGroupDesc& d : descs) {
    const int64_t cells = static_cast<int64_t>(d.p.n) * d.p.n;
    g->bytes[0] += 20 * cells;
    g->bytes[1] += 8 * cells;
    g->bytes[2] += 20 * cells;
  }
  const FvmGroupQueue q = group_queue(g);
  auto one_step = [&](hipStream_t st) {
    for (int k = 0; k < 3; ++k) {
      const int r = fvm_group_launch(k, max_n, q, g->grid[k], st);
      if (r) return r;
    }
    return static_cast<int>(JCFD_OK);
  };
  // captured launches bake in the grids (the item counts) and LDS size
  const void* count_key = reinterpret_cast<const void*>(
      (static_cast<uintptr_t>(static_cast<uint32_t>(counts[0])) << 42) |
      (static_cast<uintptr_t>(static_cast<uint32_t>(counts[1])) << 21) | static_cast<uint32_t>(counts[2]));
  return run_steps(g->graphs, g->descs, g->items, count_key, max_n, nsteps, s, one_step);
}

int jcfd_fvm_group_time_kernels(jcfd_fvm_group* g, int32_t reps, double* ms_per_launch, void* stream) {
  if (g == nullptr || ms_per_launch == nullptr) return fail(JCFD_ERR_INVALID_ARGUMENT, "null argument");
  if (reps <= 0) return fail(JCFD_ERR_INVALID_ARGUMENT, "reps must be positive");
  if (g->max_n == 0 || g->n_items[0] == 0)
    return fail(JCFD_ERR_INVALID_ARGUMENT, "no group step to time (call jcfd_fvm_group_step first)");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const FvmGroupQueue q = group_queue(g);
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  int rc = JCFD_OK;
  for (int k = 0; k < 4 && rc == JCFD_OK; ++k) rc = check_hip(hipEventCreate(&ev[k]), "event");
  double total[3] = {0, 0, 0};
  for (int r = 0; r < reps && rc == JCFD_OK; ++r) {
    rc = check_hip(hipEventRecord(ev[0], s), "event record");
    for (int k = 0; k < 3 && rc == JCFD_OK; ++k) {
      rc = fvm_group_launch(k, g->max_n, q, g->grid[k], s);
      if (rc == JCFD_OK) rc = check_hip(hipEventRecord(ev[k + 1], s), "event record");
    }
    if (rc == JCFD_OK) rc = check_hip(hipEventSynchronize(ev[3]), "event sync");
    for (int k = 0; k < 3 && rc == JCFD_OK; ++k) {
      float ms = 0.f;
      rc = check_hip(hipEventElapsedTime(&ms, ev[k], ev[k + 1]), "elapsed");
      total[k] += ms;
    }
  }
  for (int k = 0; k < 3; ++k) ms_per_launch[k] = total[k] / reps;
  for (hipEvent_t e : ev)
    if (e != nullptr) (void)hipEventDestroy(e);
  return rc;
}

int jcfd_fvm_group_launch_shape(const jcfd_fvm_group* g, int32_t* grid, int32_t* items,
                                int64_t* bytes_per_launch) {
  if (g == nullptr) return fail(JCFD_ERR_INVALID_ARGUMENT, "null group");
  for (int k = 0; k < 3; ++k) {
    if (grid) grid[k] = g->grid[k];
    if (items) items[k] = g->n_items[k];
    if (bytes_per_launch) bytes_per_launch[k] = g->bytes[k];
  }
  return JCFD_OK;
}

}  // extern "C"

// ------------------------------------------------------------ spectral path
namespace {
// nsteps CN-RK4 steps of `count` instances (slots, or 0..count-1) from
// w_in to w_out (external layout; may alias): the state is moved into the
// plan's internal-layout buffer, stepped in place there (graph replay when
// enabled) and moved back.
int ns_run(jcfd_ns2d_plan* p, const float2* w_in, float2* w_out, const int* slots, int count,
           int64_t nsteps, hipStream_t s) {
  NsParams stages[JCFD_MAX_RK_STAGES];
  for (int k = 0; k < p->n_stages; ++k) {
    stages[k] = p->stages[k];
    stages[k].slots = slots;
  }
  const int ns = p->n_stages;
  NsBuffers b{};
  b.u_in = p->ws;
  b.u_out = p->ws;
  b.h = p->h;
  b.g = p->g;
  b.pbuf = p->pbuf;
  int rc;
  if ((rc = ns_launch_state_convert(p->n, true, w_in, p->ws, slots, count, s))) return rc;
  if ((rc = ns_launch_step(p->n, stages, ns, b, count, s))) return rc;
  if ((rc = run_steps(p->graphs, b.u_out, nullptr, slots, count, nsteps - 1, s,
                      [&](hipStream_t gs) { return ns_launch_step(p->n, stages, ns, b, count, gs); })))
    return rc;
  return ns_launch_state_convert(p->n, false, p->ws, w_out, slots, count, s);
}
}  // namespace

extern "C" {

int jcfd_ns2d_plan_destroy(jcfd_ns2d_plan* p) {
  if (p == nullptr) return JCFD_OK;
  for (void* ptr : {static_cast<void*>(p->kx), static_cast<void*>(p->ky), static_cast<void*>(p->tw),
                    static_cast<void*>(p->force_idx), static_cast<void*>(p->force_val),
                    static_cast<void*>(p->h), static_cast<void*>(p->g), static_cast<void*>(p->pbuf),
                    static_cast<void*>(p->ws)})
    if (ptr != nullptr) (void)hipFree(ptr);
  p->graphs.release();
  delete p;
  return JCFD_OK;
}

int jcfd_ns2d_plan_set_graphs(jcfd_ns2d_plan* p, int32_t enabled) {
  if (p == nullptr) return fail(JCFD_ERR_INVALID_ARGUMENT, "null plan");
  p->graphs.enabled = enabled != 0;
  return JCFD_OK;
}

int jcfd_ns2d_plan_create(const jcfd_ns2d_config* config, jcfd_ns2d_plan** plan) {
  if (config == nullptr || plan == nullptr) return fail(JCFD_ERR_INVALID_ARGUMENT, "null argument");
  *plan = nullptr;
  const jcfd_ns2d_config& c = *config;
  if (!ns_size_supported(c.n))
    return fail(JCFD_ERR_UNSUPPORTED,
                "spectral grid size " + std::to_string(c.n) + " not built (32..1024, powers of two)");
  if (!(c.dx > 0) || !(c.dy > 0) || !std::isfinite(c.dx) || !std::isfinite(c.dy))
    return fail(JCFD_ERR_INVALID_ARGUMENT, "grid steps must be positive and finite");
  if (!std::isfinite(c.dt) || !std::isfinite(c.viscosity) || !std::isfinite(c.drag))
    return fail(JCFD_ERR_INVALID_ARGUMENT, "dt, viscosity and drag must be finite");
  if (c.batch < 1) return fail(JCFD_ERR_INVALID_ARGUMENT, "batch must be >= 1");
  if (c.n_forcing < 0 || (c.n_forcing > 0 && (c.forcing_index == nullptr || c.forcing_value == nullptr)))
    return fail(JCFD_ERR_INVALID_ARGUMENT, "forcing entries missing");
  if (c.rk_stages < 0 || c.rk_stages > JCFD_MAX_RK_STAGES)
    return fail(JCFD_ERR_UNSUPPORTED, "low-storage Runge-Kutta tableaux of 1 to " +
                                          std::to_string(JCFD_MAX_RK_STAGES) + " stages are built");
  for (int k = 0; k < c.rk_stages; ++k)
    if (!std::isfinite(c.rk_alpha[k]) || !std::isfinite(c.rk_alpha[k + 1]) || !std::isfinite(c.rk_beta[k]) ||
        !std::isfinite(c.rk_gamma[k]))
      return fail(JCFD_ERR_INVALID_ARGUMENT, "Runge-Kutta coefficients must be finite");
  const int n = c.n, h1 = n / 2 + 1;
  for (int t = 0; t < c.n_forcing; ++t) {
    const int fp = c.forcing_index[2 * t], fc = c.forcing_index[2 * t + 1];
    if (fp < 0 || fp >= n || fc < 0 || fc >= h1)
      return fail(JCFD_ERR_INVALID_ARGUMENT, "forcing index out of range");
  }
  auto* p = new (std::nothrow) jcfd_ns2d_plan();
  if (p == nullptr) return fail(JCFD_ERR_OUT_OF_MEMORY, "host allocation failed");
  p->cfg = c;
  p->cfg.forcing_index = nullptr;
  p->cfg.forcing_value = nullptr;
  p->n = n;
  p->batch = c.batch;
  p->graphs.enabled = c.batch > 1;  // replay pays off only once launches are long
  if (hipGetDevice(&p->device) != hipSuccess) {
    delete p;
    return fail(JCFD_ERR_DEVICE, "hipGetDevice failed (no GPU?)");
  }
  // fftfreq / rfftfreq in fp32: ordinal / fp32(d * n) (grids.py:604-658)
  std::vector<float> kx(n), ky(h1);
  const float dnx = static_cast<float>(c.dx * n), dny = static_cast<float>(c.dy * n);
  for (int i = 0; i < n; ++i) kx[i] = static_cast<float>(i < (n + 1) / 2 ? i : i - n) / dnx;
  for (int i = 0; i < h1; ++i) ky[i] = static_cast<float>(i) / dny;
  const std::vector<float2> tw = twiddles(n);
  const size_t ext = static_cast<size_t>(n) * h1, half = static_cast<size_t>(n) * (n / 2);
  const size_t B = static_cast<size_t>(c.batch);
  const int nf = std::max(c.n_forcing, 1);
  int rc = JCFD_OK;
  if ((rc = device_alloc(&p->kx, n, &p->bytes)) || (rc = device_alloc(&p->ky, h1, &p->bytes)) ||
      (rc = device_alloc(&p->tw, n, &p->bytes)) || (rc = device_alloc(&p->force_idx, 2 * nf, &p->bytes)) ||
      (rc = device_alloc(&p->force_val, nf, &p->bytes)) || (rc = device_alloc(&p->h, B * ext, &p->bytes)) ||
      (rc = device_alloc(&p->g, B * 4 * half, &p->bytes)) || (rc = device_alloc(&p->pbuf, B * half, &p->bytes)) ||
      (rc = device_alloc(&p->ws, B * ext, &p->bytes))) {
    jcfd_ns2d_plan_destroy(p);
    return rc;
  }
  std::vector<int> fidx(2 * nf, 0);
  std::vector<float2> fval(nf, make_float2(0.f, 0.f));
  for (int t = 0; t < c.n_forcing; ++t) {
    fidx[2 * t] = c.forcing_index[2 * t];
    fidx[2 * t + 1] = c.forcing_index[2 * t + 1];
    fval[t] = make_float2(c.forcing_value[2 * t], c.forcing_value[2 * t + 1]);
  }
  if ((rc = check_hip(hipMemcpy(p->kx, kx.data(), n * sizeof(float), hipMemcpyHostToDevice), "copy kx")) ||
      (rc = check_hip(hipMemcpy(p->ky, ky.data(), h1 * sizeof(float), hipMemcpyHostToDevice), "copy ky")) ||
      (rc = check_hip(hipMemcpy(p->tw, tw.data(), n * sizeof(float2), hipMemcpyHostToDevice), "copy tw")) ||
      (rc = check_hip(hipMemcpy(p->force_idx, fidx.data(), fidx.size() * sizeof(int), hipMemcpyHostToDevice),
                      "copy forcing")) ||
      (rc = check_hip(hipMemcpy(p->force_val, fval.data(), fval.size() * sizeof(float2), hipMemcpyHostToDevice),
                      "copy forcing"))) {
    jcfd_ns2d_plan_destroy(p);
    return rc;
  }
  // 2/3 rule bounds with Python int / floor-division semantics (spectral/utils.py:128-134)
  const int a = static_cast<int>(2.0 / 3.0 * n);
  const int cols = static_cast<int>(2.0 / 3.0 * h1);
  const bool cnrk4 = c.rk_stages == 0;
  p->n_stages = cnrk4 ? 5 : c.rk_stages;
  const double* alpha = cnrk4 ? kCnAlpha : c.rk_alpha;
  const double* beta = cnrk4 ? kCnBeta : c.rk_beta;
  const double* gamma = cnrk4 ? kCnGamma : c.rk_gamma;
  for (int k = 0; k < p->n_stages; ++k) {
    NsParams& q = p->stages[k];
    q.kx = p->kx;
    q.ky = p->ky;
    q.tw = p->tw;
    q.force_idx = p->force_idx;
    q.force_val = p->force_val;
    q.n_force = c.n_forcing;
    q.two_pi = static_cast<float>(2.0 * M_PI);
    q.neg_four_pi_sq = static_cast<float>(-(2.0 * M_PI) * (2.0 * M_PI));
    q.inv_nn = 1.0f / (static_cast<float>(n) * static_cast<float>(n));
    q.nu = static_cast<float>(c.viscosity);
    q.drag = static_cast<float>(c.drag);
    q.smooth = c.smooth ? 1 : 0;
    q.filt_row_lo = a / 2;
    q.filt_row_hi = n - (a + 1) / 2;
    q.filt_cols = cols;
    q.beta = static_cast<float>(beta[k]);
    q.g = static_cast<float>(gamma[k] * c.dt);
    q.mu = static_cast<float>(0.5 * c.dt * (alpha[k + 1] - alpha[k]));
    q.read_h = k > 0 ? 1 : 0;
    q.slots = nullptr;
    (void)ns_default_tiles(&q, n);
  }
  *plan = p;
  return JCFD_OK;
}

int jcfd_ns2d_plan_get_tiles(const jcfd_ns2d_plan* p, jcfd_ns2d_tiles* tiles) {
  if (p == nullptr || tiles == nullptr) return fail(JCFD_ERR_INVALID_ARGUMENT, "null argument");
  tiles->cols = p->stages[0].w_cols;
  tiles->rows = p->stages[0].rb_rows;
  tiles->radix = p->stages[0].radix;
  return JCFD_OK;
}

int jcfd_ns2d_plan_set_tiles(jcfd_ns2d_plan* p, const jcfd_ns2d_tiles* tiles) {
  if (p == nullptr || tiles == nullptr) return fail(JCFD_ERR_INVALID_ARGUMENT, "null argument");
  if (tiles->radix != 8 && tiles->radix != 4)
    return fail(JCFD_ERR_UNSUPPORTED, "radix schedule must be 8 or 4");
  if (!ns_tile_supported(p->n, 0, tiles->cols))
    return fail(JCFD_ERR_UNSUPPORTED, "column tile " + std::to_string(tiles->cols) +
                                          " not built for n = " + std::to_string(p->n));
  if (!ns_tile_supported(p->n, 1, tiles->rows))
    return fail(JCFD_ERR_UNSUPPORTED, "row tile " + std::to_string(tiles->rows) +
                                          " not built for n = " + std::to_string(p->n));
  for (NsParams& q : p->stages) {
    q.w_cols = tiles->cols;
    q.rb_rows = tiles->rows;
    q.radix = tiles->radix;
  }
  p->graphs.release();  // captured launches carry the old shapes
  return JCFD_OK;
}

int64_t jcfd_ns2d_plan_device_bytes(const jcfd_ns2d_plan* p) { return p ? p->bytes : 0; }

int jcfd_ns2d_step(jcfd_ns2d_plan* p, const jcfd_tensor* w_hat, jcfd_tensor* w_out, int64_t nsteps,
                   void* stream) {
  if (p == nullptr) return fail(JCFD_ERR_INVALID_ARGUMENT, "null plan");
  if (nsteps < 0) return fail(JCFD_ERR_INVALID_ARGUMENT, "nsteps must be >= 0");
  void *pin, *pout;
  int rc;
  if ((rc = check_spectral(w_hat, "w_hat", p->batch, p->n, p->device, &pin)) ||
      (rc = check_spectral(w_out, "w_out", p->batch, p->n, p->device, &pout)))
    return rc;
  hipStream_t s = static_cast<hipStream_t>(stream);
  const size_t bytes = sizeof(float2) * p->batch * static_cast<size_t>(p->n) * (p->n / 2 + 1);
  if (nsteps == 0) {
    if (pin != pout) return check_hip(hipMemcpyAsync(pout, pin, bytes, hipMemcpyDeviceToDevice, s), "copy");
    return JCFD_OK;
  }
  return ns_run(p, static_cast<const float2*>(pin), static_cast<float2*>(pout), nullptr, p->batch, nsteps, s);
}

int jcfd_ns2d_step_slots(jcfd_ns2d_plan* p, jcfd_tensor* w_hat, const jcfd_tensor* slots,
                         int64_t nsteps, void* stream) {
  if (p == nullptr) return fail(JCFD_ERR_INVALID_ARGUMENT, "null plan");
  if (nsteps < 0) return fail(JCFD_ERR_INVALID_ARGUMENT, "nsteps must be >= 0");
  void *pw, *ps;
  int count = 0, rc;
  if ((rc = check_spectral(w_hat, "w_hat", p->batch, p->n, p->device, &pw)) ||
      (rc = check_slots(slots, p->batch, p->device, &ps, &count)))
    return rc;
  if (nsteps == 0 || count == 0) return JCFD_OK;
  return ns_run(p, static_cast<const float2*>(pw), static_cast<float2*>(pw), static_cast<const int*>(ps),
                count, nsteps, static_cast<hipStream_t>(stream));
}

int jcfd_ns2d_explicit_terms(jcfd_ns2d_plan* p, const jcfd_tensor* w_hat, jcfd_tensor* out,
                             void* stream) {
  if (p == nullptr) return fail(JCFD_ERR_INVALID_ARGUMENT, "null plan");
  void *pin, *pout;
  int rc;
  if ((rc = check_spectral(w_hat, "w_hat", p->batch, p->n, p->device, &pin)) ||
      (rc = check_spectral(out, "out", p->batch, p->n, p->device, &pout)))
    return rc;
  if (pin == pout) return fail(JCFD_ERR_INVALID_ARGUMENT, "explicit terms: out must not alias w_hat");
  hipStream_t s = static_cast<hipStream_t>(stream);
  if ((rc = ns_launch_state_convert(p->n, true, static_cast<const float2*>(pin), p->ws, nullptr, p->batch, s)))
    return rc;
  return ns_launch_explicit_terms(p->n, p->stages[0], p->ws, static_cast<float2*>(pout), p->g, p->pbuf,
                                  p->batch, s);
}

int jcfd_ns2d_time_kernels(jcfd_ns2d_plan* p, const jcfd_tensor* w_hat, int32_t reps,
                           double* ms_per_launch, void* stream) {
  if (p == nullptr || ms_per_launch == nullptr) return fail(JCFD_ERR_INVALID_ARGUMENT, "null argument");
  if (reps <= 0) return fail(JCFD_ERR_INVALID_ARGUMENT, "reps must be positive");
  void* pin;
  int rc;
  if ((rc = check_spectral(w_hat, "w_hat", p->batch, p->n, p->device, &pin))) return rc;
  hipStream_t s = static_cast<hipStream_t>(stream);
  const size_t count = static_cast<size_t>(p->batch) * p->n * (p->n / 2 + 1);
  float2* tmp = nullptr;
  int64_t unused = 0;
  if ((rc = device_alloc(&tmp, count, &unused))) return rc;
  NsBuffers b{};
  b.u_in = p->ws;
  b.u_out = tmp;
  b.h = p->h;
  b.g = p->g;
  b.pbuf = p->pbuf;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  rc = ns_launch_state_convert(p->n, true, static_cast<const float2*>(pin), p->ws, nullptr, p->batch, s);
  if (!rc) rc = check_hip(hipEventCreate(&e0), "event");
  if (!rc) rc = check_hip(hipEventCreate(&e1), "event");
  if (!rc) rc = ns_launch_step(p->n, p->stages, p->n_stages, b, p->batch, s);  // fills every scratch buffer
  for (int k = 0; k < JCFD_NS2D_KERNEL_COUNT && rc == JCFD_OK; ++k) {
    rc = check_hip(hipEventRecord(e0, s), "event record");
    for (int r = 0; r < reps && rc == JCFD_OK; ++r) rc = ns_launch_kernel(p->n, k, p->stages, p->n_stages, b, p->batch, s);
    if (!rc) rc = check_hip(hipEventRecord(e1, s), "event record");
    if (!rc) rc = check_hip(hipEventSynchronize(e1), "event sync");
    float ms = 0.f;
    if (!rc) rc = check_hip(hipEventElapsedTime(&ms, e0, e1), "elapsed");
    ms_per_launch[k] = static_cast<double>(ms) / reps;
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipStreamSynchronize(s);
  (void)hipFree(tmp);
  return rc;
}

int jcfd_ns2d_kernel_bytes(const jcfd_ns2d_plan* p, int64_t* bytes) {
  if (p == nullptr || bytes == nullptr) return fail(JCFD_ERR_INVALID_ARGUMENT, "null argument");
  const int64_t n = p->n, B = p->batch;
  const int64_t ext = 8 * n * (n / 2 + 1);  // one complex64 half spectrum
  const int64_t packed = 8 * n * (n / 2);   // one packed half spectrum
  bytes[JCFD_NS2D_KERNEL_COLS_INVERSE] = B * (ext + 4 * packed);
  bytes[JCFD_NS2D_KERNEL_ROWS] = B * (4 * packed + packed);
  bytes[JCFD_NS2D_KERNEL_COLS_UPDATE_INVERSE] = B * (packed + 2 * ext + 2 * ext + 4 * packed);
  bytes[JCFD_NS2D_KERNEL_COLS_UPDATE] = B * (packed + 2 * ext + 2 * ext);
  return JCFD_OK;
}

int jcfd_ns2d_rfft2(jcfd_ns2d_plan* p, const jcfd_tensor* x, jcfd_tensor* out, void* stream) {
  if (p == nullptr) return fail(JCFD_ERR_INVALID_ARGUMENT, "null plan");
  void *pin, *pout;
  int rc;
  if ((rc = check_batched(x, "x", kDLFloat, 32, p->batch, p->n, p->n, p->device, &pin)) ||
      (rc = check_spectral(out, "out", p->batch, p->n, p->device, &pout)))
    return rc;
  return ns_launch_rfft2(p->n, p->stages[0], static_cast<const float*>(pin), static_cast<float2*>(pout),
                         p->pbuf, p->batch, static_cast<hipStream_t>(stream));
}

int jcfd_ns2d_irfft2(jcfd_ns2d_plan* p, const jcfd_tensor* x_hat, jcfd_tensor* out, void* stream) {
  if (p == nullptr) return fail(JCFD_ERR_INVALID_ARGUMENT, "null plan");
  void *pin, *pout;
  int rc;
  if ((rc = check_spectral(x_hat, "x_hat", p->batch, p->n, p->device, &pin)) ||
      (rc = check_batched(out, "out", kDLFloat, 32, p->batch, p->n, p->n, p->device, &pout)))
    return rc;
  return ns_launch_irfft2(p->n, p->stages[0], static_cast<const float2*>(pin), static_cast<float*>(pout),
                          p->pbuf, p->batch, static_cast<hipStream_t>(stream));
}

}  // extern "C"

// ------------------------------------------------ mixed-size spectral groups
struct jcfd_ns2d_group {
  int device = 0;
  NsGroupDesc* descs = nullptr;  // device [instance][JCFD_MAX_RK_STAGES]
  NsGroupItem* items = nullptr;  // device work lists, list k at items + k * cap_items
  int* d_slots = nullptr;        // device slot lists of the members (state conversion)
  int* d_counts = nullptr;       // device int[2]: items per list
  int cap_descs = 0, cap_items = 0, cap_slots = 0;
  char* staging = nullptr;       // pinned upload buffer
  size_t staging_bytes = 0;
  hipEvent_t staged = nullptr;
  bool staged_pending = false;
  std::vector<char> last;
  int max_n = 0;
  int grid[kNsGroupClasses] = {0, 0, 0, 0};
  int n_items[2] = {0, 0};
  int n_stages = 0;
  int64_t bytes[kNsGroupClasses] = {0, 0, 0, 0};
  GraphCache graphs;
};

namespace {

void ns_group_free_tables(jcfd_ns2d_group* g) {
  for (void* ptr : {static_cast<void*>(g->descs), static_cast<void*>(g->items), static_cast<void*>(g->d_slots)})
    if (ptr != nullptr) (void)hipFree(ptr);
  g->descs = nullptr;
  g->items = nullptr;
  g->d_slots = nullptr;
  g->cap_descs = g->cap_items = g->cap_slots = 0;
}

int ns_group_reserve(jcfd_ns2d_group* g, int n_desc, int n_item, int n_slot) {
  if (n_desc <= g->cap_descs && n_item <= g->cap_items && n_slot <= g->cap_slots) return JCFD_OK;
  int rc;
  if (g->descs != nullptr && (rc = check_hip(hipDeviceSynchronize(), "group regrow sync"))) return rc;
  const int cd = std::max({n_desc, 2 * g->cap_descs, kGroupMinDescs});
  const int ci = std::max({n_item, 2 * g->cap_items, kGroupMinItems});
  const int cs = std::max({n_slot, 2 * g->cap_slots, kGroupMinDescs});
  ns_group_free_tables(g);
  int64_t unused = 0;
  if ((rc = device_alloc(&g->descs, static_cast<size_t>(cd) * JCFD_MAX_RK_STAGES, &unused)) ||
      (rc = device_alloc(&g->items, static_cast<size_t>(2) * ci, &unused)) ||
      (rc = device_alloc(&g->d_slots, static_cast<size_t>(cs), &unused))) {
    ns_group_free_tables(g);
    return rc;
  }
  g->cap_descs = cd;
  g->cap_items = ci;
  g->cap_slots = cs;
  g->graphs.release();  // captured launches point at the old tables
  return JCFD_OK;
}

NsGroupQueue ns_group_queue(const jcfd_ns2d_group* g) {
  NsGroupQueue q{};
  q.descs = g->descs;
  q.items[0] = g->items;
  q.items[1] = g->items + g->cap_items;
  q.n_items = g->d_counts;
  return q;
}

// Upload (counts[2] | descs | items 0 | items 1 | slots) through the pinned
// staging buffer, stream-ordered on s.
int ns_group_upload(jcfd_ns2d_group* g, const std::vector<char>& table, int n_desc, const int* counts,
                    int n_slot, hipStream_t s) {
  int rc;
  if (g->staged_pending && (rc = check_hip(hipEventSynchronize(g->staged), "staging reuse"))) return rc;
  g->staged_pending = false;
  if (table.size() > g->staging_bytes) {
    if (g->staging != nullptr) (void)hipHostFree(g->staging);
    g->staging = nullptr;
    const size_t bytes = std::max({table.size(), 2 * g->staging_bytes, kGroupMinStaging});
    if ((rc = check_hip(hipHostMalloc(reinterpret_cast<void**>(&g->staging), bytes, hipHostMallocDefault),
                        "pinned group staging")))
      return rc;
    g->staging_bytes = bytes;
  }
  std::memcpy(g->staging, table.data(), table.size());
  size_t off = 2 * sizeof(int);
  const size_t desc_bytes = static_cast<size_t>(n_desc) * JCFD_MAX_RK_STAGES * sizeof(NsGroupDesc);
  if ((rc = check_hip(hipMemcpyAsync(g->d_counts, g->staging, 2 * sizeof(int), hipMemcpyHostToDevice, s),
                      "group upload")) ||
      (rc = check_hip(hipMemcpyAsync(g->descs, g->staging + off, desc_bytes, hipMemcpyHostToDevice, s),
                      "group upload")))
    return rc;
  off += desc_bytes;
  for (int k = 0; k < 2; ++k) {
    const size_t bytes = counts[k] * sizeof(NsGroupItem);
    if (bytes > 0 &&
        (rc = check_hip(hipMemcpyAsync(g->items + static_cast<size_t>(k) * g->cap_items, g->staging + off, bytes,
                                       hipMemcpyHostToDevice, s),
                        "group upload")))
      return rc;
    off += bytes;
  }
  if (n_slot > 0 && (rc = check_hip(hipMemcpyAsync(g->d_slots, g->staging + off, n_slot * sizeof(int),
                                                   hipMemcpyHostToDevice, s),
                                    "group upload")))
    return rc;
  if ((rc = check_hip(hipEventRecord(g->staged, s), "staging event"))) return rc;
  g->staged_pending = true;
  return JCFD_OK;
}

}  // namespace

extern "C" {

int jcfd_ns2d_group_create(jcfd_ns2d_group** group) {
  if (group == nullptr) return fail(JCFD_ERR_INVALID_ARGUMENT, "null argument");
  *group = nullptr;
  auto* g = new (std::nothrow) jcfd_ns2d_group();
  if (g == nullptr) return fail(JCFD_ERR_OUT_OF_MEMORY, "host allocation failed");
  int64_t unused = 0;
  int rc;
  if ((rc = check_hip(hipGetDevice(&g->device), "hipGetDevice")) ||
      (rc = device_alloc(&g->d_counts, 2, &unused)) ||
      (rc = check_hip(hipMemset(g->d_counts, 0, 2 * sizeof(int)), "zero counts")) ||
      (rc = check_hip(hipEventCreateWithFlags(&g->staged, hipEventDisableTiming), "staging event"))) {
    jcfd_ns2d_group_destroy(g);
    return rc;
  }
  *group = g;
  return JCFD_OK;
}

int jcfd_ns2d_group_destroy(jcfd_ns2d_group* g) {
  if (g == nullptr) return JCFD_OK;
  if (g->staged_pending) (void)hipEventSynchronize(g->staged);
  g->graphs.release();
  ns_group_free_tables(g);
  if (g->d_counts != nullptr) (void)hipFree(g->d_counts);
  if (g->staging != nullptr) (void)hipHostFree(g->staging);
  if (g->staged != nullptr) (void)hipEventDestroy(g->staged);
  delete g;
  return JCFD_OK;
}

int jcfd_ns2d_group_set_graphs(jcfd_ns2d_group* g, int32_t enabled) {
  if (g == nullptr) return fail(JCFD_ERR_INVALID_ARGUMENT, "null group");
  g->graphs.enabled = enabled != 0;
  return JCFD_OK;
}

int jcfd_ns2d_group_step(jcfd_ns2d_group* g, const jcfd_ns2d_group_member* members, int32_t n_members,
                         int64_t nsteps, void* stream) {
  if (g == nullptr) return fail(JCFD_ERR_INVALID_ARGUMENT, "null group");
  if (n_members < 0 || (n_members > 0 && members == nullptr))
    return fail(JCFD_ERR_INVALID_ARGUMENT, "members: null array or negative count");
  if (nsteps < 0) return fail(JCFD_ERR_INVALID_ARGUMENT, "nsteps must be >= 0");
  struct Conv {
    int n, slot_off, count;
    const jcfd_ns2d_plan* plan;
    float2* state;
  };
  std::vector<Conv> convs;
  std::vector<int> slots_all;
  struct Inst {
    int n;
    NsGroupDesc stage[JCFD_MAX_RK_STAGES];
  };
  std::vector<Inst> insts;
  std::vector<const void*> owners;
  int n_stages = 0, rc;
  for (int m = 0; m < n_members; ++m) {
    const jcfd_ns2d_group_member& mb = members[m];
    jcfd_ns2d_plan* p = mb.plan;
    const std::string who = "member " + std::to_string(m);
    if (p == nullptr) return fail(JCFD_ERR_INVALID_ARGUMENT, who + ": null plan");
    if (p->device != g->device)
      return fail(JCFD_ERR_INVALID_ARGUMENT, who + ": plan on another device than the group");
    if (p->stages[0].radix != 8)
      return fail(JCFD_ERR_UNSUPPORTED, who + ": groups run the radix-8 FFT schedule only");
    if (n_stages != 0 && p->n_stages != n_stages)
      return fail(JCFD_ERR_UNSUPPORTED, who + ": every member must step the same number of stages");
    n_stages = p->n_stages;
    const int n = p->n;
    if (!ns_group_size_supported(n))
      return fail(JCFD_ERR_UNSUPPORTED, who + ": grid size " + std::to_string(n) + " not built for groups");
    void* pw;
    if ((rc = check_spectral(mb.w_hat, "w_hat", p->batch, n, p->device, &pw)))
      return fail(rc, who + ": " + jcfd_last_error());
    if (mb.n_slots < 0 || mb.n_slots > p->batch || (mb.n_slots > 0 && mb.slots == nullptr))
      return fail(JCFD_ERR_INVALID_ARGUMENT, who + ": slots must hold 0..batch host int32 values");
    const size_t ext = static_cast<size_t>(n) * (n / 2 + 1), half = static_cast<size_t>(n) * (n / 2);
    const int slot_off = static_cast<int>(slots_all.size());
    for (int k = 0; k < mb.n_slots; ++k) {
      const int s = mb.slots[k];
      if (s < 0 || s >= p->batch)
        return fail(JCFD_ERR_INVALID_ARGUMENT, who + ": slot " + std::to_string(s) + " out of range");
      slots_all.push_back(s);
      Inst in{};
      in.n = n;
      for (int st = 0; st < p->n_stages; ++st) {
        NsGroupDesc& d = in.stage[st];
        d.q = p->stages[st];
        d.q.slots = nullptr;
        d.n = n;
        d.ws = p->ws + s * ext;
        d.h = p->h + s * ext;
        d.g = p->g + s * 4 * half;
        d.pbuf = p->pbuf + s * half;
      }
      insts.push_back(in);
      owners.push_back(p->ws + s * ext);
      owners.push_back(static_cast<float2*>(pw) + s * ext);
    }
    if (mb.n_slots > 0) convs.push_back(Conv{n, slot_off, mb.n_slots, p, static_cast<float2*>(pw)});
  }
  std::vector<const void*> sorted_owners = owners;
  std::sort(sorted_owners.begin(), sorted_owners.end());
  if (std::adjacent_find(sorted_owners.begin(), sorted_owners.end()) != sorted_owners.end())
    return fail(JCFD_ERR_INVALID_ARGUMENT,
                "group members repeat an instance (same plan slot or same state slot twice)");
  if (nsteps == 0 || insts.empty()) return JCFD_OK;
  // largest grids first (items are of equal work; this keeps an instance's
  // items contiguous in the order the fixed-stride walk deals them)
  std::stable_sort(insts.begin(), insts.end(), [](const Inst& a, const Inst& b) { return a.n > b.n; });
  const int n_desc = static_cast<int>(insts.size());
  const int max_n = insts.front().n;
  std::vector<NsGroupItem> items[2];
  for (int l = 0; l < 2; ++l)
    for (int d = 0; d < n_desc; ++d)
      for (int t = 0, nt = ns_group_tiles(insts[d].n, l); t < nt; ++t) items[l].push_back({d, t});
  const int counts[2] = {static_cast<int>(items[0].size()), static_cast<int>(items[1].size())};
  const int n_slot = static_cast<int>(slots_all.size());
  std::vector<char> table(2 * sizeof(int) + static_cast<size_t>(n_desc) * JCFD_MAX_RK_STAGES * sizeof(NsGroupDesc) +
                          (counts[0] + counts[1]) * sizeof(NsGroupItem) + n_slot * sizeof(int));
  {
    char* w = table.data();
    std::memcpy(w, counts, 2 * sizeof(int));
    w += 2 * sizeof(int);
    for (const Inst& in : insts) {
      std::memcpy(w, in.stage, JCFD_MAX_RK_STAGES * sizeof(NsGroupDesc));
      w += JCFD_MAX_RK_STAGES * sizeof(NsGroupDesc);
    }
    for (int l = 0; l < 2; ++l) {
      std::memcpy(w, items[l].data(), counts[l] * sizeof(NsGroupItem));
      w += counts[l] * sizeof(NsGroupItem);
    }
    std::memcpy(w, slots_all.data(), n_slot * sizeof(int));
  }
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (n_desc > g->cap_descs || std::max(counts[0], counts[1]) > g->cap_items || n_slot > g->cap_slots) {
    if ((rc = ns_group_reserve(g, n_desc, std::max(counts[0], counts[1]), n_slot))) return rc;
    g->last.clear();
  }
  if (table != g->last) {
    if ((rc = ns_group_upload(g, table, n_desc, counts, n_slot, s))) return rc;
    g->last = table;
  }
  g->max_n = max_n;
  for (int k = 0; k < kNsGroupClasses; ++k) g->grid[k] = counts[k == 1 ? 1 : 0];  // one workgroup per item
  g->n_items[0] = counts[0];
  g->n_items[1] = counts[1];
  g->n_stages = n_stages;
  for (int k = 0; k < kNsGroupClasses; ++k) g->bytes[k] = 0;
  for (const Inst& in : insts) {  // per instance, as jcfd_ns2d_kernel_bytes with batch 1
    const int64_t n = in.n, ext = n * (n / 2 + 1) * 8, packed = n * (n / 2) * 8;
    g->bytes[0] += ext + 4 * packed;
    g->bytes[1] += 4 * packed + packed;
    g->bytes[2] += packed + 2 * ext + 2 * ext + 4 * packed;
    g->bytes[3] += packed + 2 * ext + 2 * ext;
  }
  // external states -> the plans' internal layout, the steps, and back
  for (const Conv& c : convs)
    if ((rc = ns_launch_state_convert(c.n, true, c.state, c.plan->ws, g->d_slots + c.slot_off, c.count, s)))
      return rc;
  const NsGroupQueue q = ns_group_queue(g);
  const int S = n_stages;
  auto one_step = [&](hipStream_t st) {
    int r = ns_group_launch(0, max_n, q, 0, g->grid[0], st);
    for (int k = 0; k < S && r == JCFD_OK; ++k) {
      r = ns_group_launch(1, max_n, q, k, g->grid[1], st);
      if (r == JCFD_OK) r = ns_group_launch(k + 1 < S ? 2 : 3, max_n, q, k, g->grid[k + 1 < S ? 2 : 3], st);
    }
    return r;
  };
  // captured launches bake in the grids (the item counts) and LDS size
  const void* count_key = reinterpret_cast<const void*>(
      (static_cast<uintptr_t>(static_cast<uint32_t>(counts[0])) << 32) | static_cast<uint32_t>(counts[1]));
  if ((rc = run_steps(g->graphs, g->descs, g->items, count_key, max_n * 16 + S, nsteps, s, one_step)))
    return rc;
  for (const Conv& c : convs)
    if ((rc = ns_launch_state_convert(c.n, false, c.plan->ws, c.state, g->d_slots + c.slot_off, c.count, s)))
      return rc;
  return JCFD_OK;
}

int jcfd_ns2d_group_time_kernels(jcfd_ns2d_group* g, int32_t reps, double* ms_per_launch, void* stream) {
  if (g == nullptr || ms_per_launch == nullptr) return fail(JCFD_ERR_INVALID_ARGUMENT, "null argument");
  if (reps <= 0) return fail(JCFD_ERR_INVALID_ARGUMENT, "reps must be positive");
  if (g->max_n == 0 || g->n_items[0] == 0)
    return fail(JCFD_ERR_INVALID_ARGUMENT, "no group step to time (call jcfd_ns2d_group_step first)");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const NsGroupQueue q = ns_group_queue(g);
  const int S = g->n_stages;
  // launch sequence of one step: (class, stage)
  std::vector<std::pair<int, int>> seq{{0, 0}};
  for (int k = 0; k < S; ++k) {
    seq.push_back({1, k});
    seq.push_back({k + 1 < S ? 2 : 3, k});
  }
  std::vector<hipEvent_t> ev(seq.size() + 1, nullptr);
  int rc = JCFD_OK;
  for (size_t k = 0; k < ev.size() && rc == JCFD_OK; ++k) rc = check_hip(hipEventCreate(&ev[k]), "event");
  double total[kNsGroupClasses] = {0, 0, 0, 0};
  int count[kNsGroupClasses] = {0, 0, 0, 0};
  // whole steps on the group's internal states (the members' caller-side
  // states are untouched: nothing converts back)
  for (int r = 0; r < reps && rc == JCFD_OK; ++r) {
    rc = check_hip(hipEventRecord(ev[0], s), "event record");
    for (size_t k = 0; k < seq.size() && rc == JCFD_OK; ++k) {
      const int c = seq[k].first;
      rc = ns_group_launch(c, g->max_n, q, seq[k].second, g->grid[c], s);
      if (rc == JCFD_OK) rc = check_hip(hipEventRecord(ev[k + 1], s), "event record");
    }
    if (rc == JCFD_OK) rc = check_hip(hipEventSynchronize(ev.back()), "event sync");
    for (size_t k = 0; k < seq.size() && rc == JCFD_OK; ++k) {
      float ms = 0.f;
      rc = check_hip(hipEventElapsedTime(&ms, ev[k], ev[k + 1]), "elapsed");
      total[seq[k].first] += ms;
      ++count[seq[k].first];
    }
  }
  for (int k = 0; k < kNsGroupClasses; ++k) ms_per_launch[k] = count[k] ? total[k] / count[k] : 0.0;
  for (hipEvent_t e : ev)
    if (e != nullptr) (void)hipEventDestroy(e);
  return rc;
}

int jcfd_ns2d_group_launch_shape(const jcfd_ns2d_group* g, int32_t* grid, int32_t* items,
                                 int64_t* bytes_per_launch) {
  if (g == nullptr) return fail(JCFD_ERR_INVALID_ARGUMENT, "null group");
  for (int k = 0; k < kNsGroupClasses; ++k) {
    if (grid) grid[k] = g->grid[k];
    if (items) items[k] = g->n_items[k == 1 ? 1 : 0];
    if (bytes_per_launch) bytes_per_launch[k] = g->bytes[k];
  }
  return JCFD_OK;
}

}  // extern "C"
