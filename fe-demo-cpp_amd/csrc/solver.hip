// Mass solves on the fused operator.
//
//   FP64 : Jacobi-preconditioned CG, every vector and scalar in FP64.
//   MIXED: FP64 iterative refinement. Each step forms r = b - M x with the FP64
//          fused apply, rounds r to FP32, runs Jacobi-PCG with the FP32 fused
//          apply (FP32 vectors, FP64 dot-product accumulation) to a relative
//          reduction inner_rtol, and adds the correction to x in FP64.
//
// All CG scalars live on the device (no host round trip per iteration); a
// batch of iterations is captured once into a hipGraph and replayed, and the
// host reads the convergence flag once per batch. Once the residual test
// passes, the flag freezes x, so extra iterations of a batch change nothing.
// Reductions use a fixed grid and a fixed tree, so a solve is reproducible.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>

#include "solver.hpp"

namespace femhip {

namespace {

constexpr int kThreads = 256;
constexpr int kRedBlocks = 240;  // fixed reduction grid

// Scalar slots.
enum { kRZ = 0, kPAP, kALPHA, kBETA, kRR, kDONE, kITERS, kTHRESH, kNSCAL };

__device__ __forceinline__ double block_sum(double v, double* sh) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  const int lane = threadIdx.x % 64, w = threadIdx.x / 64;
  if (lane == 0) sh[w] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int i = 0; i < kThreads / 64; ++i) s += sh[i];
  return s;  // valid in thread 0
}

template <class T>
__global__ __launch_bounds__(kThreads) void dot_partial_kernel(size_t n, const T* __restrict__ a,
                                                               const T* __restrict__ b, double* __restrict__ part) {
  __shared__ double sh[kThreads / 64];
  double s = 0.0;
  for (size_t i = static_cast<size_t>(blockIdx.x) * kThreads + threadIdx.x; i < n;
       i += static_cast<size_t>(gridDim.x) * kThreads)
    s += static_cast<double>(a[i]) * static_cast<double>(b[i]);
  s = block_sum(s, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__device__ double sum_partials(const double* part, int nb) {
  double s = 0.0;
  for (int i = 0; i < nb; ++i) s += part[i];
  return s;
}

// r = b, z = dinv * r, p = z; partials of r.z and r.r
template <class T>
__global__ __launch_bounds__(kThreads) void cg_init_kernel(size_t n, const T* __restrict__ b,
                                                           const T* __restrict__ dinv, T* __restrict__ x,
                                                           T* __restrict__ r, T* __restrict__ z, T* __restrict__ p,
                                                           double* __restrict__ part) {
  __shared__ double sh[2][kThreads / 64];
  double srz = 0.0, srr = 0.0;
  for (size_t i = static_cast<size_t>(blockIdx.x) * kThreads + threadIdx.x; i < n;
       i += static_cast<size_t>(gridDim.x) * kThreads) {
    const T ri = b[i];
    const T zi = dinv[i] * ri;
    x[i] = T(0);
    r[i] = ri;
    z[i] = zi;
    p[i] = zi;
    srz += static_cast<double>(ri) * static_cast<double>(zi);
    srr += static_cast<double>(ri) * static_cast<double>(ri);
  }
  const double a = block_sum(srz, sh[0]);
  const double c = block_sum(srr, sh[1]);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = a;
    part[kRedBlocks + blockIdx.x] = c;
  }
}

__global__ void cg_init_scalars_kernel(const double* __restrict__ part, double* __restrict__ sc, double rtol2) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const double rz = sum_partials(part, kRedBlocks), rr = sum_partials(part + kRedBlocks, kRedBlocks);
  sc[kRZ] = rz;
  sc[kRR] = rr;
  sc[kTHRESH] = rtol2 * rr;
  sc[kDONE] = (rr <= sc[kTHRESH]) ? 1.0 : 0.0;
  sc[kITERS] = 0.0;
}

__global__ void cg_alpha_kernel(const double* __restrict__ part, double* __restrict__ sc) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const double pap = sum_partials(part, kRedBlocks);
  sc[kPAP] = pap;
  sc[kALPHA] = (sc[kDONE] != 0.0 || pap == 0.0) ? 0.0 : sc[kRZ] / pap;
}

template <class T>
__global__ __launch_bounds__(kThreads) void cg_update_kernel(size_t n, const double* __restrict__ sc,
                                                             const T* __restrict__ p, const T* __restrict__ Ap,
                                                             const T* __restrict__ dinv, T* __restrict__ x,
                                                             T* __restrict__ r, T* __restrict__ z,
                                                             double* __restrict__ part) {
  __shared__ double sh[2][kThreads / 64];
  const bool done = sc[kDONE] != 0.0;
  const T alpha = static_cast<T>(sc[kALPHA]);
  double srz = 0.0, srr = 0.0;
  for (size_t i = static_cast<size_t>(blockIdx.x) * kThreads + threadIdx.x; i < n;
       i += static_cast<size_t>(gridDim.x) * kThreads) {
    T ri = r[i];
    if (!done) {
      x[i] = x[i] + alpha * p[i];
      ri = ri - alpha * Ap[i];
      r[i] = ri;
      const T zi = dinv[i] * ri;
      z[i] = zi;
      srz += static_cast<double>(ri) * static_cast<double>(zi);
    }
    srr += static_cast<double>(ri) * static_cast<double>(ri);
  }
  const double a = block_sum(srz, sh[0]);
  const double c = block_sum(srr, sh[1]);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = a;
    part[kRedBlocks + blockIdx.x] = c;
  }
}

__global__ void cg_beta_kernel(const double* __restrict__ part, double* __restrict__ sc, int max_iter) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  if (sc[kDONE] != 0.0) {
    sc[kBETA] = 0.0;
    return;
  }
  const double rz_new = sum_partials(part, kRedBlocks), rr = sum_partials(part + kRedBlocks, kRedBlocks);
  sc[kITERS] += 1.0;
  sc[kRR] = rr;
  sc[kBETA] = (sc[kRZ] == 0.0) ? 0.0 : rz_new / sc[kRZ];
  sc[kRZ] = rz_new;
  if (rr <= sc[kTHRESH] || sc[kITERS] >= max_iter) sc[kDONE] = 1.0;
}

template <class T>
__global__ __launch_bounds__(kThreads) void cg_p_kernel(size_t n, const double* __restrict__ sc,
                                                        const T* __restrict__ z, T* __restrict__ p) {
  if (sc[kDONE] != 0.0) return;
  const T beta = static_cast<T>(sc[kBETA]);
  for (size_t i = static_cast<size_t>(blockIdx.x) * kThreads + threadIdx.x; i < n;
       i += static_cast<size_t>(gridDim.x) * kThreads)
    p[i] = z[i] + beta * p[i];
}

template <class T>
__global__ __launch_bounds__(kThreads) void recip_kernel(size_t n, const double* __restrict__ d, T* __restrict__ dinv) {
  const size_t i = static_cast<size_t>(blockIdx.x) * kThreads + threadIdx.x;
  if (i < n) dinv[i] = static_cast<T>(d[i] != 0.0 ? 1.0 / d[i] : 0.0);
}

// r = b - Mx (FP64) and its squared norm partials; r32 = (float) r.
__global__ __launch_bounds__(kThreads) void residual_kernel(size_t n, const double* __restrict__ b,
                                                            const double* __restrict__ Mx, double* __restrict__ r,
                                                            float* __restrict__ r32, double* __restrict__ part) {
  __shared__ double sh[kThreads / 64];
  double s = 0.0;
  for (size_t i = static_cast<size_t>(blockIdx.x) * kThreads + threadIdx.x; i < n;
       i += static_cast<size_t>(gridDim.x) * kThreads) {
    const double ri = b[i] - Mx[i];
    r[i] = ri;
    if (r32) r32[i] = static_cast<float>(ri);
    s += ri * ri;
  }
  s = block_sum(s, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ __launch_bounds__(kThreads) void add_correction_kernel(size_t n, const float* __restrict__ d,
                                                                  double* __restrict__ x) {
  const size_t i = static_cast<size_t>(blockIdx.x) * kThreads + threadIdx.x;
  if (i < n) x[i] += static_cast<double>(d[i]);
}

__global__ void sum_to_scalar_kernel(const double* __restrict__ part, double* __restrict__ out) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *out = sum_partials(part, kRedBlocks);
}

// One PCG instance (vectors of T) over an apply callback.
template <class T>
struct Pcg {
  size_t n;
  DevBuf<T> r, z, p, Ap, dinv;
  DevBuf<double> part, sc;
  hipGraphExec_t exec = nullptr;
  hipGraph_t graph = nullptr;
  int batch = 0;

  explicit Pcg(size_t n_) : n(n_), r(n_), z(n_), p(n_), Ap(n_), dinv(n_), part(2 * kRedBlocks), sc(kNSCAL) {}
  ~Pcg() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
  }

  template <class Apply>
  void iteration(Apply&& apply, T* x, int max_iter, hipStream_t s) {
    apply(p.data(), Ap.data(), s);
    dot_partial_kernel<T><<<kRedBlocks, kThreads, 0, s>>>(n, p.data(), Ap.data(), part.data());
    cg_alpha_kernel<<<1, 1, 0, s>>>(part.data(), sc.data());
    cg_update_kernel<T><<<kRedBlocks, kThreads, 0, s>>>(n, sc.data(), p.data(), Ap.data(), dinv.data(), x, r.data(),
                                                        z.data(), part.data());
    cg_beta_kernel<<<1, 1, 0, s>>>(part.data(), sc.data(), max_iter);
    cg_p_kernel<T><<<kRedBlocks, kThreads, 0, s>>>(n, sc.data(), z.data(), p.data());
  }

  // Solve A x = b from x = 0 until ||r|| <= rtol ||b|| or max_iter iterations.
  // The graph of `batch_iters` iterations is captured on first use (x and b
  // addresses must stay fixed between solves on this instance).
  template <class Apply>
  int solve(Apply&& apply, const T* b, T* x, double rtol, int max_iter, int batch_iters, hipStream_t s) {
    cg_init_kernel<T><<<kRedBlocks, kThreads, 0, s>>>(n, b, dinv.data(), x, r.data(), z.data(), p.data(),
                                                      part.data());
    cg_init_scalars_kernel<<<1, 1, 0, s>>>(part.data(), sc.data(), rtol * rtol);
    HIP_CHECK_LAUNCH();
    if (!exec || batch != batch_iters) {
      if (exec) HIP_CHECK(hipGraphExecDestroy(exec));
      if (graph) HIP_CHECK(hipGraphDestroy(graph));
      exec = nullptr;
      graph = nullptr;
      // One apply outside the capture, so lazily made operator state (element
      // output buffers, FP32 copies) exists before recording.
      apply(p.data(), Ap.data(), s);
      HIP_CHECK(hipStreamSynchronize(s));
      HIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      try {
        for (int k = 0; k < batch_iters; ++k) iteration(apply, x, max_iter, s);
      } catch (...) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(s, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
      }
      HIP_CHECK(hipStreamEndCapture(s, &graph));
      HIP_CHECK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
      batch = batch_iters;
    }
    double host_sc[kNSCAL];
    for (int done_iters = 0; done_iters < max_iter; done_iters += batch_iters) {
      HIP_CHECK(hipGraphLaunch(exec, s));
      HIP_CHECK(hipMemcpyAsync(host_sc, sc.data(), sizeof(host_sc), hipMemcpyDeviceToHost, s));
      HIP_CHECK(hipStreamSynchronize(s));
      if (host_sc[kDONE] != 0.0) break;
    }
    HIP_CHECK(hipMemcpyAsync(host_sc, sc.data(), sizeof(host_sc), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    return static_cast<int>(host_sc[kITERS]);
  }
};

}  // namespace

struct MassSolverState {
  size_t n = 0;
  DevBuf<double> diag, Mx, r64, part, scal;
  std::unique_ptr<Pcg<double>> pcg64;
  std::unique_ptr<Pcg<float>> pcg32;
  DevBuf<float> r32, d32;
  const double* b_bound = nullptr;  // addresses the captured graphs were built with
  double* x_bound = nullptr;
};

void mass_solver_state_destroy(MassSolverState* st) { delete st; }

static double true_residual(Operator& op, FusedMass& f, MassSolverState& st, const double* b, const double* x,
                            double bnorm, float* r32, hipStream_t s) {
  fused_mass_apply64(f, x, st.Mx.data(), s);
  residual_kernel<<<kRedBlocks, kThreads, 0, s>>>(st.n, b, st.Mx.data(), st.r64.data(), r32, st.part.data());
  sum_to_scalar_kernel<<<1, 1, 0, s>>>(st.part.data(), st.scal.data());
  HIP_CHECK_LAUNCH();
  double rr = 0.0;
  HIP_CHECK(hipMemcpyAsync(&rr, st.scal.data(), sizeof(double), hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipStreamSynchronize(s));
  (void)op;
  return bnorm > 0 ? std::sqrt(rr) / bnorm : std::sqrt(rr);
}

void mass_solve(Operator& op, MassSolverState*& state, const double* b, double* x, const SolveOptions& opt,
                SolveStats& stats, hipStream_t user_stream) {
  FusedMass& f = operator_fused_mass(op, user_stream);
  const size_t n = static_cast<size_t>(f.l_size);
  require(opt.mode == 0 || opt.mode == 1, "mass_solve: unknown mode");
  require(opt.rtol > 0 && opt.max_iter > 0, "mass_solve: rtol and max_iter must be positive");
  // Graph capture needs a non-default stream.
  hipStream_t s = user_stream;
  hipStream_t own = nullptr;
  if (s == nullptr) {
    HIP_CHECK(hipStreamCreateWithFlags(&own, hipStreamNonBlocking));
    s = own;
  }
  struct StreamGuard {
    hipStream_t s;
    ~StreamGuard() {
      if (s) {
        (void)hipStreamSynchronize(s);
        (void)hipStreamDestroy(s);
      }
    }
  } guard{own};

  if (!state || state->n != n) {
    delete state;
    state = new MassSolverState();
    state->n = n;
    state->diag.allocate(n);
    state->Mx.allocate(n);
    state->r64.allocate(n);
    state->part.allocate(2 * kRedBlocks);
    state->scal.allocate(4);
    fused_mass_diagonal(f, state->diag.data(), s);
  }
  MassSolverState& st = *state;
  if (st.b_bound != b || st.x_bound != x) {
    st.pcg64.reset();
    st.pcg32.reset();
    st.b_bound = b;
    st.x_bound = x;
  }
  // ||b||
  dot_partial_kernel<double><<<kRedBlocks, kThreads, 0, s>>>(n, b, b, st.part.data());
  sum_to_scalar_kernel<<<1, 1, 0, s>>>(st.part.data(), st.scal.data());
  double bb = 0.0;
  HIP_CHECK(hipMemcpyAsync(&bb, st.scal.data(), sizeof(double), hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipStreamSynchronize(s));
  const double bnorm = std::sqrt(bb);

  const auto t0 = std::chrono::steady_clock::now();
  stats = SolveStats{};
  const int batch = opt.batch > 0 ? opt.batch : 8;
  if (opt.mode == 0) {
    if (!st.pcg64) {
      st.pcg64 = std::make_unique<Pcg<double>>(n);
      recip_kernel<double><<<grid_for(n, kThreads), kThreads, 0, s>>>(n, st.diag.data(), st.pcg64->dinv.data());
      HIP_CHECK_LAUNCH();
    }
    auto apply = [&](const double* in, double* out, hipStream_t ss) { fused_mass_apply64(f, in, out, ss); };
    stats.iterations = st.pcg64->solve(apply, b, x, opt.rtol, opt.max_iter, batch, s);
    stats.outer = 0;
  } else {
    if (!st.pcg32) {
      st.pcg32 = std::make_unique<Pcg<float>>(n);
      st.r32.allocate(n);
      st.d32.allocate(n);
      recip_kernel<float><<<grid_for(n, kThreads), kThreads, 0, s>>>(n, st.diag.data(), st.pcg32->dinv.data());
      HIP_CHECK_LAUNCH();
      // Make the FP32 qdata copy now, outside any graph capture.
      fused_mass_apply32(f, st.d32.data(), st.r32.data(), s);
    }
    auto apply32 = [&](const float* in, float* out, hipStream_t ss) { fused_mass_apply32(f, in, out, ss); };
    zero_fill(x, n * sizeof(double), s);
    int outer = 0, inner_total = 0;
    double rel = true_residual(op, f, st, b, x, bnorm, st.r32.data(), s);
    while (rel > opt.rtol && outer < opt.max_outer) {
      inner_total += st.pcg32->solve(apply32, st.r32.data(), st.d32.data(), opt.inner_rtol, opt.max_iter, batch, s);
      add_correction_kernel<<<grid_for(n, kThreads), kThreads, 0, s>>>(n, st.d32.data(), x);
      HIP_CHECK_LAUNCH();
      ++outer;
      rel = true_residual(op, f, st, b, x, bnorm, st.r32.data(), s);
    }
    stats.iterations = inner_total;
    stats.outer = outer;
  }
  HIP_CHECK(hipStreamSynchronize(s));
  const auto t1 = std::chrono::steady_clock::now();
  stats.ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
  stats.rel_residual = true_residual(op, f, st, b, x, bnorm, nullptr, s);
}

}  // namespace femhip
