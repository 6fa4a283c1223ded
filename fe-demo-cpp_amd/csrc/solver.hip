// Mass solves on the fused operator.
//
//   FP64 : Jacobi-preconditioned CG, every vector and scalar in FP64.
//   MIXED: FP64 iterative refinement. Each step forms r = b - M x with the FP64
//          fused apply, rounds r to FP32, runs Jacobi-PCG with the FP32 fused
//          apply (FP32 vectors, FP64 dot-product accumulation) to a relative
//          reduction inner_rtol, and adds the correction to x in FP64.
//   MIXED_RELIABLE: one FP32 Jacobi-PCG recursion with reliable updates
//          (residual replacement, van der Vorst & Ye, SIAM J. Sci. Comput. 22(3),
//          2000; the mixed-precision form of Clark et al., Comput. Phys. Commun.
//          181, 2010): the FP32 iterate is an increment d; whenever the
//          recursive residual has dropped by inner_rtol since the last update,
//          x += d in FP64, r = b - M x with the FP64 apply replaces the
//          recursive residual, d = 0, and the recursion goes on with the same
//          search direction p. Without restarts the Krylov space is kept, so
//          the iteration count stays near the FP64 one.
//
// All CG scalars live on the device (no host round trip per iteration); a
// batch of iterations is captured once into a hipGraph and replayed, and the
// host reads the convergence flag once per batch (an asynchronous copy into
// pinned memory followed by an event). Once the residual test passes, the flag
// freezes x, so extra iterations of a batch change nothing. Reductions use a
// fixed grid and a fixed tree, so a solve is reproducible.
//
// A solve is a state machine (SolveJob) advanced by solve_job_step, which never
// blocks: it looks at the event of the work it enqueued last and, when that has
// finished, enqueues the next piece. mass_solve drives one job to the end; the
// serving loop drives many, each on its own stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <chrono>
#include <functional>
#include <cmath>
#include <map>
#include <mutex>
#include <vector>

#include "cg_kernels.hpp"
#include "solver.hpp"

namespace femhip {

namespace {

// One PCG instance (vectors of T) over an apply callback: the plain form
// (five dependent launches an iteration with a two-kernel apply) or, with
// `pipe`, the pipelined form (three); Ap holds n = A m in the pipelined form.
template <class T>
struct Pcg {
  size_t n;
  bool pipe;
  bool wide;  // plain form: the one-pack-per-thread kernels (long vectors)
  DevBuf<T> r, z, p, Ap, dinv;
  DevBuf<T> u, w, m, q, sv;  // pipelined form only (sv: s = A p)
  DevBuf<double> part, sc;
  DevBuf<double> partw;  // wide form: per-block partials (p.Ap, then r.z and r.r)
  // Pipelined form with the apply split (FusedSplit): the first stage writes Ap
  // and the step kernel forms the rest of A m itself.
  bool split = false;
  FusedSplit sp;
  std::function<void(const T*, T*, hipStream_t)> first;
  // Captured batches of iterations: the configured batch size and every power
  // of two below it, so any count of iterations is a few graph launches.
  struct Exec {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
  };
  std::map<int, Exec> execs;   // by iterations per launch
  int batch = 0;
  const T* b_bound = nullptr;  // the b and x the graphs were captured with
  T* x_bound = nullptr;
  uint64_t sig_bound = 0;      // and the apply's fused_signature

  int nb;   // reduction blocks
  int nbw;  // wide form: one block per kThreads packs

  // Iterations per batch graph: even, at least 2.
  static int even_batch(int b) { return std::max(2, b + (b & 1)); }

  Pcg(size_t n_, bool pipelined, bool wide_)
      : n(n_), pipe(pipelined), wide(wide_ && !pipelined), r(n_), z(n_), p(n_), Ap(n_), dinv(n_),
        part(pipelined ? 2 * kPipePart : 3 * kMaxRed), sc(kNSCAL), nb(red_blocks(n_)), nbw(wide_blocks<T>(n_)) {
    if (wide) partw.allocate(2 * static_cast<size_t>(nbw));
    if (pipe) {
      u.allocate(n_);
      w.allocate(n_);
      m.allocate(n_);
      q.allocate(n_);
      sv.allocate(n_);
    }
  }
  ~Pcg() { drop_graphs(); }
  void drop_graphs() {
    for (auto& kv : execs) {
      if (kv.second.exec) (void)hipGraphExecDestroy(kv.second.exec);
      if (kv.second.graph) (void)hipGraphDestroy(kv.second.graph);
    }
    execs.clear();
  }

  // One iteration of parity par (0 or 1; see the scalar slots): the apply,
  // p.Ap partials, the update (alpha formed in every block) and the step end
  // (beta and the state copy formed in every block). The wide form forms
  // alpha and beta once, in one-block kernels between the streaming ones.
  template <class Apply>
  void iteration(Apply&& apply, T* x, int par, hipStream_t s) {
    if (pipe) {
      if (split) {
        first(m.data(), Ap.data(), s);
        pipe_step_kernel<T, true><<<nb, kThreads, 0, s>>>(n, par, nb, sc.data(), part.data(), dinv.data(), sp,
                                                          Ap.data(), x, r.data(), u.data(), w.data(), m.data(),
                                                          z.data(), q.data(), sv.data(), p.data());
      } else {
        apply(m.data(), Ap.data(), s);
        pipe_step_kernel<T, false><<<nb, kThreads, 0, s>>>(n, par, nb, sc.data(), part.data(), dinv.data(), sp,
                                                           Ap.data(), x, r.data(), u.data(), w.data(), m.data(),
                                                           z.data(), q.data(), sv.data(), p.data());
      }
      return;
    }
    apply(p.data(), Ap.data(), s);
    double* pap = part.data() + kPartPAP;
    if (wide && aligned16(x)) {
      double* pw = partw.data();
      dot_wide_kernel<T><<<nbw, kThreads, 0, s>>>(n, p.data(), Ap.data(), pw);
      fold_partials_kernel<<<dim3(nb, 1), kThreads, 0, s>>>(pw, nbw, 0, pap, 0);
      cg_alpha_kernel<<<1, kThreads, 0, s>>>(par, nb, sc.data(), part.data());
      cg_update_wide_kernel<T><<<nbw, kThreads, 0, s>>>(n, par, sc.data(), Ap.data(), dinv.data(), r.data(), z.data(),
                                                        pw, pw + nbw);
      fold_partials_kernel<<<dim3(nb, 2), kThreads, 0, s>>>(pw, nbw, nbw, part.data() + kPartRZ, kPartRR - kPartRZ);
      cg_beta_kernel<<<1, kThreads, 0, s>>>(par, nb, sc.data(), part.data());
      cg_p_wide_kernel<T><<<nbw, kThreads, 0, s>>>(n, par, sc.data(), z.data(), p.data(), x);
    } else if (aligned16(x)) {  // every other vector is the solver's own allocation
      dot_vec_kernel<T><<<nb, kThreads, 0, s>>>(n, p.data(), Ap.data(), pap);
      cg_update_vec_kernel<T><<<nb, kThreads, 0, s>>>(n, par, nb, sc.data(), Ap.data(), dinv.data(), r.data(),
                                                      z.data(), part.data());
      cg_p_vec_kernel<T><<<nb, kThreads, 0, s>>>(n, par, nb, sc.data(), part.data(), z.data(), p.data(), x);
    } else {
      dot_partial_kernel<T><<<nb, kThreads, 0, s>>>(n, p.data(), Ap.data(), pap);
      cg_update_kernel<T><<<nb, kThreads, 0, s>>>(n, par, nb, sc.data(), Ap.data(), dinv.data(), r.data(), z.data(),
                                                  part.data());
      cg_p_kernel<T><<<nb, kThreads, 0, s>>>(n, par, nb, sc.data(), part.data(), z.data(), p.data(), x);
    }
  }

  // x = 0 and the first residual, for A x = b with ||r|| <= rtol ||b|| as the
  // goal within max_iter iterations (the pipelined form also applies A twice:
  // w = A u and the first step's n = A m comes from its first iteration).
  template <class Apply>
  void init(Apply&& apply, const T* b, T* x, double rtol, int max_iter, hipStream_t s) {
    if (pipe) {
      pipe_init_kernel<T><<<nb, kThreads, 0, s>>>(n, b, dinv.data(), x, r.data(), u.data());
      HIP_CHECK_LAUNCH();
      apply(u.data(), w.data(), s);
      pipe_init2_kernel<T><<<nb, kThreads, 0, s>>>(n, r.data(), u.data(), w.data(), dinv.data(), m.data(),
                                                   part.data());
      pipe_init_scalars_kernel<<<1, kThreads, 0, s>>>(part.data(), sc.data(), rtol * rtol, nb, max_iter);
      HIP_CHECK_LAUNCH();
      return;
    }
    cg_init_kernel<T><<<nb, kThreads, 0, s>>>(n, b, dinv.data(), x, r.data(), z.data(), p.data(), part.data());
    cg_init_scalars_kernel<<<1, kThreads, 0, s>>>(part.data(), sc.data(), rtol * rtol, nb, max_iter);
    HIP_CHECK_LAUNCH();
  }

  // Residual replacement: r = rnew, x = 0, p kept; the next segment ends when
  // ||r|| <= rtol ||rnew||.
  void replace(const T* rnew, T* x, double rtol, int max_iter, hipStream_t s) {
    require(!pipe, "mass_solve: residual replacement runs the plain CG form");
    cg_replace_kernel<T><<<nb, kThreads, 0, s>>>(n, rnew, dinv.data(), x, r.data(), z.data(), part.data());
    cg_replace_scalars_kernel<<<1, kThreads, 0, s>>>(part.data(), sc.data(), rtol * rtol, nb, max_iter);
    HIP_CHECK_LAUNCH();
  }

  // Records the graph of `batch_iters` iterations for this b and x and the
  // apply's signature `sig` (if it is not recorded yet: a changed kernel choice
  // or reallocated operator buffer records it again). No other thread does
  // device setup or teardown meanwhile.
  template <class Apply>
  void prepare_graph(Apply&& apply, const T* b, T* x, uint64_t sig, int batch_iters, hipStream_t s) {
    batch_iters = even_batch(batch_iters);
    if (split) sig ^= 0x9e3779b97f4a7c15ull;  // graphs with and without the split differ
    if (!execs.empty() && batch == batch_iters && b_bound == b && x_bound == x && sig_bound == sig) return;
    drop_graphs();
    std::vector<int> sizes{batch_iters};  // even sizes only: the state ends in copy 0
    for (int k = 2; k < batch_iters; k *= 2) sizes.push_back(k);
    ExclusiveCapture lk;
    for (int size : sizes) {
      Exec& e = execs[size];
      HIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      try {
        for (int k = 0; k < size; ++k) iteration(apply, x, k & 1, s);
      } catch (...) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(s, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
      }
      HIP_CHECK(hipStreamEndCapture(s, &e.graph));
      HIP_CHECK(hipGraphInstantiate(&e.exec, e.graph, nullptr, nullptr, 0));
    }
    batch = batch_iters;
    b_bound = b;
    x_bound = x;
    sig_bound = sig;
  }

  // Enqueues `count` iterations (at most batch_iters; the fewest graph
  // launches) and the copy of the scalars into host_sc.
  template <class Apply>
  void launch_iters(Apply&& apply, const T* b, T* x, uint64_t sig, int batch_iters, int count, double* host_sc,
                    hipStream_t s) {
    prepare_graph(apply, b, x, sig, batch_iters, s);
    batch_iters = even_batch(batch_iters);
    // Whole pairs of iterations: one more than asked costs at most one
    // iteration (and nothing once the run is done).
    count = std::max(2, std::min(count + (count & 1), batch_iters));
    for (auto it = execs.rbegin(); it != execs.rend() && count > 0; ++it) {
      while (count >= it->first) {
        HIP_CHECK(hipGraphLaunch(it->second.exec, s));
        count -= it->first;
      }
    }
    HIP_CHECK(hipMemcpyAsync(host_sc, sc.data(), kNSCAL * sizeof(double), hipMemcpyDeviceToHost, s));
  }
};

}  // namespace

struct MassSolverState {
  size_t n = 0;
  DevBuf<double> diag, Mx, r64, part, scal;
  std::unique_ptr<Pcg<double>> pcg64;
  std::unique_ptr<Pcg<float>> pcg32;
  DevBuf<float> r32, d32;
};

void mass_solver_state_destroy(MassSolverState* st) { delete st; }

namespace {

// Pinned host words the device reports into: the CG scalars, ||b||^2, ||b - Mx||^2.
enum { kHostB = kNSCAL, kHostRes, kHostN };

// Pinned host slots for those words, kept for the life of the process and
// reused across solves: hipHostFree waits for the whole device, which would
// stall every other stream's solve each time one finished.
class PinnedSlots {
 public:
  double* acquire() {
    std::lock_guard<std::mutex> lk(m_);
    if (free_.empty()) {
      char* chunk = nullptr;
      {
        SharedCapture cl;
        HIP_CHECK(hipHostMalloc(reinterpret_cast<void**>(&chunk), kChunk * kSlotBytes, hipHostMallocDefault));
      }
      for (int i = kChunk - 1; i >= 0; --i) free_.push_back(reinterpret_cast<double*>(chunk + i * kSlotBytes));
    }
    double* p = free_.back();
    free_.pop_back();
    return p;
  }
  void release(double* p) {
    std::lock_guard<std::mutex> lk(m_);
    free_.push_back(p);
  }

 private:
  static constexpr int kChunk = 256;
  static constexpr size_t kSlotBytes = 256;
  static_assert(kHostN * sizeof(double) <= kSlotBytes, "a slot holds the reported words");
  std::mutex m_;
  std::vector<double*> free_;
};

PinnedSlots& pinned_slots() {
  static PinnedSlots* slots = new PinnedSlots();  // never destroyed: the runtime may be gone at exit
  return *slots;
}

}  // namespace

struct SolveJob {
  enum Phase { kStart, kIter, kFinal, kCheck, kInner, kRelCheck, kRelIter, kRelSpec, kDone };
  Operator* op = nullptr;
  FusedMass* f = nullptr;
  MassSolverState* st = nullptr;
  const double* b = nullptr;
  double* x = nullptr;
  SolveOptions opt;
  hipStream_t s = nullptr;
  hipEvent_t ev = nullptr;
  double* hs = nullptr;  // pinned [kHostN]
  Phase phase = kStart;
  bool in_flight = false;
  int batch = 8, outer = 0, inner_total = 0;
  bool started = false;  // MIXED_RELIABLE: the FP32 recursion has begun
  int spec_iters = 0;    // MIXED_RELIABLE: iterations when the speculative segment began
  // Convergence rate of the CG recursion (factor on r.r per iteration) and the
  // last observation it was measured from; used to launch just the
  // iterations a run still needs instead of whole batches.
  double rate = 0.0, rr_seen = -1.0;
  int it_seen = 0;
  double run_red2 = 0.0;  // r.r reduction the current run stops at (its threshold over its starting r.r)
  double bnorm = 0.0;
  SolveStats stats;
  std::chrono::steady_clock::time_point t0;
  ~SolveJob() {
    if (ev) (void)hipEventDestroy(ev);
    if (hs) {
      if (s) (void)hipStreamSynchronize(s);  // no copy into the slot is still pending
      pinned_slots().release(hs);
    }
  }
};

namespace {

// Enqueues r = b - M x (FP64 fused apply), ||r||^2 into hs[kHostRes] and, if
// r32 is given, r rounded to FP32.
void enqueue_residual(SolveJob& j, float* r32) {
  MassSolverState& st = *j.st;
  fused_mass_apply64(*j.f, j.x, st.Mx.data(), j.s);
  const int nb = red_blocks(st.n);
  residual_kernel<<<nb, kThreads, 0, j.s>>>(st.n, j.b, st.Mx.data(), st.r64.data(), r32, st.part.data());
  sum_to_scalar_kernel<<<1, kThreads, 0, j.s>>>(st.part.data(), st.scal.data(), nb);
  HIP_CHECK_LAUNCH();
  HIP_CHECK(hipMemcpyAsync(j.hs + kHostRes, st.scal.data(), sizeof(double), hipMemcpyDeviceToHost, j.s));
}

void enqueue_bnorm(SolveJob& j) {
  MassSolverState& st = *j.st;
  const int nb = red_blocks(st.n);
  dot_partial_kernel<double><<<nb, kThreads, 0, j.s>>>(st.n, j.b, j.b, st.part.data());
  sum_to_scalar_kernel<<<1, kThreads, 0, j.s>>>(st.part.data(), st.scal.data(), nb);
  HIP_CHECK_LAUNCH();
  HIP_CHECK(hipMemcpyAsync(j.hs + kHostB, st.scal.data(), sizeof(double), hipMemcpyDeviceToHost, j.s));
}

// Iterations to launch after a batch that did not finish: what the observed
// rate needs to bring r.r from its current value to the threshold.
int next_count(SolveJob& j) {
  const double rr = j.hs[kRR], thresh = j.hs[kTHRESH];
  const int it = static_cast<int>(j.hs[kITERS]);
  // The first batch of a run measures from the run's starting r.r, which its
  // threshold gives (threshold = reduction^2 * r.r at the start).
  if (j.rr_seen <= 0.0 && j.run_red2 > 0.0 && thresh > 0.0) j.rr_seen = thresh / j.run_red2;
  if (j.rr_seen > 0.0 && rr > 0.0 && it > j.it_seen) j.rate = std::pow(rr / j.rr_seen, 1.0 / (it - j.it_seen));
  j.rr_seen = rr;
  j.it_seen = it;
  if (!(j.rate > 0.0 && j.rate < 1.0) || !(thresh > 0.0) || rr <= thresh) return j.batch;
  const double need = std::ceil(std::log(thresh / rr) / std::log(j.rate));
  return static_cast<int>(std::max(1.0, std::min(need, static_cast<double>(j.batch))));
}

// Iterations to launch at the start of a run that must reduce r.r by
// `reduction2` (the rate of the previous run, else kFirstProbe iterations, at
// most a batch: enough to measure the rate, few enough that a loose tolerance
// does not run a long batch of finished iterations). The next observation
// starts a new baseline.
constexpr int kFirstProbe = 8;
int start_count(SolveJob& j, double reduction2, int iters_now) {
  j.rr_seen = -1.0;
  j.it_seen = iters_now;
  j.run_red2 = reduction2;
  if (!(j.rate > 0.0 && j.rate < 1.0) || !(reduction2 > 0.0 && reduction2 < 1.0))
    return std::min(j.batch, kFirstProbe);
  const double need = std::ceil(std::log(reduction2) / std::log(j.rate));
  return static_cast<int>(std::max(1.0, std::min(need, static_cast<double>(j.batch))));
}

double rel_of(const SolveJob& j) {
  const double rr = std::sqrt(std::max(j.hs[kHostRes], 0.0));
  return j.bnorm > 0 ? rr / j.bnorm : rr;
}

// Enqueues the next piece of work for the current phase (the previous piece has finished).
void advance(SolveJob& j) {
  MassSolverState& st = *j.st;
  FusedMass& f = *j.f;
  auto apply64 = [&f](const double* in, double* out, hipStream_t ss) { fused_mass_apply64(f, in, out, ss); };
  auto apply32 = [&f](const float* in, float* out, hipStream_t ss) { fused_mass_apply32(f, in, out, ss); };
  switch (j.phase) {
    case SolveJob::kStart:
      j.t0 = std::chrono::steady_clock::now();
      enqueue_bnorm(j);
      if (j.opt.mode == 0) {
        st.pcg64->init(apply64, j.b, j.x, j.opt.rtol, j.opt.max_iter, j.s);
        st.pcg64->launch_iters(apply64, j.b, j.x, fused_signature(f, 0), j.batch,
                               start_count(j, j.opt.rtol * j.opt.rtol, 0), j.hs, j.s);
        j.phase = SolveJob::kIter;
      } else {
        zero_fill(j.x, st.n * sizeof(double), j.s);
        enqueue_residual(j, st.r32.data());
        j.phase = j.opt.mode == 1 ? SolveJob::kCheck : SolveJob::kRelCheck;
      }
      break;
    case SolveJob::kRelCheck: {  // reliable updates: the FP64 residual of x is in hs, its FP32 copy in r32
      j.bnorm = std::sqrt(j.hs[kHostB]);
      const double rel = rel_of(j);
      const int iters = j.started ? static_cast<int>(j.hs[kITERS]) : 0;
      if (rel <= j.opt.rtol || iters >= j.opt.max_iter || j.outer >= j.opt.max_outer) {
        j.stats.rel_residual = rel;
        j.stats.iterations = iters;
        j.stats.outer = j.outer;
        j.phase = SolveJob::kDone;
      } else {
        if (!j.started) {
          st.pcg32->init(apply32, st.r32.data(), st.d32.data(), j.opt.inner_rtol, j.opt.max_iter, j.s);
          j.started = true;
        } else {
          st.pcg32->replace(st.r32.data(), st.d32.data(), j.opt.inner_rtol, j.opt.max_iter, j.s);
        }
        const int count = start_count(j, j.opt.inner_rtol * j.opt.inner_rtol, iters);
        st.pcg32->launch_iters(apply32, st.r32.data(), st.d32.data(), fused_signature(f, 1), j.batch, count, j.hs, j.s);
        j.phase = SolveJob::kRelIter;
      }
      break;
    }
    case SolveJob::kRelIter:
      if (j.hs[kDONE] != 0.0) {
        add_correction_kernel<<<grid_for(st.n, kThreads), kThreads, 0, j.s>>>(st.n, st.d32.data(), j.x);
        HIP_CHECK_LAUNCH();
        ++j.outer;
        enqueue_residual(j, st.r32.data());
        // The recursive residual tracks the true one closely: when it is well
        // above the target, the replacement and the next segment go in behind
        // the FP64 residual at once (one host round trip per segment). If the
        // true residual turns out to have met the target, the speculative
        // segment is dropped: it only changes the FP32 state, never x.
        const int iters = static_cast<int>(j.hs[kITERS]);
        const double rel_rec = j.bnorm > 0 ? std::sqrt(std::max(j.hs[kRR], 0.0)) / j.bnorm : 0.0;
        if (rel_rec > 4.0 * j.opt.rtol && iters < j.opt.max_iter && j.outer < j.opt.max_outer) {
          j.spec_iters = iters;
          st.pcg32->replace(st.r32.data(), st.d32.data(), j.opt.inner_rtol, j.opt.max_iter, j.s);
          const int count = start_count(j, j.opt.inner_rtol * j.opt.inner_rtol, iters);
          st.pcg32->launch_iters(apply32, st.r32.data(), st.d32.data(), fused_signature(f, 1), j.batch, count, j.hs, j.s);
          j.phase = SolveJob::kRelSpec;
        } else {
          j.phase = SolveJob::kRelCheck;
        }
      } else {
        st.pcg32->launch_iters(apply32, st.r32.data(), st.d32.data(), fused_signature(f, 1), j.batch, next_count(j), j.hs,
                               j.s);
      }
      break;
    case SolveJob::kRelSpec: {  // the FP64 residual and a speculative segment have finished
      const double rel = rel_of(j);
      if (rel <= j.opt.rtol || j.spec_iters >= j.opt.max_iter || j.outer >= j.opt.max_outer) {
        j.stats.rel_residual = rel;
        j.stats.iterations = j.spec_iters;
        j.stats.outer = j.outer;
        j.phase = SolveJob::kDone;
        break;
      }
      j.phase = SolveJob::kRelIter;  // the segment is valid: carry on with its scalars
      advance(j);
      return;
    }
    case SolveJob::kIter:  // FP64 PCG batches
      j.bnorm = std::sqrt(j.hs[kHostB]);
      if (j.hs[kDONE] != 0.0) {
        j.stats.iterations = static_cast<int>(j.hs[kITERS]);
        enqueue_residual(j, nullptr);
        j.phase = SolveJob::kFinal;
      } else {
        st.pcg64->launch_iters(apply64, j.b, j.x, fused_signature(f, 0), j.batch, next_count(j), j.hs, j.s);
      }
      break;
    case SolveJob::kFinal:
      j.stats.rel_residual = rel_of(j);
      j.phase = SolveJob::kDone;
      break;
    case SolveJob::kCheck: {  // mixed: FP64 residual of the current x is in hs
      j.bnorm = std::sqrt(j.hs[kHostB]);
      const double rel = rel_of(j);
      if (rel <= j.opt.rtol || j.outer >= j.opt.max_outer) {
        j.stats.rel_residual = rel;
        j.stats.iterations = j.inner_total;
        j.stats.outer = j.outer;
        j.phase = SolveJob::kDone;
      } else {
        // No deeper than the target needs: the last step stops at about half of it.
        const double inner = std::max(j.opt.inner_rtol, std::min(0.5 * j.opt.rtol / rel, 0.5));
        st.pcg32->init(apply32, st.r32.data(), st.d32.data(), inner, j.opt.max_iter, j.s);
        const int count = start_count(j, inner * inner, 0);
        st.pcg32->launch_iters(apply32, st.r32.data(), st.d32.data(), fused_signature(f, 1), j.batch, count, j.hs, j.s);
        j.phase = SolveJob::kInner;
      }
      break;
    }
    case SolveJob::kInner:
      if (j.hs[kDONE] != 0.0) {
        j.inner_total += static_cast<int>(j.hs[kITERS]);
        add_correction_kernel<<<grid_for(st.n, kThreads), kThreads, 0, j.s>>>(st.n, st.d32.data(), j.x);
        HIP_CHECK_LAUNCH();
        ++j.outer;
        enqueue_residual(j, st.r32.data());
        j.phase = SolveJob::kCheck;
      } else {
        st.pcg32->launch_iters(apply32, st.r32.data(), st.d32.data(), fused_signature(f, 1), j.batch, next_count(j), j.hs,
                               j.s);
      }
      break;
    case SolveJob::kDone:
      return;
  }
  if (j.phase != SolveJob::kDone) {
    HIP_CHECK(hipEventRecord(j.ev, j.s));
    j.in_flight = true;
  } else {
    j.stats.ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - j.t0).count();
  }
}

}  // namespace

SolveJob* solve_job_create(Operator& op, MassSolverState*& state, const double* b, double* x, const SolveOptions& opt,
                           hipStream_t s) {
  require(s != nullptr, "solve job: needs a non-default stream (graph capture)");
  require(opt.mode >= 0 && opt.mode <= 2, "mass_solve: unknown mode");
  require(opt.rtol > 0 && opt.max_iter > 0, "mass_solve: rtol and max_iter must be positive");
  require(opt.inner_rtol < 1.0, "mass_solve: inner_rtol must be below 1");
  require(opt.cg_variant >= 0 && opt.cg_variant <= 3, "mass_solve: cg_variant must be 0 (automatic), 1, 2 or 3");
  require(b != nullptr && x != nullptr, "mass_solve: b and x must be device arrays");
  FusedMass& f = operator_fused_mass(op, s);
  const size_t n = static_cast<size_t>(f.l_size);
  if (!state || state->n != n) {
    delete state;
    state = nullptr;
    auto ns = std::make_unique<MassSolverState>();
    ns->n = n;
    ns->diag.allocate(n);
    ns->Mx.allocate(n);
    ns->r64.allocate(n);
    ns->part.allocate(2 * kMaxRed);
    ns->scal.allocate(4);
    fused_mass_diagonal(f, ns->diag.data(), s);
    state = ns.release();
  }
  MassSolverState& st = *state;
  // Everything the captured graphs use is made here, outside capture: the PCG
  // vectors, the Jacobi inverse, and the operator's lazily made state (one apply).
  // Automatic: the pipelined form for FP64 solves below kPipeMaxN entries
  // (where every kernel of an iteration is a few microseconds and their count
  // sets the time), the plain one otherwise: in FP32 the pipelined recurrences
  // drift further from the true residual, and an FP32-inner refinement then
  // needs an extra step (16^3, P=3: 44 iterations in 3 steps against 42 in 2).
  // Residual replacement (MIXED_RELIABLE) always runs the plain form. The
  // plain form runs its wide kernels on vectors of kWideMinBytes and more.
  const bool pipe = opt.mode != 2 && (opt.cg_variant == 2 || (opt.cg_variant == 0 && opt.mode == 0 && n <= kPipeMaxN));
  const size_t vec_bytes = n * (opt.mode == 0 ? sizeof(double) : sizeof(float));
  const bool wide = !pipe && (opt.cg_variant == 3 || (opt.cg_variant == 0 && vec_bytes >= kWideMinBytes));
  if (st.pcg64 && (st.pcg64->pipe != pipe || st.pcg64->wide != wide) && opt.mode == 0) st.pcg64.reset();
  if (st.pcg32 && (st.pcg32->pipe != pipe || st.pcg32->wide != wide) && opt.mode != 0) st.pcg32.reset();
  if (opt.mode == 0 && !st.pcg64) {
    st.pcg64 = std::make_unique<Pcg<double>>(n, pipe, wide);
    recip_kernel<double><<<grid_for(n, kThreads), kThreads, 0, s>>>(n, st.diag.data(), st.pcg64->dinv.data());
    HIP_CHECK_LAUNCH();
  }
  if (opt.mode != 0 && !st.pcg32) {
    st.pcg32 = std::make_unique<Pcg<float>>(n, pipe, wide);
    st.r32.allocate(n);
    st.d32.allocate(n);
    recip_kernel<float><<<grid_for(n, kThreads), kThreads, 0, s>>>(n, st.diag.data(), st.pcg32->dinv.data());
    HIP_CHECK_LAUNCH();
  }
  // The pipelined form folds the apply's second pass into its step kernel
  // where the apply splits (block partial sums, transpose map): one dependent
  // kernel fewer per iteration, the same arithmetic.
  FusedMass* fp = &f;
  if (opt.mode == 0 && st.pcg64->pipe) {
    FusedSplit fs;
    st.pcg64->split = !opt.separate_second_pass && fused_split(f, 0, &fs, s);
    st.pcg64->sp = st.pcg64->split ? fs : FusedSplit{};
    st.pcg64->first = [fp](const double* in, double* out, hipStream_t ss) { fused_mass_apply_first64(*fp, in, out, ss); };
  }
  if (opt.mode != 0 && st.pcg32->pipe) {
    FusedSplit fs;
    st.pcg32->split = !opt.separate_second_pass && fused_split(f, 1, &fs, s);
    st.pcg32->sp = st.pcg32->split ? fs : FusedSplit{};
    st.pcg32->first = [fp](const float* in, float* out, hipStream_t ss) { fused_mass_apply_first32(*fp, in, out, ss); };
  }
  fused_prepare(f, 0, s);  // lazily made operator state, before any capture
  if (opt.mode != 0) fused_prepare(f, 1, s);
  auto j = std::make_unique<SolveJob>();
  j->op = &op;
  j->f = &f;
  j->st = &st;
  j->b = b;
  j->x = x;
  j->opt = opt;
  if (j->opt.inner_rtol <= 0.0) j->opt.inner_rtol = default_inner_rtol(opt.mode);
  j->s = s;
  // Iterations per graph launch: small problems (iterations of a few tens of
  // microseconds) take longer batches, so a solve needs fewer host round trips
  // (the launch count follows the measured convergence rate either way).
  j->batch = opt.batch > 0 ? opt.batch : (n <= kSmallBatchMaxN ? kSmallBatch : 8);
  HIP_CHECK(hipEventCreateWithFlags(&j->ev, hipEventDisableTiming));
  j->hs = pinned_slots().acquire();
  for (int i = 0; i < kHostN; ++i) j->hs[i] = 0.0;
  return j.release();
}

void solve_job_prepare_graphs(SolveJob& j) {
  MassSolverState& st = *j.st;
  FusedMass& f = *j.f;
  if (j.opt.mode == 0) {
    auto apply64 = [&f](const double* in, double* out, hipStream_t ss) { fused_mass_apply64(f, in, out, ss); };
    st.pcg64->prepare_graph(apply64, j.b, j.x, fused_signature(f, 0), j.batch, j.s);
  } else {
    auto apply32 = [&f](const float* in, float* out, hipStream_t ss) { fused_mass_apply32(f, in, out, ss); };
    st.pcg32->prepare_graph(apply32, st.r32.data(), st.d32.data(), fused_signature(f, 1), j.batch, j.s);
  }
}

bool solve_job_step(SolveJob& j) {
  if (j.phase == SolveJob::kDone) return true;
  if (j.in_flight) {
    const hipError_t e = hipEventQuery(j.ev);
    if (e == hipErrorNotReady) return false;
    HIP_CHECK(e);
    j.in_flight = false;
  }
  advance(j);
  return j.phase == SolveJob::kDone;
}

void solve_job_wait(SolveJob& j) {
  while (j.phase != SolveJob::kDone) {
    if (j.in_flight) {
      HIP_CHECK(hipEventSynchronize(j.ev));
      j.in_flight = false;
    }
    advance(j);
  }
}

const SolveStats& solve_job_stats(const SolveJob& j) { return j.stats; }
bool solve_job_done(const SolveJob& j) { return j.phase == SolveJob::kDone; }
void solve_job_destroy(SolveJob* j) {
  if (j && j->in_flight) (void)hipEventSynchronize(j->ev);
  delete j;
}

namespace {

// Streams for solves asked for on the null stream (graph capture needs another
// one): taken from a pool per device and returned after the solve, so a solve
// does not create and destroy a stream (which costs more than a small solve's
// setup). Each comes with an event that orders it after the null stream's
// earlier work. Kept for the life of the process (the runtime may be gone at exit).
class SolveStreamPool {
 public:
  struct Entry {
    int dev;
    hipStream_t s;
    hipEvent_t ev;
  };
  Entry acquire() {
    int dev = 0;
    HIP_CHECK(hipGetDevice(&dev));
    {
      std::lock_guard<std::mutex> lk(m_);
      for (size_t i = 0; i < free_.size(); ++i)
        if (free_[i].dev == dev) {
          Entry e = free_[i];
          free_.erase(free_.begin() + static_cast<std::ptrdiff_t>(i));
          return e;
        }
    }
    Entry e{dev, nullptr, nullptr};
    HIP_CHECK(hipStreamCreateWithFlags(&e.s, hipStreamNonBlocking));
    HIP_CHECK(hipEventCreateWithFlags(&e.ev, hipEventDisableTiming));
    return e;
  }
  void release(const Entry& e) {
    std::lock_guard<std::mutex> lk(m_);
    free_.push_back(e);
  }

 private:
  std::mutex m_;
  std::vector<Entry> free_;
};

SolveStreamPool& solve_streams() {
  static SolveStreamPool* pool = new SolveStreamPool();  // never destroyed
  return *pool;
}

}  // namespace

void mass_solve(Operator& op, MassSolverState*& state, const double* b, double* x, const SolveOptions& opt,
                SolveStats& stats, hipStream_t user_stream) {
  // Graph capture needs a non-default stream: a null stream is served by a
  // pooled one, ordered after the null stream's earlier work (the caller's
  // writes of b and x). The solve is synchronous: x is complete on return.
  hipStream_t s = user_stream;
  struct PoolGuard {
    bool held = false;
    SolveStreamPool::Entry e{};
    ~PoolGuard() {
      if (held) {
        (void)hipStreamSynchronize(e.s);
        solve_streams().release(e);
      }
    }
  } guard;
  if (s == nullptr) {
    guard.e = solve_streams().acquire();
    guard.held = true;
    s = guard.e.s;
    HIP_CHECK(hipEventRecord(guard.e.ev, nullptr));
    HIP_CHECK(hipStreamWaitEvent(s, guard.e.ev, 0));
  }
  std::unique_ptr<SolveJob, void (*)(SolveJob*)> j(solve_job_create(op, state, b, x, opt, s), solve_job_destroy);
  solve_job_wait(*j);
  stats = j->stats;
}

}  // namespace femhip
