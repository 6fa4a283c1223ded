// Batched mass solves (see batch_solve.hpp).
#include "batch_solve.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <iterator>
#include <mutex>
#include <string>
#include <type_traits>

#include "cg_kernels.hpp"

namespace femhip {

bool BatchKey::operator<(const BatchKey& o) const {
  return std::tie(P, Q, NC, fp32, elems) < std::tie(o.P, o.Q, o.NC, o.fp32, o.elems);
}

namespace {

std::map<BatchKey, BatchElemKernel>& batch_registry() {
  static std::map<BatchKey, BatchElemKernel> r;
  return r;
}

const BatchElemKernel* find_batch(const BatchKey& k) {
  auto it = batch_registry().find(k);
  return it == batch_registry().end() ? nullptr : &it->second;
}

// The step kernel: work item i is reduction block i - red0 of the run whose
// range holds it, running that run's pipelined step (row sums of its element
// outputs folded in). Items split into 8 contiguous ranges, one per XCD.
template <class T>
__global__ __launch_bounds__(kThreads) void batch_step_kernel(const BatchTable* __restrict__ tab, int par) {
  __shared__ double sh[3][kThreads / 64];
  __shared__ double s_val[3];
  const int items = tab->red_items, ns = tab->nslots;
  const int per = static_cast<int>(gridDim.x) / 8, xcd = static_cast<int>(blockIdx.x) % 8;
  const int chunk = (items + 7) / 8;
  const int hi = min(items, (xcd + 1) * chunk);
  for (int it = xcd * chunk + static_cast<int>(blockIdx.x) / 8; it < hi; it += per) {
    int k = 0;
    while (k + 1 < ns && tab->slot[k + 1].red0 <= it) ++k;
    const BatchSlot& sl = tab->slot[k];
    pipe_step<T, true>(sl.n, par, sl.nb, it - sl.red0, sl.sc, sl.part, static_cast<const T*>(sl.dinv), sl.sp,
                       static_cast<const T*>(sl.nv), static_cast<T*>(sl.x), static_cast<T*>(sl.r),
                       static_cast<T*>(sl.u), static_cast<T*>(sl.w), static_cast<T*>(sl.m), static_cast<T*>(sl.z),
                       static_cast<T*>(sl.q), static_cast<T*>(sl.s), static_cast<T*>(sl.p), sh, s_val);
    __syncthreads();  // the next item reuses sh and s_val
  }
}

// Copies each run's scalar slots (state copy 0 after an even count of
// iterations) to out[slot][kNSCAL], for one read-back per replay.
__global__ void batch_gather_kernel(const BatchTable* __restrict__ tab, double* __restrict__ out) {
  const int k = blockIdx.x, i = threadIdx.x;
  if (k < tab->nslots && i < kNSCAL) out[k * kNSCAL + i] = tab->slot[k].sc[i];
}

template <class T>
int step_blocks_per_cu() {
  int n = 0;
  HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, batch_step_kernel<T>, kThreads, 0));
  return n;
}

// Pinned host words a job's FP64 residual and ||b||^2 are read into; kept for
// the life of the process (hipHostFree waits for the whole device).
class PinnedWords {
 public:
  double* acquire() {
    std::lock_guard<std::mutex> lk(m_);
    if (free_.empty()) {
      double* chunk = nullptr;
      {
        SharedCapture cl;
        HIP_CHECK(hipHostMalloc(reinterpret_cast<void**>(&chunk), kChunk * kWords * sizeof(double),
                                hipHostMallocDefault));
      }
      for (int i = kChunk - 1; i >= 0; --i) free_.push_back(chunk + i * kWords);
    }
    double* p = free_.back();
    free_.pop_back();
    return p;
  }
  void release(double* p) {
    std::lock_guard<std::mutex> lk(m_);
    free_.push_back(p);
  }
  static constexpr int kWords = 4;

 private:
  static constexpr int kChunk = 256;
  std::mutex m_;
  std::vector<double*> free_;
};
PinnedWords& pinned_words() {
  static PinnedWords* w = new PinnedWords();  // never destroyed: the runtime may be gone at exit
  return *w;
}

// The vectors and scalars of one pipelined PCG run in precision T (x and b are
// the caller's: the job's own for FP64, the refinement's correction and
// residual for FP32).
template <class T>
struct RunVecs {
  size_t n;
  int nb;
  DevBuf<T> r, u, w, m, z, q, s, p, dinv;
  DevBuf<double> part, sc;
  explicit RunVecs(size_t n_)
      : n(n_), nb(red_blocks(n_)), r(n_), u(n_), w(n_), m(n_), z(n_), q(n_), s(n_), p(n_), dinv(n_),
        part(2 * kPipePart), sc(kNSCAL) {}
};

struct Cohort;

}  // namespace

void register_batch_elem(const BatchKey& k, const BatchElemKernel& kern) { batch_registry()[k] = kern; }

std::vector<int> batch_elem_variants(int P, int Q, int NC, int fp32) {
  std::vector<int> v;
  for (const auto& kv : batch_registry())
    if (kv.first.P == P && kv.first.Q == Q && kv.first.NC == NC && kv.first.fp32 == fp32) v.push_back(kv.first.elems);
  return v;
}

bool batch_supported(int P, int Q, int NC, int mode) {
  if (mode != 0 && mode != 1) return false;
  return !batch_elem_variants(P, Q, NC, mode == 1 ? 1 : 0).empty();
}

struct BatchJob {
  enum Phase { kStart, kJoin, kRun, kCheck, kFinal, kDone };
  FusedMass* f = nullptr;
  const double* b = nullptr;
  double* x = nullptr;
  size_t n = 0;
  SolveOptions opt;
  int fp32 = 0, NE = 0;
  std::unique_ptr<RunVecs<double>> v64;
  std::unique_ptr<RunVecs<float>> v32;
  DevBuf<float> r32, d32;                 // mixed: the inner run's right-hand side and solution
  DevBuf<double> Mx, r64, rpart, rscal;  // the FP64 residual b - M x and its squared norm
  FusedSplit sp;
  Phase phase = kStart;
  bool in_flight = false;
  hipEvent_t ev = nullptr;
  hipEvent_t ready = nullptr;  // the end of the job's setup on its creation stream
  double* hs = nullptr;        // pinned: [0] ||b - M x||^2, [1] ||b||^2
  double bnorm = -1.0;         // ||b||, from hs[1] once a residual has been read back
  int outer = 0, inner_total = 0;
  // Convergence rate of the run (factor on r.r per iteration) and the last
  // observation it was measured from; sets how many iterations a replay takes.
  double rate = 0.0, rr_seen = -1.0, run_red2 = 0.0;
  int it_seen = 0;
  SolveStats stats;
  int status = 0;
  std::string error;
  Cohort* cohort = nullptr;
  std::chrono::steady_clock::time_point t0;
  ~BatchJob() {
    if (ev) (void)hipEventDestroy(ev);
    if (ready) (void)hipEventDestroy(ready);
    if (hs) pinned_words().release(hs);
  }
};

namespace {

int ipow3(int v) { return v * v * v; }

// The NE of the batched element kernel: the tuned packing of the
// transpose-map algorithm where a batched kernel has it, else the compiled
// one nearest 16.
int choose_elems(const FusedMass& f, int fp32) {
  const std::vector<int> v = batch_elem_variants(f.P, f.Q, f.NC, fp32);
  require(!v.empty(), "batch: no batched element kernel for this shape");
  {
    std::lock_guard<std::mutex> lk(tune_mutex());
    auto it = tune_alg_table().find(TuneAlgKey{f.P, f.Q, f.NC, fp32, kAlgElemTmap});
    if (it != tune_alg_table().end() && std::find(v.begin(), v.end(), it->second.elems) != v.end())
      return it->second.elems;
  }
  int best = v[0];
  for (int e : v)
    if (std::abs(e - 16) < std::abs(best - 16)) best = e;
  return best;
}

template <class T>
RunVecs<T>& vecs(BatchJob& j);
template <>
RunVecs<double>& vecs<double>(BatchJob& j) {
  return *j.v64;
}
template <>
RunVecs<float>& vecs<float>(BatchJob& j) {
  return *j.v32;
}

template <class T>
void apply_t(FusedMass& f, const T* in, T* out, hipStream_t s) {
  if constexpr (std::is_same_v<T, float>) {
    fused_mass_apply32(f, in, out, s);
  } else {
    fused_mass_apply64(f, in, out, s);
  }
}

// A run's start (pipe_init, w = A u with the full apply, the parity-0
// partials and scalars): the state a replay's first iteration reads.
template <class T>
void enqueue_run_start(BatchJob& j, const T* b, T* x, double rtol, hipStream_t s) {
  RunVecs<T>& v = vecs<T>(j);
  pipe_init_kernel<T><<<v.nb, kThreads, 0, s>>>(v.n, b, v.dinv.data(), x, v.r.data(), v.u.data());
  HIP_CHECK_LAUNCH();
  apply_t<T>(*j.f, v.u.data(), v.w.data(), s);
  pipe_init2_kernel<T><<<v.nb, kThreads, 0, s>>>(v.n, v.r.data(), v.u.data(), v.w.data(), v.dinv.data(),
                                                 v.m.data(), v.part.data());
  pipe_init_scalars_kernel<<<1, kThreads, 0, s>>>(v.part.data(), v.sc.data(), rtol * rtol, v.nb, j.opt.max_iter);
  HIP_CHECK_LAUNCH();
}

// r = b - M x in FP64 (and its FP32 copy), ||r||^2 into hs[0], then the job's event.
void enqueue_residual(BatchJob& j, float* r32, hipStream_t s) {
  fused_mass_apply64(*j.f, j.x, j.Mx.data(), s);
  const int nb = red_blocks(j.n);
  residual_kernel<<<nb, kThreads, 0, s>>>(j.n, j.b, j.Mx.data(), j.r64.data(), r32, j.rpart.data());
  sum_to_scalar_kernel<<<1, kThreads, 0, s>>>(j.rpart.data(), j.rscal.data(), nb);
  HIP_CHECK_LAUNCH();
  // With ||b||^2, which setup left in rscal[1].
  HIP_CHECK(hipMemcpyAsync(j.hs, j.rscal.data(), 2 * sizeof(double), hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipEventRecord(j.ev, s));
  j.in_flight = true;
}

double rel_of(BatchJob& j) {
  if (j.bnorm < 0) j.bnorm = std::sqrt(std::max(j.hs[1], 0.0));
  const double rr = std::sqrt(std::max(j.hs[0], 0.0));
  return j.bnorm > 0 ? rr / j.bnorm : rr;
}

template <class T>
void fill_slot(BatchJob& j, BatchSlot& sl) {
  RunVecs<T>& v = vecs<T>(j);
  FusedMass& f = *j.f;
  sl.offsets = f.rstr->offsets.data();
  sl.qdata = std::is_same_v<T, float> ? static_cast<const void*>(f.qdata32.data())
                                      : static_cast<const void*>(f.qdata64.data());
  sl.y = const_cast<void*>(j.sp.vals);
  sl.y_comp_len = static_cast<size_t>(f.num_elem) * ipow3(f.P);
  sl.num_elem = f.num_elem;
  sl.comp_stride = f.comp_stride;
  sl.nb = v.nb;
  sl.n = v.n;
  sl.sc = v.sc.data();
  sl.part = v.part.data();
  sl.dinv = v.dinv.data();
  sl.sp = j.sp;
  if constexpr (std::is_same_v<T, float>) {
    sl.x = j.d32.data();
  } else {
    sl.x = j.x;
  }
  sl.r = v.r.data();
  sl.u = v.u.data();
  sl.w = v.w.data();
  sl.m = v.m.data();
  sl.nv = nullptr;  // every row of the transpose map is a node: the step never reads it
  sl.z = v.z.data();
  sl.q = v.q.data();
  sl.s = v.s.data();
  sl.p = v.p.data();
}

// The runs of one shape and precision, the graph that advances them, and the
// stream every piece of their jobs' device work is ordered on.
struct Cohort {
  int P = 0, Q = 0, NC = 0, fp32 = 0, NE = 0;
  std::vector<double> basis;  // interp_1d of the shape (part of the key)
  std::vector<float> basis32;
  BatchElemKernel kern;
  int iters = 8, grid_elem = 0, grid_step = 0;
  hipStream_t s = nullptr;
  hipEvent_t ev = nullptr;
  BatchTable* h_tab = nullptr;  // pinned
  BatchTable* d_tab = nullptr;
  double* h_sc = nullptr;  // pinned [kBatchSlots][kNSCAL]
  double* d_sc = nullptr;
  // Graphs of 4, 8, ... up to `iters` iterations (even counts: the state ends
  // in copy 0), so a replay runs about what its members still need.
  struct Exec {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
  };
  std::map<int, Exec> execs;
  std::vector<BatchJob*> members;   // the runs in the cohort
  std::vector<BatchJob*> launched;  // the table of the replay in flight (h_sc's order)
  std::vector<BatchJob*> joining;   // started, waiting for a table entry
  std::map<const BatchJob*, int> need;  // iterations each member still needs (estimated)
  bool in_flight = false;
  long long replays = 0;

  bool matches(const BatchJob& j) const {
    return P == j.f->P && Q == j.f->Q && NC == j.f->NC && fp32 == j.fp32 && NE == j.NE && basis == j.f->B64;
  }

  Cohort(const BatchJob& j, int iters_per_launch) {
    P = j.f->P;
    Q = j.f->Q;
    NC = j.f->NC;
    fp32 = j.fp32;
    NE = j.NE;
    basis = j.f->B64;
    basis32 = j.f->B32;
    iters = std::max(2, iters_per_launch + (iters_per_launch & 1));  // even: the state ends in copy 0
    const BatchElemKernel* k = find_batch(BatchKey{P, Q, NC, fp32, NE});
    require(k != nullptr, "batch: element kernel not compiled");
    kern = *k;
    int dev = 0, cus = 0;
    HIP_CHECK(hipGetDevice(&dev));
    HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    grid_elem = std::max(8, kern.blocks_per_cu() * cus / 8 * 8);
    grid_step = std::max(8, (fp32 ? step_blocks_per_cu<float>() : step_blocks_per_cu<double>()) * cus / 8 * 8);
    {
      SharedCapture lk;
      HIP_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
      HIP_CHECK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
      HIP_CHECK(hipHostMalloc(reinterpret_cast<void**>(&h_tab), sizeof(BatchTable), hipHostMallocDefault));
      HIP_CHECK(hipHostMalloc(reinterpret_cast<void**>(&h_sc), sizeof(double) * kBatchSlots * kNSCAL,
                              hipHostMallocDefault));
      HIP_CHECK(hipMalloc(reinterpret_cast<void**>(&d_tab), sizeof(BatchTable)));
      HIP_CHECK(hipMalloc(reinterpret_cast<void**>(&d_sc), sizeof(double) * kBatchSlots * kNSCAL));
    }
    *h_tab = BatchTable();
    HIP_CHECK(hipMemcpyAsync(d_tab, h_tab, sizeof(BatchTable), hipMemcpyHostToDevice, s));
    HIP_CHECK(hipStreamSynchronize(s));
    const void* bptr = fp32 ? static_cast<const void*>(basis32.data()) : static_cast<const void*>(basis.data());
    std::vector<int> sizes{iters};
    for (int k = 4; k < iters; k *= 2) sizes.push_back(k);
    ExclusiveCapture lk;
    for (int size : sizes) {
      Exec& e = execs[size];
      HIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      try {
        for (int it = 0; it < size; ++it) {
          kern.launch(d_tab, it & 1, bptr, grid_elem, s);
          if (fp32) {
            batch_step_kernel<float><<<grid_step, kThreads, 0, s>>>(d_tab, it & 1);
          } else {
            batch_step_kernel<double><<<grid_step, kThreads, 0, s>>>(d_tab, it & 1);
          }
          HIP_CHECK_LAUNCH();
        }
        batch_gather_kernel<<<kBatchSlots, 64, 0, s>>>(d_tab, d_sc);
        HIP_CHECK_LAUNCH();
      } catch (...) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(s, &g);
        if (g) (void)hipGraphDestroy(g);
        release();
        throw;
      }
      HIP_CHECK(hipStreamEndCapture(s, &e.graph));
      HIP_CHECK(hipGraphInstantiate(&e.exec, e.graph, nullptr, nullptr, 0));
    }
  }

  void release() noexcept {
    if (s) (void)hipStreamSynchronize(s);
    for (auto& kv : execs) {
      if (kv.second.exec) (void)hipGraphExecDestroy(kv.second.exec);
      if (kv.second.graph) (void)hipGraphDestroy(kv.second.graph);
    }
    execs.clear();
    SharedCapture lk;
    if (d_tab) (void)hipFree(d_tab);
    if (d_sc) (void)hipFree(d_sc);
    if (h_tab) (void)hipHostFree(h_tab);
    if (h_sc) (void)hipHostFree(h_sc);
    if (ev) (void)hipEventDestroy(ev);
    if (s) (void)hipStreamDestroy(s);
    d_tab = nullptr;
    d_sc = nullptr;
    h_tab = nullptr;
    h_sc = nullptr;
    ev = nullptr;
    s = nullptr;
  }
  ~Cohort() { release(); }

  // Rewrites the table with the members (their item ranges in turn), then
  // enqueues it, one replay of the graph with the fewest iterations at least
  // `need` (the largest if none) and the read-back of the scalars.
  void launch(int need) {
    BatchTable& t = *h_tab;
    t.nslots = static_cast<int>(members.size());
    int elem = 0, red = 0;
    for (size_t i = 0; i < members.size(); ++i) {
      BatchJob& j = *members[i];
      BatchSlot& sl = t.slot[i];
      if (fp32) {
        fill_slot<float>(j, sl);
      } else {
        fill_slot<double>(j, sl);
      }
      sl.elem0 = elem;
      sl.red0 = red;
      elem += (sl.num_elem + NE - 1) / NE;
      red += sl.nb;
    }
    t.elem_items = elem;
    t.red_items = red;
    HIP_CHECK(hipMemcpyAsync(d_tab, h_tab, sizeof(BatchTable), hipMemcpyHostToDevice, s));
    auto it = execs.lower_bound(need);
    if (it == execs.end()) it = std::prev(execs.end());
    HIP_CHECK(hipGraphLaunch(it->second.exec, s));
    HIP_CHECK(hipMemcpyAsync(h_sc, d_sc, sizeof(double) * kBatchSlots * kNSCAL, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipEventRecord(ev, s));
    launched = members;
    in_flight = true;
    ++replays;
  }
};

}  // namespace

BatchJob* batch_job_create(Operator& op, const double* b, double* x, const SolveOptions& opt, hipStream_t s) {
  require(opt.mode == 0 || opt.mode == 1, "batch: modes 0 (FP64) and 1 (mixed refinement) only");
  require(opt.rtol > 0 && opt.max_iter > 0, "mass_solve: rtol and max_iter must be positive");
  require(opt.inner_rtol < 1.0, "mass_solve: inner_rtol must be below 1");
  require(b != nullptr && x != nullptr, "mass_solve: b and x must be device arrays");
  auto j = std::make_unique<BatchJob>();
  j->opt = opt;
  if (j->opt.inner_rtol <= 0.0) j->opt.inner_rtol = default_inner_rtol(opt.mode);
  j->fp32 = opt.mode == 1 ? 1 : 0;
  // The transpose-map algorithm, deterministic, FMA arithmetic.
  op.opt_alg = kAlgElemTmap;
  op.opt_exact = 0;
  op.opt_scatter = kScatterDeterministic;
  op.opt_elems = 0;
  FusedMass* f = &operator_fused_mass(op, s);
  if (!batch_supported(f->P, f->Q, f->NC, opt.mode)) throw Unsupported("batch: no batched element kernel for this shape");
  j->NE = choose_elems(*f, j->fp32);
  op.opt_elems = j->NE;
  f = &operator_fused_mass(op, s);
  j->f = f;
  j->b = b;
  j->x = x;
  j->n = static_cast<size_t>(f->l_size);
  const size_t n = j->n;
  fused_prepare(*f, 0, s);
  if (j->fp32) fused_prepare(*f, 1, s);
  for (int p = 0; p <= j->fp32; ++p) {
    const FusedKey k = fused_resolve(*f, p);
    require(k.alg == kAlgElemTmap && !k.exact, "batch: the operator did not take the transpose-map algorithm");
  }
  if (!fused_split(*f, j->fp32, &j->sp, s) || j->sp.node_row != nullptr)
    throw Unsupported("batch: the operator's transpose map does not cover every node in order");
  StreamTmp<double> diag(n, s);  // freed in s's order: nothing here waits for the device
  fused_mass_diagonal(*f, diag.data(), s);
  if (j->fp32) {
    j->v32 = std::make_unique<RunVecs<float>>(n);
    recip_kernel<float><<<grid_for(n, kThreads), kThreads, 0, s>>>(n, diag.data(), j->v32->dinv.data());
    j->r32.allocate(n);
    j->d32.allocate(n);
  } else {
    j->v64 = std::make_unique<RunVecs<double>>(n);
    recip_kernel<double><<<grid_for(n, kThreads), kThreads, 0, s>>>(n, diag.data(), j->v64->dinv.data());
  }
  HIP_CHECK_LAUNCH();
  j->Mx.allocate(n);
  j->r64.allocate(n);
  j->rpart.allocate(kMaxRed);
  j->rscal.allocate(2);
  HIP_CHECK(hipEventCreateWithFlags(&j->ev, hipEventDisableTiming));
  HIP_CHECK(hipEventCreateWithFlags(&j->ready, hipEventDisableTiming));
  j->hs = pinned_words().acquire();
  const int nb = red_blocks(n);
  dot_partial_kernel<double><<<nb, kThreads, 0, s>>>(n, b, b, j->rpart.data());
  sum_to_scalar_kernel<<<1, kThreads, 0, s>>>(j->rpart.data(), j->rscal.data() + 1, nb);
  HIP_CHECK_LAUNCH();
  // No host wait: the engine orders the cohort's stream after this event, and
  // ||b||^2 comes back with the job's first residual.
  HIP_CHECK(hipEventRecord(j->ready, s));
  return j.release();
}

const SolveStats& batch_job_stats(const BatchJob& j) { return j.stats; }
int batch_job_status(const BatchJob& j) { return j.status; }
const std::string& batch_job_error(const BatchJob& j) { return j.error; }
void batch_job_destroy(BatchJob* j) {
  if (j && j->in_flight) (void)hipEventSynchronize(j->ev);
  delete j;
}

class BatchEngine {
 public:
  explicit BatchEngine(int iters) : iters_(iters) {}
  ~BatchEngine() {
    for (auto& c : cohorts_)
      if (c->s) (void)hipStreamSynchronize(c->s);
  }

  void add(BatchJob* j) {
    j->phase = BatchJob::kStart;
    jobs_.push_back(j);
  }
  size_t jobs() const { return jobs_.size(); }

  bool poll(std::vector<BatchJob*>& done) {
    bool progressed = false;
    for (BatchJob* j : jobs_) {
      if (j->phase == BatchJob::kDone) continue;
      try {
        if (j->phase == BatchJob::kStart) {
          start(*j);
          progressed = true;
        } else if (j->in_flight) {
          const hipError_t e = hipEventQuery(j->ev);
          if (e == hipErrorNotReady) continue;
          HIP_CHECK(e);
          j->in_flight = false;
          on_event(*j);
          progressed = true;
        }
      } catch (const std::exception& e) {
        fail(*j, e);
        progressed = true;
      }
    }
    for (auto& c : cohorts_) {
      try {
        progressed |= poll_cohort(*c);
      } catch (const std::exception& e) {
        // A failed replay or launch: every job of the cohort fails with it.
        (void)hipStreamSynchronize(c->s);
        std::vector<BatchJob*> all = c->members;
        all.insert(all.end(), c->joining.begin(), c->joining.end());
        for (BatchJob* j : all) fail(*j, e);
        c->in_flight = false;
        progressed = true;
      }
    }
    for (size_t i = 0; i < jobs_.size();) {
      if (jobs_[i]->phase == BatchJob::kDone) {
        done.push_back(jobs_[i]);
        jobs_.erase(jobs_.begin() + static_cast<std::ptrdiff_t>(i));
      } else {
        ++i;
      }
    }
    return progressed;
  }

 private:
  Cohort& cohort_for(const BatchJob& j) {
    for (auto& c : cohorts_)
      if (c->matches(j)) return *c;
    cohorts_.push_back(std::make_unique<Cohort>(j, iters_));
    return *cohorts_.back();
  }

  static int status_of(const std::exception& e) {
    if (dynamic_cast<const std::invalid_argument*>(&e)) return 1;
    if (dynamic_cast<const DeviceError*>(&e)) return 2;
    if (dynamic_cast<const Unsupported*>(&e)) return 4;
    if (dynamic_cast<const OutOfMemory*>(&e)) return 5;
    return 3;
  }

  // Marks a job failed and takes it out of its cohort; its device work has
  // finished when the job is handed back (its buffers may then be freed).
  void fail(BatchJob& j, const std::exception& e) {
    if (j.status == 0) {
      j.status = status_of(e);
      j.error = e.what();
    }
    if (Cohort* c = j.cohort) {
      if (c->s) (void)hipStreamSynchronize(c->s);
      c->members.erase(std::remove(c->members.begin(), c->members.end(), &j), c->members.end());
      c->need.erase(&j);
      c->joining.erase(std::remove(c->joining.begin(), c->joining.end(), &j), c->joining.end());
    }
    if (j.in_flight) (void)hipEventSynchronize(j.ev);
    j.in_flight = false;
    j.phase = BatchJob::kDone;
  }

  void finish(BatchJob& j, double rel) {
    j.stats.rel_residual = rel;
    j.stats.outer = j.outer;
    j.stats.ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - j.t0).count();
    j.phase = BatchJob::kDone;
  }

  void start(BatchJob& j) {
    Cohort& c = cohort_for(j);
    j.cohort = &c;
    j.t0 = std::chrono::steady_clock::now();
    HIP_CHECK(hipStreamWaitEvent(c.s, j.ready, 0));  // the job's setup has run
    if (!j.fp32) {
      enqueue_run_start<double>(j, j.b, j.x, j.opt.rtol, c.s);
      j.run_red2 = j.opt.rtol * j.opt.rtol;
      j.phase = BatchJob::kJoin;
      c.joining.push_back(&j);
    } else {
      zero_fill(j.x, j.n * sizeof(double), c.s);
      enqueue_residual(j, j.r32.data(), c.s);
      j.phase = BatchJob::kCheck;
    }
  }

  // The job's FP64 residual has been read back.
  void on_event(BatchJob& j) {
    const double rel = rel_of(j);
    if (j.phase == BatchJob::kFinal) {
      finish(j, rel);
      return;
    }
    // kCheck (mixed): done, or the next inner FP32 run on r32.
    if (rel <= j.opt.rtol || j.outer >= j.opt.max_outer) {
      j.stats.iterations = j.inner_total;
      finish(j, rel);
      return;
    }
    // No deeper than the target needs: the last step stops at about half of it.
    const double inner = std::max(j.opt.inner_rtol, std::min(0.5 * j.opt.rtol / rel, 0.5));
    enqueue_run_start<float>(j, j.r32.data(), j.d32.data(), inner, j.cohort->s);
    j.run_red2 = inner * inner;
    j.phase = BatchJob::kJoin;
    j.cohort->joining.push_back(&j);
  }

  // A run has met its test (sc: its scalars read back after the replay).
  void on_run_done(BatchJob& j, const double* sc) {
    Cohort& c = *j.cohort;
    const int iters = static_cast<int>(sc[kITERS]);
    if (!j.fp32) {
      j.stats.iterations = iters;
      enqueue_residual(j, nullptr, c.s);
      j.phase = BatchJob::kFinal;
    } else {
      j.inner_total += iters;
      add_correction_kernel<<<grid_for(j.n, kThreads), kThreads, 0, c.s>>>(j.n, j.d32.data(), j.x);
      HIP_CHECK_LAUNCH();
      ++j.outer;
      enqueue_residual(j, j.r32.data(), c.s);
      j.phase = BatchJob::kCheck;
    }
  }

  // Iterations a run still needs, from its measured rate (kFirstProbe if none yet).
  static constexpr int kFirstProbe = 8;
  // ceil(log(reduction) / log(rate)) iterations, at least 1 and at most
  // kMostNeed (a rate within rounding of 1 would overflow an int).
  static constexpr double kMostNeed = 1 << 20;
  static int iterations_for(double reduction, double rate) {
    const double k = std::ceil(std::log(reduction) / std::log(rate));
    return static_cast<int>(std::max(1.0, std::min(k, kMostNeed)));
  }
  static int need_of(const BatchJob& j, const double* sc) {
    const double rr = sc[kRR], thresh = sc[kTHRESH];
    if (!(j.rate > 0.0 && j.rate < 1.0) || !(thresh > 0.0)) return kFirstProbe;
    if (rr <= thresh) return 1;
    return iterations_for(thresh / rr, j.rate);
  }
  // Updates a run's rate from its scalars after a replay. The first
  // observation of a run measures from its starting r.r (threshold / reduction^2).
  static void observe(BatchJob& j, const double* sc) {
    const double rr = sc[kRR], thresh = sc[kTHRESH];
    const int it = static_cast<int>(sc[kITERS]);
    if (j.rr_seen <= 0.0 && j.run_red2 > 0.0 && thresh > 0.0) j.rr_seen = thresh / j.run_red2;
    if (j.rr_seen > 0.0 && rr > 0.0 && it > j.it_seen) j.rate = std::pow(rr / j.rr_seen, 1.0 / (it - j.it_seen));
    j.rr_seen = rr;
    j.it_seen = it;
  }

  bool poll_cohort(Cohort& c) {
    bool progressed = false;
    if (c.in_flight) {
      const hipError_t e = hipEventQuery(c.ev);
      if (e == hipErrorNotReady) return false;
      HIP_CHECK(e);
      c.in_flight = false;
      progressed = true;
      // h_sc holds the scalars of the runs in the order they were launched
      // (a run that has failed since has left the members and is skipped).
      const std::vector<BatchJob*> launched = c.launched;
      for (size_t i = 0; i < launched.size(); ++i) {
        BatchJob& j = *launched[i];
        if (std::find(c.members.begin(), c.members.end(), &j) == c.members.end()) continue;
        const double* sc = c.h_sc + i * kNSCAL;
        observe(j, sc);
        if (sc[kDONE] == 0.0) {
          c.need[&j] = need_of(j, sc);
          continue;
        }
        c.members.erase(std::remove(c.members.begin(), c.members.end(), &j), c.members.end());
        c.need.erase(&j);
        try {
          on_run_done(j, sc);
        } catch (const std::exception& ex) {
          fail(j, ex);
        }
      }
    }
    while (!c.joining.empty() && c.members.size() < static_cast<size_t>(kBatchSlots)) {
      BatchJob* j = c.joining.front();
      c.joining.erase(c.joining.begin());
      j->phase = BatchJob::kRun;
      // A new run: its first replay follows the previous run's rate (refinement), else probes.
      j->rr_seen = -1.0;
      j->it_seen = 0;
      c.need[j] = (j->rate > 0.0 && j->rate < 1.0 && j->run_red2 > 0.0 && j->run_red2 < 1.0)
                      ? iterations_for(j->run_red2, j->rate)
                      : kFirstProbe;
      c.members.push_back(j);
    }
    if (!c.members.empty()) {
      // The fewest iterations any member still needs (at least 4: shorter
      // replays cost more in launches and read-backs than they save).
      int need = 1 << 30;
      for (BatchJob* j : c.members) need = std::min(need, c.need.count(j) ? c.need[j] : kFirstProbe);
      c.launch(std::max(4, need));
      progressed = true;
    }
    return progressed;
  }

  int iters_;
  std::vector<std::unique_ptr<Cohort>> cohorts_;
  std::vector<BatchJob*> jobs_;
};

BatchEngine* batch_engine_create(int iters_per_launch) {
  require(iters_per_launch >= 1, "batch: iterations per launch must be positive");
  return new BatchEngine(iters_per_launch);
}
void batch_engine_add(BatchEngine& e, BatchJob* j) { e.add(j); }
bool batch_engine_poll(BatchEngine& e, std::vector<BatchJob*>& done) { return e.poll(done); }
size_t batch_engine_jobs(const BatchEngine& e) { return e.jobs(); }
void batch_engine_destroy(BatchEngine* e) { delete e; }

}  // namespace femhip
