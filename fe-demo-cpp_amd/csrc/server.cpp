// Continuous-batching serving loop for mass solves (see server.hpp).
#include "server.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <set>
#include <stdexcept>
#include <thread>

#include "basis.hpp"
#include "batch_solve.hpp"
#include "basis_setup.hpp"
#include "fused_mass.hpp"
#include "mesh.hpp"
#include "operator.hpp"
#include "qfunction.hpp"
#include "restriction.hpp"
#include "solver.hpp"

namespace femhip {

namespace {

using Clock = std::chrono::steady_clock;
double ms_between(Clock::time_point a, Clock::time_point b) {
  return std::chrono::duration<double, std::milli>(b - a).count();
}

long long ipow3(long long v) { return v * v * v; }

// The C ABI's status code of a failure.
int status_of(const std::exception& e) {
  if (dynamic_cast<const std::invalid_argument*>(&e)) return 1;
  if (dynamic_cast<const DeviceError*>(&e)) return 2;
  if (dynamic_cast<const Unsupported*>(&e)) return 4;
  if (dynamic_cast<const OutOfMemory*>(&e)) return 5;
  return 3;
}

}  // namespace

size_t serve_request_bytes(const ServeRequest& r) {
  const long long E = ipow3(r.n), P3 = ipow3(r.P), Q3 = ipow3(r.Q), nc = r.num_comp;
  const long long N = ipow3(static_cast<long long>(r.n) * (r.P - 1) + 1), L = nc * N, nv = ipow3(r.n + 1);
  long long b = 0;
  b += E * P3 * (4 + 4 + 4 + 4);       // offsets, flagged copy, transpose map, multi-row map
  b += E * P3 * (2 + 12 * nc);         // block-local indices, element outputs (FP64 + FP32)
  b += E * Q3 * (8 + 8 + 8 + 8 + 4 + 4);  // qdata: built, operator copy, private, block order, FP32 copies
  b += E * (500 * nc);                 // block node lists and partials (build_mass is fused: no temporaries)
  b += N * 16 + L * (8 * 10 + 4 * 7);  // node maps; solver vectors (FP64 and FP32)
  b += nv * 3 * 8 * 2 + E * 8 * 4 * 2;  // coordinates, mesh restriction
  b += 16ll << 20;                      // small buffers, allocator granularity
  return static_cast<size_t>(b);
}

namespace {

// Streams reused across jobs: creating one takes milliseconds while other
// jobs' work runs on the device.
class StreamPool {
 public:
  ~StreamPool() {
    for (hipStream_t s : free_) {
      (void)hipStreamSynchronize(s);
      SharedCapture lk;
      (void)hipStreamDestroy(s);
    }
  }
  hipStream_t acquire() {
    {
      std::lock_guard<std::mutex> lk(m_);
      if (!free_.empty()) {
        hipStream_t s = free_.back();
        free_.pop_back();
        return s;
      }
    }
    hipStream_t s = nullptr;
    SharedCapture lk;
    HIP_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    return s;
  }
  void release(hipStream_t s) {
    std::lock_guard<std::mutex> lk(m_);
    free_.push_back(s);
  }

 private:
  std::mutex m_;
  std::vector<hipStream_t> free_;
};

struct Job {
  long long id = 0;
  ServeRequest req;
  size_t bytes = 0;
  Clock::time_point t_submit, t_admit, t_ready, t_done;
  hipStream_t stream = nullptr;  // from `streams`
  StreamPool* streams = nullptr;
  std::unique_ptr<Operator> op;
  // The build_mass operator, the vertex coordinates and the qdata it made:
  // kept until teardown (freeing a device block waits for its stream).
  std::unique_ptr<Operator> bop;
  DevBuf<double> coords, qdata;
  DevBuf<double> b, x;
  MassSolverState* state = nullptr;
  SolveJob* solve = nullptr;
  BatchJob* batch = nullptr;  // batched solves: driven by the server's BatchEngine
  ServeResult res;
  ~Job() {
    batch_job_destroy(batch);
    solve_job_destroy(solve);
    mass_solver_state_destroy(state);
    op.reset();
    bop.reset();
    coords.release();
    qdata.release();
    b.release();
    x.release();
    if (stream) {
      (void)hipStreamSynchronize(stream);
      streams->release(stream);
    }
  }
};

void validate(const ServeRequest& r) {
  require(r.n >= 1 && r.n <= 512, "serve: n must be in [1, 512]");
  require(r.P >= 2 && r.Q >= 1 && r.num_comp >= 1 && r.num_comp <= 3, "serve: bad P, Q or num_comp");
  require(fused_mass_supported(r.P, r.Q, r.num_comp), "serve: (P, Q, num_comp) has no compiled fused kernel");
  require(r.mode >= 0 && r.mode <= 2, "serve: mode must be 0 (FP64), 1 (mixed) or 2 (mixed, reliable updates)");
  require(r.rtol > 0 && r.jitter >= 0 && r.jitter < 0.5, "serve: bad rtol or jitter");
  require(ipow3(static_cast<long long>(r.n) * (r.P - 1) + 1) * r.num_comp < (1ll << 31), "serve: L-vector too large");
}

// The H1 Lagrange basis of a shape (GLL nodes, Gauss points), made once per
// process and shared by every request of that shape: a basis is immutable,
// and making one uploads its matrices (each upload waits for the device,
// which under load means waiting behind other requests' solves). Its device
// arrays are plain allocations, not blocks of a job's cache.
std::shared_ptr<const Basis> shape_basis(int dim, int nc, int P, int Q) {
  static std::mutex* mu = new std::mutex();  // never destroyed: the runtime may be gone at exit
  static auto* cache = new std::map<std::array<int, 4>, std::shared_ptr<const Basis>>();
  std::lock_guard<std::mutex> lk(*mu);
  std::shared_ptr<const Basis>& slot = (*cache)[{dim, nc, P, Q}];
  if (!slot) {
    const CacheScope* prev = cache_scope();
    cache_scope() = nullptr;
    try {
      slot.reset(basis_create(make_h1_lagrange(dim, nc, P, Q, 0)));
    } catch (...) {
      cache_scope() = prev;
      throw;
    }
    cache_scope() = prev;
  }
  return slot;
}

// Mesh, restrictions, qdata (build_mass on the device) and the fused mass
// operator of a request, as tests/test-operator.cpp:17-24,75-112 set them up,
// enqueued on s with no wait for the device: later work on the job's data is
// ordered after s (the build operator, coordinates and qdata go into the job).
std::unique_ptr<Operator> build_problem(const ServeRequest& r, hipStream_t s, Job& job) {
  const int n = r.n, P = r.P, Q = r.Q, nc = r.num_comp;
  const int E = n * n * n, Q3 = Q * Q * Q, nv = (n + 1) * (n + 1) * (n + 1);
  // Offsets (node and mesh restrictions) and the jittered vertex coordinates
  // are made on the device: entry for entry what box_offsets and box_coords
  // make on the host, with no host arrays, validation or uploads.
  std::shared_ptr<const Restriction> rstr(restriction_create_box(n, n, n, P, nc, kScatterDeterministic, s, false));
  std::shared_ptr<const Restriction> mrstr(restriction_create_box(n, n, n, 2, 3, kScatterAtomic, s, false));
  std::shared_ptr<const Restriction> qrstr(restriction_create_strided(E, Q3, 1, E * Q3, nullptr));
  std::shared_ptr<const Basis> basis = shape_basis(3, nc, P, Q);
  std::shared_ptr<const Basis> mbasis = shape_basis(3, 3, 2, Q);

  QFunctionDev bq;
  bq.kind = kQfBuildMass;
  qfunction_add_input(bq, "dx", 9, kEvalGrad);
  qfunction_add_input(bq, "weights", 1, kEvalWeight);
  qfunction_add_output(bq, "qdata", 1, kEvalNone);
  const int dim = 3;
  bq.ctx.resize(sizeof(int));
  std::memcpy(bq.ctx.data(), &dim, sizeof(int));
  job.bop.reset(operator_create(bq));
  Operator& bop = *job.bop;
  operator_set_field(bop, "dx", mrstr, mbasis, kVectorActive, nullptr, 0, false);
  operator_set_field(bop, "weights", nullptr, mbasis, kVectorNone, nullptr, 0, false);
  operator_set_field(bop, "qdata", qrstr, nullptr, kVectorActive, nullptr, 0, false);
  job.coords.allocate(static_cast<size_t>(3) * nv);
  job.qdata.allocate(static_cast<size_t>(E) * Q3);
  box_coords_device(n, r.jitter, r.mesh_seed, job.coords.data(), s);
  operator_apply(bop, job.coords.data(), job.coords.size(), job.qdata.data(), job.qdata.size(), s);

  QFunctionDev mq;
  mq.kind = nc == 1 ? kQfApplyMass : kQfApplyMassVector;
  if (nc > 1) {
    mq.ctx.resize(sizeof(int));
    std::memcpy(mq.ctx.data(), &nc, sizeof(int));
  }
  qfunction_add_input(mq, "u", nc, kEvalInterp);
  qfunction_add_input(mq, "qdata", 1, kEvalNone);
  qfunction_add_output(mq, "v", nc, kEvalInterp);
  std::unique_ptr<Operator> op(operator_create(mq));
  operator_set_field(*op, "u", rstr, basis, kVectorActive, nullptr, 0, false);
  // The passive copy in s's order, after the build that writes qdata.
  operator_set_field(*op, "qdata", qrstr, nullptr, kVectorPassive, job.qdata.data(), job.qdata.size(), true, s);
  operator_set_field(*op, "v", rstr, basis, kVectorActive, nullptr, 0, false);
  op->opt_scatter = kScatterDeterministic;  // the same bits alone and in any mix
  return op;
}

}  // namespace

class MassServer {
 public:
  explicit MassServer(const ServerOptions& o) : opt_(o) {
    require(o.max_active >= 1 && o.workers >= 1 && o.batch_iters >= 1, "server: bad options");
    require(o.hbm_fraction > 0 && o.hbm_fraction <= 0.95, "server: hbm_fraction must be in (0, 0.95]");
    size_t free_b = 0, total_b = 0;
    HIP_CHECK(hipMemGetInfo(&free_b, &total_b));
    budget_ = static_cast<size_t>(static_cast<double>(free_b) * o.hbm_fraction);
    HIP_CHECK(hipGetDevice(&device_));
    // The jobs' device buffers come from a block cache: at most half the budget
    // stays cached between jobs.
    cache_ = std::make_unique<BlockCache>(budget_ / 2);
    scheduler_ = std::thread([this] { scheduler_loop(); });
    for (int i = 0; i < o.workers; ++i) workers_.emplace_back([this] { worker_loop(); });
  }

  ~MassServer() {
    {
      std::unique_lock<std::mutex> lk(mu_);
      stopping_ = true;
      cv_.notify_all();
      work_cv_.notify_all();
    }
    scheduler_.join();
    {
      std::unique_lock<std::mutex> lk(mu_);
      workers_stop_ = true;
      work_cv_.notify_all();
    }
    for (auto& t : workers_) t.join();
  }

  long long submit(const ServeRequest& r) {
    auto j = std::make_unique<Job>();
    j->req = r;
    j->t_submit = Clock::now();
    std::unique_lock<std::mutex> lk(mu_);
    require(!stopping_, "server: stopping");
    j->id = next_id_++;
    ++stats_.submitted;
    const long long id = j->id;
    try {
      validate(r);
      j->bytes = serve_request_bytes(r);
      if (j->bytes > budget_) throw OutOfMemory("serve: the request needs more than the server's HBM budget");
    } catch (const std::exception& e) {
      j->res.status = status_of(e);
      j->res.error = e.what();
      finish_locked(std::move(j));
      return id;
    }
    pending_.push_back(std::move(j));
    cv_.notify_all();
    return id;
  }

  bool wait(long long id, double timeout_ms, ServeResult* out) {
    std::unique_lock<std::mutex> lk(mu_);
    auto ready = [&] { return done_.count(id) > 0; };
    if (timeout_ms < 0) {
      done_cv_.wait(lk, ready);
    } else if (!done_cv_.wait_for(lk, std::chrono::duration<double, std::milli>(timeout_ms), ready)) {
      return false;
    }
    auto it = done_.find(id);
    *out = std::move(it->second);
    done_.erase(it);
    return true;
  }

  ServerStats stats() {
    std::unique_lock<std::mutex> lk(mu_);
    ServerStats s = stats_;
    s.active = static_cast<long long>(in_setup_ + running_.size() + in_teardown_);
    s.queued = static_cast<long long>(pending_.size());
    s.budget = budget_;
    s.used = used_;
    return s;
  }

 private:
  // Admission: FIFO; the head waits for a slot and for its bytes to fit the budget.
  void admit_locked() {
    while (!pending_.empty()) {
      const long long active = static_cast<long long>(in_setup_ + running_.size() + in_teardown_);
      Job& h = *pending_.front();
      if (active >= opt_.max_active || used_ + h.bytes > budget_) break;
      used_ += h.bytes;
      stats_.used_peak = std::max(stats_.used_peak, used_);
      h.t_admit = Clock::now();
      ++in_setup_;
      stats_.active_peak = std::max(stats_.active_peak, active + 1);
      setup_q_.push_back(std::move(pending_.front()));
      pending_.pop_front();
      work_cv_.notify_one();
    }
  }

  void finish_locked(std::unique_ptr<Job> j) {
    j->t_done = Clock::now();
    ServeResult& r = j->res;
    r.latency_ms = ms_between(j->t_submit, j->t_done);
    if (r.status == 0) {
      ++stats_.completed;
    } else {
      ++stats_.failed;
    }
    r.est_bytes = j->bytes;
    done_[j->id] = std::move(r);
    done_cv_.notify_all();
  }

  // Hands a running job to the workers for teardown.
  void to_teardown(std::unique_ptr<Job> j) {
    std::unique_lock<std::mutex> lk(mu_);
    running_.erase(j.get());
    ++in_teardown_;
    teardown_q_.push_back(std::move(j));
    work_cv_.notify_one();
  }

  void scheduler_loop() {
    (void)hipSetDevice(device_);
    std::vector<std::unique_ptr<Job>> mine;  // running jobs owned by this thread
    std::map<BatchJob*, Job*> batched;       // mine's jobs in the batch engine
    while (true) {
      {
        std::unique_lock<std::mutex> lk(mu_);
        for (auto& j : ready_q_) {
          if (j->batch) {
            if (!engine_) engine_.reset(batch_engine_create(opt_.batch_iters));
            batch_engine_add(*engine_, j->batch);
            batched[j->batch] = j.get();
          }
          mine.push_back(std::move(j));
        }
        ready_q_.clear();
        admit_locked();
        const bool idle = mine.empty();
        if (idle) {
          if (stopping_ && pending_.empty() && in_setup_ == 0) break;
          cv_.wait_for(lk, std::chrono::milliseconds(2));
          continue;
        }
      }
      bool progressed = false;
      // Batched solves: one poll advances every cohort.
      if (engine_ && batch_engine_jobs(*engine_) > 0) {
        std::vector<BatchJob*> fin;
        progressed |= batch_engine_poll(*engine_, fin);
        for (BatchJob* b : fin) {
          Job* jp = batched[b];
          batched.erase(b);
          const SolveStats& st = batch_job_stats(*b);
          jp->res.iterations = st.iterations;
          jp->res.outer = st.outer;
          jp->res.rel_residual = st.rel_residual;
          jp->res.solve_ms = ms_between(jp->t_ready, Clock::now());
          if (batch_job_status(*b) != 0) {
            jp->res.status = batch_job_status(*b);
            jp->res.error = batch_job_error(*b);
          }
          for (auto& j : mine)
            if (j.get() == jp) to_teardown(std::move(j));
        }
      }
      // Advance every other running solve; hand finished ones to the workers.
      for (auto& j : mine) {
        if (!j || j->batch) continue;
        try {
          if (solve_job_step(*j->solve)) {
            progressed = true;
            const SolveStats& st = solve_job_stats(*j->solve);
            j->res.iterations = st.iterations;
            j->res.outer = st.outer;
            j->res.rel_residual = st.rel_residual;
            j->res.solve_ms = ms_between(j->t_ready, Clock::now());
            to_teardown(std::move(j));
          }
        } catch (const std::exception& e) {
          j->res.status = status_of(e);
          j->res.error = e.what();
          to_teardown(std::move(j));
          progressed = true;
        }
      }
      mine.erase(std::remove(mine.begin(), mine.end(), nullptr), mine.end());
      if (!progressed) std::this_thread::sleep_for(std::chrono::microseconds(20));
    }
  }

  void worker_loop() {
    (void)hipSetDevice(device_);
    while (true) {
      std::unique_ptr<Job> j;
      bool setup = false;
      {
        std::unique_lock<std::mutex> lk(mu_);
        work_cv_.wait(lk, [&] { return !teardown_q_.empty() || !setup_q_.empty() || workers_stop_; });
        if (!teardown_q_.empty()) {  // teardown first: it frees budget
          j = std::move(teardown_q_.front());
          teardown_q_.pop_front();
        } else if (!setup_q_.empty()) {
          j = std::move(setup_q_.front());
          setup_q_.pop_front();
          setup = true;
        } else {
          return;  // workers_stop_ and nothing left
        }
      }
      if (setup) {
        do_setup(std::move(j));
      } else {
        do_teardown(std::move(j));
      }
    }
  }

  void do_setup(std::unique_ptr<Job> j) {
    bool ok = true;
    try {
      {
        j->stream = streams_.acquire();
        j->streams = &streams_;
        CachedAllocScope cached(cache_.get(), j->stream);
        j->op = build_problem(j->req, j->stream, *j);
        const size_t L = static_cast<size_t>(j->op->active_in_l_size);
        j->res.l_size = static_cast<long long>(L);
        j->b.allocate(L);  // the rhs stream generated in place, no host copy
        rng_uniform_device(j->req.rhs_seed, L, -1.0, 1.0, j->b.data(), j->stream);
        j->x.allocate(L);
        SolveOptions so;
        so.mode = j->req.mode;
        so.rtol = j->req.rtol;
        so.max_iter = 2000;
        so.max_outer = 30;
        so.batch = opt_.batch_iters;
        so.inner_rtol = default_inner_rtol(so.mode);
        const bool mixed = j->req.mode != 0;
        const long long elem_nodes = ipow3(j->req.n) * ipow3(j->req.P) * j->req.num_comp;
        if (opt_.batched && elem_nodes <= opt_.batch_max_nodes &&
            batch_supported(j->req.P, j->req.Q, j->req.num_comp, j->req.mode)) {
          // The tuned packing of the transpose-map kernel (autotuned per shape when asked).
          if (opt_.autotune) maybe_tune(operator_fused_mass(*j->op, j->stream), mixed, j->stream);
          j->batch = batch_job_create(*j->op, j->b.data(), j->x.data(), so, j->stream);
          j->res.build_ms = ms_between(j->t_admit, Clock::now());
        }
      }
      if (!j->batch) {
        CachedAllocScope cached(cache_.get(), j->stream);
        FusedMass& f = operator_fused_mass(*j->op, j->stream);
        const bool mixed = j->req.mode != 0;
        if (opt_.autotune) maybe_tune(f, mixed, j->stream);
        // The automatic choice (tuning table, else the default order) in every
        // precision the solve applies in; with cheap_setup only where it is an
        // algorithm whose setup is all on the device (box columns, element +
        // transpose map), and without it only where the table holds a row for
        // this size. Otherwise the transpose-map algorithm (its setup is the
        // device counting sort the restriction already did; the block
        // algorithms cluster elements on the host).
        bool use_auto = opt_.cheap_setup || tuned_shape(f, mixed);
        for (int fp32 = 0; fp32 <= (mixed ? 1 : 0) && use_auto; ++fp32) {
          const FusedKey k = fused_resolve(f, fp32);
          if (opt_.cheap_setup && !(alg_box(k.alg) || k.alg == kAlgElemTmap)) use_auto = false;
        }
        if (!use_auto) {
          j->op->opt_alg = kAlgElemTmap;
          f.choice.alg = kAlgElemTmap;
        }
        SolveOptions so;
        so.mode = j->req.mode;
        so.rtol = j->req.rtol;
        so.max_iter = 2000;
        so.max_outer = 30;
        so.batch = opt_.batch_iters;
        so.inner_rtol = default_inner_rtol(so.mode);
        j->solve = solve_job_create(*j->op, j->state, j->b.data(), j->x.data(), so, j->stream);
        HIP_CHECK(hipStreamSynchronize(j->stream));
        const Clock::time_point t_built = Clock::now();
        j->res.build_ms = ms_between(j->t_admit, t_built);
        solve_job_prepare_graphs(*j->solve);  // exclusive: other threads' allocations and copies wait
        j->res.graph_ms = ms_between(t_built, Clock::now());
      }
    } catch (const std::exception& e) {
      ok = false;
      j->res.status = status_of(e);
      j->res.error = e.what();
    }
    std::unique_lock<std::mutex> lk(mu_);
    --in_setup_;
    j->t_ready = Clock::now();
    j->res.queue_ms = ms_between(j->t_submit, j->t_admit);
    j->res.setup_ms = ms_between(j->t_admit, j->t_ready);
    if (ok) {  // the scheduler thread takes it over
      running_.insert(j.get());
      ready_q_.push_back(std::move(j));
      cv_.notify_all();
    } else {
      ++in_teardown_;
      teardown_q_.push_back(std::move(j));
      work_cv_.notify_one();
    }
  }

  void do_teardown(std::unique_ptr<Job> j) {
    try {
      if (j->res.status == 0 && opt_.keep_x) {
        j->res.x.resize(j->x.size());
        if (!j->res.x.empty())
          HIP_CHECK(hipMemcpyAsync(j->res.x.data(), j->x.data(), j->res.x.size() * sizeof(double),
                                   hipMemcpyDeviceToHost, j->stream));
        HIP_CHECK(hipStreamSynchronize(j->stream));
      }
    } catch (const std::exception& e) {
      j->res.status = 3;
      j->res.error = e.what();
    }
    const size_t bytes = j->bytes;
    ServeResult res = std::move(j->res);
    j->res = ServeResult();
    const long long id = j->id;
    const Clock::time_point t_submit = j->t_submit;
    j.reset();  // frees the device state
    auto done = std::make_unique<Job>();
    done->id = id;
    done->bytes = bytes;
    done->t_submit = t_submit;
    done->res = std::move(res);
    std::unique_lock<std::mutex> lk(mu_);
    --in_teardown_;
    used_ -= bytes;
    finish_locked(std::move(done));
    cv_.notify_all();
  }

  // Mesh sizes within this factor share a tuning row.
  static constexpr double kTuneSizeRatio = 4.0;

  // Whether the tuning table holds, for every precision this solve applies
  // in, a row measured on a mesh within kTuneSizeRatio of this one's size.
  static bool tuned_shape(const FusedMass& f, bool mixed) {
    std::lock_guard<std::mutex> lk(tune_mutex());
    for (int fp32 : {0, 1}) {
      if (fp32 && !mixed) continue;
      if (!tune_lookup(TuneKey{f.P, f.Q, f.NC, fp32, 1, fused_box_operator(f) ? 1 : 0}, f.num_elem, kTuneSizeRatio))
        return false;
    }
    return true;
  }

  // Autotunes (P, Q, num_comp, precision, deterministic) once per server and
  // size class (factors of kTuneSizeRatio in elements).
  void maybe_tune(FusedMass& f, bool mixed, hipStream_t s) {
    const int size_class = static_cast<int>(std::floor(std::log(std::max(f.num_elem, 1) * 1.0) /
                                                       std::log(kTuneSizeRatio)));
    for (int fp32 : {0, 1}) {
      if (fp32 && !mixed) continue;
      const TuneKey k{f.P, f.Q, f.NC, fp32, 1, fused_box_operator(f) ? 1 : 0};
      {
        std::unique_lock<std::mutex> lk(mu_);
        if (tuned_.count({k, size_class})) continue;
        tuned_.insert({k, size_class});
      }
      fused_autotune(f, fp32, 5, s);
    }
  }

  ServerOptions opt_;
  StreamPool streams_;  // these two outlive every job (members go in reverse order)
  std::unique_ptr<BlockCache> cache_;
  int device_ = 0;
  size_t budget_ = 0, used_ = 0;
  std::mutex mu_;
  std::condition_variable cv_, work_cv_, done_cv_;
  bool stopping_ = false, workers_stop_ = false;
  long long next_id_ = 1;
  size_t in_setup_ = 0, in_teardown_ = 0;
  std::deque<std::unique_ptr<Job>> pending_, setup_q_, teardown_q_;
  std::vector<std::unique_ptr<Job>> ready_q_;
  std::set<Job*> running_;
  std::map<long long, ServeResult> done_;
  std::set<std::pair<TuneKey, int>> tuned_;  // (shape, size class) autotuned by this server
  ServerStats stats_;
  std::thread scheduler_;
  std::vector<std::thread> workers_;
  std::unique_ptr<BatchEngine, void (*)(BatchEngine*)> engine_{nullptr, batch_engine_destroy};  // scheduler thread only
};

MassServer* server_create(const ServerOptions& o) { return new MassServer(o); }
long long server_submit(MassServer& s, const ServeRequest& r) { return s.submit(r); }
bool server_wait(MassServer& s, long long id, double timeout_ms, ServeResult* out) {
  return s.wait(id, timeout_ms, out);
}
ServerStats server_stats(MassServer& s) { return s.stats(); }
void server_destroy(MassServer* s) { delete s; }

}  // namespace femhip
