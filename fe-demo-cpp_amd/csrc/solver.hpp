#pragma once

#include "operator.hpp"

namespace femhip {

struct SolveOptions {
  int mode = 0;  // 0 FP64 PCG, 1 mixed (FP32 inner PCG + FP64 refinement), 2 mixed with reliable updates
  int max_iter = 1000;
  int max_outer = 20;
  double rtol = 1e-10;
  double inner_rtol = 0.0;  // <= 0: default_inner_rtol(mode)
  int batch = 0;  // CG iterations per captured graph launch (0: kSmallBatch up to kSmallBatchMaxN entries, else 8)
  int cg_variant = 0;  // 0: automatic (FP64: pipelined up to kPipeMaxN entries; FP32 inner: plain; plain: wide
                       // from kWideMinBytes per vector), 1: plain, 2: pipelined, 3: plain, wide kernels
  int separate_second_pass = 0;  // pipelined form: 1 keeps the apply's second pass as its own kernel
};

// Vector length up to which the automatic choice takes pipelined CG (FP64
// solves). On MI355X (plain against pipelined, FP64 solves to 1e-10): 36k
// unknowns 0.83 / 0.75 ms, 118k 0.79 / 0.80 ms, 531k 1.23 / 1.41 ms, 913k
// 1.57 / 1.82 ms: the pipelined form's extra vectors cost more than its
// launches save from a few 100k unknowns.
constexpr size_t kPipeMaxN = size_t{1} << 17;
// Vector length up to which a solve takes kSmallBatch iterations per graph launch.
constexpr size_t kSmallBatchMaxN = size_t{1} << 20;
// Vector size in bytes from which the automatic choice runs the plain form's
// wide (one pack per thread) kernels: on MI355X the wide form was 2.5% faster
// on 136 MB FP64 vectors (64^3, P=5) and 1% slower on their 68 MB FP32 copies.
constexpr size_t kWideMinBytes = size_t{96} << 20;
// Iterations per graph launch for vectors of at most kSmallBatchMaxN entries.
constexpr int kSmallBatch = 32;

// Inner tolerance by mode. Refinement: 1e-5 per FP32 solve, so a 1e-10 target
// takes two refinement steps, and FP32 Jacobi-PCG on the mass matrix keeps the
// FP64 convergence rate down to there (on MI355X: 1.78x the FP64 solve at 64^3,
// P=5, against 1.37x with 1e-4, which needs three steps). Reliable updates:
// a replacement every 100x drop keeps the FP64 iteration count.
inline double default_inner_rtol(int mode) { return mode == 2 ? 1e-2 : 1e-5; }

struct SolveStats {
  int iterations = 0;
  int outer = 0;
  double rel_residual = 0.0;
  double ms = 0.0;
};

struct MassSolverState;
void mass_solver_state_destroy(MassSolverState* st);

// One solve as a non-blocking state machine on a stream of its own (not the
// default stream): create prepares everything (may synchronize the stream),
// step enqueues the next piece of work once the previous one has finished and
// returns true when the solve is complete, wait drives it to the end.
struct SolveJob;
SolveJob* solve_job_create(Operator& op, MassSolverState*& state, const double* b, double* x, const SolveOptions& opt,
                           hipStream_t s);
// Records the job's CG graphs now (else the first step does): under the
// exclusive capture lock, so call it holding no shared one.
void solve_job_prepare_graphs(SolveJob& j);
bool solve_job_step(SolveJob& j);
void solve_job_wait(SolveJob& j);
bool solve_job_done(const SolveJob& j);
const SolveStats& solve_job_stats(const SolveJob& j);
void solve_job_destroy(SolveJob* j);
// Solves M x = b on the fused mass operator. `state` caches the diagonal, work
// vectors and captured graphs between solves on the same operator.
void mass_solve(Operator& op, MassSolverState*& state, const double* b, double* x, const SolveOptions& opt,
                SolveStats& stats, hipStream_t stream);

}  // namespace femhip
