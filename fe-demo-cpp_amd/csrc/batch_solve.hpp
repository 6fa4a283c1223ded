// Batched mass solves: concurrent PCG runs on operators of one shape advanced
// together by shared persistent launches.
//
// A cohort gathers the runs of one (P, Q, num_comp, basis, precision): FP64
// Jacobi-PCG solves, or the FP32 inner solves of FP64 iterative refinement.
// Its captured hipGraph holds `iters` pipelined CG iterations (Ghysels &
// Vanroose 2014, Alg. 3), each two persistent kernels over every run in the
// cohort's table:
//   1. the element kernel: work items of NE elements across all runs (each
//      run's elements in turn), the fused gather / interp / qdata / interp^T
//      of the transpose-map algorithm writing element outputs;
//   2. the step kernel: work items of one reduction block across all runs,
//      each the pipelined step of its run with the element outputs' row sums
//      folded in (the transpose map, in the CPU scatter's order).
// The table (device memory, rewritten by the host between replays) lists the
// runs in the cohort now, so one graph serves any membership: runs join at a
// replay boundary once their start-up is on the cohort's stream, and leave
// when their done flag is read back. A run's arithmetic (element outputs,
// reduction blocks of its own vector length, its own scalar slots) does not
// depend on which runs share its launches, so its result is the same bits
// alone and in any mix.
#pragma once

#include <map>
#include <vector>

#include "fused_mass.hpp"
#include "operator.hpp"
#include "solver.hpp"

namespace femhip {

// Runs per cohort (table entries); more wait for a free entry.
constexpr int kBatchSlots = 16;

// A run's entry in its cohort's table.
struct BatchSlot {
  // element kernel: element outputs of m into y (transpose-map algorithm)
  const int* offsets = nullptr;
  const void* qdata = nullptr;  // private element layout, the run's precision
  void* y = nullptr;
  size_t y_comp_len = 0;
  int num_elem = 0, comp_stride = 0;
  int elem0 = 0;  // the run's first element item (items of NE elements)
  int red0 = 0;   // the run's first step item
  int nb = 0;     // reduction blocks (partials per dot product) of the run
  int pad = 0;
  size_t n = 0;   // vector length
  double* sc = nullptr;    // scalar slots
  double* part = nullptr;  // partials of both parities
  const void* dinv = nullptr;
  FusedSplit sp;  // row sums of y
  void *x = nullptr, *r = nullptr, *u = nullptr, *w = nullptr, *m = nullptr, *nv = nullptr, *z = nullptr,
       *q = nullptr, *s = nullptr, *p = nullptr;
};
struct BatchTable {
  int nslots = 0, elem_items = 0, red_items = 0, pad = 0;
  BatchSlot slot[kBatchSlots];
};

// The batched element kernel of one (P, Q, num_comp, precision, NE).
struct BatchKey {
  int P, Q, NC, fp32, elems;
  bool operator<(const BatchKey& o) const;
};
struct BatchElemKernel {
  void (*launch)(const BatchTable* tab, int par, const void* basis, int grid, hipStream_t s) = nullptr;
  int (*blocks_per_cu)() = nullptr;  // resident workgroups per CU (occupancy)
};
void register_batch_elem(const BatchKey& k, const BatchElemKernel& kern);
// Compiled NE of a shape and precision, ascending.
std::vector<int> batch_elem_variants(int P, int Q, int NC, int fp32);
// Whether a solve of this shape and mode can run batched (modes 0 and 1).
bool batch_supported(int P, int Q, int NC, int mode);

// One solve driven through cohorts: FP64 PCG (mode 0) or FP32-inner / FP64
// refinement (mode 1).
struct BatchJob;
// Prepares everything the solve of M x = b needs (the operator set to the
// transpose-map algorithm, the Jacobi inverse, the run vectors, ||b||^2) on
// `s` without waiting for it: the engine orders the job's work after what `s`
// holds when this returns (b and the operator's data made there included).
// b and x must stay valid until the job is destroyed.
BatchJob* batch_job_create(Operator& op, const double* b, double* x, const SolveOptions& opt, hipStream_t s);
const SolveStats& batch_job_stats(const BatchJob& j);
// 0, or the C ABI status of the failure (its message in batch_job_error).
int batch_job_status(const BatchJob& j);
const std::string& batch_job_error(const BatchJob& j);
void batch_job_destroy(BatchJob* j);

// The cohorts of one device and the jobs in them. Not thread-safe: one thread
// adds and polls.
class BatchEngine;
BatchEngine* batch_engine_create(int iters_per_launch);
void batch_engine_add(BatchEngine& e, BatchJob* j);
// Never blocks: enqueues whatever work the finished work allows (a run's
// start, a cohort replay, refinement steps). Jobs that completed or failed
// since the last call are appended to `done` (and leave the engine). Returns
// whether anything was enqueued or completed.
bool batch_engine_poll(BatchEngine& e, std::vector<BatchJob*>& done);
size_t batch_engine_jobs(const BatchEngine& e);
// Waits for the engine's streams (its jobs must have left it).
void batch_engine_destroy(BatchEngine* e);

}  // namespace femhip
