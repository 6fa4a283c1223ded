// Continuous-batching serving loop for mass solves.
//
// Requests (a jittered structured hex mesh of n^3 elements, order P, Q points,
// num_comp components, FP64 or mixed-precision solve, a seeded right-hand side)
// are admitted while a slot is free and their estimated device footprint fits
// the HBM budget (a share of the memory free when the server starts). Admitted
// requests are set up on worker threads (mesh, restrictions, qdata through the
// build_mass operator, fused mass operator with the tuned algorithm, the
// captured CG graphs), then one scheduler thread advances every running solve
// without blocking. Batched (the default): requests of one shape (P, Q,
// num_comp, precision) join a cohort whose persistent element and step
// kernels, captured once as a hipGraph, advance all of them per launch; a
// request that arrives meanwhile joins at the next launch as soon as a slot
// and memory free up. Otherwise each solve runs its own graph-captured CG
// batches on its own stream. Finished requests are copied out and torn down on
// the workers, releasing their budget.
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "common.hpp"

namespace femhip {

struct ServeRequest {
  int n = 8, P = 3, Q = 4, num_comp = 1;
  int mode = 0;  // 0 FP64 PCG, 1 FP32-inner / FP64-refined
  double jitter = 0.1, rtol = 1e-10;
  uint64_t mesh_seed = 0, rhs_seed = 1;
};

struct ServeResult {
  int status = 0;  // 0, or the error code of the failure
  std::string error;
  int iterations = 0, outer = 0;
  double rel_residual = 0.0;
  // Milliseconds: submit -> admitted, admitted -> set up, solve, submit -> done.
  double queue_ms = 0.0, setup_ms = 0.0, solve_ms = 0.0, latency_ms = 0.0;
  // Parts of setup_ms: building the problem and solver state; recording the graphs.
  double build_ms = 0.0, graph_ms = 0.0;
  long long l_size = 0;
  size_t est_bytes = 0;
  std::vector<double> x;
};

struct ServerOptions {
  int max_active = 16;        // requests set up or solving at once (one stream each)
  int workers = 8;            // host threads for setup and teardown
  double hbm_fraction = 0.5;  // share of the HBM free at start that requests may use
  int batch_iters = 8;        // most CG iterations per graph launch (fewer where the rate says so)
  bool autotune = false;      // tune each new (P, Q, num_comp, precision) on first sight
  bool keep_x = true;         // return the solution vectors
  // true: the tuned (automatic) algorithm where its setup is all on the device
  // (box columns, element + transpose map), else the transpose map; false: the
  // tuning table's fastest algorithm where it has a row for the size (block
  // algorithms cluster elements on the host), else the transpose map.
  bool cheap_setup = true;
  // true: FP64 and mixed-refinement solves of shapes with a batched element
  // kernel and at most batch_max_nodes element nodes (num_elem * P^3 *
  // num_comp) run in cohorts (batch_solve.hpp): concurrent requests of one
  // shape share each persistent element and step launch of a captured graph.
  // Every other request (and all with false): one stream and captured graph
  // per request, with the tuned algorithm (box columns for large boxes).
  bool batched = true;
  // Requests this small cannot fill the GPU alone (on MI355X the box columns,
  // which the batched kernels do not use, win from about this size up).
  long long batch_max_nodes = 1LL << 20;
};

struct ServerStats {
  long long submitted = 0, completed = 0, failed = 0, active = 0, queued = 0;
  long long active_peak = 0;
  size_t budget = 0, used = 0, used_peak = 0;
};

// Conservative device bytes a request needs while it is set up and solved.
size_t serve_request_bytes(const ServeRequest& r);

class MassServer;
MassServer* server_create(const ServerOptions& o);
// Thread-safe; returns the request id. Invalid requests fail at once (their wait reports it).
long long server_submit(MassServer& s, const ServeRequest& r);
// Waits up to timeout_ms (< 0: forever). Returns false on timeout; else moves the result out.
bool server_wait(MassServer& s, long long id, double timeout_ms, ServeResult* out);
ServerStats server_stats(MassServer& s);
// Finishes the requests already submitted, then stops the threads.
void server_destroy(MassServer* s);

}  // namespace femhip
