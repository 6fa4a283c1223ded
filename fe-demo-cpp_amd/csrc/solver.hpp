#pragma once

#include "operator.hpp"

namespace femhip {

struct SolveOptions {
  int mode = 0;  // 0 FP64 PCG, 1 mixed (FP32 inner PCG + FP64 refinement)
  int max_iter = 1000;
  int max_outer = 20;
  double rtol = 1e-10;
  double inner_rtol = 1e-4;
  int batch = 8;  // CG iterations per captured graph launch
};

struct SolveStats {
  int iterations = 0;
  int outer = 0;
  double rel_residual = 0.0;
  double ms = 0.0;
};

struct MassSolverState;
void mass_solver_state_destroy(MassSolverState* st);
// Solves M x = b on the fused mass operator. `state` caches the diagonal, work
// vectors and captured graphs between solves on the same operator.
void mass_solve(Operator& op, MassSolverState*& state, const double* b, double* x, const SolveOptions& opt,
                SolveStats& stats, hipStream_t stream);

}  // namespace femhip
