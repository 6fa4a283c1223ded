#pragma once
// ORACLE (test infrastructure only). CPU restatement of the reference interface
// include/eval-mode.hpp:11. Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may load anything built from oracle/.
namespace fem {

// None: field passed through; Interp: values at quadrature points;
// Grad: reference gradients; Weight: quadrature weights (no input vector).
enum class EvalMode { None, Interp, Grad, Weight };

}  // namespace fem
