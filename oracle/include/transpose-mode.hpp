#pragma once
// ORACLE (test infrastructure only). Restates include/transpose-mode.hpp:13.
namespace fem {

// Forward map (gather / nodes->points) or its adjoint (scatter-add / points->nodes).
enum class ContractMode { NoTranspose, Transpose };

}  // namespace fem
