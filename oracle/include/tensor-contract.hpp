#pragma once
// ORACLE (test infrastructure only). Restates the interface of include/tensor-contract.hpp:9-53.
#include <vector>
#include "basis.hpp"
#include "eval-mode.hpp"
#include "transpose-mode.hpp"

namespace fem {

void tensor_contract_apply(int A, int B, int C, int J, const std::vector<double>& t, ContractMode t_mode, bool add,
                           const std::vector<double>& u, std::vector<double>& v);
void tensor_basis_apply_interp(const TensorBasis& basis, int num_elem, ContractMode t_mode,
                               const std::vector<double>& u, std::vector<double>& v);
void tensor_basis_apply_grad(const TensorBasis& basis, int num_elem, ContractMode t_mode,
                             const std::vector<double>& u, std::vector<double>& v);
void tensor_basis_apply_weight(const TensorBasis& basis, int num_elem, ContractMode t_mode, std::vector<double>& v);
void tensor_basis_apply(const TensorBasis& basis, int num_elem, ContractMode t_mode, EvalMode eval_mode,
                        const std::vector<double>& u, std::vector<double>& v);

}  // namespace fem
