#pragma once
// ORACLE (test infrastructure only). Restates the interface of include/operator.hpp:15-64.
#include <optional>
#include <string>
#include <vector>
#include "basis.hpp"
#include "elem-restriction.hpp"
#include "qfunction.hpp"

namespace fem {

enum class VectorKind { Active, Passive, None };

struct FieldVector {
  VectorKind kind = VectorKind::None;
  std::vector<double> data;
};
inline FieldVector vector_active() { return {VectorKind::Active, {}}; }
inline FieldVector vector_none() { return {VectorKind::None, {}}; }
inline FieldVector vector_passive(std::vector<double> data) { return {VectorKind::Passive, std::move(data)}; }

struct OperatorField {
  std::string name;
  EvalMode eval_mode;
  int size;
  std::optional<ElemRestriction> restriction;
  std::optional<TensorBasis> basis;
  FieldVector vector;
};

struct Operator {
  QFunction qf;
  std::vector<std::optional<OperatorField>> inputs;
  std::vector<std::optional<OperatorField>> outputs;
  int num_elem = 0;
  int num_qpts = 0;
  int active_in_l_size = 0;
  int active_out_l_size = 0;
};

void operator_create(const QFunction& qf, Operator& op);
void operator_set_field(Operator& op, const std::string& name, std::optional<ElemRestriction> restriction,
                        std::optional<TensorBasis> basis, FieldVector vector);
void operator_check_ready(const Operator& op);
void operator_apply(Operator& op, const std::vector<double>& in, std::vector<double>& out);

}  // namespace fem
