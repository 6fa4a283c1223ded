// ORACLE (test infrastructure only). CPU restatement of src/operator.cpp:21-182:
// field binding with the same consistency checks, and operator_apply in the
// reference's field order (Weight -> weights; None -> gather; Interp/Grad ->
// gather + basis; QFunction; zero out; None -> scatter; else basis^T + scatter).
#include "operator.hpp"

#include <stdexcept>
#include "tensor-contract.hpp"

namespace fem {

namespace {
int ipow(int b, int e) {
  int r = 1;
  while (e-- > 0) r *= b;
  return r;
}
std::invalid_argument bad(const std::string& name, const std::string& msg) {
  return std::invalid_argument("operator_set_field '" + name + "': " + msg);
}
}  // namespace

void operator_create(const QFunction& qf, Operator& op) {
  if (!qf.user) throw std::invalid_argument("operator_create: QFunction has no user function");
  op = Operator{};
  op.qf = qf;
  op.inputs.resize(qf.inputs.size());
  op.outputs.resize(qf.outputs.size());
}

// src/operator.cpp:44-117
void operator_set_field(Operator& op, const std::string& name, std::optional<ElemRestriction> restriction,
                        std::optional<TensorBasis> basis, FieldVector vector) {
  const QFunctionField* field = nullptr;
  std::optional<OperatorField>* slot = nullptr;
  bool is_input = true;
  for (size_t f = 0; f < op.qf.inputs.size() && !field; ++f)
    if (op.qf.inputs[f].name == name) field = &op.qf.inputs[f], slot = &op.inputs[f];
  for (size_t f = 0; f < op.qf.outputs.size() && !field; ++f)
    if (op.qf.outputs[f].name == name) field = &op.qf.outputs[f], slot = &op.outputs[f], is_input = false;
  if (!field) throw bad(name, "the QFunction has no field with this name");
  if (slot->has_value()) throw bad(name, "already set");

  const EvalMode mode = field->eval_mode;
  const int size = field->size;
  if (mode == EvalMode::Weight) {
    if (restriction) throw bad(name, "EvalMode::Weight takes no restriction");
    if (!basis) throw bad(name, "EvalMode::Weight needs a basis (for its quadrature weights)");
    if (vector.kind != VectorKind::None) throw bad(name, "EvalMode::Weight takes vector_none()");
  } else if (mode == EvalMode::None) {
    if (!restriction) throw bad(name, "EvalMode::None needs a restriction");
    if (basis) throw bad(name, "EvalMode::None takes no basis");
    if (restriction->num_comp != size) throw bad(name, "restriction num_comp must equal the field size");
  } else {
    if (!restriction) throw bad(name, "EvalMode::Interp/Grad needs a restriction");
    if (!basis) throw bad(name, "EvalMode::Interp/Grad needs a basis");
    const int want = (mode == EvalMode::Interp) ? basis->num_comp : basis->num_comp * basis->dim;
    if (size != want)
      throw bad(name, mode == EvalMode::Interp ? "field size must equal basis num_comp"
                                              : "field size must equal basis num_comp * dim");
    if (restriction->num_comp != basis->num_comp) throw bad(name, "restriction and basis num_comp differ");
    if (restriction->elem_size != ipow(basis->P_1d, basis->dim))
      throw bad(name, "restriction elem_size must equal basis P_1d^dim");
  }
  if (mode != EvalMode::Weight && vector.kind == VectorKind::None)
    throw bad(name, "needs vector_active() or vector_passive(...)");
  if (!is_input && vector.kind != VectorKind::Active) throw bad(name, "output fields must be active");
  if (vector.kind == VectorKind::Passive && static_cast<int>(vector.data.size()) != restriction->l_size)
    throw bad(name, "passive vector size must equal restriction l_size");

  const int nq = basis ? ipow(basis->Q_1d, basis->dim) : restriction->elem_size;
  if (op.num_qpts != 0 && nq != op.num_qpts)
    throw bad(name, "number of quadrature points differs from previously set fields");
  if (restriction && op.num_elem != 0 && restriction->num_elem != op.num_elem)
    throw bad(name, "num_elem differs from previously set fields");
  if (vector.kind == VectorKind::Active) {
    int& l = is_input ? op.active_in_l_size : op.active_out_l_size;
    if (l != 0 && restriction->l_size != l)
      throw bad(name, "active fields share one vector, so their restrictions need the same l_size");
    l = restriction->l_size;
  }
  op.num_qpts = nq;
  if (restriction) op.num_elem = restriction->num_elem;
  *slot = OperatorField{name, mode, size, std::move(restriction), std::move(basis), std::move(vector)};
}

// src/operator.cpp:126-136
void operator_check_ready(const Operator& op) {
  for (size_t f = 0; f < op.inputs.size(); ++f)
    if (!op.inputs[f]) throw std::invalid_argument("operator_check_ready: field '" + op.qf.inputs[f].name + "' is not set");
  for (size_t f = 0; f < op.outputs.size(); ++f)
    if (!op.outputs[f])
      throw std::invalid_argument("operator_check_ready: field '" + op.qf.outputs[f].name + "' is not set");
  if (op.num_elem == 0) throw std::invalid_argument("operator_check_ready: no field has a restriction");
}

// src/operator.cpp:148-182
void operator_apply(Operator& op, const std::vector<double>& in, std::vector<double>& out) {
  operator_check_ready(op);
  if (op.active_in_l_size > 0 && static_cast<int>(in.size()) != op.active_in_l_size)
    throw std::invalid_argument("operator_apply: input vector size does not match active_in_l_size");
  const int Q = op.num_qpts * op.num_elem;
  std::vector<std::vector<double>> q_in(op.qf.inputs.size()), q_out;
  std::vector<double> e;
  for (size_t f = 0; f < op.inputs.size(); ++f) {
    const OperatorField& fld = *op.inputs[f];
    const std::vector<double>& src = (fld.vector.kind == VectorKind::Active) ? in : fld.vector.data;
    if (fld.eval_mode == EvalMode::Weight) {
      tensor_basis_apply_weight(*fld.basis, op.num_elem, ContractMode::NoTranspose, q_in[f]);
    } else if (fld.eval_mode == EvalMode::None) {
      elem_restriction_apply(*fld.restriction, ContractMode::NoTranspose, src, q_in[f]);
    } else {
      elem_restriction_apply(*fld.restriction, ContractMode::NoTranspose, src, e);
      tensor_basis_apply(*fld.basis, op.num_elem, ContractMode::NoTranspose, fld.eval_mode, e, q_in[f]);
    }
  }
  qfunction_apply(op.qf, Q, q_in, q_out);
  out.assign(op.active_out_l_size, 0.0);
  for (size_t f = 0; f < op.outputs.size(); ++f) {
    const OperatorField& fld = *op.outputs[f];
    if (fld.eval_mode == EvalMode::None) {
      elem_restriction_apply(*fld.restriction, ContractMode::Transpose, q_out[f], out);
    } else {
      tensor_basis_apply(*fld.basis, op.num_elem, ContractMode::Transpose, fld.eval_mode, q_out[f], e);
      elem_restriction_apply(*fld.restriction, ContractMode::Transpose, e, out);
    }
  }
}

}  // namespace fem
