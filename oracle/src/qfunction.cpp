// ORACLE (test infrastructure only). CPU restatement of src/qfunction.cpp:14-111.
#include "qfunction.hpp"

#include <stdexcept>

namespace fem {

// src/qfunction.cpp:14-20
void qfunction_create(QFunctionUser user, QFunction& qf) {
  if (!user) throw std::invalid_argument("QFunction user function pointer is null");
  qf = QFunction{};
  qf.user = user;
}

namespace {
// src/qfunction.cpp:23-36: sizes positive, Weight size 1, names unique over inputs and outputs.
void validate_field(const QFunction& qf, const std::string& name, int size, EvalMode mode, const char* who) {
  if (size <= 0) throw std::invalid_argument(std::string(who) + ": size must be positive");
  if (mode == EvalMode::Weight && size != 1)
    throw std::invalid_argument(std::string(who) + ": an EvalMode::Weight field must have size 1");
  for (const auto& f : qf.inputs)
    if (f.name == name) throw std::invalid_argument(std::string(who) + ": duplicate field name '" + name + "'");
  for (const auto& f : qf.outputs)
    if (f.name == name) throw std::invalid_argument(std::string(who) + ": duplicate field name '" + name + "'");
}
}  // namespace

void qfunction_add_input(QFunction& qf, const std::string& name, int size, EvalMode eval_mode) {
  validate_field(qf, name, size, eval_mode, "qfunction_add_input");
  qf.inputs.push_back({name, size, eval_mode});
}

void qfunction_add_output(QFunction& qf, const std::string& name, int size, EvalMode eval_mode) {
  if (eval_mode == EvalMode::Weight) throw std::invalid_argument("qfunction_add_output: EvalMode::Weight is input-only");
  validate_field(qf, name, size, eval_mode, "qfunction_add_output");
  qf.outputs.push_back({name, size, eval_mode});
}

// src/qfunction.cpp:81-111
void qfunction_apply(QFunction& qf, int Q, const std::vector<std::vector<double>>& in,
                     std::vector<std::vector<double>>& out) {
  if (Q <= 0) throw std::invalid_argument("qfunction_apply: Q must be positive");
  if (in.size() != qf.inputs.size())
    throw std::invalid_argument("qfunction_apply: number of input fields does not match QFunction");
  for (size_t f = 0; f < in.size(); ++f)
    if (in[f].size() != static_cast<size_t>(qf.inputs[f].size * Q))
      throw std::invalid_argument("qfunction_apply: input field " + qf.inputs[f].name + " has incorrect size");
  out.resize(qf.outputs.size());
  for (size_t f = 0; f < out.size(); ++f) out[f].resize(static_cast<size_t>(qf.outputs[f].size) * Q);
  std::vector<const double*> ip;
  std::vector<double*> op;
  for (const auto& x : in) ip.push_back(x.data());
  for (auto& x : out) op.push_back(x.data());
  void* ctx = qf.ctx.empty() ? nullptr : qf.ctx.data();
  if (qf.user(ctx, Q, ip.data(), op.data()) != 0)
    throw std::runtime_error("qfunction_apply: user function returned non-zero error code");
}

}  // namespace fem
