#pragma once
// ORACLE (test infrastructure only). Restates the interface of include/qfunction.hpp:20-61.
#include <cstring>
#include <string>
#include <type_traits>
#include <vector>
#include "eval-mode.hpp"

namespace fem {

// Pointwise kernel on Q points; field f holds [size][Q].
using QFunctionUser = int (*)(void* ctx, int Q, const double* const* in, double* const* out);

struct QFunctionField {
  std::string name;
  int size;
  EvalMode eval_mode;
};

struct QFunction {
  QFunctionUser user = nullptr;
  std::vector<QFunctionField> inputs;
  std::vector<QFunctionField> outputs;
  std::vector<unsigned char> ctx;
};

void qfunction_create(QFunctionUser user, QFunction& qf);
void qfunction_add_input(QFunction& qf, const std::string& name, int size, EvalMode eval_mode);
void qfunction_add_output(QFunction& qf, const std::string& name, int size, EvalMode eval_mode);

template <class T>
void qfunction_set_context(QFunction& qf, const T& data) {
  static_assert(std::is_trivially_copyable_v<T>, "context must be trivially copyable");
  qf.ctx.assign(sizeof(T), 0);
  std::memcpy(qf.ctx.data(), &data, sizeof(T));
}

void qfunction_apply(QFunction& qf, int Q, const std::vector<std::vector<double>>& in,
                     std::vector<std::vector<double>>& out);

}  // namespace fem
