// Device QFunctions: the reference's field/context model (include/qfunction.hpp:20-61)
// with a registry of device implementations in place of the host function pointer.
#pragma once

#include <string>
#include <vector>

#include "common.hpp"

namespace femhip {

enum : int { kQfBuildMass = 1, kQfApplyMass = 2, kQfApplyMassVector = 3, kQfBuildDiffusion = 4, kQfApplyDiffusion = 5 };

struct QFField {
  std::string name;
  int size;
  int eval_mode;
};

struct QFunctionDev {
  int kind = 0;
  std::vector<QFField> inputs, outputs;
  std::vector<unsigned char> ctx;
};

QFunctionDev* qfunction_create(int kind);
void qfunction_add_input(QFunctionDev& qf, const std::string& name, int size, int eval_mode);
void qfunction_add_output(QFunctionDev& qf, const std::string& name, int size, int eval_mode);
// qfunction_apply (src/qfunction.cpp:81-111): validates counts and sizes, then runs
// the device implementation of qf.kind on [size][Q] device arrays.
void qfunction_apply(const QFunctionDev& qf, int Q, int num_in, const double* const* in, const size_t* in_lens,
                     int num_out, double* const* out, const size_t* out_lens, hipStream_t s);

}  // namespace femhip
