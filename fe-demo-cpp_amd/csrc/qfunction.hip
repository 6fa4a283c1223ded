// Pointwise QFunction kernels. One thread per quadrature point p, component c
// at [c*Q + p] (coalesced). No fused multiply-add (-ffp-contract=off), and the
// determinant is the cofactor expansion along row 0 with the association of
// include/qfunctions/mass.hpp:31-33, so outputs equal the reference's bits.
#include <hip/hip_runtime.h>

#include <cstring>

#include "qfunction.hpp"

namespace femhip {

namespace {

constexpr int kThreads = 256;

template <int DIM>
__global__ __launch_bounds__(kThreads) void build_mass_kernel(int Q, const double* __restrict__ J,
                                                              const double* __restrict__ w,
                                                              double* __restrict__ qdata) {
  const int p = blockIdx.x * kThreads + threadIdx.x;
  if (p >= Q) return;
  auto j = [&](int d, int c) { return J[static_cast<size_t>(d * DIM + c) * Q + p]; };
  double det;
  if constexpr (DIM == 1) {
    det = j(0, 0);
  } else if constexpr (DIM == 2) {
    det = j(0, 0) * j(1, 1) - j(0, 1) * j(1, 0);
  } else {
    det = j(0, 0) * (j(1, 1) * j(2, 2) - j(1, 2) * j(2, 1)) - j(0, 1) * (j(1, 0) * j(2, 2) - j(1, 2) * j(2, 0)) +
          j(0, 2) * (j(1, 0) * j(2, 1) - j(1, 1) * j(2, 0));
  }
  qdata[p] = det * w[p];
}

__global__ __launch_bounds__(kThreads) void apply_mass_kernel(size_t n, size_t Q, const double* __restrict__ u,
                                                              const double* __restrict__ qd,
                                                              double* __restrict__ v) {
  const size_t i = static_cast<size_t>(blockIdx.x) * kThreads + threadIdx.x;
  if (i >= n) return;
  v[i] = qd[i % Q] * u[i];
}

int ctx_int(const QFunctionDev& qf, const char* who) {
  if (qf.ctx.size() < sizeof(int)) throw std::invalid_argument(std::string(who) + ": context missing");
  int v;
  std::memcpy(&v, qf.ctx.data(), sizeof(int));
  return v;
}

}  // namespace

QFunctionDev* qfunction_create(int kind) {
  if (kind != kQfBuildMass && kind != kQfApplyMass && kind != kQfApplyMassVector)
    throw Unsupported("qfunction_create: no device implementation for this QFunction");
  auto* q = new QFunctionDev();
  q->kind = kind;
  return q;
}

static void check_field(const QFunctionDev& qf, const std::string& name, int size, int mode, const char* who) {
  require(size > 0, std::string(who) + ": size must be positive");
  require(mode >= kEvalNone && mode <= kEvalWeight, std::string(who) + ": unknown eval_mode");
  require(!(mode == kEvalWeight && size != 1), std::string(who) + ": an EvalMode::Weight field must have size 1");
  for (const auto* fs : {&qf.inputs, &qf.outputs})
    for (const auto& f : *fs) require(f.name != name, std::string(who) + ": duplicate field name '" + name + "'");
}

void qfunction_add_input(QFunctionDev& qf, const std::string& name, int size, int mode) {
  check_field(qf, name, size, mode, "qfunction_add_input");
  qf.inputs.push_back({name, size, mode});
}

void qfunction_add_output(QFunctionDev& qf, const std::string& name, int size, int mode) {
  require(mode != kEvalWeight, "qfunction_add_output: EvalMode::Weight is input-only");
  check_field(qf, name, size, mode, "qfunction_add_output");
  qf.outputs.push_back({name, size, mode});
}

void qfunction_apply(const QFunctionDev& qf, int Q, int num_in, const double* const* in, const size_t* in_lens,
                     int num_out, double* const* out, const size_t* out_lens, hipStream_t s) {
  require(Q > 0, "qfunction_apply: Q must be positive");
  require(num_in == static_cast<int>(qf.inputs.size()),
          "qfunction_apply: number of input fields does not match QFunction");
  for (int f = 0; f < num_in; ++f)
    require(in_lens[f] == static_cast<size_t>(qf.inputs[f].size) * Q,
            "qfunction_apply: input field " + qf.inputs[f].name + " has incorrect size");
  require(num_out == static_cast<int>(qf.outputs.size()),
          "qfunction_apply: number of output fields does not match QFunction");
  for (int f = 0; f < num_out; ++f)
    require(out_lens[f] == static_cast<size_t>(qf.outputs[f].size) * Q,
            "qfunction_apply: output field " + qf.outputs[f].name + " has incorrect size");
  const unsigned grid = grid_for(static_cast<size_t>(Q), kThreads);
  switch (qf.kind) {
    case kQfBuildMass: {
      require(num_in >= 2 && num_out >= 1, "build_mass: needs inputs dx, weights and output qdata");
      const int dim = ctx_int(qf, "build_mass");
      // The reference user function returns 1 for another dim, and qfunction_apply
      // turns that into a runtime_error (include/qfunctions/mass.hpp:34, src/qfunction.cpp:108-110).
      if (dim < 1 || dim > 3 || qf.inputs[0].size < dim * dim)
        throw std::runtime_error("qfunction_apply: user function returned non-zero error code");
      if (dim == 1) build_mass_kernel<1><<<grid, kThreads, 0, s>>>(Q, in[0], in[1], out[0]);
      if (dim == 2) build_mass_kernel<2><<<grid, kThreads, 0, s>>>(Q, in[0], in[1], out[0]);
      if (dim == 3) build_mass_kernel<3><<<grid, kThreads, 0, s>>>(Q, in[0], in[1], out[0]);
      break;
    }
    case kQfApplyMass: {
      require(num_in >= 2 && num_out >= 1, "apply_mass: needs inputs u, qdata and output v");
      apply_mass_kernel<<<grid, kThreads, 0, s>>>(static_cast<size_t>(Q), static_cast<size_t>(Q), in[0], in[1],
                                                  out[0]);
      break;
    }
    case kQfApplyMassVector: {
      require(num_in >= 2 && num_out >= 1, "apply_mass_vector: needs inputs u, qdata and output v");
      const int nc = ctx_int(qf, "apply_mass_vector");
      if (nc < 1 || qf.inputs[0].size != nc || qf.outputs[0].size != nc || qf.inputs[1].size != 1)
        throw std::runtime_error("qfunction_apply: user function returned non-zero error code");
      const size_t n = static_cast<size_t>(nc) * Q;
      apply_mass_kernel<<<grid_for(n, kThreads), kThreads, 0, s>>>(n, static_cast<size_t>(Q), in[0], in[1], out[0]);
      break;
    }
    default:
      throw Unsupported("qfunction_apply: no device implementation for this QFunction");
  }
  HIP_CHECK_LAUNCH();
}

}  // namespace femhip
