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

// Diffusion QFunctions (no reference counterpart): gradient fields [d][comp][Q],
// J(d,c) = dx[(d*dim + c)*Q + p] as in build_mass. build: qdata = (w / det) C C^T
// with C the cofactor matrix of J, det expanded along row 0 as in build_mass,
// stored symmetric as (00),(11),(22),(12),(02),(01) (dim 3), (00),(11),(01)
// (dim 2), (00) (dim 1). apply: dv[a][c] = sum_b qdata(a,b) du[b][c], b
// ascending. Sums start from 0 and add left to right, without fused
// multiply-add, like the CPU statement of the same functions.
__device__ __forceinline__ int sym_index(int dim, int a, int b) {
  if (a == b) return a;
  if (dim == 2) return 2;
  const int lo = a < b ? a : b, hi = a < b ? b : a;
  return lo == 1 ? 3 : (hi == 2 ? 4 : 5);
}

template <int DIM>
__global__ __launch_bounds__(kThreads) void build_diffusion_kernel(int Q, const double* __restrict__ J,
                                                                   const double* __restrict__ w,
                                                                   double* __restrict__ qdata) {
  const int p = blockIdx.x * kThreads + threadIdx.x;
  if (p >= Q) return;
  auto A = [&](int d, int c) { return J[static_cast<size_t>(d * DIM + c) * Q + p]; };
  double C[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  double det;
  if constexpr (DIM == 1) {
    det = A(0, 0);
    C[0][0] = 1.0;
  } else if constexpr (DIM == 2) {
    det = A(0, 0) * A(1, 1) - A(0, 1) * A(1, 0);
    C[0][0] = A(1, 1);
    C[0][1] = -A(1, 0);
    C[1][0] = -A(0, 1);
    C[1][1] = A(0, 0);
  } else {
    const double m0 = A(1, 1) * A(2, 2) - A(1, 2) * A(2, 1);
    const double m1 = A(1, 0) * A(2, 2) - A(1, 2) * A(2, 0);
    const double m2 = A(1, 0) * A(2, 1) - A(1, 1) * A(2, 0);
    det = A(0, 0) * m0 - A(0, 1) * m1 + A(0, 2) * m2;
    C[0][0] = m0;
    C[0][1] = -m1;
    C[0][2] = m2;
    C[1][0] = A(0, 2) * A(2, 1) - A(0, 1) * A(2, 2);
    C[1][1] = A(0, 0) * A(2, 2) - A(0, 2) * A(2, 0);
    C[1][2] = A(0, 1) * A(2, 0) - A(0, 0) * A(2, 1);
    C[2][0] = A(0, 1) * A(1, 2) - A(0, 2) * A(1, 1);
    C[2][1] = A(0, 2) * A(1, 0) - A(0, 0) * A(1, 2);
    C[2][2] = A(0, 0) * A(1, 1) - A(0, 1) * A(1, 0);
  }
  const double scale = w[p] / det;
#pragma unroll
  for (int a = 0; a < DIM; ++a) {
#pragma unroll
    for (int b = a; b < DIM; ++b) {
      double s = 0.0;
#pragma unroll
      for (int k = 0; k < DIM; ++k) s = s + C[a][k] * C[b][k];
      qdata[static_cast<size_t>(sym_index(DIM, a, b)) * Q + p] = scale * s;
    }
  }
}

template <int DIM>
__global__ __launch_bounds__(kThreads) void apply_diffusion_kernel(int Q, int nc, const double* __restrict__ du,
                                                                   const double* __restrict__ qd,
                                                                   double* __restrict__ dv) {
  const int p = blockIdx.x * kThreads + threadIdx.x;
  if (p >= Q) return;
  for (int c = 0; c < nc; ++c) {
#pragma unroll
    for (int a = 0; a < DIM; ++a) {
      double s = 0.0;
#pragma unroll
      for (int b = 0; b < DIM; ++b)
        s = s + qd[static_cast<size_t>(sym_index(DIM, a, b)) * Q + p] * du[static_cast<size_t>(b * nc + c) * Q + p];
      dv[static_cast<size_t>(a * nc + c) * Q + p] = s;
    }
  }
}

int ctx_int(const QFunctionDev& qf, const char* who) {
  if (qf.ctx.size() < sizeof(int)) throw std::invalid_argument(std::string(who) + ": context missing");
  int v;
  std::memcpy(&v, qf.ctx.data(), sizeof(int));
  return v;
}

}  // namespace

QFunctionDev* qfunction_create(int kind) {
  if (kind < kQfBuildMass || kind > kQfApplyDiffusion)
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
    case kQfBuildDiffusion: {
      require(num_in >= 2 && num_out >= 1, "build_diffusion: needs inputs dx, weights and output qdata");
      const int dim = ctx_int(qf, "build_diffusion");
      if (dim < 1 || dim > 3 || qf.inputs[0].size != dim * dim || qf.outputs[0].size != dim * (dim + 1) / 2)
        throw std::runtime_error("qfunction_apply: user function returned non-zero error code");
      if (dim == 1) build_diffusion_kernel<1><<<grid, kThreads, 0, s>>>(Q, in[0], in[1], out[0]);
      if (dim == 2) build_diffusion_kernel<2><<<grid, kThreads, 0, s>>>(Q, in[0], in[1], out[0]);
      if (dim == 3) build_diffusion_kernel<3><<<grid, kThreads, 0, s>>>(Q, in[0], in[1], out[0]);
      break;
    }
    case kQfApplyDiffusion: {
      require(num_in >= 2 && num_out >= 1, "apply_diffusion: needs inputs du, qdata and output dv");
      require(qf.ctx.size() >= 2 * sizeof(int), "apply_diffusion: context {dim, num_comp} missing");
      int dn[2];
      std::memcpy(dn, qf.ctx.data(), sizeof(dn));
      const int dim = dn[0], nc = dn[1];
      if (dim < 1 || dim > 3 || nc < 1 || qf.inputs[0].size != dim * nc || qf.outputs[0].size != dim * nc ||
          qf.inputs[1].size != dim * (dim + 1) / 2)
        throw std::runtime_error("qfunction_apply: user function returned non-zero error code");
      if (dim == 1) apply_diffusion_kernel<1><<<grid, kThreads, 0, s>>>(Q, nc, in[0], in[1], out[0]);
      if (dim == 2) apply_diffusion_kernel<2><<<grid, kThreads, 0, s>>>(Q, nc, in[0], in[1], out[0]);
      if (dim == 3) apply_diffusion_kernel<3><<<grid, kThreads, 0, s>>>(Q, nc, in[0], in[1], out[0]);
      break;
    }
    default:
      throw Unsupported("qfunction_apply: no device implementation for this QFunction");
  }
  HIP_CHECK_LAUNCH();
}

}  // namespace femhip
