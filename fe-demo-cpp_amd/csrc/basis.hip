// Unfused tensor-basis kernels on the reference layouts (element index innermost).
//
// One thread per (a, c) column of a contraction: it loads the B inputs of its
// column into registers (coalesced across c, which is the element index for
// the first pass) and writes the J outputs. Every output is
// "sum = 0; sum += t*u over b ascending; v = v + sum", with no fused
// multiply-add (the library is built with -ffp-contract=off), which is the
// arithmetic of src/tensor-contract.cpp:48-58: the unfused path reproduces the
// reference's bits.
#include <hip/hip_runtime.h>

#include "basis.hpp"

namespace femhip {

namespace {

constexpr int kThreads = 256;

template <int MAXB>
__global__ __launch_bounds__(kThreads) void contract_kernel(int A, int B, int C, int J, const double* __restrict__ t,
                                                            int sj, int sb, int add, const double* __restrict__ u,
                                                            double* __restrict__ v) {
  const size_t idx = static_cast<size_t>(blockIdx.x) * kThreads + threadIdx.x;
  if (idx >= static_cast<size_t>(A) * C) return;
  const size_t c = idx % C, a = idx / C;
  double ub[MAXB];
  const double* ua = u + a * B * C + c;
#pragma unroll
  for (int b = 0; b < MAXB; ++b) ub[b] = (b < B) ? ua[static_cast<size_t>(b) * C] : 0.0;
  double* va = v + a * J * C + c;
  for (int j = 0; j < J; ++j) {
    double acc = 0.0;
#pragma unroll
    for (int b = 0; b < MAXB; ++b)
      if (b < B) acc += t[j * sj + b * sb] * ub[b];
    double* dst = va + static_cast<size_t>(j) * C;
    *dst = add ? *dst + acc : 0.0 + acc;
  }
}

// w[q*ne + e] with q = (qz*Q + qy)*Q + qx: the reference forms w_x, then
// w_y*(w_x), then w_z*(w_y*w_x) (src/tensor-contract.cpp:180-191).
__global__ __launch_bounds__(kThreads) void weight_kernel(int dim, int Q, int ne, const double* __restrict__ qw,
                                                          double* __restrict__ v) {
  const size_t idx = static_cast<size_t>(blockIdx.x) * kThreads + threadIdx.x;
  size_t nq = 1;
  for (int d = 0; d < dim; ++d) nq *= Q;
  if (idx >= nq * ne) return;
  size_t q = idx / ne;
  double w = 1.0;
  for (int d = 0; d < dim; ++d) {
    const int qd = static_cast<int>(q % Q);
    q /= Q;
    w = qw[qd] * w;
  }
  v[idx] = w;
}

}  // namespace

Basis* basis_create(const HostBasis& hb) {
  require(hb.dim >= 1 && hb.dim <= 3, "basis: dim must be 1, 2, or 3");
  require(hb.num_comp >= 1 && hb.P >= 1 && hb.Q >= 1, "basis: bad sizes");
  require(hb.interp.size() == static_cast<size_t>(hb.P) * hb.Q && hb.grad.size() == hb.interp.size() &&
              hb.q_weight.size() == static_cast<size_t>(hb.Q),
          "basis: 1D arrays have the wrong size");
  auto* b = new Basis();
  try {
    b->h = hb;
    b->interp.upload(hb.interp);
    b->grad.upload(hb.grad);
    b->q_weight.upload(hb.q_weight);
  } catch (...) {
    delete b;
    throw;
  }
  return b;
}

void contract(int A, int B, int C, int J, const double* t, int t_mode, bool add, const double* u, double* v,
              hipStream_t s) {
  require(A > 0 && B > 0 && C > 0 && J > 0, "tensor_contract_apply: sizes must be positive");
  require(B <= 32, "tensor_contract_apply: B > 32 is not supported on the device");
  require(t != nullptr && u != nullptr && v != nullptr, "tensor_contract_apply: null device pointer");
  const int sj = (t_mode == kNoTranspose) ? B : 1;
  const int sb = (t_mode == kNoTranspose) ? 1 : J;
  const unsigned grid = grid_for(static_cast<size_t>(A) * C, kThreads);
  if (B <= 8) {
    contract_kernel<8><<<grid, kThreads, 0, s>>>(A, B, C, J, t, sj, sb, add ? 1 : 0, u, v);
  } else if (B <= 16) {
    contract_kernel<16><<<grid, kThreads, 0, s>>>(A, B, C, J, t, sj, sb, add ? 1 : 0, u, v);
  } else {
    contract_kernel<32><<<grid, kThreads, 0, s>>>(A, B, C, J, t, sj, sb, add ? 1 : 0, u, v);
  }
  HIP_CHECK_LAUNCH();
}

size_t basis_in_len(const HostBasis& b, int ne, int t_mode, int eval_mode) {
  const size_t p = static_cast<size_t>(ipow(b.P, b.dim)), q = static_cast<size_t>(ipow(b.Q, b.dim));
  if (eval_mode == kEvalWeight) return 0;
  if (t_mode == kNoTranspose) return b.num_comp * p * ne;
  return (eval_mode == kEvalGrad ? b.dim : 1) * b.num_comp * q * ne;
}

size_t basis_out_len(const HostBasis& b, int ne, int t_mode, int eval_mode) {
  const size_t p = static_cast<size_t>(ipow(b.P, b.dim)), q = static_cast<size_t>(ipow(b.Q, b.dim));
  if (eval_mode == kEvalWeight) return q * ne;
  if (t_mode == kTranspose) return b.num_comp * p * ne;
  return (eval_mode == kEvalGrad ? b.dim : 1) * b.num_comp * q * ne;
}

void basis_apply(const Basis& bas, int ne, int t_mode, int eval_mode, const double* u, size_t u_len, double* v,
                 size_t v_len, hipStream_t s) {
  const HostBasis& b = bas.h;
  require(ne > 0, "tensor_basis_apply: num_elem must be positive");
  require(t_mode == kNoTranspose || t_mode == kTranspose, "tensor_basis_apply: unknown transpose mode");
  require(eval_mode == kEvalInterp || eval_mode == kEvalGrad || eval_mode == kEvalWeight,
          "tensor_basis_apply: unknown eval_mode");
  if (eval_mode == kEvalWeight)
    require(t_mode == kNoTranspose, "EvalMode::Weight is incompatible with ContractMode::Transpose");
  require(u_len == basis_in_len(b, ne, t_mode, eval_mode), "tensor_basis_apply: u has the wrong size");
  require(v_len == basis_out_len(b, ne, t_mode, eval_mode), "tensor_basis_apply: v has the wrong size");
  require(v != nullptr && (eval_mode == kEvalWeight || u != nullptr), "tensor_basis_apply: null device pointer");

  if (eval_mode == kEvalWeight) {
    weight_kernel<<<grid_for(v_len, kThreads), kThreads, 0, s>>>(b.dim, b.Q, ne, bas.q_weight.data(), v);
    HIP_CHECK_LAUNCH();
    return;
  }
  const bool tr = (t_mode == kTranspose);
  const int in1 = tr ? b.Q : b.P, out1 = tr ? b.P : b.Q, dim = b.dim, nc = b.num_comp;
  // Largest intermediate: pre*out1*post over the passes.
  size_t tmp_len = 0;
  {
    size_t pre = static_cast<size_t>(nc) * ipow(in1, dim - 1), post = ne;
    for (int d = 0; d < dim; ++d) {
      tmp_len = std::max(tmp_len, pre * out1 * post);
      pre /= in1;
      post *= out1;
    }
  }
  // Pass d writes tmp[(d+1)%2] (the last pass writes v) and pass d+1 reads it:
  // dim 2 uses tmp[1]; dim 3 uses tmp[1] then tmp[0].
  StreamTmp<double> t0(dim > 2 ? tmp_len : 0, s), t1(dim > 1 ? tmp_len : 0, s);
  double* tmp[2] = {t0.data(), t1.data()};

  // One sum-factorized pass sequence with matrix `mat_for(d)` per axis d.
  auto passes = [&](const double* src0, double* dst_last, bool add_last, auto mat_for) {
    int pre = nc * ipow(in1, dim - 1), post = ne;
    for (int d = 0; d < dim; ++d) {
      const double* src = (d == 0) ? src0 : tmp[d % 2];
      double* dst = (d == dim - 1) ? dst_last : tmp[(d + 1) % 2];
      contract(pre, in1, post, out1, mat_for(d), t_mode, (d == dim - 1) && add_last, src, dst, s);
      pre /= in1;
      post *= out1;
    }
  };

  if (eval_mode == kEvalInterp) {
    passes(u, v, false, [&](int) { return bas.interp.data(); });
    return;
  }
  // Grad: block d_axis uses grad_1d on axis d_axis, interp_1d elsewhere.
  const size_t q_block = static_cast<size_t>(nc) * ipow(b.Q, dim) * ne;
  if (!tr) {
    for (int da = 0; da < dim; ++da)
      passes(u, v + da * q_block, false,
             [&](int d) { return d == da ? bas.grad.data() : bas.interp.data(); });
  } else {
    // Sum of the dim adjoint blocks; v is cleared first (the reference adds into
    // whatever v held, src/tensor-contract.cpp:115-122,142).
    zero_fill(v, v_len * sizeof(double), s);
    for (int da = 0; da < dim; ++da)
      passes(u + da * q_block, v, true, [&](int d) { return d == da ? bas.grad.data() : bas.interp.data(); });
  }
}

}  // namespace femhip
