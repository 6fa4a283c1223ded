// ORACLE (test infrastructure only). CPU restatement of src/tensor-contract.cpp:35-222.
// Loop nests, the per-output "sum from 0.0 over b, then add to v" association,
// the x-axis-first pass order, the Transpose P<->Q swap, the grad block handling
// (including the Transpose accumulation into an un-zeroed v) and the weight
// product order w_z*(w_y*w_x) are kept as in the reference.
#include "tensor-contract.hpp"

#include <algorithm>
#include <stdexcept>

namespace fem {

namespace {

int ipow(int b, int e) {
  int r = 1;
  while (e-- > 0) r *= b;
  return r;
}

// v[(a*J + j)*C + c] (+)= sum_b t(j, b) * u[(a*B + b)*C + c]   (src/tensor-contract.cpp:35-59)
void contract(int A, int B, int C, int J, const double* t, ContractMode t_mode, bool add, const double* u,
              std::vector<double>& v) {
  const int sj = (t_mode == ContractMode::NoTranspose) ? B : 1;
  const int sb = (t_mode == ContractMode::NoTranspose) ? 1 : J;
  v.resize(static_cast<size_t>(A * J * C));
  if (!add) std::fill(v.begin(), v.end(), 0.0);
  for (int a = 0; a < A; ++a) {
    const double* ua = u + static_cast<size_t>(a) * B * C;
    double* va = v.data() + static_cast<size_t>(a) * J * C;
    for (int j = 0; j < J; ++j) {
      for (int c = 0; c < C; ++c) {
        double acc = 0.0;
        for (int b = 0; b < B; ++b) acc += t[j * sj + b * sb] * ua[b * C + c];
        va[j * C + c] += acc;
      }
    }
  }
}

}  // namespace

void tensor_contract_apply(int A, int B, int C, int J, const std::vector<double>& t, ContractMode t_mode, bool add,
                           const std::vector<double>& u, std::vector<double>& v) {
  contract(A, B, C, J, t.data(), t_mode, add, u.data(), v);
}

// src/tensor-contract.cpp:74-91
void tensor_basis_apply_interp(const TensorBasis& basis, int num_elem, ContractMode t_mode,
                               const std::vector<double>& u, std::vector<double>& v) {
  const bool tr = (t_mode == ContractMode::Transpose);
  const int in_1d = tr ? basis.Q_1d : basis.P_1d;
  const int out_1d = tr ? basis.P_1d : basis.Q_1d;
  const int dim = basis.dim;
  int pre = basis.num_comp * ipow(in_1d, dim - 1);
  int post = num_elem;
  std::vector<double> scratch[2];
  for (int d = 0; d < dim; ++d) {
    const double* src = (d == 0) ? u.data() : scratch[d % 2].data();
    std::vector<double>& dst = (d == dim - 1) ? v : scratch[(d + 1) % 2];
    contract(pre, in_1d, post, out_1d, basis.interp_1d.data(), t_mode, false, src, dst);
    pre /= in_1d;
    post *= out_1d;
  }
}

// src/tensor-contract.cpp:111-156
void tensor_basis_apply_grad(const TensorBasis& basis, int num_elem, ContractMode t_mode,
                             const std::vector<double>& u, std::vector<double>& v) {
  const bool tr = (t_mode == ContractMode::Transpose);
  const int dim = basis.dim, nc = basis.num_comp;
  const int in_1d = tr ? basis.Q_1d : basis.P_1d;
  const int out_1d = tr ? basis.P_1d : basis.Q_1d;
  // Block size along the "quadrature" side (reference Qdim, recomputed after the swap).
  const int q_block = nc * ipow(basis.Q_1d, dim) * num_elem;
  v.resize(static_cast<size_t>(dim) * q_block);
  if (tr) v.resize(static_cast<size_t>(nc) * ipow(out_1d, dim) * num_elem);
  for (int d_axis = 0; d_axis < dim; ++d_axis) {
    int pre = nc * ipow(in_1d, dim - 1), post = num_elem;
    std::vector<double> scratch[2], block;
    const double* u_in = tr ? u.data() + static_cast<size_t>(d_axis) * q_block : u.data();
    for (int d = 0; d < dim; ++d) {
      const double* mat = (d == d_axis) ? basis.grad_1d.data() : basis.interp_1d.data();
      const bool last = (d == dim - 1);
      const bool add = last && tr;
      const double* src = (d == 0) ? u_in : scratch[d % 2].data();
      std::vector<double>& dst = last ? (tr ? v : block) : scratch[(d + 1) % 2];
      contract(pre, in_1d, post, out_1d, mat, t_mode, add, src, dst);
      pre /= in_1d;
      post *= out_1d;
    }
    if (!tr) std::copy(block.begin(), block.end(), v.begin() + static_cast<size_t>(d_axis) * q_block);
  }
}

// src/tensor-contract.cpp:169-192: pass d multiplies in q_weight_1d along axis d,
// reading back the value written by pass d-1 at element 0.
void tensor_basis_apply_weight(const TensorBasis& basis, int num_elem, ContractMode t_mode, std::vector<double>& v) {
  if (t_mode == ContractMode::Transpose)
    throw std::invalid_argument("EvalMode::Weight is incompatible with ContractMode::Transpose");
  const int dim = basis.dim, Q = basis.Q_1d;
  v.resize(static_cast<size_t>(num_elem) * ipow(Q, dim));
  for (int d = 0; d < dim; ++d) {
    const int pre = ipow(Q, dim - d - 1), post = ipow(Q, d);
    for (int i = 0; i < pre; ++i)
      for (int j = 0; j < Q; ++j)
        for (int k = 0; k < post; ++k) {
          const size_t q = static_cast<size_t>((i * Q + j) * post + k);
          const double prev = (d == 0) ? 1.0 : v[q * num_elem];
          const double w = basis.q_weight_1d[j] * prev;
          for (int e = 0; e < num_elem; ++e) v[q * num_elem + e] = w;
        }
  }
}

// src/tensor-contract.cpp:207-222
void tensor_basis_apply(const TensorBasis& basis, int num_elem, ContractMode t_mode, EvalMode eval_mode,
                        const std::vector<double>& u, std::vector<double>& v) {
  switch (eval_mode) {
    case EvalMode::Interp: tensor_basis_apply_interp(basis, num_elem, t_mode, u, v); return;
    case EvalMode::Grad: tensor_basis_apply_grad(basis, num_elem, t_mode, u, v); return;
    case EvalMode::Weight: tensor_basis_apply_weight(basis, num_elem, t_mode, v); return;
    default: throw std::invalid_argument("tensor_basis_apply: unknown eval_mode");
  }
}

}  // namespace fem
