#pragma once
// ORACLE (test infrastructure only). Restates the interface of include/basis.hpp:7-40.
#include <vector>

namespace fem {

void gauss_legendre_quadrature(int Q, std::vector<double>& q_ref_1d, std::vector<double>& q_weight_1d);
void gauss_lobatto_points(int Q, std::vector<double>& q_ref_1d, std::vector<double>* q_weight_1d = nullptr);
void lagrange_basis_matrix(const std::vector<double>& nodes, const std::vector<double>& eval_points,
                           std::vector<double>& interp_1d, std::vector<double>& grad_1d);

enum class QuadMode { Gauss, GaussLobatto };

struct TensorBasis {
  int dim;
  int num_comp;
  int P_1d;
  int Q_1d;
  std::vector<double> q_ref_1d;
  std::vector<double> q_weight_1d;
  std::vector<double> interp_1d;  // [Q_1d][P_1d]
  std::vector<double> grad_1d;    // [Q_1d][P_1d]

  static TensorBasis create_tensor_H1_lagrange(int dim, int num_comp, int P_1d, int Q_1d,
                                               QuadMode quad_mode = QuadMode::Gauss);
};

}  // namespace fem
