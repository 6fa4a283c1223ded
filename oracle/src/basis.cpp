// ORACLE (test infrastructure only). CPU restatement of src/basis.cpp:16-228 that
// keeps every floating-point operation in the reference's order (Newton start
// values, stopping rule, odd/even middle-slot overwrites, Fornberg recurrence),
// so that compiled with -ffp-contract=off it reproduces the reference bits.
#include "basis.hpp"

#include <cmath>
#include <limits>
#include <stdexcept>

namespace fem {

namespace {

// Three-term Legendre recurrence up to degree n at x. On return p_n holds P_n(x),
// p_nm1 holds P_{n-1}(x). Operation order follows src/basis.cpp:39-43.
void legendre_upto(int n, double x, double& p_nm1, double& p_n) {
  double a = 1.0, b = x, c = 0.0;
  for (int j = 2; j <= n; ++j) {
    c = ((2.0 * j - 1.0) * x * b - (j - 1.0) * a) / j;
    a = b;
    b = c;
  }
  p_nm1 = a;
  p_n = c;  // stays 0.0 when n < 2, exactly like the reference's P2 initialisation
}

}  // namespace

// src/basis.cpp:16-68
void gauss_legendre_quadrature(int Q, std::vector<double>& q_ref_1d, std::vector<double>& q_weight_1d) {
  if (Q <= 0) throw std::invalid_argument("gauss_legendre_quadrature: Q must be positive");
  q_ref_1d.resize(Q);
  q_weight_1d.resize(Q);
  const double pi = 4.0 * atan(1.0);
  const double tol = 10 * std::numeric_limits<double>::epsilon();
  for (int i = 0; i <= Q / 2; ++i) {
    double x = std::cos(pi * (2.0 * i + 1.0) / (2.0 * Q));
    double pm1, pq;
    legendre_upto(Q, x, pm1, pq);
    double dpq = (x * pq - pm1) * Q / (x * x - 1.0);
    x = x - pq / dpq;
    for (int k = 0; k < 100 && std::fabs(pq) > tol; ++k) {
      legendre_upto(Q, x, pm1, pq);
      dpq = (x * pq - pm1) * Q / (x * x - 1.0);
      x -= pq / dpq;
    }
    const double w = 2.0 / ((1.0 - x * x) * dpq * dpq);
    q_weight_1d[i] = w;
    q_weight_1d[Q - 1 - i] = w;
    q_ref_1d[i] = -x;
    q_ref_1d[Q - 1 - i] = x;
  }
}

// src/basis.cpp:81-144 (recurrence to degree Q-1, Newton on the derivative).
void gauss_lobatto_points(int Q, std::vector<double>& q_ref_1d, std::vector<double>* q_weight_1d) {
  if (Q <= 1) throw std::invalid_argument("Cannot create Lobatto quadrature with Q <= 1");
  q_ref_1d.resize(Q);
  if (q_weight_1d) q_weight_1d->resize(Q);
  const double pi = 4.0 * atan(1.0);
  const double tol = 10 * std::numeric_limits<double>::epsilon();
  q_ref_1d[0] = -1.0;
  q_ref_1d[Q - 1] = 1.0;
  if (q_weight_1d) {
    const double w = 2.0 / (Q * (Q - 1));
    (*q_weight_1d)[0] = w;
    (*q_weight_1d)[Q - 1] = w;
  }
  for (int i = 1; i <= (Q - 1) / 2; ++i) {
    double x = std::cos(pi * i / (Q - 1));
    double pm1, pq;
    legendre_upto(Q - 1, x, pm1, pq);
    double d1 = (x * pq - pm1) * Q / (x * x - 1.0);
    double d2 = (2 * x * d1 - (Q * (Q - 1)) * pq) / (1.0 - x * x);
    x -= d1 / d2;
    for (int k = 0; k < 100 && std::fabs(d1) > tol; ++k) {
      legendre_upto(Q - 1, x, pm1, pq);
      d1 = (x * pq - pm1) * Q / (x * x - 1.0);
      d2 = (2 * x * d1 - (Q * (Q - 1)) * pq) / (1.0 - x * x);
      x -= d1 / d2;
    }
    q_ref_1d[i] = -x;
    q_ref_1d[Q - 1 - i] = x;
    if (q_weight_1d) {
      const double w = 2.0 / ((Q * (Q - 1)) * pq * pq);
      (*q_weight_1d)[i] = w;
      (*q_weight_1d)[Q - 1 - i] = w;
    }
  }
}

// src/basis.cpp:156-186, Fornberg (1998) for the value and first derivative.
void lagrange_basis_matrix(const std::vector<double>& nodes, const std::vector<double>& eval_points,
                           std::vector<double>& interp_1d, std::vector<double>& grad_1d) {
  const size_t P = nodes.size();
  const size_t M = eval_points.size();
  interp_1d.resize(M * P);
  grad_1d.resize(M * P);
  for (size_t m = 0; m < M; ++m) {
    double* L = &interp_1d[m * P];
    double* D = &grad_1d[m * P];
    const double xe = eval_points[m];
    double prod_prev = 1.0;
    double dist_cur = nodes[0] - xe;
    L[0] = 1.0;
    for (size_t j = 1; j < P; ++j) {
      double prod = 1.0;
      const double dist_prev = dist_cur;
      dist_cur = nodes[j] - xe;
      for (size_t k = 0; k < j; ++k) {
        const double h = nodes[j] - nodes[k];
        prod *= h;
        if (k == j - 1) {
          D[j] = prod_prev * (L[k] - dist_prev * D[k]) / prod;
          L[j] = -prod_prev * dist_prev * L[k] / prod;
        }
        D[k] = (dist_cur * D[k] - L[k]) / h;
        L[k] = dist_cur * L[k] / h;
      }
      prod_prev = prod;
    }
  }
}

// src/basis.cpp:202-228
TensorBasis TensorBasis::create_tensor_H1_lagrange(int dim, int num_comp, int P_1d, int Q_1d, QuadMode quad_mode) {
  if (dim < 1 || dim > 3) throw std::invalid_argument("create_tensor_H1_lagrange: dim must be 1, 2, or 3");
  if (num_comp < 1 || num_comp > 3) throw std::invalid_argument("create_tensor_H1_lagrange: num_comp must be 1, 2, or 3");
  if (P_1d < 2) throw std::invalid_argument("create_tensor_H1_lagrange: P_1d must be >= 2");
  if (Q_1d < 1) throw std::invalid_argument("create_tensor_H1_lagrange: Q_1d must be >= 1");
  std::vector<double> nodes;
  gauss_lobatto_points(P_1d, nodes);
  TensorBasis b{dim, num_comp, P_1d, Q_1d, {}, {}, {}, {}};
  if (quad_mode == QuadMode::Gauss) {
    gauss_legendre_quadrature(Q_1d, b.q_ref_1d, b.q_weight_1d);
  } else if (quad_mode == QuadMode::GaussLobatto) {
    gauss_lobatto_points(Q_1d, b.q_ref_1d, &b.q_weight_1d);
  } else {
    throw std::invalid_argument("create_tensor_H1_lagrange: unknown quad_mode");
  }
  lagrange_basis_matrix(nodes, b.q_ref_1d, b.interp_1d, b.grad_1d);
  return b;
}

}  // namespace fem
