// Host-side 1D basis setup: Gauss-Legendre and Gauss-Lobatto rules by Newton
// iteration and Lagrange value/derivative matrices by Fornberg's recurrence.
// These produce the constant matrices the device kernels contract with, so they
// must carry exactly the reference's bits: every floating-point operation is in
// the order of src/basis.cpp:16-186 (start values cos(pi(2i+1)/(2Q)) and
// cos(pi i/(Q-1)), the |f| > 10 eps stopping rule with at most 100 steps, the
// symmetric slot writes that overwrite the middle slot for odd Q, and
// Fornberg's update order). The library is built with -ffp-contract=off for the
// host pass, so no fused multiply-add changes a rounding.
#include "basis_setup.hpp"

#include <cmath>
#include <limits>
#include <stdexcept>

namespace femhip {

namespace {

struct Legendre {
  double below;  // P_{n-1}(x)
  double at;     // P_n(x); 0.0 when n < 2
};

Legendre legendre(int n, double x) {
  double p0 = 1.0, p1 = x, p2 = 0.0;
  for (int j = 2; j <= n; ++j) {
    p2 = ((2.0 * j - 1.0) * x * p1 - (j - 1.0) * p0) / j;
    p0 = p1;
    p1 = p2;
  }
  return {p0, p2};
}

const double kPi = 4.0 * atan(1.0);
const double kTol = 10 * std::numeric_limits<double>::epsilon();

}  // namespace

void gauss_legendre(int Q, double* x_out, double* w_out) {
  require(Q > 0, "gauss_legendre_quadrature: Q must be positive");
  for (int i = 0; i <= Q / 2; ++i) {
    double x = std::cos(kPi * (2.0 * i + 1.0) / (2.0 * Q));
    Legendre L = legendre(Q, x);
    double dp = (x * L.at - L.below) * Q / (x * x - 1.0);
    x = x - L.at / dp;
    for (int k = 0; k < 100 && std::fabs(L.at) > kTol; ++k) {
      L = legendre(Q, x);
      dp = (x * L.at - L.below) * Q / (x * x - 1.0);
      x -= L.at / dp;
    }
    const double w = 2.0 / ((1.0 - x * x) * dp * dp);
    w_out[i] = w;
    w_out[Q - 1 - i] = w;
    x_out[i] = -x;
    x_out[Q - 1 - i] = x;
  }
}

void gauss_lobatto(int Q, double* x_out, double* w_out) {
  require(Q > 1, "Cannot create Lobatto quadrature with Q <= 1");
  x_out[0] = -1.0;
  x_out[Q - 1] = 1.0;
  if (w_out) {
    w_out[0] = w_out[Q - 1] = 2.0 / (Q * (Q - 1));
  }
  for (int i = 1; i <= (Q - 1) / 2; ++i) {
    double x = std::cos(kPi * i / (Q - 1));
    Legendre L = legendre(Q - 1, x);
    double d1 = (x * L.at - L.below) * Q / (x * x - 1.0);
    double d2 = (2 * x * d1 - (Q * (Q - 1)) * L.at) / (1.0 - x * x);
    x -= d1 / d2;
    for (int k = 0; k < 100 && std::fabs(d1) > kTol; ++k) {
      L = legendre(Q - 1, x);
      d1 = (x * L.at - L.below) * Q / (x * x - 1.0);
      d2 = (2 * x * d1 - (Q * (Q - 1)) * L.at) / (1.0 - x * x);
      x -= d1 / d2;
    }
    x_out[i] = -x;
    x_out[Q - 1 - i] = x;
    if (w_out) {
      const double w = 2.0 / ((Q * (Q - 1)) * L.at * L.at);
      w_out[i] = w;
      w_out[Q - 1 - i] = w;
    }
  }
}

void lagrange_matrices(int P, const double* nodes, int M, const double* pts, double* interp, double* grad) {
  for (int m = 0; m < M; ++m) {
    double* L = interp + static_cast<size_t>(m) * P;
    double* D = grad + static_cast<size_t>(m) * P;
    for (int j = 0; j < P; ++j) D[j] = 0.0;  // a fresh matrix, as the reference's resize gives
    double c1 = 1.0;
    double c3 = nodes[0] - pts[m];
    L[0] = 1.0;
    for (int j = 1; j < P; ++j) {
      double c2 = 1.0;
      const double c4 = c3;
      c3 = nodes[j] - pts[m];
      for (int k = 0; k < j; ++k) {
        const double dx = nodes[j] - nodes[k];
        c2 *= dx;
        if (k == j - 1) {
          D[j] = c1 * (L[k] - c4 * D[k]) / c2;
          L[j] = -c1 * c4 * L[k] / c2;
        }
        D[k] = (c3 * D[k] - L[k]) / dx;
        L[k] = c3 * L[k] / dx;
      }
      c1 = c2;
    }
  }
}

HostBasis make_h1_lagrange(int dim, int num_comp, int P, int Q, int quad_mode) {
  require(dim >= 1 && dim <= 3, "create_tensor_H1_lagrange: dim must be 1, 2, or 3");
  require(num_comp >= 1 && num_comp <= 3, "create_tensor_H1_lagrange: num_comp must be 1, 2, or 3");
  require(P >= 2, "create_tensor_H1_lagrange: P_1d must be >= 2");
  require(Q >= 1, "create_tensor_H1_lagrange: Q_1d must be >= 1");
  require(quad_mode == 0 || quad_mode == 1, "create_tensor_H1_lagrange: unknown quad_mode");
  HostBasis b;
  b.dim = dim;
  b.num_comp = num_comp;
  b.P = P;
  b.Q = Q;
  std::vector<double> nodes(P);
  gauss_lobatto(P, nodes.data(), nullptr);
  b.q_ref.assign(Q, 0.0);
  b.q_weight.assign(Q, 0.0);
  if (quad_mode == 0) {
    gauss_legendre(Q, b.q_ref.data(), b.q_weight.data());
  } else {
    gauss_lobatto(Q, b.q_ref.data(), b.q_weight.data());
  }
  b.interp.assign(static_cast<size_t>(Q) * P, 0.0);
  b.grad.assign(static_cast<size_t>(Q) * P, 0.0);
  lagrange_matrices(P, nodes.data(), Q, b.q_ref.data(), b.interp.data(), b.grad.data());
  return b;
}

}  // namespace femhip
