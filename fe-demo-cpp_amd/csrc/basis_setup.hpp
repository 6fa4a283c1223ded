// Host-side 1D basis setup (quadrature rules, Lagrange matrices) and the host
// copy of a tensor basis.
#pragma once

#include <vector>

#include "common.hpp"

namespace femhip {

struct HostBasis {
  int dim = 0, num_comp = 0, P = 0, Q = 0;
  std::vector<double> q_ref, q_weight;  // [Q]
  std::vector<double> interp, grad;     // [Q][P]
};

void gauss_legendre(int Q, double* x, double* w);
void gauss_lobatto(int Q, double* x, double* w /* nullable */);
void lagrange_matrices(int P, const double* nodes, int M, const double* pts, double* interp, double* grad);
HostBasis make_h1_lagrange(int dim, int num_comp, int P, int Q, int quad_mode);

}  // namespace femhip
