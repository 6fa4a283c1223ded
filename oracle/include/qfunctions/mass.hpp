#pragma once
// ORACLE (test infrastructure only). Restates include/qfunctions/mass.hpp:8-51:
// build_mass (qdata = det(J) * w, cofactor expansion along row 0) and apply_mass
// (v = qdata * u, size-1 fields).

namespace fem::qfunctions {

struct BuildMassContext {
  int dim;
};

// in[0] = dx, size dim*dim, entry (d, c) of point p at in[0][(d*dim + c)*Q + p];
// in[1] = weights; out[0] = qdata.
inline int build_mass(void* ctx, int Q, const double* const* in, double* const* out) {
  const int dim = static_cast<const BuildMassContext*>(ctx)->dim;
  const double* jac = in[0];
  const double* wts = in[1];
  double* qd = out[0];
  if (dim < 1 || dim > 3) return 1;
  for (int p = 0; p < Q; ++p) {
    const auto J = [&](int d, int c) { return jac[(d * dim + c) * Q + p]; };
    double det;
    if (dim == 1) {
      det = J(0, 0);
    } else if (dim == 2) {
      det = J(0, 0) * J(1, 1) - J(0, 1) * J(1, 0);
    } else {
      // Same association as mass.hpp:31-33: a*(..) - b*(..) + c*(..), left to right.
      const double m0 = J(1, 1) * J(2, 2) - J(1, 2) * J(2, 1);
      const double m1 = J(1, 0) * J(2, 2) - J(1, 2) * J(2, 0);
      const double m2 = J(1, 0) * J(2, 1) - J(1, 1) * J(2, 0);
      det = J(0, 0) * m0 - J(0, 1) * m1 + J(0, 2) * m2;
    }
    qd[p] = det * wts[p];
  }
  return 0;
}

// in[0] = u, in[1] = qdata, out[0] = v; all size 1.
inline int apply_mass(void*, int Q, const double* const* in, double* const* out) {
  for (int p = 0; p < Q; ++p) out[0][p] = in[1][p] * in[0][p];
  return 0;
}

}  // namespace fem::qfunctions
