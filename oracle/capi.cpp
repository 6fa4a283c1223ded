// ORACLE (test infrastructure only): a C entry-point layer over the fem:: CPU
// interface, so Python tests can drive it through ctypes. The same file is
// compiled twice:
//   * against oracle/include + oracle/src (the restatement)  -> oracle/liboracle.so
//   * against the reference's own include/ + src/*.cpp       -> oracle/_ref/libfemref.so
// and only uses the interface both share (include/basis.hpp, elem-restriction.hpp,
// tensor-contract.hpp, qfunction.hpp, qfunctions/mass.hpp, operator.hpp). The
// vector-mass and diffusion QFunctions, which the reference lacks, are defined
// here, so both builds run the same code for them.
// Nothing in the product links or loads this.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <exception>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "basis.hpp"
#include "elem-restriction.hpp"
#include "operator.hpp"
#include "qfunction.hpp"
#include "qfunctions/mass.hpp"
#include "tensor-contract.hpp"
#ifdef ORACLE_RESTATEMENT
#include "transpose-map.hpp"
#endif

namespace {

thread_local std::string g_err;

template <class F>
int guarded(F&& f) {
  try {
    f();
    return 0;
  } catch (const std::invalid_argument& e) {
    g_err = e.what();
    return 1;
  } catch (const std::exception& e) {
    g_err = e.what();
    return 2;
  }
}

fem::ContractMode cmode(int t) { return t ? fem::ContractMode::Transpose : fem::ContractMode::NoTranspose; }
fem::QuadMode qmode(int m) { return m ? fem::QuadMode::GaussLobatto : fem::QuadMode::Gauss; }

fem::EvalMode emode(int m) {
  switch (m) {
    case 0: return fem::EvalMode::None;
    case 1: return fem::EvalMode::Interp;
    case 2: return fem::EvalMode::Grad;
    default: return fem::EvalMode::Weight;
  }
}

// Vector mass for num_comp > 1 (no reference counterpart; apply_mass is size-1
// only, include/qfunctions/mass.hpp:43-51): v[c][p] = qdata[p] * u[c][p].
struct VecMassCtx {
  int num_comp;
};
int vector_mass(void* ctx, int Q, const double* const* in, double* const* out) {
  const int nc = static_cast<VecMassCtx*>(ctx)->num_comp;
  for (int c = 0; c < nc; ++c)
    for (int p = 0; p < Q; ++p) out[0][c * Q + p] = in[1][p] * in[0][c * Q + p];
  return 0;
}

// Diffusion (Poisson) QFunctions: no reference counterpart. They follow the
// reference's QFunction signature (include/qfunction.hpp:17-20) and layouts:
// gradient fields are [d][comp][Q] (src/tensor-contract.cpp:115,153), J is read
// as in build_mass, J(d,c) = dx[(d*dim + c)*Q + p] = dx_c/dX_d
// (include/qfunctions/mass.hpp:13-14).
//   build_diffusion: qdata = w det(J) J^-1 J^-T in reference coordinates, with
//     C the cofactor matrix of A(d,c) = J(d,c): qdata = (w / det) C C^T, stored
//     symmetric in the order (00), (11), (22), (12), (02), (01) for dim 3,
//     (00), (11), (01) for dim 2, (00) for dim 1. det is expanded along row 0
//     with the association of include/qfunctions/mass.hpp:31-33.
//   apply_diffusion: dv[a][c] = sum_b qdata(a,b) du[b][c], b ascending.
// The device QFunctions perform the same operations in the same order.
struct DiffusionCtx {
  int dim, num_comp;
};
inline int sym_index(int dim, int a, int b) {
  if (a == b) return a;
  if (dim == 2) return 2;
  const int lo = a < b ? a : b, hi = a < b ? b : a;
  return lo == 1 ? 3 : (hi == 2 ? 4 : 5);  // (1,2) -> 3, (0,2) -> 4, (0,1) -> 5
}
int build_diffusion(void* ctx, int Q, const double* const* in, double* const* out) {
  const int dim = static_cast<DiffusionCtx*>(ctx)->dim;
  if (dim < 1 || dim > 3) return 1;
  const int ns = dim * (dim + 1) / 2;
  for (int p = 0; p < Q; ++p) {
    auto A = [&](int d, int c) { return in[0][(d * dim + c) * Q + p]; };
    const double w = in[1][p];
    double C[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    double det;
    if (dim == 1) {
      det = A(0, 0);
      C[0][0] = 1.0;
    } else if (dim == 2) {
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
    const double scale = w / det;
    for (int a = 0; a < dim; ++a) {
      for (int b = a; b < dim; ++b) {
        double s = 0.0;
        for (int k = 0; k < dim; ++k) s = s + C[a][k] * C[b][k];
        out[0][sym_index(dim, a, b) * Q + p] = scale * s;
      }
    }
    (void)ns;
  }
  return 0;
}
int apply_diffusion(void* ctx, int Q, const double* const* in, double* const* out) {
  const int dim = static_cast<DiffusionCtx*>(ctx)->dim, nc = static_cast<DiffusionCtx*>(ctx)->num_comp;
  if (dim < 1 || dim > 3) return 1;
  for (int p = 0; p < Q; ++p) {
    for (int c = 0; c < nc; ++c) {
      for (int a = 0; a < dim; ++a) {
        double s = 0.0;
        for (int b = 0; b < dim; ++b) s = s + in[1][sym_index(dim, a, b) * Q + p] * in[0][(b * nc + c) * Q + p];
        out[0][(a * nc + c) * Q + p] = s;
      }
    }
  }
  return 0;
}

// The mass operator of the reference tests (tests/test-operator.cpp:17-24,75-82):
// u Interp active, qdata None passive (backend strides), v Interp active.
struct MassOp {
  fem::Operator op;
  int l_size = 0;
};

std::vector<int> to_vec(const int* p, size_t n) { return std::vector<int>(p, p + n); }
std::vector<double> to_vec(const double* p, size_t n) { return std::vector<double>(p, p + n); }

int ipow(int b, int e) {
  int r = 1;
  while (e-- > 0) r *= b;
  return r;
}

}  // namespace

extern "C" {

const char* orc_last_error() { return g_err.c_str(); }

int orc_is_reference() {
#ifdef ORACLE_RESTATEMENT
  return 0;
#else
  return 1;
#endif
}

int orc_gauss(int Q, double* x, double* w) {
  return guarded([&] {
    std::vector<double> a, b;
    fem::gauss_legendre_quadrature(Q, a, b);
    std::memcpy(x, a.data(), Q * sizeof(double));
    std::memcpy(w, b.data(), Q * sizeof(double));
  });
}

int orc_lobatto(int Q, double* x, double* w) {
  return guarded([&] {
    std::vector<double> a, b;
    fem::gauss_lobatto_points(Q, a, w ? &b : nullptr);
    std::memcpy(x, a.data(), Q * sizeof(double));
    if (w) std::memcpy(w, b.data(), Q * sizeof(double));
  });
}

int orc_lagrange(int P, const double* nodes, int M, const double* pts, double* interp, double* grad) {
  return guarded([&] {
    std::vector<double> I, G;
    fem::lagrange_basis_matrix(to_vec(nodes, P), to_vec(pts, M), I, G);
    std::memcpy(interp, I.data(), I.size() * sizeof(double));
    std::memcpy(grad, G.data(), G.size() * sizeof(double));
  });
}

int orc_basis(int dim, int nc, int P, int Q, int quad_mode, double* qref, double* qw, double* interp, double* grad) {
  return guarded([&] {
    auto b = fem::TensorBasis::create_tensor_H1_lagrange(dim, nc, P, Q, qmode(quad_mode));
    std::memcpy(qref, b.q_ref_1d.data(), Q * sizeof(double));
    std::memcpy(qw, b.q_weight_1d.data(), Q * sizeof(double));
    std::memcpy(interp, b.interp_1d.data(), Q * P * sizeof(double));
    std::memcpy(grad, b.grad_1d.data(), Q * P * sizeof(double));
  });
}

// v has A*J*C entries on entry (used when add != 0) and on return.
int orc_contract(int A, int B, int C, int J, const double* t, int t_size, int t_mode, int add, const double* u,
                 double* v) {
  return guarded([&] {
    std::vector<double> vv = to_vec(v, static_cast<size_t>(A) * J * C);
    fem::tensor_contract_apply(A, B, C, J, to_vec(t, t_size), cmode(t_mode), add != 0,
                               to_vec(u, static_cast<size_t>(A) * B * C), vv);
    std::memcpy(v, vv.data(), vv.size() * sizeof(double));
  });
}

// Basis apply with a basis built by create_tensor_H1_lagrange. v_len entries are
// copied in first (so the Transpose-grad accumulation into stale v can be seen)
// and v_len entries are copied back.
int orc_basis_apply(int dim, int nc, int P, int Q, int quad_mode, int num_elem, int t_mode, int eval_mode,
                    const double* u, size_t u_len, double* v, size_t v_len, int v_in) {
  return guarded([&] {
    auto b = fem::TensorBasis::create_tensor_H1_lagrange(dim, nc, P, Q, qmode(quad_mode));
    std::vector<double> vv;
    if (v_in) vv = to_vec(v, v_len);
    fem::tensor_basis_apply(b, num_elem, cmode(t_mode), emode(eval_mode), to_vec(u, u_len), vv);
    if (vv.size() != v_len) throw std::runtime_error("orc_basis_apply: output size " + std::to_string(vv.size()));
    std::memcpy(v, vv.data(), v_len * sizeof(double));
  });
}

// Offset restriction when strides == nullptr, else strided (backend != 0 uses backend strides).
// NoTranspose: out (e_size) overwritten. Transpose: out (l_size) accumulated into.
int orc_restriction_apply(int ne, int es, int nc, int cs, int l_size, const int* offsets, const int* strides,
                          int backend, int t_mode, const double* in, double* out) {
  return guarded([&] {
    fem::ElemRestriction r;
    if (backend) {
      fem::elem_restriction_create_strided(ne, es, nc, l_size, r);
    } else if (strides) {
      fem::elem_restriction_create_strided(ne, es, nc, l_size, {strides[0], strides[1], strides[2]}, r);
    } else {
      fem::elem_restriction_create(ne, es, nc, cs, l_size, to_vec(offsets, static_cast<size_t>(ne) * es), r);
    }
    const bool tr = t_mode != 0;
    const size_t in_len = tr ? static_cast<size_t>(r.e_size) : static_cast<size_t>(l_size);
    const size_t out_len = tr ? static_cast<size_t>(l_size) : static_cast<size_t>(r.e_size);
    std::vector<double> o = tr ? to_vec(out, out_len) : std::vector<double>();
    fem::elem_restriction_apply(r, cmode(t_mode), to_vec(in, in_len), o);
    std::memcpy(out, o.data(), out_len * sizeof(double));
  });
}

int orc_restriction_multiplicity(int ne, int es, int nc, int cs, int l_size, const int* offsets, double* mult) {
  return guarded([&] {
    fem::ElemRestriction r;
    fem::elem_restriction_create(ne, es, nc, cs, l_size, to_vec(offsets, static_cast<size_t>(ne) * es), r);
    std::vector<double> m;
    fem::elem_restriction_get_multiplicity(r, m);
    std::memcpy(mult, m.data(), m.size() * sizeof(double));
  });
}

int orc_qf_build_mass(int dim, int Q, const double* J, const double* w, double* qdata) {
  return guarded([&] {
    fem::QFunction qf;
    fem::qfunction_create(fem::qfunctions::build_mass, qf);
    fem::qfunction_add_input(qf, "dx", dim * dim, fem::EvalMode::Grad);
    fem::qfunction_add_input(qf, "weights", 1, fem::EvalMode::Weight);
    fem::qfunction_add_output(qf, "qdata", 1, fem::EvalMode::None);
    fem::qfunction_set_context(qf, fem::qfunctions::BuildMassContext{dim});
    std::vector<std::vector<double>> out;
    fem::qfunction_apply(qf, Q, {to_vec(J, static_cast<size_t>(dim) * dim * Q), to_vec(w, Q)}, out);
    std::memcpy(qdata, out[0].data(), Q * sizeof(double));
  });
}

int orc_qf_apply_mass(int Q, const double* u, const double* qd, double* v) {
  return guarded([&] {
    fem::QFunction qf;
    fem::qfunction_create(fem::qfunctions::apply_mass, qf);
    fem::qfunction_add_input(qf, "u", 1, fem::EvalMode::Interp);
    fem::qfunction_add_input(qf, "qdata", 1, fem::EvalMode::None);
    fem::qfunction_add_output(qf, "v", 1, fem::EvalMode::Interp);
    std::vector<std::vector<double>> out;
    fem::qfunction_apply(qf, Q, {to_vec(u, Q), to_vec(qd, Q)}, out);
    std::memcpy(v, out[0].data(), Q * sizeof(double));
  });
}

// qdata = operator_apply(build operator, coords): dx = Grad of the coordinates
// (mesh basis num_comp = dim, P = mesh_P), weights = Weight, qdata = None through a
// backend-strided restriction (tests/test-operator.cpp:103-112).
int orc_build_qdata(int dim, int ne, int mesh_P, int Q, int quad_mode, int coord_l_size, int coord_comp_stride,
                    const int* mesh_offsets, const double* coords, double* qdata) {
  return guarded([&] {
    const int es = ipow(mesh_P, dim), nq = ipow(Q, dim);
    fem::TensorBasis mb = fem::TensorBasis::create_tensor_H1_lagrange(dim, dim, mesh_P, Q, qmode(quad_mode));
    fem::ElemRestriction cr, qr;
    fem::elem_restriction_create(ne, es, dim, coord_comp_stride, coord_l_size,
                                 to_vec(mesh_offsets, static_cast<size_t>(ne) * es), cr);
    fem::elem_restriction_create_strided(ne, nq, 1, ne * nq, qr);
    fem::QFunction qf;
    fem::qfunction_create(fem::qfunctions::build_mass, qf);
    fem::qfunction_add_input(qf, "dx", dim * dim, fem::EvalMode::Grad);
    fem::qfunction_add_input(qf, "weights", 1, fem::EvalMode::Weight);
    fem::qfunction_add_output(qf, "qdata", 1, fem::EvalMode::None);
    fem::qfunction_set_context(qf, fem::qfunctions::BuildMassContext{dim});
    fem::Operator op;
    fem::operator_create(qf, op);
    fem::operator_set_field(op, "dx", cr, mb, fem::vector_active());
    fem::operator_set_field(op, "weights", std::nullopt, mb, fem::vector_none());
    fem::operator_set_field(op, "qdata", qr, std::nullopt, fem::vector_active());
    std::vector<double> out;
    fem::operator_apply(op, to_vec(coords, coord_l_size), out);
    std::memcpy(qdata, out.data(), out.size() * sizeof(double));
  });
}

// Mass operator handle. nc == 1 uses the reference's apply_mass; nc > 1 the
// vector mass above (blocked or interleaved layout given by comp_stride).
void* orc_mass_create(int dim, int ne, int P, int Q, int nc, int quad_mode, int l_size, int comp_stride,
                      const int* offsets, const double* qdata) {
  MassOp* m = nullptr;
  const int rc = guarded([&] {
    auto h = std::make_unique<MassOp>();
    const int es = ipow(P, dim), nq = ipow(Q, dim);
    fem::TensorBasis b = fem::TensorBasis::create_tensor_H1_lagrange(dim, nc, P, Q, qmode(quad_mode));
    fem::ElemRestriction r, qr;
    fem::elem_restriction_create(ne, es, nc, comp_stride, l_size, to_vec(offsets, static_cast<size_t>(ne) * es), r);
    fem::elem_restriction_create_strided(ne, nq, 1, ne * nq, qr);
    fem::QFunction qf;
    if (nc == 1) {
      fem::qfunction_create(fem::qfunctions::apply_mass, qf);
    } else {
      fem::qfunction_create(vector_mass, qf);
      fem::qfunction_set_context(qf, VecMassCtx{nc});
    }
    fem::qfunction_add_input(qf, "u", nc, fem::EvalMode::Interp);
    fem::qfunction_add_input(qf, "qdata", 1, fem::EvalMode::None);
    fem::qfunction_add_output(qf, "v", nc, fem::EvalMode::Interp);
    fem::operator_create(qf, h->op);
    fem::operator_set_field(h->op, "u", r, b, fem::vector_active());
    fem::operator_set_field(h->op, "qdata", qr, std::nullopt,
                            fem::vector_passive(to_vec(qdata, static_cast<size_t>(ne) * nq)));
    fem::operator_set_field(h->op, "v", r, b, fem::vector_active());
    h->l_size = l_size;
    m = h.release();
  });
  return rc == 0 ? m : nullptr;
}

int orc_mass_apply(void* h, const double* u, double* v) {
  return guarded([&] {
    auto* m = static_cast<MassOp*>(h);
    std::vector<double> out;
    fem::operator_apply(m->op, to_vec(u, m->l_size), out);
    std::memcpy(v, out.data(), out.size() * sizeof(double));
  });
}

void orc_mass_destroy(void* h) { delete static_cast<MassOp*>(h); }

// Jacobi diagonal of the mass operator, assembled as
// diag = E^T [ (B o B)^T (x) (B o B)^T (x) (B o B)^T qdata ] with B o B the
// entry-wise square of interp_1d, through the same interface calls.
int orc_mass_diagonal(void* h, double* diag) {
  return guarded([&] {
    auto* m = static_cast<MassOp*>(h);
    const fem::OperatorField& uf = *m->op.inputs[0];
    const fem::OperatorField& qf = *m->op.inputs[1];
    fem::TensorBasis b2 = *uf.basis;
    for (double& x : b2.interp_1d) x = x * x;
    b2.num_comp = 1;
    std::vector<double> eq, e1;
    fem::elem_restriction_apply(*qf.restriction, fem::ContractMode::NoTranspose, qf.vector.data, eq);
    fem::tensor_basis_apply_interp(b2, m->op.num_elem, fem::ContractMode::Transpose, eq, e1);
    const fem::ElemRestriction& r = *uf.restriction;
    const size_t per_comp = e1.size();
    std::vector<double> e(per_comp * r.num_comp);
    for (int c = 0; c < r.num_comp; ++c) std::memcpy(&e[c * per_comp], e1.data(), per_comp * sizeof(double));
    std::vector<double> d(m->l_size, 0.0);
    fem::elem_restriction_apply(r, fem::ContractMode::Transpose, e, d);
    std::memcpy(diag, d.data(), d.size() * sizeof(double));
  });
}

// FP64 Jacobi-preconditioned CG on M x = b from x = 0 (diag == nullptr: no
// preconditioner). Stops when ||b - M x|| / ||b|| <= rtol (recurrence residual).
int orc_mass_cg(void* h, const double* b, const double* diag, double* x, int max_iter, double rtol, int* iters,
                double* rel_res) {
  return guarded([&] {
    auto* m = static_cast<MassOp*>(h);
    const size_t n = m->l_size;
    std::vector<double> r(b, b + n), z(n), p(n), Ap;
    std::vector<double> xx(n, 0.0);
    auto dot = [&](const std::vector<double>& a, const std::vector<double>& c) {
      double s = 0.0;
      for (size_t i = 0; i < n; ++i) s += a[i] * c[i];
      return s;
    };
    auto precond = [&] {
      for (size_t i = 0; i < n; ++i) z[i] = diag ? r[i] / diag[i] : r[i];
    };
    const double bnorm = std::sqrt(dot(r, r));
    precond();
    p = z;
    double rz = dot(r, z);
    int it = 0;
    double rel = bnorm > 0 ? 1.0 : 0.0;
    while (it < max_iter && rel > rtol) {
      fem::operator_apply(m->op, p, Ap);
      const double alpha = rz / dot(p, Ap);
      for (size_t i = 0; i < n; ++i) {
        xx[i] += alpha * p[i];
        r[i] -= alpha * Ap[i];
      }
      ++it;
      rel = std::sqrt(dot(r, r)) / bnorm;
      if (rel <= rtol) break;
      precond();
      const double rz_new = dot(r, z);
      const double beta = rz_new / rz;
      rz = rz_new;
      for (size_t i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
    }
    std::memcpy(x, xx.data(), n * sizeof(double));
    *iters = it;
    *rel_res = rel;
  });
}

#ifdef ORACLE_RESTATEMENT
// Fills rows (number of used L-nodes); the arrays must hold l_size, l_size + 1 and ne*es ints.
int orc_transpose_map(int ne, int es, int l_size, const int* offsets, int* rows, int* l_vec_indices, int* t_offsets,
                      int* t_indices) {
  return guarded([&] {
    fem::ElemRestriction r;
    fem::elem_restriction_create(ne, es, 1, 1, l_size, to_vec(offsets, static_cast<size_t>(ne) * es), r);
    fem::TransposeMap map;
    fem::oracle_build_transpose_map(r, map);
    *rows = static_cast<int>(map.l_vec_indices.size());
    std::memcpy(l_vec_indices, map.l_vec_indices.data(), map.l_vec_indices.size() * sizeof(int));
    std::memcpy(t_offsets, map.t_offsets.data(), map.t_offsets.size() * sizeof(int));
    std::memcpy(t_indices, map.t_indices.data(), map.t_indices.size() * sizeof(int));
  });
}
#endif

// qdata of the diffusion operator on a mesh given by coordinates (as orc_build_qdata):
// ne * Q^dim points, dim*(dim+1)/2 values each, field-major [s][q*ne + e].
int orc_build_diffusion_qdata(int dim, int ne, int mesh_P, int Q, int quad_mode, int coord_l_size,
                              int coord_comp_stride, const int* mesh_offsets, const double* coords, double* qdata) {
  return guarded([&] {
    const int es = ipow(mesh_P, dim), nq = ipow(Q, dim), ns = dim * (dim + 1) / 2;
    fem::TensorBasis mb = fem::TensorBasis::create_tensor_H1_lagrange(dim, dim, mesh_P, Q, qmode(quad_mode));
    fem::ElemRestriction cr, qr;
    fem::elem_restriction_create(ne, es, dim, coord_comp_stride, coord_l_size,
                                 to_vec(mesh_offsets, static_cast<size_t>(ne) * es), cr);
    fem::elem_restriction_create_strided(ne, nq, ns, ns * ne * nq, qr);
    fem::QFunction qf;
    fem::qfunction_create(build_diffusion, qf);
    fem::qfunction_add_input(qf, "dx", dim * dim, fem::EvalMode::Grad);
    fem::qfunction_add_input(qf, "weights", 1, fem::EvalMode::Weight);
    fem::qfunction_add_output(qf, "qdata", ns, fem::EvalMode::None);
    fem::qfunction_set_context(qf, DiffusionCtx{dim, 1});
    fem::Operator op;
    fem::operator_create(qf, op);
    fem::operator_set_field(op, "dx", cr, mb, fem::vector_active());
    fem::operator_set_field(op, "weights", std::nullopt, mb, fem::vector_none());
    fem::operator_set_field(op, "qdata", qr, std::nullopt, fem::vector_active());
    std::vector<double> out;
    fem::operator_apply(op, to_vec(coords, coord_l_size), out);
    std::memcpy(qdata, out.data(), out.size() * sizeof(double));
  });
}

// Diffusion apply v = K u, composed from the interface calls: gather, gradient,
// apply_diffusion, transposed gradient into a zeroed E-vector, scatter-add into a
// zeroed v. (operator_apply would add the transposed gradient to the gathered
// E-values it reuses as scratch: src/operator.cpp:178 with
// src/tensor-contract.cpp:115-122,142.)
int orc_diffusion_apply(int dim, int ne, int P, int Q, int nc, int quad_mode, int l_size, int comp_stride,
                        const int* offsets, const double* qdata, const double* u, double* v) {
  return guarded([&] {
    const int es = ipow(P, dim), nq = ipow(Q, dim), ns = dim * (dim + 1) / 2;
    fem::TensorBasis b = fem::TensorBasis::create_tensor_H1_lagrange(dim, nc, P, Q, qmode(quad_mode));
    fem::ElemRestriction r;
    fem::elem_restriction_create(ne, es, nc, comp_stride, l_size, to_vec(offsets, static_cast<size_t>(ne) * es), r);
    std::vector<double> e, du, dv, ev;
    fem::elem_restriction_apply(r, fem::ContractMode::NoTranspose, to_vec(u, l_size), e);
    fem::tensor_basis_apply(b, ne, fem::ContractMode::NoTranspose, fem::EvalMode::Grad, e, du);
    const int Qt = ne * nq;
    dv.assign(static_cast<size_t>(dim) * nc * Qt, 0.0);
    DiffusionCtx ctx{dim, nc};
    const double* ins[2] = {du.data(), qdata};
    double* outs[1] = {dv.data()};
    if (apply_diffusion(&ctx, Qt, ins, outs) != 0) throw std::runtime_error("apply_diffusion failed");
    (void)ns;
    ev.assign(static_cast<size_t>(nc) * es * ne, 0.0);
    fem::tensor_basis_apply(b, ne, fem::ContractMode::Transpose, fem::EvalMode::Grad, dv, ev);
    std::vector<double> out(l_size, 0.0);
    fem::elem_restriction_apply(r, fem::ContractMode::Transpose, ev, out);
    std::memcpy(v, out.data(), out.size() * sizeof(double));
  });
}

}  // extern "C"
