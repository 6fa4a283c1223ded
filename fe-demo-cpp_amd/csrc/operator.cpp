// Device operator_apply. The unfused path follows src/operator.cpp:148-182 step
// by step on device arrays (inputs in field order: Weight -> weights, None ->
// gather, Interp/Grad -> gather + basis; QFunction; zero out; outputs: None ->
// scatter, else basis^T + scatter). A mass operator (apply_mass or the vector
// mass with u Interp active, qdata None passive on a backend-strided
// restriction, v Interp active on the same restriction and basis, dim 3, a
// compiled (P, Q, num_comp)) runs the fused per-element kernel instead. A
// build_mass operator on trilinear hex geometry (dx Grad active on 8-node
// elements with 3 components and a P=2 basis, weights Weight on the same basis,
// qdata None on a backend-strided restriction, dim 3) runs the fused geometry
// kernel, which computes the same bits without the gradient Q-vector.
#include "operator.hpp"

#include <cstring>

#include "fused_build.hpp"

namespace femhip {

namespace {

std::invalid_argument field_error(const std::string& name, const std::string& msg) {
  return std::invalid_argument("operator_set_field '" + name + "': " + msg);
}

bool same_basis(const Basis& a, const Basis& b) {
  return a.h.dim == b.h.dim && a.h.num_comp == b.h.num_comp && a.h.P == b.h.P && a.h.Q == b.h.Q &&
         a.h.interp == b.h.interp && a.h.grad == b.h.grad && a.h.q_weight == b.h.q_weight;
}

// build_mass (include/qfunctions/mass.hpp:18-40) with dim 3 on trilinear hexes.
bool fused_build_ok(const Operator& op) {
  if (op.qf.kind != kQfBuildMass || op.inputs.size() != 2 || op.outputs.size() != 1) return false;
  int dim = 0;
  if (op.qf.ctx.size() < sizeof(int)) return false;
  std::memcpy(&dim, op.qf.ctx.data(), sizeof(int));
  const OpField& dx = op.inputs[0];
  const OpField& w = op.inputs[1];
  const OpField& q = op.outputs[0];
  if (dim != 3 || dx.eval_mode != kEvalGrad || dx.vkind != kVectorActive || !dx.rstr || !dx.basis || dx.size != 9)
    return false;
  if (!fused_build_mass_supported(*dx.rstr, *dx.basis)) return false;
  if (w.eval_mode != kEvalWeight || !w.basis || w.size != 1) return false;
  if (!(w.basis == dx.basis || (w.basis->h.dim == 3 && w.basis->h.Q == dx.basis->h.Q &&
                                 w.basis->h.q_weight == dx.basis->h.q_weight)))
    return false;
  if (q.eval_mode != kEvalNone || q.vkind != kVectorActive || !q.rstr || q.size != 1) return false;
  const Restriction& qr = *q.rstr;
  const int Q = dx.basis->h.Q;
  // The fused kernel writes exactly num_elem*Q^3 entries; a strided qdata
  // restriction with a larger l_size (legal) keeps the generic path, which
  // zeroes the whole output as src/operator.cpp:172 does.
  return qr.type == 1 && qr.backend_strides && qr.num_comp == 1 && qr.num_elem == dx.rstr->num_elem &&
         qr.elem_size == Q * Q * Q && static_cast<long long>(qr.l_size) == static_cast<long long>(qr.num_elem) * Q * Q * Q;
}

void decide_fused(Operator& op) {
  op.fused_checked = true;
  op.fused_kind = 0;
  if (fused_build_ok(op)) {
    op.fused_kind = 2;
    return;
  }
  const bool mass_qf = op.qf.kind == kQfApplyMass || op.qf.kind == kQfApplyMassVector;
  if (!mass_qf || op.inputs.size() != 2 || op.outputs.size() != 1) return;
  const OpField& u = op.inputs[0];
  const OpField& q = op.inputs[1];
  const OpField& v = op.outputs[0];
  if (u.eval_mode != kEvalInterp || u.vkind != kVectorActive || !u.rstr || !u.basis) return;
  if (q.eval_mode != kEvalNone || q.vkind != kVectorPassive || !q.rstr) return;
  if (v.eval_mode != kEvalInterp || v.vkind != kVectorActive || !v.rstr || !v.basis) return;
  const Basis& b = *u.basis;
  const int nc = b.h.num_comp;
  if (b.h.dim != 3 || u.rstr->type != 0) return;
  if (!(u.rstr == v.rstr || u.rstr->same_layout(*v.rstr))) return;
  if (!(u.basis == v.basis || same_basis(b, *v.basis))) return;
  if (!q.rstr->backend_strides || q.rstr->num_comp != 1) return;
  if (op.qf.kind == kQfApplyMass && (nc != 1 || u.size != 1 || v.size != 1 || q.size != 1)) return;
  if (op.qf.kind == kQfApplyMassVector) {
    int ctx_nc = 0;
    if (op.qf.ctx.size() < sizeof(int)) return;
    std::memcpy(&ctx_nc, op.qf.ctx.data(), sizeof(int));
    if (ctx_nc != nc || u.size != nc || v.size != nc || q.size != 1) return;
  }
  if (!fused_mass_supported(b.h.P, b.h.Q, nc)) return;
  op.fused_kind = 1;
}

}  // namespace

Operator* operator_create(const QFunctionDev& qf) {
  auto* op = new Operator();
  op->qf = qf;
  op->inputs.resize(qf.inputs.size());
  op->outputs.resize(qf.outputs.size());
  return op;
}

// src/operator.cpp:44-117
void operator_set_field(Operator& op, const std::string& name, std::shared_ptr<const Restriction> r,
                        std::shared_ptr<const Basis> b, int vkind, const double* passive, size_t passive_len,
                        bool passive_on_device, hipStream_t copy_stream) {
  const QFField* qf_field = nullptr;
  OpField* slot = nullptr;
  bool is_input = true;
  for (size_t f = 0; f < op.qf.inputs.size() && !qf_field; ++f)
    if (op.qf.inputs[f].name == name) qf_field = &op.qf.inputs[f], slot = &op.inputs[f];
  for (size_t f = 0; f < op.qf.outputs.size() && !qf_field; ++f)
    if (op.qf.outputs[f].name == name) qf_field = &op.qf.outputs[f], slot = &op.outputs[f], is_input = false;
  if (!qf_field) throw field_error(name, "the QFunction has no field with this name");
  if (slot->set) throw field_error(name, "already set");
  if (vkind != kVectorActive && vkind != kVectorPassive && vkind != kVectorNone)
    throw field_error(name, "unknown vector kind");

  const int mode = qf_field->eval_mode, size = qf_field->size;
  switch (mode) {
    case kEvalWeight:
      if (r) throw field_error(name, "EvalMode::Weight takes no restriction");
      if (!b) throw field_error(name, "EvalMode::Weight needs a basis (for its quadrature weights)");
      if (vkind != kVectorNone) throw field_error(name, "EvalMode::Weight takes vector_none()");
      break;
    case kEvalNone:
      if (!r) throw field_error(name, "EvalMode::None needs a restriction");
      if (b) throw field_error(name, "EvalMode::None takes no basis");
      if (r->num_comp != size) throw field_error(name, "restriction num_comp must equal the field size");
      break;
    default: {
      if (!r) throw field_error(name, "EvalMode::Interp/Grad needs a restriction");
      if (!b) throw field_error(name, "EvalMode::Interp/Grad needs a basis");
      const int want = (mode == kEvalInterp) ? b->h.num_comp : b->h.num_comp * b->h.dim;
      if (size != want)
        throw field_error(name, mode == kEvalInterp ? "field size must equal basis num_comp"
                                                    : "field size must equal basis num_comp * dim");
      if (r->num_comp != b->h.num_comp) throw field_error(name, "restriction and basis num_comp differ");
      if (r->elem_size != ipow(b->h.P, b->h.dim))
        throw field_error(name, "restriction elem_size must equal basis P_1d^dim");
    }
  }
  if (mode != kEvalWeight && vkind == kVectorNone) throw field_error(name, "needs vector_active() or vector_passive(...)");
  if (!is_input && vkind != kVectorActive) throw field_error(name, "output fields must be active");
  if (vkind == kVectorPassive && static_cast<long long>(passive_len) != r->l_size)
    throw field_error(name, "passive vector size must equal restriction l_size");
  if (vkind == kVectorPassive && passive == nullptr) throw field_error(name, "passive vector data missing");

  const int nq = b ? ipow(b->h.Q, b->h.dim) : r->elem_size;
  if (op.num_qpts != 0 && nq != op.num_qpts)
    throw field_error(name, "number of quadrature points differs from previously set fields");
  if (r && op.num_elem != 0 && r->num_elem != op.num_elem)
    throw field_error(name, "num_elem differs from previously set fields");
  if (vkind == kVectorActive) {
    int& l = is_input ? op.active_in_l_size : op.active_out_l_size;
    if (l != 0 && r->l_size != l)
      throw field_error(name, "active fields share one vector, so their restrictions need the same l_size");
  }

  OpField f;
  f.set = true;
  f.name = name;
  f.eval_mode = mode;
  f.size = size;
  f.rstr = std::move(r);
  f.basis = std::move(b);
  f.vkind = vkind;
  if (vkind == kVectorPassive) {
    f.passive.allocate(passive_len);
    if (passive_on_device && copy_stream) {
      HIP_CHECK(hipMemcpyAsync(f.passive.data(), passive, passive_len * sizeof(double), hipMemcpyDeviceToDevice,
                               copy_stream));
    } else {
      SharedCapture lk;
      HIP_CHECK(hipMemcpy(f.passive.data(), passive, passive_len * sizeof(double),
                          passive_on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice));
      // A device-to-device hipMemcpy may return before the copy is done: finish it,
      // so the caller may reuse the source and other streams may read the copy.
      if (passive_on_device) HIP_CHECK(hipStreamSynchronize(nullptr));
    }
  }
  if (vkind == kVectorActive) (is_input ? op.active_in_l_size : op.active_out_l_size) = f.rstr->l_size;
  op.num_qpts = nq;
  if (f.rstr) op.num_elem = f.rstr->num_elem;
  *slot = std::move(f);
  op.fused_checked = false;
  op.fused.reset();
}

// src/operator.cpp:126-136
void operator_check_ready(const Operator& op) {
  for (size_t f = 0; f < op.inputs.size(); ++f)
    require(op.inputs[f].set, "operator_check_ready: field '" + op.qf.inputs[f].name + "' is not set");
  for (size_t f = 0; f < op.outputs.size(); ++f)
    require(op.outputs[f].set, "operator_check_ready: field '" + op.qf.outputs[f].name + "' is not set");
  require(op.num_elem != 0, "operator_check_ready: no field has a restriction");
}

FusedMass& operator_fused_mass(Operator& op, hipStream_t s) {
  operator_check_ready(op);
  if (!op.fused_checked) decide_fused(op);
  if (op.fused_kind != 1) throw Unsupported("operator: not a fused mass operator");
  if (!op.fused) {
    const OpField& u = op.inputs[0];
    op.fused = fused_mass_create(*u.rstr, u.basis->h, op.inputs[1].passive.data(), s);
  }
  FusedMass& f = *op.fused;
  f.scatter = op.opt_scatter >= 0 ? op.opt_scatter : op.inputs[0].rstr->scatter_mode;
  if (f.scatter == kScatterDeterministic && op.inputs[0].rstr->scatter_mode != kScatterDeterministic)
    throw std::invalid_argument("operator: deterministic scatter needs a restriction created with ScatterMode::Deterministic");
  f.exact = op.opt_exact;
  f.choice.alg = op.opt_alg == -2 ? (f.scatter == kScatterDeterministic ? kAlgElemTmap : kAlgElemAtomic) : op.opt_alg;
  f.choice.elems = op.opt_elems;
  f.box_lz = op.opt_box_lz;
  return f;
}

void operator_apply(Operator& op, const double* in, size_t in_len, double* out, size_t out_len, hipStream_t s) {
  operator_check_ready(op);
  if (op.active_in_l_size > 0)
    require(in_len == static_cast<size_t>(op.active_in_l_size),
            "operator_apply: input vector size does not match active_in_l_size");
  require(out_len == static_cast<size_t>(op.active_out_l_size),
          "operator_apply: output vector size does not match active_out_l_size");
  if (!op.fused_checked) decide_fused(op);
  if (op.fused_kind == 1) {
    FusedMass& f = operator_fused_mass(op, s);
    fused_mass_apply64(f, in, out, s);
    return;
  }
  if (op.fused_kind == 2) {
    fused_build_mass(*op.inputs[0].rstr, *op.inputs[0].basis, in, out, s);
    return;
  }

  const size_t Q = static_cast<size_t>(op.num_qpts) * op.num_elem;
  require(Q < (1ull << 31), "operator_apply: num_qpts * num_elem must fit in int32");
  // E-vector scratch shared by the fields (the reference's `e`).
  size_t e_len = 0;
  for (const auto* fs : {&op.inputs, &op.outputs})
    for (const auto& f : *fs)
      if (f.rstr && f.basis) e_len = std::max(e_len, f.rstr->e_size);
  StreamTmp<double> e(e_len, s);

  std::vector<std::unique_ptr<StreamTmp<double>>> q_in_buf(op.inputs.size()), q_out_buf(op.outputs.size());
  std::vector<const double*> q_in(op.inputs.size());
  std::vector<double*> q_out(op.outputs.size());
  std::vector<size_t> in_lens(op.inputs.size()), out_lens(op.outputs.size());

  for (size_t fi = 0; fi < op.inputs.size(); ++fi) {
    const OpField& f = op.inputs[fi];
    const double* vec = (f.vkind == kVectorActive) ? in : f.passive.data();
    const size_t len = static_cast<size_t>(f.size) * Q;
    in_lens[fi] = len;
    if (f.eval_mode == kEvalNone && f.rstr->type == 1 && f.rstr->backend_strides) {
      q_in[fi] = vec;  // backend strides: the E-vector is the L-vector itself
      continue;
    }
    q_in_buf[fi] = std::make_unique<StreamTmp<double>>(len, s);
    double* dst = q_in_buf[fi]->data();
    q_in[fi] = dst;
    if (f.eval_mode == kEvalWeight) {
      basis_apply(*f.basis, op.num_elem, kNoTranspose, kEvalWeight, nullptr, 0, dst, len, s);
    } else if (f.eval_mode == kEvalNone) {
      restriction_apply<double>(*f.rstr, kNoTranspose, vec, f.rstr->l_size, dst, len, s);
    } else {
      restriction_apply<double>(*f.rstr, kNoTranspose, vec, f.rstr->l_size, e.data(), f.rstr->e_size, s);
      basis_apply(*f.basis, op.num_elem, kNoTranspose, f.eval_mode, e.data(), f.rstr->e_size, dst, len, s);
    }
  }
  for (size_t fo = 0; fo < op.outputs.size(); ++fo) {
    out_lens[fo] = static_cast<size_t>(op.outputs[fo].size) * Q;
    q_out_buf[fo] = std::make_unique<StreamTmp<double>>(out_lens[fo], s);
    q_out[fo] = q_out_buf[fo]->data();
  }
  qfunction_apply(op.qf, static_cast<int>(Q), static_cast<int>(q_in.size()), q_in.data(), in_lens.data(),
                  static_cast<int>(q_out.size()), q_out.data(), out_lens.data(), s);
  zero_fill(out, out_len * sizeof(double), s);
  for (size_t fo = 0; fo < op.outputs.size(); ++fo) {
    const OpField& f = op.outputs[fo];
    if (f.eval_mode == kEvalNone) {
      restriction_apply<double>(*f.rstr, kTranspose, q_out[fo], out_lens[fo], out, out_len, s);
    } else {
      basis_apply(*f.basis, op.num_elem, kTranspose, f.eval_mode, q_out[fo], out_lens[fo], e.data(), f.rstr->e_size, s);
      restriction_apply<double>(*f.rstr, kTranspose, e.data(), f.rstr->e_size, out, out_len, s);
    }
  }
}

void operator_apply_f32(Operator& op, const float* in, size_t in_len, float* out, size_t out_len, hipStream_t s) {
  operator_check_ready(op);
  require(in_len == static_cast<size_t>(op.active_in_l_size) && out_len == static_cast<size_t>(op.active_out_l_size),
          "operator_apply_f32: vector sizes do not match the operator");
  FusedMass& f = operator_fused_mass(op, s);
  fused_mass_apply32(f, in, out, s);
}

}  // namespace femhip
