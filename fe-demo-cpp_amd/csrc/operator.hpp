// Device operator: the reference's field model (include/operator.hpp:15-64) over
// device restrictions, bases and QFunctions, with a fused path for mass operators.
#pragma once

#include <memory>
#include <string>
#include <vector>

#include "basis.hpp"
#include "fused_mass.hpp"
#include "qfunction.hpp"
#include "restriction.hpp"

namespace femhip {

enum : int { kVectorActive = 0, kVectorPassive = 1, kVectorNone = 2 };

struct OpField {
  bool set = false;
  std::string name;
  int eval_mode = 0;
  int size = 0;
  std::shared_ptr<const Restriction> rstr;  // immutable after create: sharing equals the reference's by-value copy
  std::shared_ptr<const Basis> basis;
  int vkind = kVectorNone;
  DevBuf<double> passive;
};

struct Operator {
  QFunctionDev qf;
  std::vector<OpField> inputs, outputs;
  int num_elem = 0, num_qpts = 0, active_in_l_size = 0, active_out_l_size = 0;

  // Fused path (decided at the first apply after the fields are complete).
  bool fused_checked = false;
  int fused_kind = 0;  // 0 none, 1 mass, 2 build_mass on trilinear hexes
  std::unique_ptr<FusedMass> fused;
  // Fused options: scatter (-1: the restriction's mode), exact arithmetic, and the
  // algorithm (-1 automatic, -2 the element kernels of the scatter mode, else a
  // FusedAlg) with its elements per workgroup/block (0 = default).
  int opt_scatter = -1, opt_exact = 0, opt_alg = -1, opt_elems = 0, opt_box_lz = 0;
};

Operator* operator_create(const QFunctionDev& qf);
// A passive device array is copied at once (complete on return), or with
// copy_stream set, in that stream's order (the caller keeps the source alive
// until the copy has run and orders later users of the operator after it).
void operator_set_field(Operator& op, const std::string& name, std::shared_ptr<const Restriction> r,
                        std::shared_ptr<const Basis> b, int vkind, const double* passive, size_t passive_len,
                        bool passive_on_device, hipStream_t copy_stream = nullptr);
void operator_check_ready(const Operator& op);
void operator_apply(Operator& op, const double* in, size_t in_len, double* out, size_t out_len, hipStream_t s);
void operator_apply_f32(Operator& op, const float* in, size_t in_len, float* out, size_t out_len, hipStream_t s);
// Fused state (creating it on first use); throws if the operator is not a fused mass operator.
FusedMass& operator_fused_mass(Operator& op, hipStream_t s);

}  // namespace femhip
