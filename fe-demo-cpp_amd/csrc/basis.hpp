// Device tensor-product basis on the reference E/Q-vector layouts.
#pragma once

#include "basis_setup.hpp"
#include "common.hpp"

namespace femhip {

struct Basis {
  HostBasis h;
  DevBuf<double> interp, grad, q_weight;  // device copies of the 1D data
};

Basis* basis_create(const HostBasis& hb);

// v[(a*J + j)*C + c] (+)= sum_b t(j, b) u[(a*B + b)*C + c] on device arrays
// (src/tensor-contract.cpp:35-59); t(j,b) = t[j*B + b], or t[b*J + j] for Transpose.
void contract(int A, int B, int C, int J, const double* t, int t_mode, bool add, const double* u, double* v,
              hipStream_t s);

// tensor_basis_apply (src/tensor-contract.cpp:74-222). Sizes are checked.
void basis_apply(const Basis& b, int num_elem, int t_mode, int eval_mode, const double* u, size_t u_len, double* v,
                 size_t v_len, hipStream_t s);

size_t basis_in_len(const HostBasis& b, int num_elem, int t_mode, int eval_mode);
size_t basis_out_len(const HostBasis& b, int num_elem, int t_mode, int eval_mode);

}  // namespace femhip
