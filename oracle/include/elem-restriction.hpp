#pragma once
// ORACLE (test infrastructure only). Restates the interface of include/elem-restriction.hpp:22-59.
#include <array>
#include <vector>
#include "transpose-mode.hpp"

namespace fem {

enum class RestrictionType { Offset, Strided };

struct ElemRestriction {
  RestrictionType type = RestrictionType::Offset;
  int num_elem;
  int elem_size;
  int num_comp;
  int comp_stride;
  int l_size;
  int e_size;
  std::vector<int> offsets;  // [num_elem][elem_size]
  std::array<int, 3> strides{};
  bool has_backend_strides = false;
};

void elem_restriction_apply(const ElemRestriction& restriction, ContractMode t_mode, const std::vector<double>& in,
                            std::vector<double>& out);
void elem_restriction_create(const int num_elem, const int elem_size, const int num_comp, const int comp_stride,
                             const int l_size, const std::vector<int>& offsets, ElemRestriction& restriction);
void elem_restriction_create_strided(const int num_elem, const int elem_size, const int num_comp, const int l_size,
                                     ElemRestriction& restriction);
void elem_restriction_create_strided(const int num_elem, const int elem_size, const int num_comp, const int l_size,
                                     const std::array<int, 3>& strides, ElemRestriction& restriction);
void elem_restriction_get_multiplicity(const ElemRestriction& restriction, std::vector<double>& multiplicity);

}  // namespace fem
