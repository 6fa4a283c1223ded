#pragma once
// ORACLE (test infrastructure only). Restates the TransposeMap of
// include/cuda/cuda-elem-restriction.hpp:33-40 (host CSR inverse of the offsets).
#include <vector>
#include "elem-restriction.hpp"

namespace fem {

struct TransposeMap {
  std::vector<int> l_vec_indices;  // used L-nodes, ascending
  std::vector<int> t_offsets;      // row starts, size rows + 1
  std::vector<int> t_indices;      // t = i + e*elem_size, ascending within a row
};

void oracle_build_transpose_map(const ElemRestriction& r, TransposeMap& map);

}  // namespace fem
