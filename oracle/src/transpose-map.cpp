// ORACLE (test infrastructure only). CPU restatement of build_transpose_map,
// src/cuda/cuda-elem-restriction.cu:149-177: the CSR inverse of the offsets over
// used L-nodes (ascending L index), each row listing t = i + e*elem_size in
// ascending order. That source is CUDA and cannot be compiled here, so this
// restatement is pinned by the reference's known answers
// (tests/cuda/test-cuda-elem-restriction.cu:87-128) and a brute-force check.
#include "transpose-map.hpp"

#include <stdexcept>

namespace fem {

void oracle_build_transpose_map(const ElemRestriction& r, TransposeMap& map) {
  if (r.type != RestrictionType::Offset) throw std::invalid_argument("build_transpose_map: only for Offset restrictions");
  const int n_t = r.num_elem * r.elem_size;
  std::vector<int> row_of(r.l_size, -1);
  for (int t = 0; t < n_t; ++t) row_of[r.offsets[t]] = 0;
  map.l_vec_indices.clear();
  for (int l = 0; l < r.l_size; ++l) {
    if (row_of[l] == 0) {
      row_of[l] = static_cast<int>(map.l_vec_indices.size());
      map.l_vec_indices.push_back(l);
    }
  }
  const int rows = static_cast<int>(map.l_vec_indices.size());
  map.t_offsets.assign(rows + 1, 0);
  for (int t = 0; t < n_t; ++t) ++map.t_offsets[row_of[r.offsets[t]] + 1];
  for (int n = 0; n < rows; ++n) map.t_offsets[n + 1] += map.t_offsets[n];
  map.t_indices.assign(n_t, 0);
  std::vector<int> fill(map.t_offsets.begin(), map.t_offsets.end() - 1);
  for (int t = 0; t < n_t; ++t) map.t_indices[fill[row_of[r.offsets[t]]]++] = t;
}

}  // namespace fem
