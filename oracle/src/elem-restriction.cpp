// ORACLE (test infrastructure only). CPU restatement of src/elem-restriction.cpp:34-227.
// Gather overwrites; scatter-add accumulates into `out` without zeroing it, in
// the e -> k -> i loop order, so each L entry receives its contributions in
// ascending t = i + e*elem_size (the order the deterministic GPU scatter must match).
#include "elem-restriction.hpp"

#include <stdexcept>

namespace fem {

void elem_restriction_apply(const ElemRestriction& r, ContractMode t_mode, const std::vector<double>& in,
                            std::vector<double>& out) {
  const int ne = r.num_elem, es = r.elem_size, nc = r.num_comp;
  const bool gather = (t_mode == ContractMode::NoTranspose);
  if (gather) {
    out.resize(r.e_size);
  } else {
    out.resize(r.l_size);
  }
  for (int e = 0; e < ne; ++e) {
    for (int k = 0; k < nc; ++k) {
      for (int i = 0; i < es; ++i) {
        const size_t ei = static_cast<size_t>(k * es + i) * ne + e;
        size_t li;
        if (r.type == RestrictionType::Strided) {
          li = static_cast<size_t>(i) * r.strides[0] + static_cast<size_t>(k) * r.strides[1] +
               static_cast<size_t>(e) * r.strides[2];
        } else {
          li = static_cast<size_t>(r.offsets[i + e * es]) + static_cast<size_t>(k) * r.comp_stride;
        }
        if (gather) {
          out[ei] = in[li];
        } else {
          out[li] += in[ei];
        }
      }
    }
  }
}

// src/elem-restriction.cpp:102-138
void elem_restriction_create(const int num_elem, const int elem_size, const int num_comp, const int comp_stride,
                             const int l_size, const std::vector<int>& offsets, ElemRestriction& r) {
  if (num_elem <= 0) throw std::invalid_argument("elem_restriction_create: num_elem must be positive");
  if (elem_size <= 0) throw std::invalid_argument("elem_restriction_create: elem_size must be at least 1");
  if (num_comp <= 0) throw std::invalid_argument("elem_restriction_create: num_comp must be at least 1");
  if (num_comp > 1 && comp_stride <= 0)
    throw std::invalid_argument("elem_restriction_create: comp_stride must be at least 1 when num_comp > 1");
  if (l_size <= 0) throw std::invalid_argument("elem_restriction_create: l_size must be positive");
  if (offsets.size() != static_cast<size_t>(num_elem * elem_size))
    throw std::invalid_argument("elem_restriction_create: offsets size must be num_elem * elem_size");
  for (const int o : offsets) {
    if (o < 0 || o + (num_comp - 1) * comp_stride >= l_size)
      throw std::invalid_argument("elem_restriction_create: offsets (including component offsets) out of range");
  }
  r.type = RestrictionType::Offset;
  r.num_elem = num_elem;
  r.elem_size = elem_size;
  r.num_comp = num_comp;
  r.comp_stride = comp_stride;
  r.l_size = l_size;
  r.e_size = num_comp * elem_size * num_elem;
  r.offsets = offsets;
  r.strides = {};
  r.has_backend_strides = false;
}

// src/elem-restriction.cpp:155-181
void elem_restriction_create_strided(const int num_elem, const int elem_size, const int num_comp, const int l_size,
                                     const std::array<int, 3>& strides, ElemRestriction& r) {
  if (num_elem <= 0) throw std::invalid_argument("elem_restriction_create_strided: num_elem must be positive");
  if (elem_size <= 0) throw std::invalid_argument("elem_restriction_create_strided: elem_size must be at least 1");
  if (num_comp <= 0) throw std::invalid_argument("elem_restriction_create_strided: num_comp must be at least 1");
  if (l_size <= 0) throw std::invalid_argument("elem_restriction_create_strided: l_size must be positive");
  for (const int s : strides)
    if (s < 0) throw std::invalid_argument("elem_restriction_create_strided: strides must be non-negative");
  const long long top = static_cast<long long>(elem_size - 1) * strides[0] +
                        static_cast<long long>(num_comp - 1) * strides[1] +
                        static_cast<long long>(num_elem - 1) * strides[2];
  if (top >= l_size)
    throw std::invalid_argument("elem_restriction_create_strided: strided indices must be in [0, l_size - 1]");
  r.type = RestrictionType::Strided;
  r.num_elem = num_elem;
  r.elem_size = elem_size;
  r.num_comp = num_comp;
  r.comp_stride = 0;
  r.l_size = l_size;
  r.e_size = num_comp * elem_size * num_elem;
  r.offsets.clear();
  r.strides = strides;
  r.has_backend_strides = false;
}

// src/elem-restriction.cpp:200-204: strides {num_elem, elem_size*num_elem, 1}.
void elem_restriction_create_strided(const int num_elem, const int elem_size, const int num_comp, const int l_size,
                                     ElemRestriction& r) {
  elem_restriction_create_strided(num_elem, elem_size, num_comp, l_size, {num_elem, elem_size * num_elem, 1}, r);
  r.has_backend_strides = true;
}

// src/elem-restriction.cpp:215-227: E^T (E 1).
void elem_restriction_get_multiplicity(const ElemRestriction& r, std::vector<double>& multiplicity) {
  std::vector<double> e_vec(r.e_size);
  multiplicity.assign(r.l_size, 1.0);
  elem_restriction_apply(r, ContractMode::NoTranspose, multiplicity, e_vec);
  multiplicity.assign(r.l_size, 0.0);
  elem_restriction_apply(r, ContractMode::Transpose, e_vec, multiplicity);
}

}  // namespace fem
