// Device element restriction: offset (gather / atomic scatter-add /
// deterministic transpose-map scatter-add) and strided, on the reference layouts
//   E-vector  E[(k*elem_size + i)*num_elem + e]
//   L-vector  L[offsets[i + e*elem_size] + k*comp_stride]   (offset)
//             L[i*s0 + k*s1 + e*s2]                         (strided)
// (src/elem-restriction.cpp:34-83, src/cuda/cuda-elem-restriction.cu:10-14).
#pragma once

#include <array>
#include <mutex>
#include <vector>

#include "common.hpp"

namespace femhip {

struct TransposeMapHost {
  std::vector<int> l_vec_indices;  // used L-nodes (component 0), ascending
  std::vector<int> t_offsets;      // CSR row starts, rows + 1 entries
  std::vector<int> t_indices;      // t = i + e*elem_size, ascending within each row
};

// CSR inverse of the offsets (src/cuda/cuda-elem-restriction.cu:149-177): rows are
// the used L-nodes in ascending order, each row the t's that map to it, ascending.
void build_transpose_map(int num_elem, int elem_size, int l_size, const int* offsets, TransposeMapHost& map);

struct Restriction {
  int type = 0;  // 0 offset, 1 strided
  int num_elem = 0, elem_size = 0, num_comp = 0, comp_stride = 0, l_size = 0;
  size_t e_size = 0;
  std::array<int, 3> strides{};
  bool backend_strides = false;
  int scatter_mode = kScatterAtomic;

  DevBuf<int> offsets;  // [num_elem][elem_size]
  // Host copy of the offsets (block plans, host checks): made at create from
  // host offsets, or downloaded on first use for restrictions made on the device.
  const std::vector<int>& host_offsets() const;

  int max_offset = 0;  // largest offset (offset restrictions)
  // Made by restriction_create_box: the box's elements per direction (the
  // offsets are the box formula by construction); 0 otherwise.
  int box_n[3] = {0, 0, 0};

  // Deterministic offset restrictions only (built on the device: the same map
  // as build_transpose_map).
  int num_nodes = 0;
  bool identity_nodes = false;  // every L-node 0..l_size-1 used: l_vec_indices not read
  DevBuf<int> l_vec_indices, t_offsets, t_indices;

  bool same_layout(const Restriction& o) const;

  mutable std::vector<int> h_offsets;
  mutable bool h_offsets_valid = false;
  mutable std::mutex h_mu;
};

// The deterministic map is built on stream s (nullptr: the null stream), which
// is synchronized before the call returns.
Restriction* restriction_create_offset(int num_elem, int elem_size, int num_comp, int comp_stride, int l_size,
                                       const int* offsets, int scatter_mode, hipStream_t s = nullptr);
Restriction* restriction_create_strided(int num_elem, int elem_size, int num_comp, int l_size, const int* strides);
// The offset restriction of an nx*ny*nz box of P^3-node hexes with
// lexicographic node numbering and blocked components (comp_stride = node
// count, l_size = num_comp * node count), its offsets made on the device (valid
// by construction: no host validation or upload); the deterministic map is
// built on s, which is synchronized before the call returns unless `sync` is
// false (then the device arrays are complete in s's order only: a caller
// that uses them on another stream orders that stream after s).
Restriction* restriction_create_box(int nx, int ny, int nz, int P, int num_comp, int scatter_mode, hipStream_t s,
                                    bool sync = true);
// Deep copy (device arrays duplicated), for operators that store restrictions by value.
Restriction* restriction_clone(const Restriction& r);

template <class T>
void restriction_apply(const Restriction& r, int t_mode, const T* in, size_t in_len, T* out, size_t out_len,
                       hipStream_t s);

// What one apply reads: the sizes and the device index arrays, owned by
// whoever made the view (a Restriction, or a caller holding its own device
// arrays as include/cuda/cuda-elem-restriction.hpp:46-58 does).
struct RestrictionView {
  int type = 0;  // 0 offset, 1 strided
  int num_elem = 0, elem_size = 0, num_comp = 0, comp_stride = 0, l_size = 0;
  size_t e_size = 0;
  std::array<int, 3> strides{};
  int scatter_mode = kScatterAtomic;
  const int* offsets = nullptr;        // offset: [num_elem][elem_size]
  int num_nodes = 0;                   // deterministic offset: rows of the transpose map
  const int* l_vec_indices = nullptr;  // null: row m is L-node m
  const int* t_offsets = nullptr;
  const int* t_indices = nullptr;
};
RestrictionView restriction_view(const Restriction& r);
template <class T>
void restriction_apply_view(const RestrictionView& r, int t_mode, const T* in, size_t in_len, T* out, size_t out_len,
                            hipStream_t s);

// Deterministic transpose-map sum of an element-output array Y laid out
// [comp][num_elem*elem_size] (index t = i + e*elem_size): for every used L-node,
// acc = (zero_start ? 0 : L[l]); acc += Y[t] over its row ascending; L[l] = acc.
template <class T>
void csr_scatter(const Restriction& r, const T* Y, T* L, bool zero_start, hipStream_t s);

}  // namespace femhip
