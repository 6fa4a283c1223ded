#pragma once

#include <cstdint>
#include <vector>

#include "common.hpp"

namespace femhip {

struct BlockPlan {
  int NB = 0, es = 0, nblk = 0, max_nodes = 0, max_colors = 0;
  std::vector<int> elem;          // [nblk][NB] element of each slot, -1 for padding
  std::vector<int> nelem;         // [nblk] live slots
  std::vector<int> node_ptr;      // [nblk + 1] into nodes
  std::vector<int> nodes;         // per block: interior nodes then boundary nodes, each ascending
  std::vector<int> ninterior;     // [nblk]
  std::vector<uint16_t> loc;      // [nblk][NB][es] index of each element node in its block's list
  std::vector<uint8_t> color;     // [nblk][NB] colour of each slot: slots of one colour share no node
  std::vector<int> ncolor;        // [nblk] colours used by the block
  std::vector<int> part_ptr;      // [nblk + 1] prefix of boundary-node counts (partial-sum slots)
};

// Greedy clustering of elements into blocks of at most NB elements and at most
// max_nodes unique nodes (0: no node limit).
BlockPlan build_block_plan(const int* offsets, int num_elem, int elem_size, int l_size, int NB, int max_nodes = 0);

// The second pass of the partial-sum scatter: one row per block-boundary node
// (ascending node), listing its partial-sum slots in ascending block order.
struct PartialRows {
  std::vector<int> node;  // [nrows] comp-0 L index
  std::vector<int> ptr;   // [nrows + 1]
  std::vector<int> part;  // [npart] partial slot, ascending within a row
  bool covers_all = false;  // interior nodes and rows, over all components, write every L entry
};

PartialRows build_partial_rows(const BlockPlan& plan, int l_size, int num_comp, int comp_stride);

}  // namespace femhip
