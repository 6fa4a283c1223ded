#pragma once

#include <cstdint>
#include <vector>

#include "common.hpp"

namespace femhip {

struct BlockPlan {
  int NB = 0, es = 0, nblk = 0, max_nodes = 0;
  std::vector<int> elem;          // [nblk][NB] element of each slot, -1 for padding
  std::vector<int> nelem;         // [nblk] live slots
  std::vector<int> node_ptr;      // [nblk + 1] into nodes
  std::vector<int> nodes;         // per block: interior nodes then boundary nodes, each ascending
  std::vector<int> ninterior;     // [nblk]
  std::vector<uint16_t> loc;      // [nblk][NB][es] index of each element node in its block's list
};

BlockPlan build_block_plan(const int* offsets, int num_elem, int elem_size, int l_size, int NB);

}  // namespace femhip
