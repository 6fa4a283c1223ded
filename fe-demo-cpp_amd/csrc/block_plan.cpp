// Element blocks for the fused mass kernel's block-accumulated scatters.
//
// Elements are grouped into blocks of at most NB by a greedy clustering on
// shared nodes: a block starts at the lowest-numbered free element and
// repeatedly takes the free element sharing the most nodes with the block (ties:
// lowest index), so a block is a compact patch of the mesh whatever the element
// numbering. A block also stops growing when the next element would take its
// unique-node count past max_nodes (the kernel keeps one value per node in LDS).
//
// For each block the unique L-nodes it touches are listed, block-interior nodes
// (all of whose element references lie in the block) first, then boundary
// nodes, each group ascending; every element node gets its 16-bit index into
// that list. Slots are coloured greedily (lowest colour not used by an earlier
// slot sharing a node), so the slots of one colour can add their element
// outputs into the block's node sums without conflicts, colour after colour, in
// a fixed order.
//
// The kernels sum a block's contributions per node in LDS. Interior nodes are
// then complete and stored; boundary nodes get one atomic add each (atomic
// scatter) or one partial-sum slot each (part_ptr), which a second pass adds up
// per node in ascending block order (deterministic scatter).
#include "block_plan.hpp"

#include <algorithm>

namespace femhip {

BlockPlan build_block_plan(const int* off, int E, int es, int l_size, int NB, int max_nodes) {
  require(E > 0 && es > 0 && NB > 0, "block plan: bad sizes");
  require(static_cast<long long>(NB) * es < 65536, "block plan: NB * elem_size must fit 16-bit local indices");
  require(max_nodes == 0 || max_nodes >= es, "block plan: max_nodes below one element");
  require(NB <= 255, "block plan: NB must fit 8-bit colours");
  const size_t nt = static_cast<size_t>(E) * es;
  // node -> element incidence (CSR)
  std::vector<int> mult(static_cast<size_t>(l_size) + 1, 0);
  for (size_t t = 0; t < nt; ++t) ++mult[off[t] + 1];
  std::vector<int> nptr(mult);
  for (int l = 0; l < l_size; ++l) nptr[l + 1] += nptr[l];
  std::vector<int> nel(nt);
  {
    std::vector<int> cur(nptr.begin(), nptr.end() - 1);
    for (size_t t = 0; t < nt; ++t) nel[cur[off[t]]++] = static_cast<int>(t / es);
  }
  BlockPlan plan;
  plan.NB = NB;
  plan.es = es;
  std::vector<char> used(E, 0);
  std::vector<int> score(E, 0), cand;
  std::vector<int> stamp(l_size, -1), in_blk(l_size, 0), local(l_size, 0);
  // node -> (block id, slot-colour mask) while colouring the current block
  std::vector<int> cstamp(l_size, -1);
  std::vector<uint32_t> cmask(l_size, 0);
  int seed = 0, blk_id = 0, nnodes = 0;
  std::vector<int> members, nodes;
  plan.node_ptr.push_back(0);
  plan.part_ptr.push_back(0);
  // New nodes element e would bring into the current block.
  auto new_nodes = [&](int e) {
    int k = 0;
    const int* oe = off + static_cast<size_t>(e) * es;
    for (int i = 0; i < es; ++i) k += (stamp[oe[i]] != blk_id) ? 1 : 0;
    return k;
  };
  while (true) {
    while (seed < E && used[seed]) ++seed;
    if (seed >= E) break;
    members.clear();
    nnodes = 0;
    auto add = [&](int e) {
      used[e] = 1;
      members.push_back(e);
      const int* oe = off + static_cast<size_t>(e) * es;
      for (int i = 0; i < es; ++i) {
        const int node = oe[i];
        if (stamp[node] != blk_id) {
          stamp[node] = blk_id;
          ++nnodes;
        }
        for (int k = nptr[node]; k < nptr[node + 1]; ++k) {
          const int f = nel[k];
          if (used[f]) continue;
          if (score[f] == 0) cand.push_back(f);
          ++score[f];
        }
      }
    };
    add(seed);
    while (static_cast<int>(members.size()) < NB) {
      int best = -1, bs = 0;
      for (int f : cand) {
        if (used[f]) continue;
        if (score[f] > bs || (score[f] == bs && f < best)) {
          if (max_nodes && nnodes + new_nodes(f) > max_nodes) continue;
          bs = score[f];
          best = f;
        }
      }
      if (best < 0) {
        // No connected candidate fits: start a disconnected piece only when nothing
        // connected was ever available (the element set is disconnected).
        bool any_connected = false;
        for (int f : cand) any_connected = any_connected || !used[f];
        if (any_connected) break;
        while (seed < E && used[seed]) ++seed;
        if (seed >= E) break;
        if (max_nodes && nnodes + new_nodes(seed) > max_nodes) break;
        best = seed;
      }
      add(best);
    }
    for (int f : cand) score[f] = 0;
    cand.clear();
    // Unique nodes of the block, with in-block reference counts.
    // add() stamped the block's nodes with blk_id; collect each once.
    nodes.clear();
    for (int e : members) {
      const int* oe = off + static_cast<size_t>(e) * es;
      for (int i = 0; i < es; ++i) {
        const int node = oe[i];
        if (stamp[node] == blk_id) {
          stamp[node] = -2 - blk_id;  // collected
          in_blk[node] = 0;
          nodes.push_back(node);
        }
        ++in_blk[node];
      }
    }
    std::sort(nodes.begin(), nodes.end());
    // interior first (stable: ascending within each group)
    std::stable_partition(nodes.begin(), nodes.end(), [&](int node) { return in_blk[node] == mult[node + 1]; });
    int n_int = 0;
    for (int node : nodes) n_int += (in_blk[node] == mult[node + 1]) ? 1 : 0;
    require(nodes.size() < 65536, "block plan: too many nodes in a block");
    for (size_t j = 0; j < nodes.size(); ++j) local[nodes[j]] = static_cast<int>(j);
    plan.nodes.insert(plan.nodes.end(), nodes.begin(), nodes.end());
    plan.node_ptr.push_back(static_cast<int>(plan.nodes.size()));
    plan.part_ptr.push_back(plan.part_ptr.back() + static_cast<int>(nodes.size()) - n_int);
    plan.ninterior.push_back(n_int);
    plan.nelem.push_back(static_cast<int>(members.size()));
    plan.max_nodes = std::max(plan.max_nodes, static_cast<int>(nodes.size()));
    // Greedy slot colouring: lowest colour unused by earlier slots sharing a node.
    int ncol = 0;
    for (int j = 0; j < NB; ++j) {
      const int e = j < static_cast<int>(members.size()) ? members[j] : -1;
      int col = 0;
      if (e >= 0) {
        const int* oe = off + static_cast<size_t>(e) * es;
        uint32_t taken = 0;
        for (int i = 0; i < es; ++i)
          if (cstamp[oe[i]] == blk_id) taken |= cmask[oe[i]];
        while (col < 32 && (taken >> col & 1u)) ++col;
        require(col < 32, "block plan: more than 32 colours in a block");
        for (int i = 0; i < es; ++i) {
          if (cstamp[oe[i]] != blk_id) {
            cstamp[oe[i]] = blk_id;
            cmask[oe[i]] = 0;
          }
          cmask[oe[i]] |= 1u << col;
        }
        ncol = std::max(ncol, col + 1);
      }
      plan.color.push_back(static_cast<uint8_t>(col));
      plan.elem.push_back(e);
      for (int i = 0; i < es; ++i)
        plan.loc.push_back(e >= 0 ? static_cast<uint16_t>(local[off[static_cast<size_t>(e) * es + i]]) : 0);
    }
    plan.ncolor.push_back(ncol);
    plan.max_colors = std::max(plan.max_colors, ncol);
    ++blk_id;
  }
  plan.nblk = static_cast<int>(plan.nelem.size());
  require(plan.nodes.size() < (1ull << 31), "block plan: node list exceeds int32");
  return plan;
}

PartialRows build_partial_rows(const BlockPlan& plan, int l_size, int num_comp, int comp_stride) {
  PartialRows r;
  // (node, slot) pairs in block order, then a stable counting sort by node.
  std::vector<int> count(static_cast<size_t>(l_size) + 1, 0);
  std::vector<char> interior(l_size, 0);
  for (int b = 0; b < plan.nblk; ++b) {
    const int n0 = plan.node_ptr[b], n1 = plan.node_ptr[b + 1], ni = plan.ninterior[b];
    for (int j = n0; j < n0 + ni; ++j) interior[plan.nodes[j]] = 1;
    for (int j = n0 + ni; j < n1; ++j) ++count[plan.nodes[j] + 1];
  }
  for (int l = 0; l < l_size; ++l) {
    if (count[l + 1] > 0) {
      r.node.push_back(l);
      r.ptr.push_back(static_cast<int>(r.part.size()));
      r.part.resize(r.part.size() + count[l + 1]);
    }
  }
  r.ptr.push_back(static_cast<int>(r.part.size()));
  std::vector<int> row_of(l_size, -1), fill;
  for (size_t m = 0; m < r.node.size(); ++m) row_of[r.node[m]] = static_cast<int>(m);
  fill.assign(r.ptr.begin(), r.ptr.end() - 1);
  for (int b = 0; b < plan.nblk; ++b) {
    const int n0 = plan.node_ptr[b], n1 = plan.node_ptr[b + 1], ni = plan.ninterior[b];
    for (int j = n0 + ni; j < n1; ++j) r.part[fill[row_of[plan.nodes[j]]]++] = plan.part_ptr[b] + (j - n0 - ni);
  }
  // Does every L entry get written (so v needs no zeroing)?
  std::vector<char> hit(l_size, 0);
  bool ok = true;
  for (int l = 0; l < l_size && ok; ++l) {
    if (!interior[l] && row_of[l] < 0) continue;
    for (int c = 0; c < num_comp; ++c) {
      const long long g = static_cast<long long>(l) + static_cast<long long>(c) * comp_stride;
      if (g >= l_size) {
        ok = false;
        break;
      }
      hit[g] = 1;
    }
  }
  if (ok)
    for (int l = 0; l < l_size && ok; ++l) ok = hit[l] != 0;
  r.covers_all = ok;
  return r;
}

}  // namespace femhip
