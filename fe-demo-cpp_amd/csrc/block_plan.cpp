// Element blocks for the fused mass kernel's block-accumulated scatter.
//
// Elements are grouped into blocks of NB by a greedy clustering on shared nodes:
// a block starts at the lowest-numbered free element and repeatedly takes the
// free element sharing the most nodes with the block (ties: lowest index), so a
// block is a compact patch of the mesh whatever the element numbering. For each
// block the unique L-nodes it touches are listed, block-interior nodes (all of
// whose element references lie in the block) first, then boundary nodes, each
// group ascending; every element node gets its 16-bit index into that list.
// The kernel sums a block's contributions per node in LDS and then writes each
// interior node with a plain store and each boundary node with one atomic add.
#include "block_plan.hpp"

#include <algorithm>

namespace femhip {

BlockPlan build_block_plan(const int* off, int E, int es, int l_size, int NB) {
  require(E > 0 && es > 0 && NB > 0, "block plan: bad sizes");
  require(static_cast<long long>(NB) * es < 65536, "block plan: NB * elem_size must fit 16-bit local indices");
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
  int seed = 0;
  std::vector<int> members, nodes;
  plan.node_ptr.push_back(0);
  while (true) {
    while (seed < E && used[seed]) ++seed;
    if (seed >= E) break;
    members.clear();
    auto add = [&](int e) {
      used[e] = 1;
      members.push_back(e);
      const int* oe = off + static_cast<size_t>(e) * es;
      for (int i = 0; i < es; ++i) {
        const int node = oe[i];
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
          bs = score[f];
          best = f;
        }
      }
      if (best < 0) {
        while (seed < E && used[seed]) ++seed;
        if (seed >= E) break;
        best = seed;
      }
      add(best);
    }
    for (int f : cand) score[f] = 0;
    cand.clear();
    // Unique nodes of the block, with in-block reference counts.
    const int blk = static_cast<int>(plan.nelem.size());
    nodes.clear();
    for (int e : members) {
      const int* oe = off + static_cast<size_t>(e) * es;
      for (int i = 0; i < es; ++i) {
        const int node = oe[i];
        if (stamp[node] != blk) {
          stamp[node] = blk;
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
    plan.ninterior.push_back(n_int);
    plan.nelem.push_back(static_cast<int>(members.size()));
    plan.max_nodes = std::max(plan.max_nodes, static_cast<int>(nodes.size()));
    for (int j = 0; j < NB; ++j) {
      const int e = j < static_cast<int>(members.size()) ? members[j] : -1;
      plan.elem.push_back(e);
      for (int i = 0; i < es; ++i)
        plan.loc.push_back(e >= 0 ? static_cast<uint16_t>(local[off[static_cast<size_t>(e) * es + i]]) : 0);
    }
  }
  plan.nblk = static_cast<int>(plan.nelem.size());
  require(plan.nodes.size() < (1ull << 31), "block plan: node list exceeds int32");
  return plan;
}

}  // namespace femhip
