// The batched element kernel (see batch_solve.hpp): one persistent launch over
// the element work of every run in a cohort's table. Included by the
// fused_inst_p*.hip units, which register the compiled shapes.
#pragma once

#include <hip/hip_runtime.h>

#include "batch_solve.hpp"
#include "cg_slots.hpp"
#include "fused_mass_kernel.hpp"

namespace femhip {

// Work item i is the group of NE elements i - elem0 of the run whose range
// holds it. The grid (a multiple of 8 workgroups, every one resident) splits
// the items into 8 contiguous ranges, one per XCD (workgroup w runs on XCD
// w mod 8), so neighbouring elements, which share nodes of u, meet in one L2.
// A run whose done flag (copy `par`) is set is skipped: its step only copies
// its state on.
template <int P, int Q, int NC, class T, int NE>
__global__ __launch_bounds__(round64(NE* Shape<P, Q>::TPE)) void batch_elem_kernel(const BatchTable* __restrict__ tab,
                                                                                   int par, MatArg<T, P * Q> Bm) {
  __shared__ T s_b[NE][Shape<P, Q>::FBUF];
  const int items = tab->elem_items, ns = tab->nslots;
  const int per = static_cast<int>(gridDim.x) / 8, xcd = static_cast<int>(blockIdx.x) % 8;
  const int chunk = (items + 7) / 8;
  const int hi = min(items, (xcd + 1) * chunk);
  for (int it = xcd * chunk + static_cast<int>(blockIdx.x) / 8; it < hi; it += per) {
    int k = 0;
    while (k + 1 < ns && tab->slot[k + 1].elem0 <= it) ++k;
    const BatchSlot& sl = tab->slot[k];
    if (sl.sc[slot(kDONE, par)] != 0.0) continue;  // the same for the whole workgroup
    fast_elem_group<P, Q, NC, T, NE, true>(it - sl.elem0, sl.num_elem, sl.comp_stride, sl.offsets,
                                           static_cast<const T*>(sl.qdata), static_cast<const T*>(sl.m),
                                           static_cast<T*>(sl.y), sl.y_comp_len, Bm, s_b, nullptr);
    __syncthreads();  // the next item overwrites the element tiles
  }
}

template <int P, int Q, int NC, class T, int NE>
void launch_batch_elem(const BatchTable* tab, int par, const void* basis, int grid, hipStream_t s) {
  require(grid > 0 && grid % 8 == 0, "batch: the element grid must be a positive multiple of 8");
  batch_elem_kernel<P, Q, NC, T, NE>
      <<<grid, round64(NE * Shape<P, Q>::TPE), 0, s>>>(tab, par, basis_arg<P, Q, NC, T>(basis));
  HIP_CHECK_LAUNCH();
}

template <int P, int Q, int NC, class T, int NE>
int batch_elem_blocks_per_cu() {
  int n = 0;
  HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, batch_elem_kernel<P, Q, NC, T, NE>,
                                                         round64(NE * Shape<P, Q>::TPE), 0));
  return n;
}

// FP64 and FP32 batched element kernels of one shape (registered at load).
template <int P, int Q, int NC, int NE64, int NE32>
struct BatchShapeReg {
  BatchShapeReg() {
    register_batch_elem(BatchKey{P, Q, NC, 0, NE64}, BatchElemKernel{&launch_batch_elem<P, Q, NC, double, NE64>,
                                                                     &batch_elem_blocks_per_cu<P, Q, NC, double, NE64>});
    register_batch_elem(BatchKey{P, Q, NC, 1, NE32}, BatchElemKernel{&launch_batch_elem<P, Q, NC, float, NE32>,
                                                                     &batch_elem_blocks_per_cu<P, Q, NC, float, NE32>});
  }
};

}  // namespace femhip
