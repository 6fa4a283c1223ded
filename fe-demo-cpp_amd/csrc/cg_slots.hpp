// CG scalar slots and the reduction-grid limit, shared by the solver kernels
// and the batched element kernel (which reads a run's done flag).
#pragma once

#include <hip/hip_runtime.h>

namespace femhip {

constexpr int kMaxRed = 2048;  // most reduction blocks (partials per dot product)

// Scalar slots. The state a CG iteration starts from (r.z, r.r, the iteration
// count, the done flag) has two copies: an iteration of parity k reads copy k
// and its last kernel writes copy 1-k, so no kernel writes a slot that its own
// blocks read. Graphs hold even numbers of iterations, so between launches the
// state is in copy 0 (kRZ, kRR, kITERS, kDONE), where the host reads it.
// An iteration that starts done leaves x, r, z and p alone; the iteration that
// meets the residual test still forms the next search direction (a residual
// replacement continues from it).
// kMAXIT (the iteration limit) and kTHRESH are written by init / replace, so
// a captured graph follows the limits of the run it is replayed in.
// kAP (pipelined CG: the previous step's alpha) has two copies like kRZ.
enum {
  kRZ = 0, kPAP, kALPHA, kBETA, kRR, kDONE, kITERS, kTHRESH, kRZ1, kRR1, kDONE1, kITERS1, kMAXIT, kAP, kAP1, kNSCAL
};
__host__ __device__ constexpr int slot(int s, int par) {
  return par == 0 ? s
                  : (s == kRZ ? kRZ1 : s == kRR ? kRR1 : s == kDONE ? kDONE1 : s == kAP ? kAP1 : kITERS1);
}

}  // namespace femhip
