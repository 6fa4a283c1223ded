// Fused per-element mass operator: gather -> interp (x,y,z) -> qdata * u ->
// interp^T (x,y,z) -> scatter, one launch per apply (plus the transpose-map
// pass for the deterministic scatter).
#pragma once

#include <cstdint>
#include <map>
#include <memory>
#include <tuple>
#include <vector>

#include "basis.hpp"
#include "common.hpp"
#include "restriction.hpp"

namespace femhip {

// What a fused launch computes, chosen per apply.
struct FusedKey {
  int P, Q, NC;
  int fp32;       // 0 FP64, 1 FP32
  int det;        // 0: atomic scatter-add into v; 1: element outputs to Y, then the transpose-map pass
  int exact;      // 1: no FMA, reference association (FP64 only)
  int elems;      // elements per workgroup
  bool operator<(const FusedKey& o) const {
    return std::tie(P, Q, NC, fp32, det, exact, elems) < std::tie(o.P, o.Q, o.NC, o.fp32, o.det, o.exact, o.elems);
  }
};

// Launch parameters shared by every instantiation.
struct FusedArgs {
  int num_elem;
  int comp_stride;
  const int* offsets;          // [E][P^3]
  const int* flagged_offsets;  // the same with bit 31 set on single-contributor nodes
  const void* qdata;   // private layout [E][Q^2][Q] (qz innermost), T
  const void* u;       // L-vector, T
  void* out;           // atomic: L-vector v (pre-zeroed); det: Y [NC][E*P^3]
  size_t y_comp_len;   // E*P^3
  const void* basis;   // host pointer to Q*P values of T (interp_1d), copied into the kernel arguments
};

using FusedLauncher = void (*)(const FusedArgs&, hipStream_t);

// Launch parameters of the block-accumulated kernel (atomic scatter mode).
struct FusedBlockArgs {
  int nblk;
  int comp_stride;
  const int* node_ptr;   // [nblk + 1]
  const int* nodes;      // block node lists (comp-0 L indices)
  const int* ninterior;  // [nblk]
  const int* nelem;      // [nblk]
  const uint16_t* loc;   // [nblk*NB][P^3]
  const void* qdata;     // [nblk*NB][Q^3], T
  const void* u;
  void* v;               // pre-zeroed L-vector
  const void* basis;     // host [Q][P] of T
};
using FusedBlockLauncher = void (*)(const FusedBlockArgs&, hipStream_t);
std::map<FusedKey, FusedBlockLauncher>& fused_block_registry();
void register_fused_block(const FusedKey& k, FusedBlockLauncher f);

// Registry of compiled instantiations (filled by the fused_mass_*.hip units).
std::map<FusedKey, FusedLauncher>& fused_registry();
void register_fused(const FusedKey& k, FusedLauncher f);
// Default workgroup packing for a shape (elements per workgroup).
int fused_default_elems(int P, int Q, int NC, int fp32);
// Every compiled elements-per-workgroup variant for a shape (for the tuner).
std::vector<int> fused_elem_variants(int P, int Q, int NC, int fp32, int det, int exact);

// The device state of a fused mass operator.
struct FusedMass {
  int P = 0, Q = 0, NC = 0, num_elem = 0, comp_stride = 0, l_size = 0;
  const Restriction* rstr = nullptr;  // owned by the operator (node restriction)
  DevBuf<int> flagged_offsets;        // offsets with bit 31 set where the node has one contributor
  DevBuf<double> qdata64;             // private layout
  DevBuf<float> qdata32;              // lazily made from qdata64
  std::vector<double> B64;            // interp_1d [Q][P]
  std::vector<float> B32;
  std::vector<double> Bsq;            // interp_1d squared entry-wise (Jacobi diagonal)
  // Block plan of the block-accumulated atomic kernel (built on first use).
  int block_nb = 0, nblk = 0;
  bool block_tried = false;
  DevBuf<int> b_node_ptr, b_nodes, b_ninterior, b_nelem;
  DevBuf<uint16_t> b_loc;
  DevBuf<double> b_q64;  // qdata in block-slot order [slot][Q^3]
  DevBuf<float> b_q32;
  DevBuf<double> Y64;                 // deterministic element outputs (lazy)
  DevBuf<float> Y32;
  int scatter = kScatterAtomic;
  int exact = 0;
  int elems = 0;  // 0 = default / tuned

  // Optional per-apply stage timing with HIP events on the apply's stream:
  // a..b = pre-pass (output zeroing), b..c = the fused kernel, c..d = post-pass
  // (transpose-map sum). Events are pooled and read back by timing_end.
  struct StageEvents {
    hipEvent_t a, b, c, d;
  };
  bool timing = false;
  std::vector<StageEvents> events;
  size_t events_used = 0;
  ~FusedMass();
};

void fused_timing_begin(FusedMass& f);
// Sums of the stage times (ms) over the applies since timing_begin.
void fused_timing_end(FusedMass& f, double* kernel_ms, double* aux_ms, int* count);

// qdata_ref is the backend-strided layout qdata[q*E + e] (device).
std::unique_ptr<FusedMass> fused_mass_create(const Restriction& r, const HostBasis& b, const double* qdata_ref,
                                             hipStream_t s);
bool fused_mass_supported(int P, int Q, int NC);
void fused_mass_apply64(FusedMass& f, const double* u, double* v, hipStream_t s);
// Builds lazily made state (block plan) so later applies allocate nothing (graph capture).
void fused_prepare(FusedMass& f, hipStream_t s);
void fused_mass_apply32(FusedMass& f, const float* u, float* v, hipStream_t s);
// Jacobi diagonal: diag = E^T [ (B o B)^T x3 qdata ] (FP64).
void fused_mass_diagonal(FusedMass& f, double* diag, hipStream_t s);

}  // namespace femhip
