#pragma once
#include <cstddef>
#include <cstdint>

#include "common.hpp"

namespace femhip {

struct SplitMix64 {
  uint64_t state;
  double uniform01();  // (x >> 11) * 2^-53, in [0, 1)
};

void rng_uniform(uint64_t seed, size_t n, double lo, double hi, double* out);
// The same stream into device memory out[0..n), enqueued on s.
void rng_uniform_device(uint64_t seed, size_t n, double lo, double hi, double* out, hipStream_t s);
void box_offsets(int nx, int ny, int nz, int P, int node_stride, int* offsets);
void box_coords(int n, double jitter, uint64_t seed, double* coords);
// box_offsets / box_coords written into device memory by kernels on s, entry
// for entry the host functions' output.
void box_offsets_device(int nx, int ny, int nz, int P, int node_stride, int* offsets, hipStream_t s);
void box_coords_device(int n, double jitter, uint64_t seed, double* coords, hipStream_t s);

}  // namespace femhip
