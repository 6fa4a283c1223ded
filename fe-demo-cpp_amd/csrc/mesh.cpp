// Synthetic structured hex meshes for the benchmarks and tests: a splitmix64
// stream, lexicographic box offsets, and jittered trilinear vertex coordinates.
// (The reference has neither an RNG nor a mesh generator; its structured-offset
// helper is box_restriction in tests/cuda/test-cuda-elem-restriction.cu:24-42,
// whose numbering these offsets follow.)
#include "mesh.hpp"

namespace femhip {

double SplitMix64::uniform01() {
  state += 0x9E3779B97F4A7C15ull;
  uint64_t z = state;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  return static_cast<double>(z >> 11) * 0x1.0p-53;
}

void rng_uniform(uint64_t seed, size_t n, double lo, double hi, double* out) {
  SplitMix64 g{seed};
  for (size_t i = 0; i < n; ++i) out[i] = lo + (hi - lo) * g.uniform01();
}

void box_offsets(int nx, int ny, int nz, int P, int node_stride, int* offsets) {
  require(nx > 0 && ny > 0 && nz > 0 && P >= 2 && node_stride >= 1, "box_offsets: bad sizes");
  const long long Nx = static_cast<long long>(nx) * (P - 1) + 1, Ny = static_cast<long long>(ny) * (P - 1) + 1,
                  Nz = static_cast<long long>(nz) * (P - 1) + 1;
  require(Nx * Ny * Nz * node_stride < (1ll << 31), "box_offsets: node indices exceed int32");
  size_t t = 0;
  for (int ez = 0; ez < nz; ++ez)
    for (int ey = 0; ey < ny; ++ey)
      for (int ex = 0; ex < nx; ++ex)
        for (int iz = 0; iz < P; ++iz)
          for (int iy = 0; iy < P; ++iy)
            for (int ix = 0; ix < P; ++ix) {
              const long long gx = static_cast<long long>(ex) * (P - 1) + ix;
              const long long gy = static_cast<long long>(ey) * (P - 1) + iy;
              const long long gz = static_cast<long long>(ez) * (P - 1) + iz;
              offsets[t++] = static_cast<int>((gx + Nx * (gy + Ny * gz)) * node_stride);
            }
}

void box_coords(int n, double jitter, uint64_t seed, double* coords) {
  require(n > 0, "box_coords: n must be positive");
  const size_t nv1 = static_cast<size_t>(n) + 1;
  const size_t nv = nv1 * nv1 * nv1;
  const double h = 1.0 / n;
  SplitMix64 g{seed};
  for (size_t gz = 0; gz < nv1; ++gz)
    for (size_t gy = 0; gy < nv1; ++gy)
      for (size_t gx = 0; gx < nv1; ++gx) {
        const size_t v = gx + nv1 * (gy + nv1 * gz);
        double x = static_cast<double>(gx) / n, y = static_cast<double>(gy) / n, z = static_cast<double>(gz) / n;
        const bool interior = gx > 0 && gx < nv1 - 1 && gy > 0 && gy < nv1 - 1 && gz > 0 && gz < nv1 - 1;
        if (interior) {
          x += jitter * h * (2.0 * g.uniform01() - 1.0);
          y += jitter * h * (2.0 * g.uniform01() - 1.0);
          z += jitter * h * (2.0 * g.uniform01() - 1.0);
        }
        coords[v] = x;
        coords[nv + v] = y;
        coords[2 * nv + v] = z;
      }
}

}  // namespace femhip
