// The splitmix64 uniform stream of rng_uniform, generated on the device. The
// generator's state after i + 1 draws is seed + (i + 1) * 0x9E3779B97F4A7C15
// (mod 2^64), so element i is a pure function of (seed, i): one thread per
// element, no sequential state. Integer mixing is exact; the conversion
// (x >> 11) * 2^-53 is exact; lo + (hi - lo) * u is two rounded operations with
// no contraction (the library is built with -ffp-contract=off), as on the host.
// The output is therefore the host stream's, bit for bit. The box offsets and
// jittered vertex coordinates of mesh.cpp are made here too, entry for entry
// the host's (the coordinates from the same counter stream).
#include <algorithm>

#include "mesh.hpp"

namespace femhip {

namespace {

constexpr int kThreads = 256;

__global__ __launch_bounds__(kThreads) void rng_uniform_kernel(uint64_t seed, size_t n, double lo, double span,
                                                               double* __restrict__ out) {
  const size_t stride = static_cast<size_t>(gridDim.x) * kThreads;
  for (size_t i = static_cast<size_t>(blockIdx.x) * kThreads + threadIdx.x; i < n; i += stride) {
    uint64_t z = seed + (static_cast<uint64_t>(i) + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    const double u = static_cast<double>(z >> 11) * 0x1.0p-53;
    out[i] = lo + span * u;
  }
}

__device__ __forceinline__ double splitmix_u01(uint64_t seed, uint64_t draw) {  // draw 1, 2, ...
  uint64_t z = seed + draw * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  return static_cast<double>(z >> 11) * 0x1.0p-53;
}

// offsets[t], t = ((e*P + iz)*P + iy)*P + ix with e = ex + nx*(ey + ny*ez):
// (gx + Nx*(gy + Ny*gz)) * node_stride, g = e*(P-1) + i per direction.
__global__ __launch_bounds__(kThreads) void box_offsets_kernel(int nx, int ny, int P, long long Nx, long long Ny,
                                                               int node_stride, size_t nt, int* __restrict__ off) {
  const size_t stride = static_cast<size_t>(gridDim.x) * kThreads;
  const size_t P3 = static_cast<size_t>(P) * P * P;
  for (size_t t = static_cast<size_t>(blockIdx.x) * kThreads + threadIdx.x; t < nt; t += stride) {
    const size_t e = t / P3, i = t - e * P3;
    const int ix = static_cast<int>(i % P), iy = static_cast<int>((i / P) % P), iz = static_cast<int>(i / (P * P));
    const long long ex = static_cast<long long>(e % nx), ey = static_cast<long long>((e / nx) % ny),
                    ez = static_cast<long long>(e / (static_cast<size_t>(nx) * ny));
    const long long gx = ex * (P - 1) + ix, gy = ey * (P - 1) + iy, gz = ez * (P - 1) + iz;
    off[t] = static_cast<int>((gx + Nx * (gy + Ny * gz)) * node_stride);
  }
}

// Vertex v = (gx, gy, gz) of the (n+1)^3 grid: x = gx/n (+ jitter*h*(2u - 1)
// for interior vertices, u the draws 3k+1, 3k+2, 3k+3 of the stream, k the
// vertex's rank among the interior vertices in lexicographic order).
__global__ __launch_bounds__(kThreads) void box_coords_kernel(int n, double jitter, uint64_t seed,
                                                              double* __restrict__ coords) {
  const size_t nv1 = static_cast<size_t>(n) + 1, nv = nv1 * nv1 * nv1, m = static_cast<size_t>(n) - 1;
  const double h = 1.0 / n;
  const size_t stride = static_cast<size_t>(gridDim.x) * kThreads;
  for (size_t v = static_cast<size_t>(blockIdx.x) * kThreads + threadIdx.x; v < nv; v += stride) {
    const size_t gx = v % nv1, gy = (v / nv1) % nv1, gz = v / (nv1 * nv1);
    double x = static_cast<double>(gx) / n, y = static_cast<double>(gy) / n, z = static_cast<double>(gz) / n;
    if (gx > 0 && gx < nv1 - 1 && gy > 0 && gy < nv1 - 1 && gz > 0 && gz < nv1 - 1) {
      const uint64_t k = ((gz - 1) * m + (gy - 1)) * m + (gx - 1);
      x += jitter * h * (2.0 * splitmix_u01(seed, 3 * k + 1) - 1.0);
      y += jitter * h * (2.0 * splitmix_u01(seed, 3 * k + 2) - 1.0);
      z += jitter * h * (2.0 * splitmix_u01(seed, 3 * k + 3) - 1.0);
    }
    coords[v] = x;
    coords[nv + v] = y;
    coords[2 * nv + v] = z;
  }
}

unsigned grid_of(size_t n) { return static_cast<unsigned>(std::min<size_t>((n + kThreads - 1) / kThreads, 16384)); }

}  // namespace

void box_offsets_device(int nx, int ny, int nz, int P, int node_stride, int* offsets, hipStream_t s) {
  require(nx > 0 && ny > 0 && nz > 0 && P >= 2 && node_stride >= 1, "box_offsets: bad sizes");
  const long long Nx = static_cast<long long>(nx) * (P - 1) + 1, Ny = static_cast<long long>(ny) * (P - 1) + 1,
                  Nz = static_cast<long long>(nz) * (P - 1) + 1;
  require(Nx * Ny * Nz * node_stride < (1ll << 31), "box_offsets: node indices exceed int32");
  const size_t nt = static_cast<size_t>(nx) * ny * nz * P * P * P;
  require(offsets != nullptr, "box_offsets: offsets must not be null");
  box_offsets_kernel<<<grid_of(nt), kThreads, 0, s>>>(nx, ny, P, Nx, Ny, node_stride, nt, offsets);
  HIP_CHECK_LAUNCH();
}

void box_coords_device(int n, double jitter, uint64_t seed, double* coords, hipStream_t s) {
  require(n > 0, "box_coords: n must be positive");
  require(coords != nullptr, "box_coords: coords must not be null");
  const size_t nv1 = static_cast<size_t>(n) + 1;
  box_coords_kernel<<<grid_of(nv1 * nv1 * nv1), kThreads, 0, s>>>(n, jitter, seed, coords);
  HIP_CHECK_LAUNCH();
}

void rng_uniform_device(uint64_t seed, size_t n, double lo, double hi, double* out, hipStream_t s) {
  if (n == 0) return;
  require(out != nullptr, "rng_uniform_device: out must not be null");
  const size_t want = (n + kThreads - 1) / kThreads;
  const unsigned grid = static_cast<unsigned>(std::min<size_t>(want, 16384));
  rng_uniform_kernel<<<grid, kThreads, 0, s>>>(seed, n, lo, hi - lo, out);
  HIP_CHECK_LAUNCH();
}

}  // namespace femhip
