// The splitmix64 uniform stream of rng_uniform, generated on the device. The
// generator's state after i + 1 draws is seed + (i + 1) * 0x9E3779B97F4A7C15
// (mod 2^64), so element i is a pure function of (seed, i): one thread per
// element, no sequential state. Integer mixing is exact; the conversion
// (x >> 11) * 2^-53 is exact; lo + (hi - lo) * u is two rounded operations with
// no contraction (the library is built with -ffp-contract=off), as on the host.
// The output is therefore the host stream's, bit for bit.
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

}  // namespace

void rng_uniform_device(uint64_t seed, size_t n, double lo, double hi, double* out, hipStream_t s) {
  if (n == 0) return;
  require(out != nullptr, "rng_uniform_device: out must not be null");
  const size_t want = (n + kThreads - 1) / kThreads;
  const unsigned grid = static_cast<unsigned>(std::min<size_t>(want, 16384));
  rng_uniform_kernel<<<grid, kThreads, 0, s>>>(seed, n, lo, hi - lo, out);
  HIP_CHECK_LAUNCH();
}

}  // namespace femhip
