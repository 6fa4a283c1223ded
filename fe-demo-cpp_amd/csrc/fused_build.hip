// Fused build_mass on trilinear hex geometry. The unfused composition
// (src/operator.cpp:148-182 with include/qfunctions/mass.hpp:18-40) gathers the
// vertex coordinates into an E-vector, contracts them into the gradient
// Q-vector dx[d][c][q][e] (src/tensor-contract.cpp:111-156), forms the weights
// w[q][e] (:169-192), computes det(J) w per point and scatters it into a zeroed
// output. Here one thread owns the quadrature points (qx = 0..Q-1, qy, qz) of
// one element and computes all of that in registers:
//   * every 1D product is "acc = 0; acc += t(j,0) u(0); acc += t(j,1) u(1);
//     v = 0 + acc", axis x first, then y, then z, with grad_1d on the axis of
//     the derivative and interp_1d on the others (the unfused kernels'
//     arithmetic, which is the reference's: src/tensor-contract.cpp:48-58);
//   * det(J) by cofactor expansion along row 0 in the association of
//     include/qfunctions/mass.hpp:31-33, times w = w_z (w_y (w_x * 1));
//   * the output is 0 + det w (the scatter into the zeroed L-vector).
// The library is built with -ffp-contract=off, so none of it is fused into
// multiply-adds and the result is the unfused path's, bit for bit.
#include "fused_build.hpp"

namespace femhip {

namespace {

constexpr int kThreads = 256;

// 2-point product of the unfused contraction: 0 + ((0 + t0*u0) + t1*u1).
__device__ __forceinline__ double dot2(double t0, double t1, double u0, double u1) {
  double acc = 0.0;
  acc += t0 * u0;
  acc += t1 * u1;
  return 0.0 + acc;
}

// Thread (qz, qy, e), e fastest: the Q points qx of one (qy, qz) row of element
// e. qdata[q*E + e], q = (qz*Q + qy)*Q + qx, so a wavefront's stores for one qx
// are contiguous.
__global__ __launch_bounds__(kThreads) void build_mass_trilinear_kernel(
    int E, int Q, const int* __restrict__ off, int cs, const double* __restrict__ x, const double* __restrict__ B,
    const double* __restrict__ G, const double* __restrict__ qw, double* __restrict__ qdata) {
  const long long idx = static_cast<long long>(blockIdx.x) * kThreads + threadIdx.x;
  if (idx >= static_cast<long long>(Q) * Q * E) return;
  const int e = static_cast<int>(idx % E);
  const int r = static_cast<int>(idx / E);
  const int qy = r % Q, qz = r / Q;
  // u[c][(iz*2 + iy)*2 + ix]: vertex coordinates of element e.
  double u[3][8];
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const int o = off[static_cast<size_t>(e) * 8 + i];
#pragma unroll
    for (int c = 0; c < 3; ++c) u[c][i] = x[o + static_cast<size_t>(c) * cs];
  }
  const double by0 = B[qy * 2], by1 = B[qy * 2 + 1], gy0 = G[qy * 2], gy1 = G[qy * 2 + 1];
  const double bz0 = B[qz * 2], bz1 = B[qz * 2 + 1], gz0 = G[qz * 2], gz1 = G[qz * 2 + 1];
  double* out = qdata + static_cast<size_t>(r) * Q * E + e;
#pragma unroll 1
  for (int qx = 0; qx < Q; ++qx) {
    const double bx0 = B[qx * 2], bx1 = B[qx * 2 + 1], gx0 = G[qx * 2], gx1 = G[qx * 2 + 1];
    double J[3][3];  // J[d][c] = d x_c / d X_d
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const double mx0 = d == 0 ? gx0 : bx0, mx1 = d == 0 ? gx1 : bx1;
      const double my0 = d == 1 ? gy0 : by0, my1 = d == 1 ? gy1 : by1;
      const double mz0 = d == 2 ? gz0 : bz0, mz1 = d == 2 ? gz1 : bz1;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        double a2[2];
#pragma unroll
        for (int iz = 0; iz < 2; ++iz) {
          const double a10 = dot2(mx0, mx1, u[c][iz * 4 + 0], u[c][iz * 4 + 1]);  // iy = 0
          const double a11 = dot2(mx0, mx1, u[c][iz * 4 + 2], u[c][iz * 4 + 3]);  // iy = 1
          a2[iz] = dot2(my0, my1, a10, a11);
        }
        J[d][c] = dot2(mz0, mz1, a2[0], a2[1]);
      }
    }
    const double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                       J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                       J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
    const double w = qw[qz] * (qw[qy] * (qw[qx] * 1.0));
    out[static_cast<size_t>(qx) * E] = 0.0 + det * w;
  }
}

}  // namespace

bool fused_build_mass_supported(const Restriction& r, const Basis& b) {
  return r.type == 0 && r.elem_size == 8 && r.num_comp == 3 && b.h.dim == 3 && b.h.num_comp == 3 && b.h.P == 2 &&
         b.h.Q >= 1 && r.num_elem > 0;
}

void fused_build_mass(const Restriction& r, const Basis& b, const double* coords, double* qdata, hipStream_t s) {
  require(fused_build_mass_supported(r, b), "fused build_mass: needs 8-node hexes with 3 components and a P=2 basis");
  const int E = r.num_elem, Q = b.h.Q;
  const long long threads = static_cast<long long>(Q) * Q * E;
  require(threads * Q < (1LL << 31), "fused build_mass: num_elem * Q^3 must fit in int32");
  build_mass_trilinear_kernel<<<static_cast<unsigned>((threads + kThreads - 1) / kThreads), kThreads, 0, s>>>(
      E, Q, r.offsets.data(), r.comp_stride, coords, b.interp.data(), b.grad.data(), b.q_weight.data(), qdata);
  HIP_CHECK_LAUNCH();
}

}  // namespace femhip
