// CG building blocks shared by the solver (solver.hip) and the batched solve
// engine (batch_solve.hip): reduction grids and the fixed-order reductions,
// the scalar slots, the plain and pipelined Jacobi-PCG kernels, and the
// residual / correction kernels of mixed-precision refinement. Included by
// those two translation units only (everything here has internal linkage; the
// plain form's scalar and fold kernels, which only the solver launches, are
// marked maybe_unused).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "cg_slots.hpp"
#include "fused_mass.hpp"

namespace femhip {
namespace {

constexpr int kThreads = 256;

// Reduction grid for vectors of n entries: about 2 entries per thread (small
// vectors are latency-bound), at most kMaxRed blocks (large ones stream with
// every CU busy). Fixed for a given n, so every solve of that size sums in the
// same order.
int red_blocks(size_t n) {
  const size_t per = static_cast<size_t>(kThreads) * 2;
  return static_cast<int>(std::min<size_t>(kMaxRed, std::max<size_t>(1, (n + per - 1) / per)));
}

// Partial-sum regions of `part`: r.z and r.r (update and init), p.Ap (dot).
constexpr int kPartRZ = 0, kPartRR = kMaxRed, kPartPAP = 2 * kMaxRed;

__device__ __forceinline__ double block_sum(double v, double* sh) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  const int lane = threadIdx.x % 64, w = threadIdx.x / 64;
  if (lane == 0) sh[w] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int i = 0; i < kThreads / 64; ++i) s += sh[i];
  return s;  // valid in thread 0
}

template <class T>
__global__ __launch_bounds__(kThreads) void dot_partial_kernel(size_t n, const T* __restrict__ a,
                                                               const T* __restrict__ b, double* __restrict__ part) {
  __shared__ double sh[kThreads / 64];
  double s = 0.0;
  for (size_t i = static_cast<size_t>(blockIdx.x) * kThreads + threadIdx.x; i < n;
       i += static_cast<size_t>(gridDim.x) * kThreads)
    s += static_cast<double>(a[i]) * static_cast<double>(b[i]);
  s = block_sum(s, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// Sum of nb partials by one block: thread t adds partials t, t + kThreads, ...
// in order, then the fixed tree of block_sum. The result is valid in thread 0.
__device__ double sum_partials(const double* part, int nb, double* sh) {
  double s = 0.0;
  for (int i = threadIdx.x; i < nb; i += kThreads) s += part[i];
  return block_sum(s, sh);
}

// 16-byte packs of T (double2 / float4 loads and stores) for the per-iteration
// kernels; the last n % W entries are handled by block 0.
template <class T>
struct alignas(16) Pack {
  static constexpr int W = 16 / sizeof(T);
  T v[W];
};
template <class T>
__device__ __forceinline__ Pack<T> ldp(const T* a, size_t iv) {
  return reinterpret_cast<const Pack<T>*>(a)[iv];
}
template <class T>
__device__ __forceinline__ void stp(T* a, size_t iv, const Pack<T>& x) {
  reinterpret_cast<Pack<T>*>(a)[iv] = x;
}
inline bool aligned16(const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0; }

template <class T>
__global__ __launch_bounds__(kThreads) void dot_vec_kernel(size_t n, const T* __restrict__ a, const T* __restrict__ b,
                                                           double* __restrict__ part) {
  constexpr int W = Pack<T>::W;
  __shared__ double sh[kThreads / 64];
  double s = 0.0;
  const size_t nv = n / W;
  for (size_t iv = static_cast<size_t>(blockIdx.x) * kThreads + threadIdx.x; iv < nv;
       iv += static_cast<size_t>(gridDim.x) * kThreads) {
    const Pack<T> x = ldp(a, iv), y = ldp(b, iv);
#pragma unroll
    for (int k = 0; k < W; ++k) s += static_cast<double>(x.v[k]) * static_cast<double>(y.v[k]);
  }
  if (blockIdx.x == 0 && threadIdx.x < n - nv * W) {
    const size_t i = nv * W + threadIdx.x;
    s += static_cast<double>(a[i]) * static_cast<double>(b[i]);
  }
  s = block_sum(s, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// r = b, z = dinv * r, p = z; partials of r.z and r.r
template <class T>
__global__ __launch_bounds__(kThreads) void cg_init_kernel(size_t n, const T* __restrict__ b,
                                                           const T* __restrict__ dinv, T* __restrict__ x,
                                                           T* __restrict__ r, T* __restrict__ z, T* __restrict__ p,
                                                           double* __restrict__ part) {
  __shared__ double sh[2][kThreads / 64];
  double srz = 0.0, srr = 0.0;
  for (size_t i = static_cast<size_t>(blockIdx.x) * kThreads + threadIdx.x; i < n;
       i += static_cast<size_t>(gridDim.x) * kThreads) {
    const T ri = b[i];
    const T zi = dinv[i] * ri;
    x[i] = T(0);
    r[i] = ri;
    z[i] = zi;
    p[i] = zi;
    srz += static_cast<double>(ri) * static_cast<double>(zi);
    srr += static_cast<double>(ri) * static_cast<double>(ri);
  }
  const double a = block_sum(srz, sh[0]);
  const double c = block_sum(srr, sh[1]);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = a;
    part[kMaxRed + blockIdx.x] = c;
  }
}

[[maybe_unused]] __global__ __launch_bounds__(kThreads) void cg_init_scalars_kernel(const double* __restrict__ part,
                                                                   double* __restrict__ sc, double rtol2, int nb,
                                                                   int max_iter) {
  __shared__ double sh[2][kThreads / 64];
  const double rz = sum_partials(part, nb, sh[0]), rr = sum_partials(part + kMaxRed, nb, sh[1]);
  if (threadIdx.x != 0) return;
  sc[kRZ] = rz;
  sc[kRR] = rr;
  sc[kTHRESH] = rtol2 * rr;
  sc[kDONE] = (rr <= sc[kTHRESH]) ? 1.0 : 0.0;
  sc[kITERS] = 0.0;
  sc[kMAXIT] = static_cast<double>(max_iter);
}

// Residual replacement: r = rnew, z = dinv * r, x = 0 (the increment restarts);
// p is kept. Partials of r.z and r.r.
template <class T>
__global__ __launch_bounds__(kThreads) void cg_replace_kernel(size_t n, const T* __restrict__ rnew,
                                                              const T* __restrict__ dinv, T* __restrict__ x,
                                                              T* __restrict__ r, T* __restrict__ z,
                                                              double* __restrict__ part) {
  __shared__ double sh[2][kThreads / 64];
  double srz = 0.0, srr = 0.0;
  for (size_t i = static_cast<size_t>(blockIdx.x) * kThreads + threadIdx.x; i < n;
       i += static_cast<size_t>(gridDim.x) * kThreads) {
    const T ri = rnew[i];
    const T zi = dinv[i] * ri;
    x[i] = T(0);
    r[i] = ri;
    z[i] = zi;
    srz += static_cast<double>(ri) * static_cast<double>(zi);
    srr += static_cast<double>(ri) * static_cast<double>(ri);
  }
  const double a = block_sum(srz, sh[0]);
  const double c = block_sum(srr, sh[1]);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = a;
    part[kMaxRed + blockIdx.x] = c;
  }
}

// After a replacement: new r.z and r.r, the next segment's threshold; the
// iteration count runs on.
[[maybe_unused]] __global__ __launch_bounds__(kThreads) void cg_replace_scalars_kernel(const double* __restrict__ part,
                                                                      double* __restrict__ sc, double rtol2, int nb,
                                                                      int max_iter) {
  __shared__ double sh[2][kThreads / 64];
  const double rz = sum_partials(part, nb, sh[0]), rr = sum_partials(part + kMaxRed, nb, sh[1]);
  if (threadIdx.x != 0) return;
  sc[kRZ] = rz;
  sc[kRR] = rr;
  sc[kTHRESH] = rtol2 * rr;
  sc[kDONE] = (rr <= sc[kTHRESH]) ? 1.0 : 0.0;
  sc[kMAXIT] = static_cast<double>(max_iter);
}

// The sum of the nb partials at part, the same in every thread and every block
// (sum_partials' fixed order, then a broadcast through LDS).
__device__ double all_partials(const double* part, int nb, double* sh, double* out) {
  const double v = sum_partials(part, nb, sh);
  if (threadIdx.x == 0) *out = v;
  __syncthreads();
  return *out;
}

// One CG step's update, with alpha = r.z / p.Ap formed by every block from the
// p.Ap partials: r -= alpha Ap, z = dinv r, and the partials of the new r.z and
// r.r. x += alpha p is left to the direction kernel, which reads p anyway (one
// pass over x and p fewer an iteration, the same arithmetic). An iteration
// that starts done changes nothing (alpha 0) and only sums r.r.
template <class T>
__global__ __launch_bounds__(kThreads) void cg_update_kernel(size_t n, int par, int nb, double* __restrict__ sc,
                                                             const T* __restrict__ Ap, const T* __restrict__ dinv,
                                                             T* __restrict__ r, T* __restrict__ z,
                                                             double* __restrict__ part) {
  __shared__ double sh[3][kThreads / 64];
  __shared__ double s_pap;
  const bool done = sc[slot(kDONE, par)] != 0.0;
  const double pap = all_partials(part + kPartPAP, nb, sh[2], &s_pap);
  const double alpha_d = (done || pap == 0.0) ? 0.0 : sc[slot(kRZ, par)] / pap;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    sc[kPAP] = pap;
    sc[kALPHA] = alpha_d;
  }
  const T alpha = static_cast<T>(alpha_d);
  double srz = 0.0, srr = 0.0;
  for (size_t i = static_cast<size_t>(blockIdx.x) * kThreads + threadIdx.x; i < n;
       i += static_cast<size_t>(gridDim.x) * kThreads) {
    T ri = r[i];
    if (!done) {
      ri = ri - alpha * Ap[i];
      r[i] = ri;
      const T zi = dinv[i] * ri;
      z[i] = zi;
      srz += static_cast<double>(ri) * static_cast<double>(zi);
    }
    srr += static_cast<double>(ri) * static_cast<double>(ri);
  }
  const double a = block_sum(srz, sh[0]);
  const double c = block_sum(srr, sh[1]);
  if (threadIdx.x == 0) {
    part[kPartRZ + blockIdx.x] = a;
    part[kPartRR + blockIdx.x] = c;
  }
}

// The update kernel on 16-byte packs (same arithmetic per entry).
template <class T>
__global__ __launch_bounds__(kThreads) void cg_update_vec_kernel(size_t n, int par, int nb, double* __restrict__ sc,
                                                                 const T* __restrict__ Ap, const T* __restrict__ dinv,
                                                                 T* __restrict__ r, T* __restrict__ z,
                                                                 double* __restrict__ part) {
  constexpr int W = Pack<T>::W;
  __shared__ double sh[3][kThreads / 64];
  __shared__ double s_pap;
  const bool done = sc[slot(kDONE, par)] != 0.0;
  const double pap = all_partials(part + kPartPAP, nb, sh[2], &s_pap);
  const double alpha_d = (done || pap == 0.0) ? 0.0 : sc[slot(kRZ, par)] / pap;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    sc[kPAP] = pap;
    sc[kALPHA] = alpha_d;
  }
  const T alpha = static_cast<T>(alpha_d);
  double srz = 0.0, srr = 0.0;
  auto one = [&](T& ri, T api, T di, T& zi) {
    if (!done) {
      ri = ri - alpha * api;
      zi = di * ri;
      srz += static_cast<double>(ri) * static_cast<double>(zi);
    }
    srr += static_cast<double>(ri) * static_cast<double>(ri);
  };
  const size_t nv = n / W;
  for (size_t iv = static_cast<size_t>(blockIdx.x) * kThreads + threadIdx.x; iv < nv;
       iv += static_cast<size_t>(gridDim.x) * kThreads) {
    Pack<T> rv = ldp(r, iv);
    if (!done) {
      Pack<T> zv;
      const Pack<T> av = ldp(Ap, iv), dv = ldp(dinv, iv);
#pragma unroll
      for (int k = 0; k < W; ++k) one(rv.v[k], av.v[k], dv.v[k], zv.v[k]);
      stp(r, iv, rv);
      stp(z, iv, zv);
    } else {
#pragma unroll
      for (int k = 0; k < W; ++k) srr += static_cast<double>(rv.v[k]) * static_cast<double>(rv.v[k]);
    }
  }
  if (blockIdx.x == 0 && threadIdx.x < n - nv * W) {
    const size_t i = nv * W + threadIdx.x;
    T ri = r[i], zi = T(0);
    one(ri, Ap[i], dinv[i], zi);
    if (!done) {
      r[i] = ri;
      z[i] = zi;
    }
  }
  const double a = block_sum(srz, sh[0]);
  const double c = block_sum(srr, sh[1]);
  if (threadIdx.x == 0) {
    part[kPartRZ + blockIdx.x] = a;
    part[kPartRR + blockIdx.x] = c;
  }
}

// The end of a CG step, with beta = r.z(new) / r.z(old) formed by every block
// from the update's partials: x += alpha p, p = z + beta p, and (block 0) the state copy
// 1-par: the iteration count, r.r, r.z and the done test r.r <= threshold or
// the iteration limit (both read from the scalar slots). An iteration that started done copies the state on and
// leaves p alone (*frozen). Returns beta.
[[maybe_unused]] __device__ double cg_step_end(int par, int nb, double* sc, const double* part, double (*sh)[kThreads / 64],
                              double* s_val, bool* frozen) {
  const bool done0 = sc[slot(kDONE, par)] != 0.0;
  const double rz_new = all_partials(part + kPartRZ, nb, sh[0], &s_val[0]);
  const double rr = all_partials(part + kPartRR, nb, sh[1], &s_val[1]);
  const double rz_old = sc[slot(kRZ, par)];
  const double beta = done0 ? 0.0 : (rz_old == 0.0 ? 0.0 : rz_new / rz_old);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    const int q = 1 - par;
    if (done0) {
      sc[slot(kRZ, q)] = rz_old;
      sc[slot(kRR, q)] = sc[slot(kRR, par)];
      sc[slot(kITERS, q)] = sc[slot(kITERS, par)];
      sc[slot(kDONE, q)] = 1.0;
    } else {
      const double iters = sc[slot(kITERS, par)] + 1.0;
      sc[slot(kITERS, q)] = iters;
      sc[slot(kRR, q)] = rr;
      sc[slot(kRZ, q)] = rz_new;
      sc[slot(kDONE, q)] = (rr <= sc[kTHRESH] || iters >= sc[kMAXIT]) ? 1.0 : 0.0;
    }
    sc[kBETA] = beta;
  }
  *frozen = done0;
  return beta;
}

// x += alpha p (alpha from its slot, as the update used it), then p = z + beta p.
template <class T>
__global__ __launch_bounds__(kThreads) void cg_p_vec_kernel(size_t n, int par, int nb,
                                                            double* __restrict__ sc, const double* __restrict__ part,
                                                            const T* __restrict__ z, T* __restrict__ p,
                                                            T* __restrict__ x) {
  constexpr int W = Pack<T>::W;
  __shared__ double sh[2][kThreads / 64];
  __shared__ double s_val[2];
  bool frozen = false;
  const T beta = static_cast<T>(cg_step_end(par, nb, sc, part, sh, s_val, &frozen));
  if (frozen) return;  // x and p stay
  const T alpha = static_cast<T>(sc[kALPHA]);
  const size_t nv = n / W;
  for (size_t iv = static_cast<size_t>(blockIdx.x) * kThreads + threadIdx.x; iv < nv;
       iv += static_cast<size_t>(gridDim.x) * kThreads) {
    const Pack<T> zv = ldp(z, iv);
    Pack<T> pv = ldp(p, iv), xv = ldp(x, iv);
#pragma unroll
    for (int k = 0; k < W; ++k) {
      xv.v[k] = xv.v[k] + alpha * pv.v[k];
      pv.v[k] = zv.v[k] + beta * pv.v[k];
    }
    stp(x, iv, xv);
    stp(p, iv, pv);
  }
  if (blockIdx.x == 0 && threadIdx.x < n - nv * W) {
    const size_t i = nv * W + threadIdx.x;
    const T pi = p[i];
    x[i] = x[i] + alpha * pi;
    p[i] = z[i] + beta * pi;
  }
}

template <class T>
__global__ __launch_bounds__(kThreads) void cg_p_kernel(size_t n, int par, int nb,
                                                        double* __restrict__ sc, const double* __restrict__ part,
                                                        const T* __restrict__ z, T* __restrict__ p,
                                                        T* __restrict__ x) {
  __shared__ double sh[2][kThreads / 64];
  __shared__ double s_val[2];
  bool frozen = false;
  const T beta = static_cast<T>(cg_step_end(par, nb, sc, part, sh, s_val, &frozen));
  if (frozen) return;  // x and p stay
  const T alpha = static_cast<T>(sc[kALPHA]);
  for (size_t i = static_cast<size_t>(blockIdx.x) * kThreads + threadIdx.x; i < n;
       i += static_cast<size_t>(gridDim.x) * kThreads) {
    const T pi = p[i];
    x[i] = x[i] + alpha * pi;
    p[i] = z[i] + beta * pi;
  }
}

// ---- The plain form on long vectors ("wide"): one 16-byte pack per thread
// instead of a grid stride over at most kMaxRed blocks, so that every CU has
// many independent loads in flight. Each block writes one partial; a
// fold kernel sums them into the red_blocks(n) partials of the standard form
// (block b of nb takes a fixed contiguous range, so the order is fixed), and
// one-block kernels form alpha and beta (and the step-end state) into the
// scalar slots, which every block of the next wide kernel reads. An iteration
// is the apply and seven kernels instead of three: the form for vectors long
// enough that streaming, not launches, sets the time.

// Wide blocks of a vector of n entries (one pack per thread).
template <class T>
int wide_blocks(size_t n) {
  const size_t nv = std::max<size_t>(1, n / Pack<T>::W);
  return static_cast<int>((nv + kThreads - 1) / kThreads);
}

template <class T>
__global__ __launch_bounds__(kThreads) void dot_wide_kernel(size_t n, const T* __restrict__ a, const T* __restrict__ b,
                                                            double* __restrict__ part) {
  constexpr int W = Pack<T>::W;
  __shared__ double sh[kThreads / 64];
  const size_t nv = n / W, iv = static_cast<size_t>(blockIdx.x) * kThreads + threadIdx.x;
  double s = 0.0;
  if (iv < nv) {
    const Pack<T> x = ldp(a, iv), y = ldp(b, iv);
#pragma unroll
    for (int k = 0; k < W; ++k) s += static_cast<double>(x.v[k]) * static_cast<double>(y.v[k]);
  }
  if (blockIdx.x == 0 && threadIdx.x < n - nv * W) {
    const size_t i = nv * W + threadIdx.x;
    s += static_cast<double>(a[i]) * static_cast<double>(b[i]);
  }
  s = block_sum(s, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// out[y * out_stride + b] = the sum of in[y * in_stride + i] over block b's
// contiguous share of i < nin (nout = gridDim.x blocks), in ascending order
// per thread and then the fixed tree of block_sum.
[[maybe_unused]] __global__ __launch_bounds__(kThreads) void fold_partials_kernel(const double* __restrict__ in, int nin,
                                                                 size_t in_stride, double* __restrict__ out,
                                                                 size_t out_stride) {
  __shared__ double sh[kThreads / 64];
  const int nout = static_cast<int>(gridDim.x);
  const int per = (nin + nout - 1) / nout;
  const int lo = static_cast<int>(blockIdx.x) * per, hi = min(nin, lo + per);
  const double* src = in + blockIdx.y * in_stride;
  double s = 0.0;
  for (int i = lo + static_cast<int>(threadIdx.x); i < hi; i += kThreads) s += src[i];
  s = block_sum(s, sh);
  if (threadIdx.x == 0) out[blockIdx.y * out_stride + blockIdx.x] = s;
}

// One block: alpha = r.z / p.Ap from the folded p.Ap partials (0 once done).
[[maybe_unused]] __global__ __launch_bounds__(kThreads) void cg_alpha_kernel(int par, int nb, double* __restrict__ sc,
                                                            const double* __restrict__ part) {
  __shared__ double sh[kThreads / 64];
  __shared__ double s_pap;
  const double pap = all_partials(part + kPartPAP, nb, sh, &s_pap);
  if (threadIdx.x == 0) {
    const bool done = sc[slot(kDONE, par)] != 0.0;
    sc[kPAP] = pap;
    sc[kALPHA] = (done || pap == 0.0) ? 0.0 : sc[slot(kRZ, par)] / pap;
  }
}

// r -= alpha Ap, z = dinv r, the block's r.z and r.r partials (alpha from its
// slot; x += alpha p is the direction kernel's; an iteration that starts done
// only sums r.r).
template <class T>
__global__ __launch_bounds__(kThreads) void cg_update_wide_kernel(size_t n, int par, const double* __restrict__ sc,
                                                                  const T* __restrict__ Ap, const T* __restrict__ dinv,
                                                                  T* __restrict__ r, T* __restrict__ z,
                                                                  double* __restrict__ part_rz,
                                                                  double* __restrict__ part_rr) {
  constexpr int W = Pack<T>::W;
  __shared__ double sh[2][kThreads / 64];
  const bool done = sc[slot(kDONE, par)] != 0.0;
  const T alpha = static_cast<T>(sc[kALPHA]);
  double srz = 0.0, srr = 0.0;
  auto one = [&](T& ri, T api, T di, T& zi) {
    if (!done) {
      ri = ri - alpha * api;
      zi = di * ri;
      srz += static_cast<double>(ri) * static_cast<double>(zi);
    }
    srr += static_cast<double>(ri) * static_cast<double>(ri);
  };
  const size_t nv = n / W, iv = static_cast<size_t>(blockIdx.x) * kThreads + threadIdx.x;
  if (iv < nv) {
    Pack<T> rv = ldp(r, iv);
    if (!done) {
      Pack<T> zv;
      const Pack<T> av = ldp(Ap, iv), dv = ldp(dinv, iv);
#pragma unroll
      for (int k = 0; k < W; ++k) one(rv.v[k], av.v[k], dv.v[k], zv.v[k]);
      stp(r, iv, rv);
      stp(z, iv, zv);
    } else {
#pragma unroll
      for (int k = 0; k < W; ++k) srr += static_cast<double>(rv.v[k]) * static_cast<double>(rv.v[k]);
    }
  }
  if (blockIdx.x == 0 && threadIdx.x < n - nv * W) {
    const size_t i = nv * W + threadIdx.x;
    T ri = r[i], zi = T(0);
    one(ri, Ap[i], dinv[i], zi);
    if (!done) {
      r[i] = ri;
      z[i] = zi;
    }
  }
  const double a = block_sum(srz, sh[0]);
  const double c = block_sum(srr, sh[1]);
  if (threadIdx.x == 0) {
    part_rz[blockIdx.x] = a;
    part_rr[blockIdx.x] = c;
  }
}

// One block: the step end (beta and the state copy 1-par) from the folded
// r.z and r.r partials.
[[maybe_unused]] __global__ __launch_bounds__(kThreads) void cg_beta_kernel(int par, int nb, double* __restrict__ sc,
                                                           const double* __restrict__ part) {
  __shared__ double sh[2][kThreads / 64];
  __shared__ double s_val[2];
  bool frozen = false;
  (void)cg_step_end(par, nb, sc, part, sh, s_val, &frozen);
}

// x += alpha p, then p = z + beta p (alpha and beta from their slots); an
// iteration that started done leaves both.
template <class T>
__global__ __launch_bounds__(kThreads) void cg_p_wide_kernel(size_t n, int par, const double* __restrict__ sc,
                                                             const T* __restrict__ z, T* __restrict__ p,
                                                             T* __restrict__ x) {
  constexpr int W = Pack<T>::W;
  if (sc[slot(kDONE, par)] != 0.0) return;
  const T alpha = static_cast<T>(sc[kALPHA]), beta = static_cast<T>(sc[kBETA]);
  const size_t nv = n / W, iv = static_cast<size_t>(blockIdx.x) * kThreads + threadIdx.x;
  if (iv < nv) {
    const Pack<T> zv = ldp(z, iv);
    Pack<T> pv = ldp(p, iv), xv = ldp(x, iv);
#pragma unroll
    for (int k = 0; k < W; ++k) {
      xv.v[k] = xv.v[k] + alpha * pv.v[k];
      pv.v[k] = zv.v[k] + beta * pv.v[k];
    }
    stp(x, iv, xv);
    stp(p, iv, pv);
  }
  if (blockIdx.x == 0 && threadIdx.x < n - nv * W) {
    const size_t i = nv * W + threadIdx.x;
    const T pi = p[i];
    x[i] = x[i] + alpha * pi;
    p[i] = z[i] + beta * pi;
  }
}

// ---- Pipelined Jacobi-PCG (Ghysels & Vanroose, "Hiding global
// synchronization latency in the preconditioned Conjugate Gradient
// algorithm", Parallel Computing 40(7), 2014, Alg. 3, with M = diag(A)).
// Vectors: x, r, u = M^-1 r, w = A u, m = M^-1 w, n = A m, and the recurrences
// p, s = A p, q = M^-1 s, z = A q. An iteration is the apply n = A m and one
// kernel (pipe_step_kernel), in place of the apply, dot, update and direction
// kernels of the plain form: fewer dependent launches for problems whose
// kernels are too short to fill the GPU, at the price of more vector traffic
// (about 20 vector passes an iteration instead of 13). The dot products of an
// iteration are formed from partials the previous kernel wrote, double-buffered
// by parity (regions gamma = r.u, delta = w.u, rho = r.r).
constexpr int kPipePart = 3 * kMaxRed;

// x = 0, r = b, u = M^-1 r
template <class T>
__global__ __launch_bounds__(kThreads) void pipe_init_kernel(size_t n, const T* __restrict__ b,
                                                             const T* __restrict__ dinv, T* __restrict__ x,
                                                             T* __restrict__ r, T* __restrict__ u) {
  for (size_t i = static_cast<size_t>(blockIdx.x) * kThreads + threadIdx.x; i < n;
       i += static_cast<size_t>(gridDim.x) * kThreads) {
    const T ri = b[i];
    x[i] = T(0);
    r[i] = ri;
    u[i] = dinv[i] * ri;
  }
}

// m = M^-1 w; the parity-0 partials of r.u, w.u and r.r.
template <class T>
__global__ __launch_bounds__(kThreads) void pipe_init2_kernel(size_t n, const T* __restrict__ r,
                                                              const T* __restrict__ u, const T* __restrict__ w,
                                                              const T* __restrict__ dinv, T* __restrict__ m,
                                                              double* __restrict__ part) {
  __shared__ double sh[3][kThreads / 64];
  double g = 0.0, d = 0.0, q = 0.0;
  for (size_t i = static_cast<size_t>(blockIdx.x) * kThreads + threadIdx.x; i < n;
       i += static_cast<size_t>(gridDim.x) * kThreads) {
    const T ri = r[i], ui = u[i], wi = w[i];
    m[i] = dinv[i] * wi;
    g += static_cast<double>(ri) * static_cast<double>(ui);
    d += static_cast<double>(wi) * static_cast<double>(ui);
    q += static_cast<double>(ri) * static_cast<double>(ri);
  }
  const double a = block_sum(g, sh[0]), bb = block_sum(d, sh[1]), c = block_sum(q, sh[2]);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = a;
    part[kMaxRed + blockIdx.x] = bb;
    part[2 * kMaxRed + blockIdx.x] = c;
  }
}

__global__ __launch_bounds__(kThreads) void pipe_init_scalars_kernel(const double* __restrict__ part,
                                                                     double* __restrict__ sc, double rtol2, int nb,
                                                                     int max_iter) {
  __shared__ double sh[kThreads / 64];
  const double rr = sum_partials(part + 2 * kMaxRed, nb, sh);
  if (threadIdx.x != 0) return;
  sc[kRR] = rr;
  sc[kTHRESH] = rtol2 * rr;
  sc[kDONE] = (rr <= sc[kTHRESH]) ? 1.0 : 0.0;
  sc[kITERS] = 0.0;
  sc[kMAXIT] = static_cast<double>(max_iter);
  sc[kRZ] = 0.0;  // the previous gamma and alpha (none before the first step)
  sc[kAP] = 0.0;
}

// One pipelined step of parity par: gamma, delta, rho from the parity-par
// partials (every block sums them in one fixed order); the residual test on
// rho (r of this step) and the iteration limit; then, unless stopped,
//   beta = gamma / gamma_prev (0 first), alpha = gamma / (delta - beta gamma / alpha_prev)
//   z = n + beta z, q = m + beta q, s = w + beta s, p = u + beta p,
//   x += alpha p, r -= alpha s, u -= alpha q, w -= alpha z, m = M^-1 w,
// and the parity-(1-par) partials of the new r.u, w.u and r.r. Block 0
// writes the state copy 1-par (a stopped step copies it on unchanged).
// With SPLIT the apply's second pass is folded in: n_i is formed here from
// what the element kernel left (FusedSplit), in the second pass's order.
// pipe_step is block `bid` of `nb` (pipe_step_kernel runs it once per block,
// the batched step kernel once per block of each run it serves); sh and
// s_val are the block's LDS scratch.
template <class T, bool SPLIT>
__device__ __forceinline__ void pipe_step(size_t n, int par, int nb, int bid, double* __restrict__ sc,
                                          double* __restrict__ part, const T* __restrict__ dinv, const FusedSplit& sp,
                                          const T* __restrict__ nv, T* __restrict__ x, T* __restrict__ r,
                                          T* __restrict__ u, T* __restrict__ w, T* __restrict__ m,
                                          T* __restrict__ z, T* __restrict__ q, T* __restrict__ s,
                                          T* __restrict__ p, double (*sh)[kThreads / 64], double* s_val) {
  const double* pin = part + static_cast<size_t>(par) * kPipePart;
  double* pout = part + static_cast<size_t>(1 - par) * kPipePart;
  const double gamma = all_partials(pin, nb, sh[0], &s_val[0]);
  const double delta = all_partials(pin + kMaxRed, nb, sh[1], &s_val[1]);
  const double rho = all_partials(pin + 2 * kMaxRed, nb, sh[2], &s_val[2]);
  const bool done0 = sc[slot(kDONE, par)] != 0.0;
  const double iters = sc[slot(kITERS, par)];
  const bool stop = done0 || rho <= sc[kTHRESH] || iters >= sc[kMAXIT];
  const double g_prev = sc[slot(kRZ, par)], a_prev = sc[slot(kAP, par)];
  double beta = 0.0, alpha = 0.0;
  if (!stop) {
    if (iters > 0.0 && g_prev != 0.0) beta = gamma / g_prev;
    const double den = (iters > 0.0 && a_prev != 0.0) ? delta - beta * gamma / a_prev : delta;
    alpha = den != 0.0 ? gamma / den : 0.0;
  }
  if (bid == 0 && threadIdx.x == 0) {
    const int o = 1 - par;
    sc[slot(kDONE, o)] = stop ? 1.0 : 0.0;
    sc[slot(kITERS, o)] = stop ? iters : iters + 1.0;
    sc[slot(kRR, o)] = done0 ? sc[slot(kRR, par)] : rho;
    sc[slot(kRZ, o)] = stop ? g_prev : gamma;
    sc[slot(kAP, o)] = stop ? a_prev : alpha;
    sc[kALPHA] = alpha;
    sc[kBETA] = beta;
  }
  if (stop) return;  // the same in every block
  // The first step of a run starts the recurrences (z = n, q = m, s = w,
  // p = u) without reading them: they hold whatever an earlier run or the
  // allocation left, which beta = 0 would not clear if it is Inf or NaN.
  const bool first = iters == 0.0;
  const T a = static_cast<T>(alpha), bt = static_cast<T>(beta);
  double g = 0.0, d = 0.0, rr = 0.0;
  const T* __restrict__ spv = static_cast<const T*>(sp.vals);
  for (size_t i = static_cast<size_t>(bid) * kThreads + threadIdx.x; i < n;
       i += static_cast<size_t>(nb) * kThreads) {
    T ni;
    if constexpr (SPLIT) {
      const unsigned ui = static_cast<unsigned>(i), ucs = static_cast<unsigned>(sp.cs);
      const unsigned c = ui >= ucs ? ui / ucs : 0u;
      const int node = static_cast<int>(ui - c * ucs);
      const int row = sp.node_row ? sp.node_row[node] : node;
      if (row < 0) {
        ni = nv[i];
      } else {
        const T* vc = spv + static_cast<size_t>(c) * sp.cl;
        T acc = T(0);
        for (int k = sp.ptr[row], ke = sp.ptr[row + 1]; k < ke; ++k) acc += vc[sp.idx[k]];
        ni = acc;
      }
    } else {
      ni = nv[i];
    }
    T zi = ni, qi = m[i], si = w[i], pi = u[i];
    if (!first) {
      zi += bt * z[i];
      qi += bt * q[i];
      si += bt * s[i];
      pi += bt * p[i];
    }
    z[i] = zi;
    q[i] = qi;
    s[i] = si;
    p[i] = pi;
    x[i] = x[i] + a * pi;
    const T ri = r[i] - a * si, ui = u[i] - a * qi, wi = w[i] - a * zi;
    r[i] = ri;
    u[i] = ui;
    w[i] = wi;
    m[i] = dinv[i] * wi;
    g += static_cast<double>(ri) * static_cast<double>(ui);
    d += static_cast<double>(wi) * static_cast<double>(ui);
    rr += static_cast<double>(ri) * static_cast<double>(ri);
  }
  const double sg = block_sum(g, sh[0]), sd = block_sum(d, sh[1]), sr = block_sum(rr, sh[2]);
  if (threadIdx.x == 0) {
    pout[bid] = sg;
    pout[kMaxRed + bid] = sd;
    pout[2 * kMaxRed + bid] = sr;
  }
}

template <class T, bool SPLIT>
__global__ __launch_bounds__(kThreads) void pipe_step_kernel(size_t n, int par, int nb, double* __restrict__ sc,
                                                             double* __restrict__ part, const T* __restrict__ dinv,
                                                             FusedSplit sp, const T* __restrict__ nv, T* __restrict__ x,
                                                             T* __restrict__ r, T* __restrict__ u,
                                                             T* __restrict__ w, T* __restrict__ m,
                                                             T* __restrict__ z, T* __restrict__ q,
                                                             T* __restrict__ s, T* __restrict__ p) {
  __shared__ double sh[3][kThreads / 64];
  __shared__ double s_val[3];
  pipe_step<T, SPLIT>(n, par, nb, static_cast<int>(blockIdx.x), sc, part, dinv, sp, nv, x, r, u, w, m, z, q, s, p, sh,
                      s_val);
}

template <class T>
__global__ __launch_bounds__(kThreads) void recip_kernel(size_t n, const double* __restrict__ d, T* __restrict__ dinv) {
  const size_t i = static_cast<size_t>(blockIdx.x) * kThreads + threadIdx.x;
  if (i < n) dinv[i] = static_cast<T>(d[i] != 0.0 ? 1.0 / d[i] : 0.0);
}

// r = b - Mx (FP64) and its squared norm partials; r32 = (float) r.
__global__ __launch_bounds__(kThreads) void residual_kernel(size_t n, const double* __restrict__ b,
                                                            const double* __restrict__ Mx, double* __restrict__ r,
                                                            float* __restrict__ r32, double* __restrict__ part) {
  __shared__ double sh[kThreads / 64];
  double s = 0.0;
  for (size_t i = static_cast<size_t>(blockIdx.x) * kThreads + threadIdx.x; i < n;
       i += static_cast<size_t>(gridDim.x) * kThreads) {
    const double ri = b[i] - Mx[i];
    r[i] = ri;
    if (r32) r32[i] = static_cast<float>(ri);
    s += ri * ri;
  }
  s = block_sum(s, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ __launch_bounds__(kThreads) void add_correction_kernel(size_t n, const float* __restrict__ d,
                                                                  double* __restrict__ x) {
  const size_t i = static_cast<size_t>(blockIdx.x) * kThreads + threadIdx.x;
  if (i < n) x[i] += static_cast<double>(d[i]);
}

__global__ __launch_bounds__(kThreads) void sum_to_scalar_kernel(const double* __restrict__ part,
                                                                 double* __restrict__ out, int nb) {
  __shared__ double sh[kThreads / 64];
  const double v = sum_partials(part, nb, sh);
  if (threadIdx.x == 0) *out = v;
}

}  // namespace
}  // namespace femhip
