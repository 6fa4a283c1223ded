// The fused per-element mass kernel template (gfx950, wave64).
//
// Workgroup = NE elements x TPE threads (TPE = max(P^2, PQ, Q^2), padded to a
// whole number of wavefronts). Per element and component the data stays in LDS
// between the six 1D contractions; each contraction is done by one thread per
// 1D line (row or column of the tile) holding that line in registers:
//   gather  U[iz][iy][ix]      <- u[offsets[e][i] + c*comp_stride]
//   F1 (x)  A [iz][iy][qx]     rows (iz,iy)
//   F2 (y)  Bf[iz][qy][qx]     columns (iz,qx)
//   F3 (z)  D [qz][qy][qx] = qdata * (sum over iz)   columns (qy,qx), in place
//   T1 (x)  A1[qz][qy][ix]     rows (qz,qy)
//   T2 (y)  A2[qz][iy][ix]     columns (qz,ix)
//   T3 (z)  V [iz][iy][ix]     columns (iy,ix) -> scatter
// This is the axis order of src/tensor-contract.cpp:74-91 in both directions
// (x first), so with EXACT (no fused multiply-add, sums from 0 in ascending
// order, 0 + sum) every element output equals the reference's bit for bit.
//
// qdata is read in a private layout [e][qy][qx][qz] so that each column thread
// loads its Q values contiguously. Workgroup ids are remapped so that each of
// the 8 XCDs (which receive workgroups round-robin) walks one contiguous range
// of elements, keeping shared nodes of neighbouring elements in its own L2.
#pragma once

#include <hip/hip_runtime.h>

#include "fused_mass.hpp"

namespace femhip {

template <class T, int N>
struct MatArg {
  T m[N];
};

constexpr int cmax(int a, int b) { return a > b ? a : b; }
constexpr int round64(int n) { return (n + 63) / 64 * 64; }

template <int P, int Q>
struct Shape {
  static constexpr int P2 = P * P, P3 = P * P * P, Q2 = Q * Q, Q3 = Q * Q * Q;
  static constexpr int TPE = cmax(cmax(P2, P * Q), Q2);
  static constexpr int BUF = cmax(cmax(P3, P2 * Q), cmax(P * Q2, Q3));
};

template <bool EXACT, class T>
__device__ __forceinline__ T mac(T a, T b, T acc) {
  if constexpr (EXACT) {
    return acc + a * b;  // separate rounding (the library is built with -ffp-contract=off)
  } else {
    return __builtin_fma(a, b, acc);
  }
}

template <bool EXACT, class T>
__device__ __forceinline__ T finish(T acc) {
  if constexpr (EXACT) {
    return T(0) + acc;  // the reference adds the sum into a zeroed output
  } else {
    return acc;
  }
}

template <int P, int Q, int NC, class T, int NE, bool EXACT, bool DET>
__global__ __launch_bounds__(round64(NE* Shape<P, Q>::TPE)) void fused_mass_kernel(
    int num_elem, int comp_stride, const int* __restrict__ offsets, const T* __restrict__ qdata,
    const T* __restrict__ u, T* __restrict__ out, size_t y_comp_len, MatArg<T, P * Q> Bm) {
  using S = Shape<P, Q>;
  constexpr int P2 = S::P2, P3 = S::P3, Q2 = S::Q2, TPE = S::TPE, BUF = S::BUF;
  __shared__ T s_buf[2][NE][BUF];
  __shared__ int s_off[NE * P3];

  // XCD-aware block order: hardware id h runs on XCD h % 8; give XCD x the
  // contiguous logical range [x*per, (x+1)*per). The grid is a multiple of 8.
  const int per = gridDim.x / 8;
  const int blk = (blockIdx.x % 8) * per + blockIdx.x / 8;
  const long long e0 = static_cast<long long>(blk) * NE;
  if (e0 >= num_elem) return;  // whole workgroup leaves together: no barrier is skipped
  const int n_here = static_cast<int>(min(static_cast<long long>(NE), num_elem - e0));

  const int el = threadIdx.x / TPE;
  const int t = threadIdx.x - el * TPE;
  const bool live = el < n_here;
  const long long e = e0 + el;

  for (int idx = threadIdx.x; idx < n_here * P3; idx += blockDim.x) s_off[idx] = offsets[e0 * P3 + idx];

  T qd[Q];
  if (live && t < Q2) {
    const T* qp = qdata + (static_cast<size_t>(e) * Q2 + t) * Q;
#pragma unroll
    for (int qz = 0; qz < Q; ++qz) qd[qz] = qp[qz];
  }
  __syncthreads();

  auto B = [&](int q, int p) -> T { return Bm.m[q * P + p]; };
  T* b0 = live ? s_buf[0][el] : nullptr;
  T* b1 = live ? s_buf[1][el] : nullptr;
  const int* off = s_off + el * P3;

#pragma unroll 1
  for (int c = 0; c < NC; ++c) {
    const size_t coff = static_cast<size_t>(c) * comp_stride;
    if (live) {
      for (int i = t; i < P3; i += TPE) b0[i] = u[static_cast<size_t>(off[i]) + coff];
    }
    __syncthreads();
    // F1: x-lines (iz,iy): b0[r*P + ix] -> b1[r*Q + qx]
    if (live && t < P2) {
      T x[P];
#pragma unroll
      for (int ix = 0; ix < P; ++ix) x[ix] = b0[t * P + ix];
#pragma unroll
      for (int qx = 0; qx < Q; ++qx) {
        T acc = T(0);
#pragma unroll
        for (int ix = 0; ix < P; ++ix) acc = mac<EXACT>(B(qx, ix), x[ix], acc);
        b1[t * Q + qx] = finish<EXACT>(acc);
      }
    }
    __syncthreads();
    // F2: y-lines (iz,qx): b1[(iz*P + iy)*Q + qx] -> b0[(iz*Q + qy)*Q + qx]
    if (live && t < P * Q) {
      const int iz = t / Q, qx = t - iz * Q;
      T x[P];
#pragma unroll
      for (int iy = 0; iy < P; ++iy) x[iy] = b1[(iz * P + iy) * Q + qx];
#pragma unroll
      for (int qy = 0; qy < Q; ++qy) {
        T acc = T(0);
#pragma unroll
        for (int iy = 0; iy < P; ++iy) acc = mac<EXACT>(B(qy, iy), x[iy], acc);
        b0[(iz * Q + qy) * Q + qx] = finish<EXACT>(acc);
      }
    }
    __syncthreads();
    // F3: z-lines (qy,qx) = t: b0[iz*Q2 + t] -> D = qdata * sum -> b0[qz*Q2 + t] (own line, in place)
    if (live && t < Q2) {
      T x[P];
#pragma unroll
      for (int iz = 0; iz < P; ++iz) x[iz] = b0[iz * Q2 + t];
#pragma unroll
      for (int qz = 0; qz < Q; ++qz) {
        T acc = T(0);
#pragma unroll
        for (int iz = 0; iz < P; ++iz) acc = mac<EXACT>(B(qz, iz), x[iz], acc);
        b0[qz * Q2 + t] = qd[qz] * finish<EXACT>(acc);
      }
    }
    __syncthreads();
    // T1: x-lines (qz,qy) = t: b0[t*Q + qx] -> b1[t*P + ix]
    if (live && t < Q2) {
      T x[Q];
#pragma unroll
      for (int qx = 0; qx < Q; ++qx) x[qx] = b0[t * Q + qx];
#pragma unroll
      for (int ix = 0; ix < P; ++ix) {
        T acc = T(0);
#pragma unroll
        for (int qx = 0; qx < Q; ++qx) acc = mac<EXACT>(B(qx, ix), x[qx], acc);
        b1[t * P + ix] = finish<EXACT>(acc);
      }
    }
    __syncthreads();
    // T2: y-lines (qz,ix): b1[(qz*Q + qy)*P + ix] -> b0[(qz*P + iy)*P + ix]
    if (live && t < Q * P) {
      const int qz = t / P, ix = t - qz * P;
      T x[Q];
#pragma unroll
      for (int qy = 0; qy < Q; ++qy) x[qy] = b1[(qz * Q + qy) * P + ix];
#pragma unroll
      for (int iy = 0; iy < P; ++iy) {
        T acc = T(0);
#pragma unroll
        for (int qy = 0; qy < Q; ++qy) acc = mac<EXACT>(B(qy, iy), x[qy], acc);
        b0[(qz * P + iy) * P + ix] = finish<EXACT>(acc);
      }
    }
    __syncthreads();
    // T3: z-lines (iy,ix) = t: b0[qz*P2 + t] -> V[iz] -> scatter
    if (live && t < P2) {
      T x[Q];
#pragma unroll
      for (int qz = 0; qz < Q; ++qz) x[qz] = b0[qz * P2 + t];
#pragma unroll
      for (int iz = 0; iz < P; ++iz) {
        T acc = T(0);
#pragma unroll
        for (int qz = 0; qz < Q; ++qz) acc = mac<EXACT>(B(qz, iz), x[qz], acc);
        const T val = finish<EXACT>(acc);
        const int i = iz * P2 + t;
        if constexpr (DET) {
          out[static_cast<size_t>(c) * y_comp_len + static_cast<size_t>(e) * P3 + i] = val;
        } else {
          unsafeAtomicAdd(&out[static_cast<size_t>(off[i]) + coff], val);
        }
      }
    }
    if (c + 1 < NC) __syncthreads();
  }
}

template <int P, int Q, int NC, class T, int NE, bool EXACT, bool DET>
void launch_fused(const FusedArgs& a, hipStream_t s) {
  using S = Shape<P, Q>;
  MatArg<T, P * Q> Bm;
  const T* src = static_cast<const T*>(a.basis);
  for (int i = 0; i < P * Q; ++i) Bm.m[i] = src[i];
  const long long blocks = (static_cast<long long>(a.num_elem) + NE - 1) / NE;
  const long long grid = (blocks + 7) / 8 * 8;
  require(grid < (1ll << 31), "fused mass: too many elements for one launch");
  constexpr int threads = round64(NE * S::TPE);
  fused_mass_kernel<P, Q, NC, T, NE, EXACT, DET><<<static_cast<unsigned>(grid), threads, 0, s>>>(
      a.num_elem, a.comp_stride, a.offsets, static_cast<const T*>(a.qdata), static_cast<const T*>(a.u),
      static_cast<T*>(a.out), a.y_comp_len, Bm);
  HIP_CHECK_LAUNCH();
}

template <int P, int Q, int NC, class T, int NE, bool EXACT, bool DET>
struct FusedReg {
  FusedReg() {
    register_fused(FusedKey{P, Q, NC, std::is_same<T, float>::value ? 1 : 0, DET ? 1 : 0, EXACT ? 1 : 0, NE},
                   &launch_fused<P, Q, NC, T, NE, EXACT, DET>);
  }
};

// All variants of one shape: FP64 {atomic, det} fast, FP64 det exact, FP32 {atomic, det}.
template <int P, int Q, int NC, int NE64, int NE32>
struct FusedShapeReg {
  FusedReg<P, Q, NC, double, NE64, false, false> a;
  FusedReg<P, Q, NC, double, NE64, false, true> b;
  FusedReg<P, Q, NC, double, NE64, true, true> c;
  FusedReg<P, Q, NC, double, NE64, true, false> d;
  FusedReg<P, Q, NC, float, NE32, false, false> e;
  FusedReg<P, Q, NC, float, NE32, false, true> f;
};

}  // namespace femhip
